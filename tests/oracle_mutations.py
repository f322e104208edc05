"""Mutation sweep of the oracle pins (test infrastructure, run by hand:
`python tests/oracle_mutations.py`).  Each entry plants one plausible
restatement mistake in a copy of oracle/ built in a temporary directory, then
runs the CPU pin suites against that build (WAVES_ORACLE_LIB); every mutation
must make at least one pin fail.  The real oracle/liboracle.so is not touched."""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = ["tests/test_oracle_state_pins.py", "tests/test_oracle_pins.py", "tests/test_oracle_motion_pins.py"]

MUTATIONS = [
 ("wo_motion.c", "float correction = s->motion[N + i] * (face_courant(s, r, axis) - face_courant(s, q, axis));",
                 "float correction = s->motion[N + i] * (face_courant(s, q, axis) - face_courant(s, r, axis));", "VOF correction sign"),
 ("wo_motion.c", "I3 donor = q; if (velocity > 0) donor.v[axis]--;", "I3 donor = q; if (velocity < 0) donor.v[axis]--;", "flux donor side"),
 ("wo_motion.c", "if (velocity > 0) low.v[axis] = 1 - width;", "if (velocity < 0) low.v[axis] = 1 - width;", "flux slab"),
 ("wo_motion.c", "amount *= signf(velocity);", "amount *= 1;", "flux sign"),
 ("wo_motion.c", "c[20] = fminf_(PC_DTMAX(s), .45f / fmaxf_(rate, 1e-20f));", "c[20] = fminf_(PC_DTMAX(s), .5f / fmaxf_(rate, 1e-20f));", "CFL number"),
 ("wo_motion.c", "if (accelerate) speed.v[1] += 9.81f * PC_DTMAX(s) * PC_FIX(s, 1);", "", "CFL gravity term"),
 ("wo_motion.c", "c[30] = c[20] * .5f;", "c[30] = c[20];", "retry cap"),
 ("wo_motion.c", "if (retry && c[29] < 2)", "if (retry && c[29] < 3)", "retry budget"),
 ("wo_motion.c", "float value = old + (1 - expf(-rate * timestep(s))) * (target - old);", "float value = old + expf(-rate * timestep(s)) * (target - old);", "relaxation factor"),
 ("wo_motion.c", "r = 1000 / dt * divergence(s, q);", "r = 1000 * dt * divergence(s, q);", "true residual scale"),
 ("wo_state.h", "if (axis == 1 && (l.v[1] >= s->ny || r.v[1] >= s->ny)) return .5f * h;", "if (axis == 1 && (l.v[1] >= s->ny || r.v[1] >= s->ny)) return h;", "top half face"),
 ("wo_state.h", "return h * fmaxf_((wl ? pl : pr) / fmaxf_(pl + pr, 1e-30f), 1e-6f);", "return h * fmaxf_((wl ? pr : pl) / fmaxf_(pl + pr, 1e-30f), 1e-6f);", "ghost fraction side"),
 ("wo_state.h", "if (axis == 0 && nb.v[0] < 0) return 0.0f;", "", "ocean-side Neumann"),
 ("wo_mg.c", "(mode == 1 || st[10] + 1 >= (uint32_t)s->lag)", "(mode == 1 || st[10] >= (uint32_t)s->lag)", "audit lag"),
 ("wo_mg.c", "int incompatible = st[0] == 0 || st[1] > 0 || st[2] > 0;", "int incompatible = st[0] == 0 || st[1] > 0;", "audit signature"),
 ("wo_pressure.c", "c[19] < .5f && c[58] < 4) { c[57] = 1;", "c[19] < .5f && c[58] < 3) { c[57] = 1;", "defect budget"),
 ("wo_pressure.c", "if (c[27] > .450001f) { c[3]++; c[7] = 20; return; }", "", "acceptance CFL"),
 ("wo_geometry.c", "r.v[b] = -(heights[2] - heights[1]) / (2 * PC_H(s, b));", "r.v[b] = -(heights[2] - heights[1]) / (PC_H(s, b));", "height-normal difference"),
 ("wo_geometry.c", "heights[column] = sum * PC_H(s, axis);", "heights[column] = sum;", "height units"),
 ("wo_motion.c", "dv += (old_face(s, r, axis) - old_face(s, q, axis)) / PC_H(s, axis);", "dv += (old_face(s, r, axis) - old_face(s, q, axis)) / PC_H(s, 0);", "audit divergence spacing"),
 ("wo_motion.c", "v = f4(f > .5f ? dt * dv * cell : 0,", "v = f4(f > .4f ? dt * dv * cell : 0,", "audit full-cell gate"),
 ("wo_motion.c", "ch.v[axis] = .5f * (s->motion[forward_base(s) + face_local(s, q, axis)] + s->motion[forward_base(s) + face_local(s, r, axis)]);",
                 "ch.v[axis] = (s->motion[forward_base(s) + face_local(s, q, axis)] + s->motion[forward_base(s) + face_local(s, r, axis)]);", "momentum face average"),
 ("wo_motion.c", "c[88] = mn; c[89] = mx; if (mn < 0 || mx > 1) { c[7] = 23; c[3]++; }", "c[88] = mn; c[89] = mx; if (mn < 0 || mx > 1.5f) { c[7] = 23; c[3]++; }", "forcing bounds"),
]


def main() -> int:
    missed = 0
    with tempfile.TemporaryDirectory() as tmp:
        for fname, old, new, name in MUTATIONS:
            work = os.path.join(tmp, "oracle")
            shutil.rmtree(work, ignore_errors=True)
            shutil.copytree(os.path.join(ROOT, "oracle"), work,
                            ignore=shutil.ignore_patterns("*.so", "__pycache__", "_ref"))
            path = os.path.join(work, fname)
            text = open(path).read()
            assert text.count(old) == 1, (name, text.count(old))
            open(path, "w").write(text.replace(old, new))
            subprocess.run(["make", "-s", "-C", work], check=True)
            env = dict(os.environ, WAVES_ORACLE_LIB=os.path.join(work, "liboracle.so"))
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", *PINS], cwd=ROOT, env=env,
                               capture_output=True, text=True)
            failed = [line[7:] for line in r.stdout.splitlines() if line.startswith("FAILED")]
            caught = r.returncode != 0 and bool(failed)
            missed += not caught
            print(f"{name:28s} {'caught by ' + failed[0].split(' ')[0] if caught else 'MISSED'}", flush=True)
    print(f"{len(MUTATIONS) - missed} of {len(MUTATIONS)} mutations caught")
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main())
