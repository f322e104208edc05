"""GPU: the learned warm-start network's HIP kernels against the FP64 oracle
(oracle/unet.py, with bf16 rounding emulated at the points where the device
stores bf16): forward outputs, loss and every parameter gradient; training
reduces the loss; and the solver's mode 2 (learned warm start inside the
flexible mixed-precision PCG) converges to the same FP64 solution even with an
untrained network."""
import numpy as np
import pytest

from oracle import unet, pressure64
from waves_amd import net as NN
from waves_amd import solver as S
from waves_amd.settings import Settings
from waves_amd.scenarios import load_pressure_case

pytestmark = pytest.mark.gpu


def random_case(n, B, seed):
    rng = np.random.default_rng(seed)
    params = unet.init_params(seed).astype(np.float64)
    # non-zero biases so the bias column is exercised
    for (name, co, ci), w in zip(unet.layer_shapes(), unet.split_params(params).values()):
        w[:, 27 * ci] = rng.normal(0, 0.1, co)
    cells = n[0] * n[1] * n[2]
    x = np.zeros((B, cells, 8))
    x[..., 0] = rng.random((B, cells))
    x[..., 1] = rng.normal(size=(B, cells))
    x[..., 2] = rng.normal(size=(B, cells))
    x[..., 3] = rng.random((B, cells)) > 0.3
    xb = NN.to_bf16_bits(x.astype(np.float32))
    x16 = NN.from_bf16_bits(xb).astype(np.float64)
    t = rng.normal(size=(B, cells)).astype(np.float32)
    return params.astype(np.float32), xb, x16, t


def oracle_shape(x, n, B):
    return x.reshape(B, n[2], n[1], n[0], -1)


@pytest.mark.parametrize("n,B", [((12, 8, 6), 2), ((9, 7, 5), 1)])
def test_forward_matches_oracle(gpu, n, B):
    params, xb, x16, _ = random_case(n, B, 7)
    net = NN.Net(n, B)
    net.set_params(params)
    qf, qc = net.forward(xb)
    of, oc, _ = unet.forward(params, oracle_shape(x16, n, B), use_bf16=True)
    of, oc = of.reshape(B, -1), oc.reshape(B, -1)
    ef = np.linalg.norm(qf - of) / np.linalg.norm(of)
    ec = np.linalg.norm(qc - oc) / np.linalg.norm(oc)
    print(f"forward n={n} B={B}: rel err fine {ef:.2e} coarse {ec:.2e}", flush=True)
    # bf16 re-rounding of a few activations is the only expected difference
    assert ef < 2e-3 and ec < 2e-3
    net.close()


def test_gradients_match_oracle(gpu):
    n, B = (12, 8, 6), 2
    params, xb, x16, t = random_case(n, B, 11)
    net = NN.Net(n, B)
    net.set_params(params)
    loss, g = net.loss_grad(xb, t)
    mask = x16[..., 3].reshape(B, n[2], n[1], n[0]) > 0.5
    oloss, og = unet.loss_and_grad(params, oracle_shape(x16, n, B), t.reshape(B, n[2], n[1], n[0]).astype(np.float64),
                                   mask, use_bf16=True)
    print(f"loss device {loss:.6f} oracle {oloss:.6f}", flush=True)
    assert abs(loss - oloss) <= 1e-3 * abs(oloss)
    gs, ogs = unet.split_params(g.astype(np.float64)), unet.split_params(og)
    for name in gs:
        a, b = gs[name].ravel(), ogs[name].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(f"  grad {name}: rel {rel:.2e} cos {cos:.6f}", flush=True)
        assert rel < 2e-3 and cos > 0.99999, name
    net.close()


def test_training_reduces_loss(gpu):
    """Fit a smooth target that depends on the inputs (a fixed linear filter of
    channel 1 plus channel 2): the loss falls well below its starting value."""
    n, B, count = (16, 8, 8), 4, 32
    rng = np.random.default_rng(5)
    cells = n[0] * n[1] * n[2]
    x = np.zeros((count, cells, 8), np.float32)
    x[..., 0] = rng.random((count, cells))
    x[..., 1] = rng.normal(size=(count, cells))
    x[..., 2] = rng.normal(size=(count, cells))
    x[..., 3] = 1
    g = x[..., 1].reshape(count, n[2], n[1], n[0])
    sm = (g + np.roll(g, 1, 3) + np.roll(g, -1, 3) + np.roll(g, 1, 2) + np.roll(g, -1, 2)) / 5
    t = (sm.reshape(count, cells) + 0.5 * x[..., 2]).astype(np.float32)
    net = NN.Net(n, B)
    net.set_params(unet.init_params(1))
    net.set_dataset(NN.to_bf16_bits(x), t)
    before = net.eval(0, count)
    losses = net.train(300, count, lr=2e-3, seed=3)
    after = net.eval(0, count)
    print(f"train: eval loss {before:.4f} -> {after:.4f}; step losses {losses[0]:.4f} .. {losses[-1]:.4f}", flush=True)
    assert np.all(np.isfinite(losses))
    assert after < 0.3 * before
    # deterministic: same seed, same start -> same losses
    net.set_params(unet.init_params(1))
    again = net.train(20, count, lr=2e-3, seed=3)
    assert np.array_equal(again, losses[:20])
    net.close()


def test_mode2_untrained_network_still_converges(gpu):
    s = Settings(n=(96, 48, 16), origin=(0.5, -0.6, -0.4), zone_length=0.75, history_capacity=64)
    g = S.Solver(s)
    g.start()
    net = NN.Net(s.n, 1)
    net.set_params(unet.init_params(2))
    g.attach_net(net)
    for tol in (1e-6, 1e-8):
        load_pressure_case(g, s)
        r = g.pressure_solve(2, tol)
        work, geo = g.read(S.FIELD_WORK), g.read(S.FIELD_GEO)
        A, b, wet = pressure64.assemble(work, geo, s.n)
        p = g.read(S.FIELD_PRESSURE64, np.float64)[wet]
        p64 = pressure64.solve(A, b)
        err = float(np.linalg.norm(p - p64) / np.linalg.norm(p64))
        relres = float(np.linalg.norm(b - A @ p) / np.linalg.norm(b))
        print(f"mode 2 tol {tol:g}: iters {r['iterations']} restarts {r['restarts']} relres {relres:.2e} err {err:.2e} "
              f"device {r['device_ms']:.2f} ms", flush=True)
        assert r["converged"] == 1 and r["failure"] == 0
        assert relres <= tol * 1.01 and err < 1e3 * tol
        assert abs(r["true_rel_residual"] - relres) <= 1e-3 * relres + 1e-15
    g.close()
