"""The multilayer perceptron kernels of libnexd_gpu.so (nexd_mlp_*) against
the CPU oracle (oracle/mlp_oracle.py), whose gradients are pinned by finite
differences (tests/test_mlp_oracle.py).

The oracle rounds to bfloat16 what the GPU rounds (GEMM operands) and
accumulates in float64; the GPU accumulates in FP32 on the matrix cores.
The two differ by accumulation order and by the occasional activation
that rounds to the neighbouring bfloat16 value, so outputs and gradients
are compared in relative L2 with a bound of 5e-3.
"""
import numpy as np
import pytest

from nexd_amd import mlp
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

TOL = 5e-3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("widths,B", [([9, 128, 128, 192], 64), ([9, 512, 512, 512, 512, 24576], 64),
                                      ([70, 64, 128], 128)])
def test_forward_matches_oracle(widths, B):
    print(f"[gpu mlp forward] widths={widths} B={B}", flush=True)
    rng = np.random.default_rng(3)
    params = mlp.init_params(widths, seed=4)
    params = [(w, (rng.standard_normal(w.shape[0]) * 0.1).astype(np.float32)) for w, _ in params]
    x = rng.standard_normal((B, widths[0])).astype(np.float32)
    with mlp.MLP(widths, max_batch=256) as net:
        net.set_params(params)
        y = net.forward(x)
        back = net.params()
    yo, _ = mo.forward(params, x, bf16_ops=True)
    err = rel(y, yo)
    print(f"  rel L2 vs oracle {err:.2e}", flush=True)
    assert err <= TOL
    for (w, b), (w2, b2) in zip(params, back):
        assert np.array_equal(w, w2) and np.array_equal(b, b2)


def test_gradients_match_oracle():
    widths = [9, 128, 128, 128, 192]
    B = 128
    rng = np.random.default_rng(5)
    params = [(w, (rng.standard_normal(w.shape[0]) * 0.1).astype(np.float32)) for w, _ in mlp.init_params(widths, 6)]
    x = rng.standard_normal((300, widths[0])).astype(np.float32)
    t = rng.standard_normal((300, widths[-1])).astype(np.float32)
    idx = rng.permutation(300)[:B].astype(np.int32)
    with mlp.MLP(widths, max_batch=B) as net:
        net.set_params(params)
        net.upload_dataset(x, t)
        loss = net.gradients(idx)
        g = net.grads()
    lo, go = mo.loss_and_grads(params, x[idx], t[idx], bf16_ops=True)
    print(f"[gpu mlp gradients] loss {loss:.6f} vs oracle {lo:.6f}", flush=True)
    assert abs(loss - lo) <= 1e-4 * lo
    for l in range(len(params)):
        ew, eb = rel(g[l][0], go[l][0]), rel(g[l][1], go[l][1])
        print(f"  layer {l}: dW rel L2 {ew:.2e}, db {eb:.2e}", flush=True)
        assert ew <= TOL and eb <= TOL


def test_adam_steps_match_oracle_and_training_learns():
    """Three Adam steps equal the oracle's (same gradients up to rounding),
    and 300 steps on a teacher network's outputs cut the loss tenfold."""
    widths = [9, 64, 64, 64]
    B = 64
    rng = np.random.default_rng(7)
    params = mlp.init_params(widths, 8)
    teacher = mlp.init_params(widths, 9)
    x = rng.standard_normal((512, 9)).astype(np.float32)
    t, _ = mo.forward(teacher, x, bf16_ops=False)
    t = t.astype(np.float32)
    idx = np.concatenate([rng.permutation(512) for _ in range(40)]).astype(np.int32)
    with mlp.MLP(widths, max_batch=B) as net:
        net.set_params(params)
        net.upload_dataset(x, t)
        net.train(idx[:3 * B], B, lr=1e-3)
        p3 = net.params()
        net.set_params(params)
        losses = net.train(idx[:300 * B], B, lr=3e-3)
        print(f"[gpu mlp train] loss {losses[0]:.4e} -> {losses[-10:].mean():.4e} in {net.last_run_ms():.1f} ms",
              flush=True)
    state, po = {}, [(w.astype(np.float64), b.astype(np.float64)) for w, b in params]
    for s in range(3):
        _, go = mo.loss_and_grads([(w.astype(np.float32), b.astype(np.float32)) for w, b in po],
                                  x[idx[s * B:(s + 1) * B]], t[idx[s * B:(s + 1) * B]], bf16_ops=True)
        po = mo.adam_step(po, go, state, lr=1e-3)
    for (w, b), (wo, bo) in zip(p3, po):
        d0 = np.abs(np.asarray(params[0][0], np.float64)).max()
        assert np.max(np.abs(w - wo)) <= 2e-5 * max(1.0, d0) and np.max(np.abs(b - bo)) <= 2e-5
    assert losses[-10:].mean() < 0.1 * losses[0]


def test_bad_shapes_are_refused():
    with pytest.raises(Exception, match="multiple of 64"):
        mlp.MLP([9, 64, 64], max_batch=100)
    with mlp.MLP([9, 64, 64], max_batch=64) as net:
        with pytest.raises(Exception, match="batch"):
            net.forward(np.zeros((10, 9), np.float32))
