"""The matrix-core MLP kernels against the FP64 oracle (oracle/mlp.py).

* forward: against the oracle with the same BF16 rounding points, and
  against exact FP64 (tolerances in the asserts);
* gradient: the backward kernels against the oracle's analytic gradient
  (itself pinned by central differences), relative L2 error stated;
* training: Adam on a smooth regression target lowers the loss."""
import numpy as np
import pytest

from gpucell_amd.surrogate import Mlp
from oracle import mlp as omlp
from tests import gpu_util as gu

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("batch", [1, 63, 64, 1000, 4099])
def test_forward_matches_oracle(gpu, batch):
    net = Mlp(seed=11)
    rng = np.random.default_rng(batch)
    p = omlp.glorot(5) + 0.02 * rng.normal(size=omlp.NPARAM)
    net.set_params(p)
    x = rng.normal(size=(batch, omlp.IN)).astype(np.float32)
    y = gu.host(net.forward(gu.dev(x)))
    ye, _ = omlp.forward(p.astype(np.float32).astype(np.float64), x, emulate=True)
    yx, _ = omlp.forward(p.astype(np.float32).astype(np.float64), x, emulate=False)
    e_emul, e_exact = _rel(y, ye), _rel(y, yx)
    print(f"batch {batch}: forward rel L2 vs BF16-emulating oracle {e_emul:.2e}, vs FP64 {e_exact:.2e}")
    assert e_emul < 2e-3
    assert e_exact < 2e-2
    # row-wise: every sample within tolerance (no misplaced tile or row)
    err = np.abs(y - ye).max(axis=1) / (np.abs(ye).max(axis=1) + 1e-3)
    assert err.max() < 2e-2


@pytest.mark.parametrize("batch", [256, 777])
def test_gradient_matches_oracle(gpu, batch):
    net = Mlp(seed=3)
    rng = np.random.default_rng(10 + batch)
    p = (omlp.glorot(9) + 0.02 * rng.normal(size=omlp.NPARAM)).astype(np.float32)
    net.set_params(p)
    x = rng.normal(size=(batch, omlp.IN)).astype(np.float32)
    dy = rng.normal(size=(batch, omlp.OUT)).astype(np.float32)
    g = gu.host(net.gradient(gu.dev(x), gu.dev(dy)))
    ge = omlp.backward(p.astype(np.float64), x, dy, emulate=True)
    gx = omlp.backward(p.astype(np.float64), x, dy, emulate=False)
    e_emul, e_exact = _rel(g, ge), _rel(g, gx)
    print(f"batch {batch}: gradient rel L2 vs BF16-emulating oracle {e_emul:.2e}, vs FP64 {e_exact:.2e}")
    assert e_emul < 1e-2
    assert e_exact < 5e-2
    # each parameter block separately (a transposed or misplaced block fails here)
    G, Ge = omlp.unflatten(g), omlp.unflatten(ge)
    for k in G:
        assert _rel(G[k], Ge[k]) < 3e-2, k


def test_training_reduces_loss(gpu):
    import torch
    rng = np.random.default_rng(0)
    n = 32768
    x = rng.normal(size=(n, omlp.IN)).astype(np.float32)
    A = rng.normal(size=(omlp.IN, omlp.OUT)) / np.sqrt(omlp.IN)
    y = np.tanh(x @ A).astype(np.float32) + 0.1 * np.sin(x[:, :omlp.OUT]).astype(np.float32)
    net = Mlp(seed=1)
    xd, yd = gu.dev(x), gu.dev(y)
    loss = net.train(xd, yd, epochs=6, batch=2048, lr=3e-3)
    print("loss per epoch", loss)
    assert np.all(np.isfinite(loss))
    assert loss[-1] < 0.3 * loss[0] and loss[-1] < loss[1]
    pred = net.forward(xd[:4096])
    mse = float(((pred - yd[:4096]) ** 2).mean())
    assert mse < 2 * loss[-1] + 1e-4
    # training is deterministic
    net2 = Mlp(seed=1)
    loss2 = net2.train(xd, yd, epochs=6, batch=2048, lr=3e-3)
    assert np.array_equal(loss, loss2)
    assert np.array_equal(net.params(), net2.params())
    torch.cuda.synchronize()
