"""GPU checks of the learned traveltime surrogate (kmig_nn.hip) against the
CPU oracle (oracle/mlp.py): BF16-MFMA forward pass, loss and parameter
gradient on the same BF16 operands, and training on v(z) traveltimes."""
import numpy as np
import pytest

from oracle import mlp as M

pytestmark = pytest.mark.gpu

CFG = dict(s_min=0.0, s_max=3.0, z_min=0.0, z_max=2.0, x_min=0.0, x_max=3.0, z_src=0.0,
           slowness=0.5, sigma_t=0.01, sigma_g=0.1)


def _net(**kw):
    from kirch_datuming_amd import nn
    c = dict(CFG, batch=1024)
    c.update(kw)
    return nn.Surrogate(**c)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _records(n, seed):
    rng = np.random.default_rng(seed)
    s, z, x = rng.uniform(0, 3, n), rng.uniform(0.02, 2, n), rng.uniform(0, 3, n)
    return M.vz_records(s, z, x, 1.5, 0.5)


def test_forward_matches_oracle():
    net = _net()
    p = M.init_params(11, scale_out=1.0).astype(np.float32)
    net.params = p
    pts = _records(1000, 1)[:, :3]
    out = net.eval(pts)
    ref, _, _ = M.forward(p.astype(np.float64), CFG, pts, bf16_ops=True)
    ref64, _, _ = M.forward(p.astype(np.float64), CFG, pts)
    print("forward relL2 vs bf16 oracle", _rel(out, ref), "vs fp64", _rel(out, ref64))
    assert _rel(out, ref) < 2e-3
    assert _rel(out, ref64) < 3e-2
    net.close()


@pytest.mark.parametrize("n", [64, 200])
def test_gradient_matches_oracle(n):
    net = _net()
    p = M.init_params(12, scale_out=1.0).astype(np.float32)
    net.params = p
    rec = _records(n, 2)
    loss, g = net.grad(rec)
    lref, gref = M.loss_grad(p.astype(np.float64), CFG, rec, bf16_ops=True)
    l64, g64 = M.loss_grad(p.astype(np.float64), CFG, rec)
    print(f"n={n} loss {loss} oracle {lref} grad relL2 {_rel(g, gref)} vs fp64 {_rel(g, g64)}")
    assert abs(loss - lref) <= 1e-3 * abs(lref)
    assert _rel(g, gref) < 5e-3
    assert _rel(g, g64) < 5e-2
    # per layer, so a transposed or misplaced block cannot hide in the total
    o = 0
    for m, k in zip(M.LM, M.LK):
        sl = slice(o, o + m * k)
        assert _rel(g[sl], gref[sl]) < 1e-2
        o += m * k + m
    net.close()


def test_training_fits_vz_traveltimes():
    net = _net(batch=4096, lr=2e-3, lr_final=1e-4, sigma_t=0.002)
    rec = _records(200000, 3)
    net.set_data(rec)
    test = _records(4000, 4)
    before = net.eval(test[:, :3])
    loss = net.train(600)
    after = net.eval(test[:, :3])
    e0 = np.abs(before[:, 0] - test[:, 3]).max()
    e1 = np.abs(after[:, 0] - test[:, 3]).max()
    print("loss", loss[:3], loss[-3:], "max |dT| before", e0, "after", e1,
          net.stats())
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < 0.02 * loss[:10].mean()
    assert e1 < 0.1 * e0
    # a second call continues from the trained weights (graph replays)
    loss2 = net.train(50)
    assert loss2.mean() < 2.0 * loss[-50:].mean()
    net.close()
