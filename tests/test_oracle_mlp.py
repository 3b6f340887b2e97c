"""The order-selector network's CPU oracle, pinned by central finite
differences in float64: a wrong sign, a transposed weight or a dropped
ReLU mask in the backward pass fails these checks."""
import numpy as np
import pytest

from oracle import mlp


def _case(seed, n=96):
    rng = np.random.default_rng(seed)
    W = mlp.init(rng, ymean=3.0)
    # non-zero biases so that the ReLU masks differ between units
    W[mlp.NH * mlp.NIN:mlp.NH * mlp.NIN + mlp.NH] = rng.normal(0.0, 0.5, mlp.NH)
    o2 = mlp.NH * mlp.NIN + mlp.NH + mlp.NH * mlp.NH
    W[o2:o2 + mlp.NH] = rng.normal(0.0, 0.5, mlp.NH)
    norm = np.concatenate([rng.normal(0.0, 1.0, 4), rng.uniform(0.5, 2.0, 4)])
    X = rng.normal(0.0, 1.5, (n, 4))
    Y = rng.uniform(0.0, 11.0, n)
    return W, norm, X, Y


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_finite_differences(seed):
    W, norm, X, Y = _case(seed)
    loss, g = mlp.loss_grad(W, norm, X, Y, asym=4.0)
    assert np.isclose(loss, mlp.loss_only(W, norm, X, Y, asym=4.0), rtol=1e-14)
    rng = np.random.default_rng(100 + seed)
    idx = np.concatenate([rng.choice(mlp.NPARAM, 200, replace=False), [mlp.NPARAM - 1]])
    h = 1e-6
    for k in idx:
        Wp, Wm = W.copy(), W.copy()
        Wp[k] += h
        Wm[k] -= h
        fd = (mlp.loss_only(Wp, norm, X, Y) - mlp.loss_only(Wm, norm, X, Y)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_gradient_of_each_block_is_exercised():
    """Every parameter block carries gradient (no block silently zero)."""
    W, norm, X, Y = _case(5)
    _, g = mlp.loss_grad(W, norm, X, Y)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(g)
    for blk in (W1, b1, W2, b2, W3):
        assert np.count_nonzero(blk) > 0
    assert b3 != 0.0


def test_asymmetry_weights_under_prediction():
    W, norm, X, _ = _case(7, n=8)
    y = mlp.forward(W, norm, X)
    over = mlp.loss_only(W, norm, X, y - 1.0, asym=4.0)    # predictions 1 above the labels
    under = mlp.loss_only(W, norm, X, y + 1.0, asym=4.0)   # predictions 1 below
    assert np.isclose(over, 1.0) and np.isclose(under, 4.0)


def test_pack_unpack_layout():
    W = np.arange(mlp.NPARAM, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(W)
    assert W1[0, 1] == 1 and b1[0] == 256 and W2[0, 0] == 320 and b2[0] == 4416 and W3[0] == 4480
    assert b3 == 4544 and mlp.NPARAM == 4545
    assert np.array_equal(mlp.pack(W1, b1, W2, b2, W3, b3), W)
