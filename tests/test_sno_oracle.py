"""CPU checks of the spin-up proposal network's restatement (oracle/sno_oracle.py):
transform pairs, closed forms of the axis operators, and the hand-written
reverse-mode gradient against central finite differences (FP64)."""
import math

import numpy as np
import pytest

from oracle import sno_oracle as S


@pytest.mark.parametrize("n", [24, 8, 6])
def test_dft_pair_is_identity(n):
    D, Di = S.dft_mats(n)
    np.testing.assert_allclose(Di @ D, np.eye(n), atol=1e-13)


@pytest.mark.parametrize("n", [33, 9, 5])
def test_chebyshev_pair_is_identity(n):
    D, Di = S.cheb_mats(n)
    np.testing.assert_allclose(Di @ D, np.eye(n), atol=1e-12)
    # coefficients of a known polynomial: 2 z^2 - 1 = T_2
    z = -np.cos(math.pi * np.arange(n) / (n - 1))
    a = D @ (2 * z * z - 1)
    ref = np.zeros(n)
    ref[2] = 1.0
    np.testing.assert_allclose(a, ref, atol=1e-12)


def _small(L=2, C=3, cin=2, cout=2, nx=6, ny=4, nz=5):
    return S.Net(nx=nx, ny=ny, nz=nz, C=C, L=L, cin=cin, cout=cout)


def test_axis_identity_mixing_is_identity():
    net = _small()
    rng = np.random.default_rng(0)
    H = rng.standard_normal(net.shape5(2, 3))
    C = 3
    Rx = np.zeros((S.modes(6), 2, C, C)); Rx[:, 0] = np.eye(C)
    Ry = np.zeros((S.modes(4), 2, C, C)); Ry[:, 0] = np.eye(C)
    Rz = np.zeros((5, C, C)); Rz[:] = np.eye(C)
    np.testing.assert_allclose(net._axis(H, Rx, 0), H, atol=1e-13)
    np.testing.assert_allclose(net._axis(H, Ry, 1), H, atol=1e-13)
    np.testing.assert_allclose(net._axis(H, Rz, 2), H, atol=1e-12)


def test_axis_single_mode_rotation():
    # H = cos(2 pi m x / n) in channel 0; mixing a + ib on mode m gives
    # a cos - b sin (closed form), zero on the other modes
    net = _small(C=1)
    n, m, a, b = 6, 2, 0.7, -0.3
    x = np.arange(n)
    H = np.zeros(net.shape5(1, 1))
    H[..., 0] = np.cos(2 * math.pi * m * x / n)[None, None, None, :]
    R = np.zeros((S.modes(n), 2, 1, 1))
    R[m, 0, 0, 0] = a
    R[m, 1, 0, 0] = b
    out = net._axis(H, R, 0)[0, 0, 0, :, 0]
    th = 2 * math.pi * m * x / n
    np.testing.assert_allclose(out, a * np.cos(th) - b * np.sin(th), atol=1e-13)


def test_gradient_matches_finite_differences():
    net = _small()
    lay = net.lay
    p = S.init_params(lay, seed=1)
    rng = np.random.default_rng(2)
    nb, pts = 2, 6 * 4 * 5
    x = rng.standard_normal((nb, pts, lay.cin))
    t = rng.standard_normal((nb, pts, lay.cout))
    w = np.array([1.0, 0.5])
    loss, g = net.loss_grad(p, x, t, w)
    assert abs(loss - net.loss(p, x, t, w)) < 1e-14 * max(1.0, loss)
    # every parameter block, a few entries each
    blocks = [lay.P, lay.pb, lay.Q, lay.qb] + lay.W + lay.b + lay.Rx + lay.Ry + lay.Rz
    idx = sorted(set(int(o + k) for o in blocks for k in (0, 1, 5)) & set(range(lay.total)))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (net.loss(p + e, x, t, w) - net.loss(p - e, x, t, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
