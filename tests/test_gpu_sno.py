"""Spin-up proposal network on the device against the FP64 restatement
(oracle/sno_oracle.py).  Tolerances: the GEMM is compared with numpy on the
same BF16-rounded operands (FP32 accumulation, 1e-5 relative); the network
forward and gradient in BF16 are compared with the FP64 restatement at 3e-2 /
6e-2 relative L2 (operands rounded to 8 significant bits)."""
import numpy as np
import pytest

from oracle import sno_oracle as S
from turbid_amd import runtime as RT

from helpers import rel_l2

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.view(np.float32)


@pytest.mark.parametrize("m,n,k", [(64, 32, 16), (37, 45, 70), (5, 3, 4096), (200, 1, 33)])
def test_bf16_gemm(m, n, k):
    rng = np.random.default_rng(m + n + k)
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((k, n)).astype(np.float32)
    ref = bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)
    got = RT.bf16_gemm(a, b)
    assert rel_l2(got, ref) < 1e-5


def _pair(nx=8, ny=6, nz=9, C=16, L=2, cin=3, cout=2, batch=2, seed=0):
    net = S.Net(nx=nx, ny=ny, nz=nz, C=C, L=L, cin=cin, cout=cout)
    dev = RT.SpectralOperatorNet(nx=nx, ny=ny, nz=nz, channels=C, layers=L, cin=cin, cout=cout, batch=batch)
    assert dev.nparams == net.lay.total
    p = S.init_params(net.lay, seed=seed)
    dev.set_params(p)
    return net, dev, p


def test_forward_matches_oracle():
    net, dev, p = _pair()
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, dev.points, 3))
    ref = net.forward(p, x)
    got = dev.forward(x)
    assert rel_l2(got, ref) < 3e-2


def test_gradient_matches_oracle():
    net, dev, p = _pair()
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, dev.points, 3))
    t = rng.standard_normal((2, dev.points, 2))
    w = np.array([1.0, 0.5])
    loss_ref, g_ref = net.loss_grad(p, x, t, w)
    loss = dev.loss_grad(x, t, w)
    g = dev.grads()
    assert abs(loss - loss_ref) < 3e-2 * loss_ref
    assert rel_l2(g, g_ref) < 6e-2
    lay = net.lay
    for name, off, size in [("lift", lay.P, lay.cin * lay.C), ("Rx0", lay.Rx[0], 5 * 2 * 16 * 16),
                            ("Rz1", lay.Rz[1], 9 * 16 * 16), ("Q", lay.Q, 16 * 2)]:
        assert rel_l2(g[off:off + size], g_ref[off:off + size]) < 6e-2, name


def test_training_reduces_loss():
    # fit a fixed smooth target map with Adam: the loss falls by 10x
    net, dev, p = _pair(seed=5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((2, dev.points, 3)).astype(np.float32)
    t = np.tanh(x[..., :2] * 0.8) + 0.3 * x[..., 1:3] ** 2
    w = np.array([1.0, 1.0])
    first = dev.loss_grad(x, t, w)
    for _ in range(150):
        last = dev.train_step(x, t, w, 3e-3)
    assert last < 0.1 * first
