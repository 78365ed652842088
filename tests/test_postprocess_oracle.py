"""Pins of the swirling-strength restatement (oracle.swirling_strength), CPU
only: tensors with known eigenvalues, and a band-limited field whose
velocity-gradient eigenvalues have a closed form."""
import numpy as np

from oracle import oracle as O

from helpers import small_config


def test_swirl_known_tensors():
    w = 0.8
    rot = np.array([[0.0, -w, 0.0], [w, 0.0, 0.0], [0.0, 0.0, 0.0]])
    sym = np.array([[1.0, 2.0, 0.5], [2.0, -0.3, 0.1], [0.5, 0.1, 0.7]])
    shear = np.array([[0.0, 3.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    # rotation about z plus stretching along z: eigenvalues s, +-i w
    rs = rot + np.diag([0.0, 0.0, 0.4])
    got = O.swirling_strength_tensors(np.stack([rot, sym, shear, rs]))
    np.testing.assert_allclose(got, [w, 0.0, 0.0, w], atol=1e-14)


def test_swirl_of_a_cellular_flow():
    # u = a sin(ky y), v = -b sin(kx x): G has lambda^2 = -a b kx ky cos(kx x) cos(ky y)
    cfg = small_config(nx=24, ny=18, nz=9)
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    x = cfg["lx"] / nx * np.arange(nx)
    y = cfg["ly"] / ny * np.arange(ny)
    X, Y = np.meshgrid(x, y)
    kx = 2 * np.pi / cfg["lx"] * 2
    ky = 2 * np.pi / cfg["ly"]
    a, b = 0.7, 1.3
    phys = np.zeros((3, nz, ny, nx))
    phys[0] = a * np.sin(ky * Y)
    phys[1] = -b * np.sin(kx * X)
    sw = O.swirling_strength(O.dft_forward(phys), cfg)
    c = a * b * kx * ky * np.cos(ky * Y) * np.cos(kx * X)
    expect = np.broadcast_to(np.sqrt(np.maximum(c, 0.0)), sw.shape)
    # sqrt(eps) where the discriminant crosses zero, round-off elsewhere
    assert np.abs(sw - expect).max() < 1e-7 * expect.max()
    away = np.broadcast_to(np.abs(c) > 1e-3 * np.abs(c).max(), sw.shape)
    assert np.abs(sw - expect)[away].max() < 1e-12 * expect.max()


def test_velocity_gradient_z_derivative():
    # w-independent check of the D1 column: u = z^2 on every column
    cfg = small_config(nx=12, ny=12, nz=17)
    g, _, _ = O.chebyshev(cfg["nz"])
    phys = np.zeros((3, cfg["nz"], cfg["ny"], cfg["nx"]))
    phys[0] = (np.asarray(g) ** 2)[:, None, None]
    G = O.velocity_gradient(O.dft_forward(phys), cfg)
    np.testing.assert_allclose(G[0, 2], 2 * np.asarray(g)[:, None, None] * np.ones_like(phys[0]), atol=1e-11)
    assert np.abs(G[0, 0]).max() < 1e-12 and np.abs(G[0, 1]).max() < 1e-12
