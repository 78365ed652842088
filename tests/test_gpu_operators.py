"""Operators and factorisations supplied by the caller (tbd_set_operators,
tbd_set_factorisation): the drop-in host may hand over the arrays its own
ChebyshevSpaceStepper::Configure (stepper_space_chebyshev.cc:101-228) and
CmDz2DecompositionCplx::Construct (cmdz2_decomposition.cc:19-297) built."""
import numpy as np
import pytest

from turbid_amd import runtime as RT
from turbid_amd import workloads as W
from helpers import rel_l2, small_config

pytestmark = pytest.mark.gpu


def _solver(cfg):
    g = RT.Solver(cfg)
    uvw, p, c = W.initial_state(cfg)
    g.upload(uvw, p, c if cfg["nphase"] > 1 else None)
    return g


def _nfac(cfg):
    return 5 if cfg["nphase"] > 1 else 4


@pytest.mark.parametrize("precision", [0, 1])
def test_own_operators_set_back_change_nothing(precision):
    # the context's own D1, D2 and factorisations handed back, before the
    # first step and again after captured steps: bit-identical states
    cfg = small_config(precision=precision)
    a, b = _solver(cfg), _solver(cfg)
    _, d1, d2 = a.operators()
    b.set_operators(d1, d2)
    for k in range(_nfac(cfg)):
        b.set_factorisation(k, *a.factorisation(k))
    sa, sb = RT.Stepper(cfg["dt"]), RT.Stepper(cfg["dt"])
    RT.run_steps(a, sa, 2)
    RT.run_steps(b, sb, 2)
    for k in range(_nfac(cfg)):
        b.set_factorisation(k, *a.factorisation(k))
    RT.run_steps(a, sa, 3)
    RT.run_steps(b, sb, 3)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.features(), b.features())


def _reordered(V, Vi, lam, rng):
    # the same factorisation with its interior eigenvectors permuted and
    # rescaled (V' = V P D, V'^-1 = D^-1 P^T V^-1, lambda' = P^T lambda), as an
    # eigen-solver may return them
    n = V.shape[0]
    perm = 1 + rng.permutation(n - 2)
    scale = rng.uniform(0.5, 2.0, n - 2) * rng.choice([-1.0, 1.0], n - 2)
    V2, Vi2, lam2 = V.copy(), Vi.copy(), lam.copy()
    V2[:, 1:n - 1] = V[:, perm] * scale[None, :]
    Vi2[1:n - 1, :] = Vi[perm, :] / scale[:, None]
    lam2[1:n - 1] = lam[perm]
    return V2, Vi2, lam2


@pytest.mark.parametrize("nz", [33, 65])
def test_factorisation_in_any_eigenvector_order(nz):
    # eigenvectors in another order and scaling give the same solves (the
    # even/odd form is rebuilt from them); perturbed eigenvalues do not, so
    # the supplied arrays are the ones in use
    cfg = small_config(nz=nz)
    rng = np.random.default_rng(nz)
    a, b, c = _solver(cfg), _solver(cfg), _solver(cfg)
    for k in range(_nfac(cfg)):
        V, Vi, lam, fix = a.factorisation(k)
        V2, Vi2, lam2 = _reordered(V, Vi, lam, rng)
        b.set_factorisation(k, V2, Vi2, lam2, fix)
        lam3 = lam.copy()
        lam3[1:nz - 1] *= 1.0 + 1e-3 * rng.standard_normal(nz - 2)
        c.set_factorisation(k, V, Vi, lam3, fix)
    for g in (a, b, c):
        RT.run_steps(g, RT.Stepper(cfg["dt"]), 5)
    (ua, pa, ca), (ub, pb, cb), (uc, pc, cc) = a.download(), b.download(), c.download()
    assert rel_l2(ub, ua) < 1e-9
    assert rel_l2(pb, pa) < 1e-9
    assert rel_l2(cb, ca) < 1e-9
    assert rel_l2(uc, ua) > 1e-8


def test_set_factorisation_rejects_bad_input():
    cfg = small_config()
    g = _solver(cfg)
    V, Vi, lam, fix = g.factorisation(0)
    with pytest.raises(RT.TurbidError, match="bad index"):
        g.set_factorisation(7, V, Vi, lam, fix)
    V[3, 4] = np.nan
    with pytest.raises(RT.TurbidError, match="not finite"):
        g.set_factorisation(0, V, Vi, lam, fix)
    # the context still steps after the rejected calls
    RT.run_steps(g, RT.Stepper(cfg["dt"]), 2)
    assert np.all(np.isfinite(g.features()))
