"""Kernel variants of a step and their autotuner (tbd_autotune,
tbd_get_variant, tbd_set_variant): every variant steps to the same result as
the default one, within the precision's tolerance; tuning leaves the
context's state alone and keeps the fastest variant it tried."""
import math

import numpy as np
import pytest

from turbid_amd import runtime as RT
from turbid_amd import workloads as W
from helpers import rel_l2, small_config

pytestmark = pytest.mark.gpu

DEFAULT = RT.Solver.VARIANT_BITS["rhs"] | RT.Solver.VARIANT_BITS["overlap"]


def _solver(cfg, bits=None):
    g = RT.Solver(cfg)
    uvw, p, c = W.initial_state(cfg)
    g.upload(uvw, p, c if cfg["nphase"] > 1 else None)
    if bits is not None:
        g.set_variant(bits)
    return g


def _steps(g, cfg, n):
    RT.run_steps(g, RT.Stepper(cfg["dt"]), n)
    return g.download()


@pytest.mark.parametrize("precision,tol", [(0, 1e-9), (1, 1e-5)])
def test_every_variant_steps_like_the_default(precision, tol):
    cfg = small_config(precision=precision)
    ref = _steps(_solver(cfg, DEFAULT), cfg, 6)
    wide = [RT.Solver.VARIANT_BITS[k] for k in ("wide_yb", "wide_yf", "wide_xb", "wide_xf")]
    swap = [RT.Solver.VARIANT_BITS[k] for k in ("swap_yb", "swap_yf", "swap_xb", "swap_xf")]
    # 24 x 24: both lengths have a swapped radix order (4 x 6)
    for bits in (list(range(8)) + [DEFAULT | w for w in wide] + [DEFAULT | RT.Solver.WIDE_FFT]
                 + [DEFAULT | w for w in swap] + [DEFAULT | RT.Solver.WIDE_FFT | RT.Solver.SWAP_FFT]
                 + [DEFAULT | RT.Solver.VARIANT_BITS["wide_zgemm"], 1 | RT.Solver.VARIANT_BITS["wide_zgemm"]]):
        g = _solver(cfg, bits)
        assert g.variant() == bits
        got = _steps(g, cfg, 6)
        for a, b in zip(got, ref):
            assert rel_l2(a, b) < tol, bits


def test_autotune_keeps_state_and_the_fastest_variant():
    cfg = small_config()
    g = _solver(cfg)
    RT.run_steps(g, RT.Stepper(cfg["dt"]), 2)
    before = g.download()
    bits, ms = g.autotune(3)
    after = g.download()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    tried = [i for i in range(RT.Solver.N_VARIANTS) if not math.isnan(ms[i])]
    # the eight step-level combinations, then each of the four widths, each
    # of the four radix orders (24 x 24 has both) and the contraction panel
    # width
    assert len(tried) == 17
    assert all(i in tried for i in range(8))
    assert all(ms[i] > 0 for i in tried)
    assert bits == min(tried, key=lambda i: ms[i])
    assert g.variant() == bits
    # the tuned context keeps stepping like an untuned one
    ref = _solver(cfg)
    RT.run_steps(ref, RT.Stepper(cfg["dt"]), 2)
    for a, b in zip(_steps(g, cfg, 3), _steps(ref, cfg, 3)):
        assert rel_l2(a, b) < 1e-9


def test_environment_pins_a_switch(monkeypatch):
    monkeypatch.setenv("TBD_FUSED", "0")
    cfg = small_config()
    g = _solver(cfg)
    g.set_variant(15)
    assert g.variant() & RT.Solver.VARIANT_BITS["fused"] == 0
    bits, ms = g.autotune(2)
    assert bits & RT.Solver.VARIANT_BITS["fused"] == 0
    tried = [i for i in range(RT.Solver.N_VARIANTS) if not math.isnan(ms[i])]
    assert not any(i & 1 for i in tried)
    # the four admissible step-level combinations (at the widths set above),
    # then the four widths, the four radix orders and the panel width
    assert {i & 7 for i in tried} == {0, 2, 4, 6} and len(tried) == 13


def test_set_variant_rejects_bad_bits():
    g = _solver(small_config())
    with pytest.raises(RT.TurbidError, match="out of range"):
        g.set_variant(RT.Solver.N_VARIANTS)


def test_radix_order_bit_clear_where_no_swap():
    # 32 x 32: the length 32 has the swapped order 4 x 8; 16 (4 x 4) has none
    cfg = small_config(nx=16, ny=32, nz=17)
    g = _solver(cfg)
    g.set_variant(DEFAULT | RT.Solver.SWAP_FFT)
    b = RT.Solver.VARIANT_BITS
    assert g.variant() == DEFAULT | b["swap_yb"] | b["swap_yf"]
