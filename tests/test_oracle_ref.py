"""Pins of the oracle's neuron models against the reference neuron classes.

HH2_test, HH2_buffer and HHI_buffer are compiled unmodified from the reference
sources into oracle/_ref (oracle/Makefile) with a replay of Poisson_noise that
serves the same counter-based streams.  With the oracle's exp/log linked in
place of libm's the reference must reproduce the oracle bit for bit; with
glibc's exp/log the spike trains must agree and states stay within 1e-8.
The committed fixtures in tests/golden (made by tests/golden/make_golden.py
from oracle/_ref) carry the same check to machines without the reference.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle_lib import i_inputs, oracle, ra_inputs, ref, ref_built, run_i_trace, run_ra_trace

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N, STEPS, EVERY = 8, 2000, 100
RA_CASES = [(0, 0), (1, 0), (1, 1)]


@pytest.mark.parametrize("model,pois", RA_CASES)
def test_oracle_ra_matches_golden(model, pois):
    st, fi = run_ra_trace(oracle(), "ora", model, ra_inputs(N, STEPS), poisson=pois, seed=(7, 0), trial=2)
    g = np.load(os.path.join(GOLDEN, "ref_ora_ra_m%d_p%d.npz" % (model, pois)))
    np.testing.assert_array_equal(fi, g["fired"])
    np.testing.assert_array_equal(st[::EVERY], g["states"])
    assert (fi & 1).sum() > 0 and (fi & 2).sum() > 0
    gl = np.load(os.path.join(GOLDEN, "ref_glibc_ra_m%d_p%d.npz" % (model, pois)))
    np.testing.assert_array_equal(fi, gl["fired"])
    np.testing.assert_allclose(st[::EVERY], gl["states"], rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("first", [0, 1])
def test_oracle_i_matches_golden(first):
    st, fi = run_i_trace(oracle(), "ora", 0, i_inputs(N, STEPS, exc_scale=0.3), first_trial=first,
                         seed=(7, 0), trial=3)
    g = np.load(os.path.join(GOLDEN, "ref_ora_i_m0_f%d.npz" % first))
    np.testing.assert_array_equal(fi, g["fired"])
    np.testing.assert_array_equal(st[::EVERY], g["states"])
    assert fi.sum() > 0
    gl = np.load(os.path.join(GOLDEN, "ref_glibc_i_m0_f%d.npz" % first))
    np.testing.assert_array_equal(fi, gl["fired"])
    np.testing.assert_allclose(st[::EVERY], gl["states"], rtol=1e-8, atol=1e-12)


def test_poisson_formulas_match_reference_poisson_noise():
    """random(G) = G*u/max and get_spike_time = -log(1-u/max)/lambda on the same
    mt19937 stream as the reference's poisson_noise.cpp (one draw per call)."""
    g = np.load(os.path.join(GOLDEN, "ref_poisson_noise.npz"))
    raw = g["mt_raw"].astype(np.float64)
    # kind 0: random(0.45) consumes one draw per call
    np.testing.assert_array_equal(g["draws0"], 0.45 * raw / 4294967295.0)
    # kind 1: get_spike_time(250) with glibc log
    expect = np.array([-math.log(1 - u / 4294967295.0) / 250.0 for u in raw])
    np.testing.assert_array_equal(g["draws1"], expect)
    # the oracle's log agrees within 1 ulp on these arguments
    lib = oracle()
    for u in raw[:256]:
        a = -lib.ora_log(1 - u / 4294967295.0) / 250.0
        b = -math.log(1 - u / 4294967295.0) / 250.0
        assert abs(a - b) <= 2 * math.ulp(b)


def test_hhi_rk4_diverges_at_reference_dt():
    """HHI_buffer::Runge4_step (3/8 rule) is unstable at dt = 0.02 ms: every
    neuron that fires ends in NaN, which is why the RK4 form pairs HH2_buffer's
    RK4 with HHI_buffer's DP8.  The reference class shows the same."""
    inp = i_inputs(16, 25000, exc_scale=0.2)
    st, fi = run_i_trace(oracle(), "ora", 1, inp, first_trial=0, seed=(7, 0), trial=3)
    dead = np.isnan(st[-1, :, 0])
    assert dead.sum() >= 8
    r = ref("ora")
    if r is not None:
        st2, _ = run_i_trace(r, "ref", 1, inp, first_trial=0, seed=(7, 0), trial=3)
        np.testing.assert_array_equal(np.isnan(st2[-1, :, 0]), dead)
    st0, _ = run_i_trace(oracle(), "ora", 0, inp, first_trial=0, seed=(7, 0), trial=3)
    assert np.isfinite(st0).all()


needs_ref = pytest.mark.skipif(not ref_built("ora"), reason="oracle/_ref not built (reference checkout absent)")


@needs_ref
@pytest.mark.parametrize("model,pois", RA_CASES)
def test_oracle_ra_bit_exact_vs_reference_classes(model, pois):
    inp = ra_inputs(48, 5000, seed=11)
    a, fa = run_ra_trace(oracle(), "ora", model, inp, poisson=pois, seed=(3, 9), trial=5)
    b, fb = run_ra_trace(ref("ora"), "ref", model, inp, poisson=pois, seed=(3, 9), trial=5)
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(a, b)
    c, fc = run_ra_trace(ref("glibc"), "ref", model, inp, poisson=pois, seed=(3, 9), trial=5)
    np.testing.assert_array_equal(fa, fc)
    np.testing.assert_allclose(a, c, rtol=1e-8, atol=1e-12)


@needs_ref
@pytest.mark.parametrize("model,first", [(0, 0), (0, 1), (1, 0)])
def test_oracle_i_bit_exact_vs_reference_classes(model, first):
    steps = 5000 if model == 0 else 400
    inp = i_inputs(48, steps, seed=12, exc_scale=0.3 if model == 0 else 0.02)
    a, fa = run_i_trace(oracle(), "ora", model, inp, first_trial=first, seed=(3, 9), trial=5)
    b, fb = run_i_trace(ref("ora"), "ref", model, inp, first_trial=first, seed=(3, 9), trial=5)
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(a, b)
