"""Pins of the oracle's trial logic (oracle/ora_trial.c).

The reference trial (NetworkGrowthSimulator.cpp:8839-9571, 12747-13118) needs
Boost and libconfig++, which this image lacks, so it is not built.  Its logic
is pinned here by
  * an independent line-by-line Python restatement with the reference's data
    structures (tests/pyref_trial.py), compared bit for bit on small networks
    that drive every branch (activation, supersynapse, saturation,
    desaturation, axon remodeling, LTD on delivery, LTP at sync);
  * closed-form known answers of LTP / LTD (NGS:16615-16652);
  * invariants (a network without synapses behaves as isolated neurons; a
    chain test leaves synapses untouched; the threaded neuron steps equal one
    thread).
Known answers derived by hand for tiny networks (arrival steps, zero delay,
sync application, every STDP branch, update_synapse, remodeling, decay) are in
test_oracle_pins.py.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import OraSynParams, OracleSim, oracle
from pyref_trial import PyRefSim


def _net(seed=3, rr="dense", n_ra=14, n_i=5, spread=20.0, w_rr=0.0, n_train=7):
    from unsupervised_growth_amd.network import random_network
    return random_network(n_ra, n_i, n_train=n_train, p_ra2i=0.4, p_i2ra=0.4, gie_max=1.0, gei_max=0.5,
                          spread_max=spread, seed=seed, rr=rr, w_rr=w_rr)


def _syn(**kw):
    from unsupervised_growth_amd.network import SynapticParams
    return SynapticParams(**kw)


def _cmp_trial(a, b):
    assert a["rc"] == b["rc"]
    for key in ("soma", "dend", "interneuron"):
        assert len(a[key]) == len(b[key])
        for x, y in zip(a[key], b[key]):
            np.testing.assert_array_equal(np.asarray(x, float), np.asarray(y, float), err_msg=key)
    assert a["warnings"] == b["warnings"]
    np.testing.assert_array_equal(a["ra_state"], b["ra_state"])
    np.testing.assert_array_equal(a["i_state"], b["i_state"])


def _cmp_syn(a, b):
    for k in ("w", "act_cnt", "act_ids", "sup_cnt", "sup_ids", "remodeled"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def _net_desaturating():
    """Neuron 0 saturated (Nss = 1) with supersynapse 0 -> 1; neuron 1 fires 20 ms
    before neuron 0, so the delivery 0 -> 1 meets a recent post event of 1: LTD in
    the saturated branch drops the weight below the supersynapse threshold and the
    row desaturates (NGS:9104-9146)."""
    net = _net(seed=8, n_ra=8, n_i=3, spread=0.0, n_train=2)
    net.train_ids = np.array([0, 1], np.int32)
    net.spread = np.array([20.0, 0.0])
    e01 = int(net.rr_ptr[0]) + 1
    net.rr_w[e01] = 1.5
    net.act_cnt[0] = 1
    net.act_ids = np.array([1], np.int32)
    net.sup_cnt[0] = 1
    net.sup_ids = np.array([1], np.int32)
    return net


SCENARIOS = [
    # (name, net kwargs or builder, synaptic params, trials)
    ("dense-default", dict(seed=3), dict(), 3),
    ("saturated-ltd-desaturation", _net_desaturating,
     dict(Nss=1, SUPERSYNAPSE_THRESHOLD=1.0, ACTIVATION_THRESHOLD=0.5, A_D=0.9, T_0=0.0, T_D=50.0), 2),
    ("saturating", dict(seed=4, w_rr=0.4), dict(Nss=1, SUPERSYNAPSE_THRESHOLD=1.0, ACTIVATION_THRESHOLD=0.3), 3),
    ("candidates", dict(seed=5, rr=6), dict(Nss=2), 3),
    ("desaturating", dict(seed=6, w_rr=1.2), dict(Nss=2, SUPERSYNAPSE_THRESHOLD=1.0, ACTIVATION_THRESHOLD=0.5,
                                                   A_D=0.9, T_0=-2.0, BETA=0.3, BETA_SUPERSYNAPSE=0.3), 3),
]


@pytest.mark.parametrize("name,netkw,synkw,trials", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_oracle_trial_matches_python_restatement(name, netkw, synkw, trials):
    net = netkw() if callable(netkw) else _net(**netkw)
    syn = _syn(**synkw)
    trial_ms = 200.0
    a = OracleSim(net, trial_ms=trial_ms, syn=syn)
    b = PyRefSim(net, syn, trial_ms=trial_ms)
    events = 0
    for tr in range(trials):
        kw = dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0), mean_injection_time=50.0 + 7 * tr)
        ra, rb = a.run_trial(**kw), b.run_trial(**kw)
        _cmp_trial(ra, rb)
        _cmp_syn(a.synapses(), b.synapses())
        events += sum(map(len, ra["soma"]))
    ra, rb = a.run_trial(mode=1, trial_index=trials), b.run_trial(mode=1, trial_index=trials)
    _cmp_trial(ra, rb)
    assert events > 0


def test_scenarios_reach_every_branch():
    """Together the scenarios drive every branch of the trial's STDP logic, so a
    mistake in any of them shows up as a mismatch above."""
    hits = {}
    for name, netkw, synkw, trials in SCENARIOS:
        net = netkw() if callable(netkw) else _net(**netkw)
        b = PyRefSim(net, _syn(**synkw), trial_ms=200.0)
        for tr in range(trials):
            b.run_trial(mode=0, trial_index=tr, first_trial=(tr == 0), mean_injection_time=50.0 + 7 * tr)
        for k, v in b.hits.items():
            hits[k] = hits.get(k, 0) + v
    for branch in ("activate", "super", "unsuper", "deactivate", "remodel", "ltd_delivery",
                   "ltd_delivery_saturated", "desaturation", "sync_ltp", "sync_ltd", "sync_saturated"):
        assert hits.get(branch, 0) > 0, branch


def _p(**kw):
    p = OraSynParams()
    base = dict(Nss=2, A_P=2.0, A_D=0.2, T_0=1.0, T_P=5.0, T_D=5.0, TAU_P=20.0, TAU_D=20.0, BETA=0.05,
                BETA_SUPERSYNAPSE=0.02, ACTIVATION_THRESHOLD=0.5, SUPERSYNAPSE_THRESHOLD=2.0, WEIGHT_MAX=10.0)
    base.update(kw)
    for k, v in base.items():
        setattr(p, k, v)
    return p


def _ltp(p, w, dt):
    rej = ctypes.c_int()
    return oracle().ora_ltp(ctypes.byref(p), w, dt, ctypes.byref(rej)), rej.value


def _ltd(p, w, dt):
    rej = ctypes.c_int()
    return oracle().ora_ltd(ctypes.byref(p), w, dt, ctypes.byref(rej)), rej.value


def test_ltp_known_answers():
    p = _p()
    # dt <= T_0: rejected, weight unchanged (NGS:16618-16622)
    assert _ltp(p, 1.5, 1.0) == (1.5, 1)
    assert _ltp(p, 1.5, -3.0) == (1.5, 1)
    # linear branch: dt = T_0 + T_P/2 adds A_P/2
    assert _ltp(p, 1.0, 3.5) == (2.0, 0)
    # boundary dt = T_0 + T_P is still linear: adds A_P
    assert _ltp(p, 1.0, 6.0) == (3.0, 0)
    # exponential branch: one TAU_P past the boundary adds A_P / e
    w, r = _ltp(p, 1.0, 26.0)
    assert r == 0 and abs(w - (1.0 + 2.0 / np.e)) < 1e-15
    # clamp to WEIGHT_MAX
    assert _ltp(p, 9.5, 6.0) == (10.0, 0)


def test_ltd_known_answers():
    p = _p()
    # dt > T_0: rejected (NGS:16639-16643)
    assert _ltd(p, 1.5, 1.5) == (1.5, 1)
    # dt = T_0 is the linear branch with zero change
    assert _ltd(p, 1.5, 1.0) == (1.5, 0)
    # linear: dt = T_0 - T_D/2 scales w by (1 - A_D/2)
    assert _ltd(p, 2.0, 1.0 - 2.5) == (2.0 * (1 - 0.1), 0)
    # boundary dt = T_0 - T_D: w (1 - A_D)
    w, _ = _ltd(p, 2.0, -4.0)
    assert abs(w - 2.0 * (1 - 0.2)) < 1e-15
    # exponential: one TAU_D below the boundary removes w A_D / e
    w, r = _ltd(p, 2.0, -24.0)
    assert r == 0 and abs(w - (2.0 - 2.0 * 0.2 / np.e)) < 1e-15
    # clamp at 0
    assert _ltd(_p(A_D=3.0), 1.0, -4.0) == (0.0, 0)


def test_clock_closed_form():
    """internal_time += 0.02 for 24999 steps; syncs every 0.25 ms (NGS:8906-9563)."""
    lib = oracle()
    n = lib.ora_num_steps(500.0, 0.02)
    assert n == 25000
    clock = np.zeros(n)
    sync = np.zeros(n, np.int32)
    lib.ora_clock(500.0, 0.02, 0.25, clock.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                  sync.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    t = 0.0
    for k in range(1, n):
        t += 0.02
        assert clock[k] == t
    steps = np.nonzero(sync)[0]
    assert len(steps) == 2000 and steps[0] == 13 and steps[-1] == 24999
    gaps = np.diff(np.concatenate([[0], steps]))
    assert sorted(set(gaps.tolist())) == [11, 12, 13]
    assert clock[-1] == 499.97999999983523


def test_isolated_neurons_equal_traces():
    """Without connections a trial is the neurons' own traces (kick at the kick step)."""
    from oracle_lib import run_i_trace
    from unsupervised_growth_amd.network import random_network
    net = random_network(6, 3, n_train=2, p_ra2i=0.0, p_i2ra=0.0, rr=1, seed=9)
    sim = OracleSim(net, trial_ms=60.0)
    out = sim.run_trial(mode=0, trial_index=4, first_trial=False, mean_injection_time=20.0)
    steps = 2999
    st, fi = run_i_trace(oracle(), "ora", 0, (np.zeros((steps, 3)), np.zeros((steps, 3))), trial=4, seed=(1, 0))
    np.testing.assert_array_equal(out["i_state"], st[-1])
    t = np.cumsum(np.full(steps, 0.02))
    for k in range(3):
        np.testing.assert_array_equal(out["interneuron"][k], t[fi[:, k] == 1])


def test_chain_trial_leaves_synapses_unchanged():
    net = _net(seed=7, w_rr=0.6)
    a = OracleSim(net, trial_ms=150.0, syn=_syn(ACTIVATION_THRESHOLD=0.5))
    a.run_trial(mode=0, trial_index=0, first_trial=True)
    before = a.synapses()
    a.run_trial(mode=1, trial_index=1)
    _cmp_syn(before, a.synapses())


def test_threaded_neuron_steps_equal_one_thread():
    """Above 512 neurons the oracle steps its neurons on OpenMP threads before
    the step's deliveries; the trial must equal the one-thread run bit for bit."""
    from unsupervised_growth_amd.configs import sphere_network
    net = sphere_network(np.random.default_rng(2), 700, 190, 5.0, k_rr=30, n_train=12, spread_max=5.0)
    lib = oracle()
    threads = lib.ora_threads()
    out = []
    for n in (1, max(threads, 2)):
        lib.ora_set_threads(n)
        sim = OracleSim(net, trial_ms=40.0, seed=(4, 2))
        r = [sim.run_trial(mode=0, trial_index=k, first_trial=(k == 0), mean_injection_time=10.0) for k in range(2)]
        out.append((r, sim.synapses()))
        sim.close()
    lib.ora_set_threads(threads)
    (ra, sa), (rb, sb) = out
    assert sum(map(len, ra[1]["soma"])) > 0
    for x, y in zip(ra, rb):
        _cmp_trial(x, y)
    _cmp_syn(sa, sb)
