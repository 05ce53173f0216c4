"""FP64 parity of the HIP engine against the oracle, through the C ABI.

The FP64 form evaluates the reference's expression trees operation by
operation with the same exp/log, noise streams and accumulation order as the
oracle, so every comparison here is bit-exact: spike times, neuron states,
weights, active / super lists (in insertion order) and remodeled flags.
"""
import numpy as np
import pytest

from oracle_lib import OracleSim, i_inputs, oracle, ra_inputs, run_i_trace, run_ra_trace

pytestmark = pytest.mark.gpu


def _engine(**kw):
    from unsupervised_growth_amd.engine import Engine
    return Engine(**kw)


def _same_spikes(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("model,poisson", [(0, 0), (1, 0), (1, 1)])
def test_trace_ra_bit_exact(model, poisson):
    n, steps = 64, 3000
    inp = ra_inputs(n, steps)
    print("trace_ra model=%d poisson=%d n=%d steps=%d" % (model, poisson, n, steps), flush=True)
    ref_state, ref_fired = run_ra_trace(oracle(), "ora", model, inp, poisson=poisson, seed=(7, 0), trial=2)
    with _engine(ra_model=model, ra_poisson=poisson, seed=(7, 0)) as eng:
        st, fi = eng.trace_ra(*inp, trial=2)
    assert (ref_fired & 1).sum() > 0
    np.testing.assert_array_equal(fi, ref_fired)
    np.testing.assert_array_equal(st, ref_state)


@pytest.mark.parametrize("model,first", [(0, 0), (0, 1), (1, 0)])
def test_trace_i_bit_exact(model, first):
    n, steps = 64, 2000 if model == 0 else 300
    inp = i_inputs(n, steps, exc_scale=0.3 if model == 0 else 0.05)
    print("trace_i model=%d first=%d n=%d steps=%d" % (model, first, n, steps), flush=True)
    ref_state, ref_fired = run_i_trace(oracle(), "ora", model, inp, first_trial=first, seed=(7, 0), trial=3)
    with _engine(i_model=model, seed=(7, 0)) as eng:
        st, fi = eng.trace_i(*inp, trial=3, first_trial=first)
    np.testing.assert_array_equal(fi, ref_fired)
    np.testing.assert_array_equal(st, ref_state)


def _small_net(rr="dense", seed=3, n_ra=40, n_i=10):
    from unsupervised_growth_amd.network import random_network
    return random_network(n_ra, n_i, n_train=8, p_ra2i=0.3, p_i2ra=0.3, gie_max=1.0, gei_max=0.5,
                          spread_max=20.0, seed=seed, rr=rr)


def _compare_trial(eng_res, ora_res):
    _same_spikes(eng_res.soma, ora_res["soma"])
    _same_spikes(eng_res.dend, ora_res["dend"])
    _same_spikes(eng_res.interneuron, ora_res["interneuron"])
    assert eng_res.warnings == ora_res["warnings"]
    np.testing.assert_array_equal(eng_res.ra_state, ora_res["ra_state"])
    np.testing.assert_array_equal(eng_res.i_state, ora_res["i_state"])


def _compare_synapses(syn, osyn):
    np.testing.assert_array_equal(syn.w, osyn["w"])
    for k in ("act_cnt", "act_ids", "sup_cnt", "sup_ids", "remodeled"):
        np.testing.assert_array_equal(getattr(syn, k), osyn[k], err_msg=k)


@pytest.mark.parametrize("rr,ra_model,ra_lanes,i_lanes", [("dense", 0, 0, 1), (12, 0, 0, 2), ("dense", 1, 1, 2),
                                                             (12, 1, 0, 1), (12, 1, 1, 0), ("dense", 0, 0, 3),
                                                             (12, 1, 1, 3)])
def test_growth_trials_bit_exact(rr, ra_model, ra_lanes, i_lanes):
    """Full growth trials in every integration layout: HVC-RA (RK4 form) with
    four lanes or one thread per neuron, HVC-I (DP8) with four, five or eight
    lanes or one thread per neuron."""
    net = _small_net(rr=rr)
    trial_ms = 300.0
    print("growth rr=%s ra_model=%d ra_lanes=%d i_lanes=%d n_ra=%d n_i=%d trial_ms=%g" % (
        rr, ra_model, ra_lanes, i_lanes, net.n_ra, net.n_i, trial_ms), flush=True)
    ora = OracleSim(net, trial_ms=trial_ms, ra_model=ra_model)
    with _engine(trial_ms=trial_ms, ra_model=ra_model) as eng:
        eng.upload(net)
        eng.set_tuning(ra_lanes=ra_lanes, i_lanes=i_lanes)
        for tr in range(3):
            kw = dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0),
                      mean_injection_time=100.0 + 10 * tr)
            o = ora.run_trial(**kw)
            r = eng.run_trial(want_state=True, **kw)
            print("  trial %d spikes soma=%d I=%d" % (tr, sum(map(len, r.soma)), sum(map(len, r.interneuron))),
                  flush=True)
            _compare_trial(r, o)
            _compare_synapses(eng.download_synapses(), ora.synapses())
        assert ora.synapses()["sup_cnt"].sum() > 0, "test network should reach supersynapses"


def test_growth_rows_beyond_the_stream_cache_bit_exact():
    """RA->RA rows with more spikes in flight than the drain keeps in LDS (8):
    tonic firing (mean somatic drive 0.6) with 100-200 ms RA->RA delays holds
    9-14 spikes per row in flight, so those rows drain through their stream
    state in memory and drop exhausted streams there."""
    from unsupervised_growth_amd.network import NoiseParams, random_network
    net = random_network(30, 8, n_train=6, p_ra2i=0.3, p_i2ra=0.3, gie_max=1.0, gei_max=0.5, spread_max=20.0,
                         seed=5, rr=8, rr_delays=("uniform", 100.0, 200.0))
    noise = NoiseParams(white_noise_mean_soma=0.6)
    trial_ms = 400.0
    print("growth many streams n_ra=%d n_i=%d trial_ms=%g" % (net.n_ra, net.n_i, trial_ms), flush=True)
    ora = OracleSim(net, trial_ms=trial_ms, noise=noise)
    most = 0
    with _engine(trial_ms=trial_ms, noise=noise) as eng:
        eng.upload(net)
        for tr in range(2):
            kw = dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0), mean_injection_time=100.0)
            o = ora.run_trial(**kw)
            r = eng.run_trial(want_state=True, **kw)
            _compare_trial(r, o)
            _compare_synapses(eng.download_synapses(), ora.synapses())
            for i, sp in enumerate(o["soma"]):
                sp = np.sort(np.asarray(sp))
                dmax = net.rr_d[net.rr_ptr[i]:net.rr_ptr[i + 1]].max()
                for t in sp:
                    most = max(most, int(np.sum((sp >= t) & (sp <= t + dmax))))
    print("  most spikes of a row within its longest delay: %d" % most, flush=True)
    assert 8 < most <= 16


def test_chain_trial_bit_exact():
    net = _small_net()
    trial_ms = 300.0
    ora = OracleSim(net, trial_ms=trial_ms)
    with _engine(trial_ms=trial_ms) as eng:
        eng.upload(net)
        for tr in range(2):   # grow first so that active synapses exist
            ora.run_trial(mode=0, trial_index=tr, first_trial=(tr == 0))
            eng.run_trial(mode=0, trial_index=tr, first_trial=(tr == 0))
        for tr in range(2, 4):
            o = ora.run_trial(mode=1, trial_index=tr, mean_injection_time=100.0)
            r = eng.run_trial(mode=1, trial_index=tr, mean_injection_time=100.0, want_state=True)
            _compare_trial(r, o)
        assert sum(map(len, r.soma)) > 0
