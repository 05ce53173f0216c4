"""FP32 form against the FP64 oracle.

The FP32 form integrates membrane, gating and conductance state in float (the
clock, spike times, noise event times and STDP stay FP64), so its results are
judged by the spikes it produces, with the tolerance of the precision form:
every spike of the oracle is matched by a spike of the FP32 run within one
time step (0.02 ms) and vice versa -- at most 1 % of the spikes may be
unmatched -- and the largest |dt| over matched spikes is at most one step.
Synapse state is compared exactly when the spike trains are identical (STDP
then sees the same FP64 spike times).
"""
import numpy as np
import pytest

from oracle_lib import OracleSim, i_inputs, oracle, ra_inputs, run_i_trace, run_ra_trace

pytestmark = pytest.mark.gpu

DT = 0.02
MAX_DT = DT + 1e-9          # one time step
MAX_UNMATCHED = 0.01        # fraction of all spikes


def _engine(**kw):
    from unsupervised_growth_amd.engine import Engine
    return Engine(precision=1, **kw)


def _trains(fired, bit):
    t = np.arange(1, fired.shape[0] + 1) * DT
    return [t[(fired[:, k] & bit) != 0] for k in range(fired.shape[1])]


def _check(ref, test, what):
    from unsupervised_growth_amd.spikes import compare_trains
    c = compare_trains(ref, test, window=1.0)
    print("%s: %s" % (what, c), flush=True)
    assert c["spikes_ref"] > 0, what
    assert c["unpaired"] <= MAX_UNMATCHED * (c["spikes_ref"] + c["spikes_test"]), (what, c)
    assert c["max_abs_dt_ms"] <= MAX_DT, (what, c)
    return c


@pytest.mark.parametrize("model,poisson", [(0, 0), (1, 0), (1, 1)])
def test_fp32_trace_ra_spikes_within_one_step(model, poisson):
    n, steps = 64, 6000
    inp = ra_inputs(n, steps)
    print("fp32 trace_ra model=%d poisson=%d n=%d steps=%d" % (model, poisson, n, steps), flush=True)
    _, ref_fired = run_ra_trace(oracle(), "ora", model, inp, poisson=poisson, seed=(7, 0), trial=2)
    with _engine(ra_model=model, ra_poisson=poisson, seed=(7, 0)) as eng:
        st, fi = eng.trace_ra(*inp, trial=2)
    assert np.isfinite(st).all()
    _check(_trains(ref_fired, 1), _trains(fi, 1), "soma")
    _check(_trains(ref_fired, 2), _trains(fi, 2), "dend")


def test_fp32_trace_i_spikes_within_one_step():
    n, steps = 64, 6000
    inp = i_inputs(n, steps, exc_scale=0.3)
    print("fp32 trace_i n=%d steps=%d" % (n, steps), flush=True)
    _, ref_fired = run_i_trace(oracle(), "ora", 0, inp, first_trial=0, seed=(7, 0), trial=3)
    with _engine(i_model=0, seed=(7, 0)) as eng:
        st, fi = eng.trace_i(*inp, trial=3, first_trial=0)
    assert np.isfinite(st).all()
    _check(_trains(ref_fired, 1), _trains(fi, 1), "interneuron")


def _net(seed=3, n_ra=60, n_i=16, rr="dense"):
    from unsupervised_growth_amd.network import random_network
    return random_network(n_ra, n_i, n_train=10, p_ra2i=0.3, p_i2ra=0.3, gie_max=1.0, gei_max=0.5,
                          spread_max=20.0, seed=seed, rr=rr)


@pytest.mark.parametrize("ra_model", [0, 1])
def test_fp32_growth_trials_against_oracle(ra_model):
    net = _net()
    trial_ms = 300.0
    print("fp32 growth ra_model=%d n_ra=%d n_i=%d" % (ra_model, net.n_ra, net.n_i), flush=True)
    ora = OracleSim(net, trial_ms=trial_ms, ra_model=ra_model)
    with _engine(trial_ms=trial_ms, ra_model=ra_model) as eng:
        eng.upload(net)
        for tr in range(3):
            kw = dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0),
                      mean_injection_time=100.0 + 10 * tr)
            o = ora.run_trial(**kw)
            r = eng.run_trial(**kw)
            same = True
            for kind in ("soma", "dend", "interneuron"):
                c = _check(o[kind], getattr(r, kind), "trial %d %s" % (tr, kind))
                same = same and c["unpaired"] == 0 and c["max_abs_dt_ms"] == 0.0
            syn, osyn = eng.download_synapses(), ora.synapses()
            if same:
                np.testing.assert_array_equal(syn.w, osyn["w"])
                np.testing.assert_array_equal(syn.act_ids, osyn["act_ids"])
                np.testing.assert_array_equal(syn.sup_ids, osyn["sup_ids"])
            else:
                np.testing.assert_allclose(syn.w, osyn["w"], rtol=0, atol=0.5)


def test_fp32_chain_trial_against_oracle():
    net = _net(seed=4)
    trial_ms = 300.0
    ora = OracleSim(net, trial_ms=trial_ms)
    with _engine(trial_ms=trial_ms) as eng:
        eng.upload(net)
        for tr in range(2):
            ora.run_trial(mode=0, trial_index=tr, first_trial=(tr == 0))
        eng.upload_synapses(_synapses_of(ora.synapses()))
        for tr in range(2, 4):
            o = ora.run_trial(mode=1, trial_index=tr, mean_injection_time=100.0)
            r = eng.run_trial(mode=1, trial_index=tr, mean_injection_time=100.0)
            for kind in ("soma", "dend", "interneuron"):
                _check(o[kind], getattr(r, kind), "chain %d %s" % (tr, kind))


def _synapses_of(osyn):
    from unsupervised_growth_amd.engine import SynapseState
    return SynapseState(w=osyn["w"], act_cnt=osyn["act_cnt"], act_ids=osyn["act_ids"], sup_cnt=osyn["sup_cnt"],
                        sup_ids=osyn["sup_ids"], remodeled=osyn["remodeled"])
