"""Error paths of the engine's trial, through the C ABI.

* check_bad_values (HH2_test.cpp:244-260, NetworkGrowthSimulator.cpp:4538-4551):
  an HVC-RA neuron whose membrane potential leaves [-200, 100] mV or turns NaN
  ends the trial with HVC_E_NUMERIC and bad_values = -1 for its network.  A NaN
  resting potential (set_Erest through hvc_patch_neurons) makes the neuron's
  voltage NaN from the first step on.  In a batch only the network holding that
  neuron is flagged, and the other network's results are those of its run alone.
* A spike log smaller than the trial's spikes returns HVC_E_CAPACITY; the
  trial's spikes are then refused (HVC_E_STATE) until the next trial, and the
  next trial that fits runs normally.  The kernels check the log's capacity
  before every write, so the overflow itself writes nothing out of bounds.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _net(seed=3, n_ra=40, n_i=10):
    from unsupervised_growth_amd.network import random_network
    return random_network(n_ra, n_i, n_train=8, p_ra2i=0.3, p_i2ra=0.3, gie_max=1.0, gei_max=0.5,
                          spread_max=20.0, seed=seed, rr="dense")


def test_nan_resting_potential_flags_bad_values():
    from unsupervised_growth_amd._abi import HVC_E_NUMERIC, HVC_OK, Trial, TrialOut
    from unsupervised_growth_amd.engine import Engine
    net = _net()
    print("bad values: n_ra=%d n_i=%d" % (net.n_ra, net.n_i), flush=True)
    with Engine(trial_ms=60.0) as eng:
        eng.upload(net)
        t = Trial(0, 1, 0, 1, 20.0)
        o = TrialOut()
        assert eng.lib.hvc_run_trial(eng.ctx, ctypes.byref(t), ctypes.byref(o)) == HVC_OK
        assert o.bad_values == 0
        eng.patch_neurons([5], [float("nan")], [net.gca[5]])
        st = eng.lib.hvc_run_trial(eng.ctx, ctypes.byref(t), ctypes.byref(o))
        assert st == HVC_E_NUMERIC, st
        assert o.bad_values == -1
        assert b"check_bad_values" in eng.lib.hvc_last_error(eng.ctx)
        r = eng.run_trial(mode=0, trial_index=1, mean_injection_time=20.0, want_state=True)
        assert r.bad_values == -1
        assert np.isnan(r.ra_state[5, 0]) and np.isfinite(np.delete(r.ra_state[:, :2], 5, axis=0)).all()
        # the neuron's spikes stop; the state of the others stays in range
        assert len(r.soma[5]) == 0


def test_bad_values_flag_only_the_failing_network_of_a_batch():
    from unsupervised_growth_amd.engine import Engine
    nets = [_net(seed=3), _net(seed=4, n_ra=30, n_i=8)]
    keys = [(70, 1), (71, 1)]
    kw = [dict(mode=0, training=True, trial_index=2, first_trial=True, mean_injection_time=20.0)] * 2
    with Engine(trial_ms=60.0, seed=keys[0]) as e:
        e.upload(nets[0])
        alone = e.run_trial(**kw[0])
    with Engine(trial_ms=60.0) as e:
        e.upload_batch(nets, keys)
        ids = np.array([nets[0].n_ra + 7], np.int32)   # neuron 7 of network 1 in the merged ids
        e.patch_neurons(ids, [float("nan")], [nets[1].gca[7]])
        res = e.run_batch(kw)
    assert res[0].bad_values == 0 and res[1].bad_values == -1
    for kind in ("soma", "dend", "interneuron"):
        for a, b in zip(getattr(alone, kind), getattr(res[0], kind)):
            np.testing.assert_array_equal(a, b)


def test_spike_log_overflow_returns_capacity_and_refuses_spikes():
    from unsupervised_growth_amd._abi import HVC_E_CAPACITY, HVC_E_STATE, HVC_OK, Trial, TrialOut
    from unsupervised_growth_amd.engine import Engine
    net = _net()
    kicked = dict(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=20.0)
    quiet = dict(mode=1, trial_index=1, mean_injection_time=1000.0)   # no kick inside the trial
    with Engine(trial_ms=150.0) as ref:
        ref.upload(net)
        full = ref.run_trial(**kicked)
        calm = ref.run_trial(**quiet)
    count = lambda r: [sum(map(len, getattr(r, k))) for k in ("soma", "dend", "interneuron")]
    cap = max(count(calm)) + 1
    print("capacity: kicked trial %s spikes, quiet trial %s, log capacity %d" % (count(full), count(calm), cap),
          flush=True)
    assert count(full)[0] > cap
    with Engine(trial_ms=150.0, spike_capacity=cap) as eng:
        eng.upload(net)
        t = Trial(0, 1, 0, 1, 20.0)
        o = TrialOut()
        st = eng.lib.hvc_run_trial(eng.ctx, ctypes.byref(t), ctypes.byref(o))
        assert st == HVC_E_CAPACITY, st
        assert b"capacity" in eng.lib.hvc_last_error(eng.ctx)
        tot = ctypes.c_int()
        assert eng.lib.hvc_trial_spikes(eng.ctx, 0, None, None, 0, ctypes.byref(tot)) == HVC_E_STATE
        # the next trial that fits the log runs normally (the failed trial left
        # the synapse state undefined, so only the log's bound is checked)
        t2 = Trial(1, 0, 1, 0, 1000.0)
        assert eng.lib.hvc_run_trial(eng.ctx, ctypes.byref(t2), ctypes.byref(o)) == HVC_OK
        assert eng.lib.hvc_trial_spikes(eng.ctx, 0, None, None, 0, ctypes.byref(tot)) == HVC_OK
        assert tot.value <= cap
