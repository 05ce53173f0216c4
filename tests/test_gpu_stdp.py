"""The STDP branch scenarios on the HIP engine, bit-exact against the oracle.

The five scenarios of test_oracle_trial.py together drive every branch of the
growth trial's synapse-state machine (NetworkGrowthSimulator.cpp:9091-9191 LTD
on delivery with saturation and desaturation, 9375-9554 sync-point LTP / LTD
and axon remodeling, 16546-16592 update_synapse).  Here the FP64 engine runs
them through the C ABI: spikes, neuron states, weights, active / super lists
and remodeled flags equal the oracle's after every trial, and the engine's own
branch counters (hvc_branch_counts) equal the pass counts of the independent
Python restatement on the same trials, so the GPU is shown to have taken the
saturated-LTD and desaturation branches, not only to end in the same state.
"""
import numpy as np
import pytest

from oracle_lib import OracleSim
from pyref_trial import PyRefSim
from test_oracle_trial import SCENARIOS, _net, _syn

pytestmark = pytest.mark.gpu

TRIAL_MS = 200.0


def _kw(tr):
    return dict(mode=0, training=True, trial_index=tr, first_trial=(tr == 0), mean_injection_time=50.0 + 7 * tr)


@pytest.mark.parametrize("name,netkw,synkw,trials", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_stdp_scenario_bit_exact_with_branch_counts(name, netkw, synkw, trials):
    from unsupervised_growth_amd.engine import Engine
    net = netkw() if callable(netkw) else _net(**netkw)
    syn = _syn(**synkw)
    print("stdp scenario %s n_ra=%d n_i=%d trials=%d" % (name, net.n_ra, net.n_i, trials), flush=True)
    ora = OracleSim(net, trial_ms=TRIAL_MS, syn=syn)
    py = PyRefSim(net, syn, trial_ms=TRIAL_MS)
    with Engine(trial_ms=TRIAL_MS, syn=syn) as eng:
        eng.upload(net)
        eng.set_branch_counting(True)
        for tr in range(trials):
            o = ora.run_trial(**_kw(tr))
            py.run_trial(**_kw(tr))
            r = eng.run_trial(want_state=True, **_kw(tr))
            for kind in ("soma", "dend", "interneuron"):
                a, b = getattr(r, kind), o[kind]
                assert len(a) == len(b)
                for x, y in zip(a, b):
                    np.testing.assert_array_equal(np.asarray(x), np.asarray(y), err_msg=kind)
            assert r.warnings == o["warnings"]
            np.testing.assert_array_equal(r.ra_state, o["ra_state"])
            np.testing.assert_array_equal(r.i_state, o["i_state"])
            s, os_ = eng.download_synapses(), ora.synapses()
            for k in ("w", "act_cnt", "act_ids", "sup_cnt", "sup_ids", "remodeled"):
                np.testing.assert_array_equal(getattr(s, k), os_[k], err_msg="trial %d %s" % (tr, k))
        hits = eng.branch_counts()
    print("  engine branch counts: %s" % hits, flush=True)
    print("  restatement counts:   %s" % dict(sorted(py.hits.items())), flush=True)
    for branch, n in hits.items():
        assert n == py.hits.get(branch, 0), branch
    if name == "saturated-ltd-desaturation":
        assert hits["ltd_delivery_saturated"] > 0 and hits["desaturation"] > 0 and hits["unsuper"] > 0


def test_scenarios_reach_every_branch_on_the_engine():
    from unsupervised_growth_amd.engine import Engine
    total = {}
    for name, netkw, synkw, trials in SCENARIOS:
        net = netkw() if callable(netkw) else _net(**netkw)
        with Engine(trial_ms=TRIAL_MS, syn=_syn(**synkw)) as eng:
            eng.upload(net)
            eng.set_branch_counting(True)
            for tr in range(trials):
                eng.run_trial(**_kw(tr))
            for k, v in eng.branch_counts().items():
                total[k] = total.get(k, 0) + v
    print("engine branch counts over the scenarios: %s" % total, flush=True)
    for branch, n in total.items():
        assert n > 0, branch
