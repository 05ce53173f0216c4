"""Batched form: several networks of different sizes in one context.

Each network's noise streams are keyed by its own seed and its own neuron ids,
and every ordered sum runs over its own neurons only, so network b of a batch
must produce exactly what a context holding network b alone produces -- spike
trains, neuron states, weights, active / super lists -- in both precision
forms.  The FP64 single-network form is itself bit-exact with the oracle
(test_gpu_parity.py), so this pins the batch to the oracle as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _nets():
    from unsupervised_growth_amd.network import random_network
    sizes = [(40, 10, "dense"), (25, 7, 8), (60, 16, 12), (33, 9, "dense")]
    return [random_network(nra, ni, n_train=6, p_ra2i=0.3, p_i2ra=0.3, gie_max=1.0, gei_max=0.5,
                           spread_max=20.0, seed=11 + k, rr=rr) for k, (nra, ni, rr) in enumerate(sizes)]


def _trial(b, tr, mode=0):
    return dict(mode=mode, training=True, trial_index=tr + 3 * b, first_trial=(tr == 0),
                mean_injection_time=100.0 + 7 * b + 3 * tr)


def _same(a, b, what):
    assert len(a) == len(b), what
    for x, y in zip(a, b):
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y), err_msg=what)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_each_network_alone(precision):
    from unsupervised_growth_amd.engine import Engine
    nets = _nets()
    keys = [(100 + b, 7) for b in range(len(nets))]
    trial_ms = 250.0
    print("batch precision=%d sizes=%s" % (precision, [(n.n_ra, n.n_i) for n in nets]), flush=True)
    alone = []
    for b, net in enumerate(nets):
        with Engine(precision=precision, trial_ms=trial_ms, seed=keys[b]) as e:
            e.upload(net)
            res = [e.run_trial(**_trial(b, tr)) for tr in range(3)]
            syn = e.download_synapses()
            res.append(e.run_trial(**_trial(b, 3, mode=1)))
            alone.append((res, syn))
    with Engine(precision=precision, trial_ms=trial_ms) as e:
        e.upload_batch(nets, keys)
        batch = [e.run_batch([_trial(b, tr) for b in range(len(nets))]) for tr in range(3)]
        syns = [e.download_synapses(b) for b in range(len(nets))]
        batch.append(e.run_batch([_trial(b, 3, mode=1) for b in range(len(nets))]))
    spikes = 0
    for b in range(len(nets)):
        res, syn = alone[b]
        for tr in range(4):
            for kind in ("soma", "dend", "interneuron"):
                _same(getattr(batch[tr][b], kind), getattr(res[tr], kind), "net %d trial %d %s" % (b, tr, kind))
            spikes += sum(map(len, res[tr].soma))
        for k in ("w", "act_cnt", "act_ids", "sup_cnt", "sup_ids", "remodeled"):
            np.testing.assert_array_equal(getattr(syns[b], k), getattr(syn, k), err_msg="net %d %s" % (b, k))
    assert spikes > 0
