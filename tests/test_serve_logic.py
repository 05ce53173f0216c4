"""Host logic of the serving loop with a stand-in pool (no device): first come,
first served admission by capacity, rounds over every resident request,
release when complete, latency statistics."""
from types import SimpleNamespace

import numpy as np

from unsupervised_growth_amd.serve import Request, Server


class _Params:
    trial_ms = 10.0
    dt = 0.02


class FakePool:
    """Capacity in neurons; run() records which handles ran together."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.used = 0
        self.next = 1
        self.resident = {}
        self.params = _Params()
        self.rounds = []

    def admit(self, net, key):
        n = net.n_ra + net.n_i
        if self.used + n > self.capacity:
            return None
        h = self.next
        self.next += 1
        self.resident[h] = n
        self.used += n
        return h

    def release(self, h):
        self.used -= self.resident.pop(h)

    def run(self, trials):
        self.rounds.append(sorted(trials))
        out = {}
        for h, kw in trials.items():
            n = self.resident[h]
            # neuron 0 fires at the kick time + trial index, the others stay silent
            soma = [np.array([kw["mean_injection_time"] + kw["trial_index"] % 100])] + [np.zeros(0)] * (n - 1)
            out[h] = SimpleNamespace(tag=("result", h, kw["trial_index"]), soma=soma, dend=[np.zeros(0)] * n)
        return out

    def download_synapses(self, h):
        return ("syn", h)


class _Net:
    def __init__(self, n_ra, n_i):
        self.n_ra, self.n_i = n_ra, n_i


def _req(rid, arrival, n, trials, mode=1):
    return Request(rid=rid, arrival=arrival, net=_Net(n, 0), key=(rid, 0), onsets=np.full(trials, 100.0),
                   mode=mode, first_trial_index=100 * rid)


def test_fifo_admission_and_release():
    pool = FakePool(capacity=10)
    reqs = [_req(0, 0, 6, 2), _req(1, 0, 6, 1), _req(2, 0, 3, 1), _req(3, 5, 2, 1, mode=0)]
    stats = Server(pool).run(reqs, realtime=False)
    # request 1 does not fit beside 0, and request 2 waits behind it (first come, first served)
    assert pool.rounds[0] == [1] and pool.rounds[1] == [1]
    assert reqs[1].admitted == 2 and reqs[2].admitted == 2
    assert pool.rounds[2] == [2, 3]
    # every request ran its trials with consecutive trial indices
    for r in reqs:
        assert [x.tag[2] for x in r.results] == [r.first_trial_index + k for k in range(r.n_trials)]
    assert reqs[3].synapses == ("syn", reqs[3].handle) and reqs[3].stats is None
    # chain-test payload: neuron 0 fired once per trial at 100 + trial index % 100 (0 and 1 for request 0)
    st = reqs[0].stats
    assert st["num_test_trials"] == 2 and st["probability_soma_spike"].tolist() == [1.0] + [0.0] * 5
    assert st["mean_first_soma_spike_time"][0] == 100.5 and st["mean_first_soma_spike_time"][1] == -1.0
    assert st["std_first_soma_spike_time"][0] == np.std([100.0, 101.0], ddof=1)
    assert pool.used == 0 and not pool.resident
    assert stats["requests"] == 4 and stats["rounds"] == len(pool.rounds) + 2


def test_latency_statistics():
    pool = FakePool(capacity=100)
    reqs = [_req(k, k, 5, 3) for k in range(5)]
    stats = Server(pool).run(reqs, realtime=False)
    lat = [r.finished - r.arrival for r in reqs]
    assert stats["latency_p50"] == np.percentile(lat, 50)
    assert stats["latency_max"] == max(lat)


class FakeAsyncPool(FakePool):
    """submit / poll on top of FakePool: a poll runs the next round (the oldest
    queued trial of every handle that has one) when the polled ticket is not
    done yet; results carry the round that ran them."""

    def __init__(self, capacity):
        super().__init__(capacity)
        self.queue = []
        self.done = {}
        self.ticket = 0
        self.round = 0

    def submit(self, h, **kw):
        assert h in self.resident
        self.ticket += 1
        self.queue.append((self.ticket, h, kw))
        return self.ticket

    def _round(self):
        picked, seen = [], set()
        for item in self.queue:
            if item[1] not in seen:
                seen.add(item[1])
                picked.append(item)
        for item in picked:
            self.queue.remove(item)
        res = self.run({h: kw for _, h, kw in picked})
        for t, h, _ in picked:
            self.done[t] = (self.round, res[h])
        self.round += 1

    def poll(self, t):
        if t not in self.done and self.queue:
            self._round()
        return self.done.pop(t, None)

    def release(self, h):
        assert not any(item[1] == h for item in self.queue), "released with trials queued"
        super().release(h)


def test_async_server_submits_in_order_and_releases_when_done():
    from unsupervised_growth_amd.serve import AsyncServer
    pool = FakeAsyncPool(capacity=10)
    reqs = [_req(0, 0.0, 6, 3), _req(1, 0.0, 6, 2), _req(2, 0.0, 3, 2), _req(3, 0.0, 2, 1, mode=0)]
    stats = AsyncServer(pool).run(reqs, poll_s=0.0001, timeout_s=30)
    assert stats["requests"] == 4 and pool.used == 0 and not pool.resident
    for r in reqs:
        # every trial of a request, in order, with consecutive trial indices
        assert [x.tag[2] for x in r.results] == [r.first_trial_index + k for k in range(r.n_trials)]
        assert r.finished is not None and r.finished >= r.admitted
    # request 1 (6 neurons) does not fit beside request 0: it and request 2 behind it wait (first come, first served)
    assert reqs[1].admitted >= reqs[0].finished
    assert reqs[2].admitted >= reqs[1].admitted
    # a network never ran two trials in one round
    assert all(len(r) == len(set(r)) for r in pool.rounds)
    assert reqs[3].synapses == ("syn", reqs[3].handle)
    assert reqs[0].stats["num_test_trials"] == 3
