"""Serving loop: trial and chain-test requests on networks of mixed size,
continuously batched on one device.

Requests wait in arrival order; at every round boundary the loop admits the
waiting requests whose networks fit into the pool (its capacities fix the HBM
footprint), then runs one trial of every resident network together
(hvc_pool_run: one launch per kernel and coupling window for all of them) and
releases the networks whose requests are complete.  A network stays on the
device from admission to release, so a growth request's synapse state evolves
in place; a request that arrives while others run joins at the next round.
Admission is first come, first served: a request that does not fit keeps the
ones behind it waiting, so large networks are not starved.  Server drives the
rounds itself (hvc_pool_run); AsyncServer hands the trials to the pool's own
serving thread (hvc_submit / hvc_poll), so that arrivals join mid-round work
at the next round without waiting for the caller.

Each request's result equals what the same network and trials give alone (the
pool keys every network's noise by its own key and ids and sums over its own
neurons only; see tests/test_gpu_serve.py).  A chain-test request's payload
also carries the chain-test statistics of test_chain / calculate_and_write_jitter
(c++/src/NetworkGrowthSimulator.cpp:1313-1482, 1641-1782): first-spike
probability, mean and jitter per neuron, accumulated round by round.
"""
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .chain import ChainStats
from .engine import Pool


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start (rounds with a virtual clock)
    net: object                    # Network
    key: tuple                     # noise-stream key of the request
    onsets: np.ndarray             # mean_injection_time of each trial (ms)
    mode: int = 1                  # 1: chain test, 0: growth trials
    first_trial_index: int = 0
    handle: int = None
    admitted: float = None
    finished: float = None
    results: list = field(default_factory=list)
    synapses: object = None        # growth requests: synapse state at release
    stats: dict = None             # chain-test requests: jitter statistics at release
    _chain: object = None

    @property
    def n_trials(self):
        return len(self.onsets)

    def trial_kw(self, k):
        return dict(mode=self.mode, training=True, trial_index=self.first_trial_index + k,
                    first_trial=(self.mode == 0 and k == 0), mean_injection_time=float(self.onsets[k]))


class Server:
    """Continuous batching of requests on a Pool."""

    def __init__(self, pool: Pool):
        self.pool = pool

    def run(self, requests, realtime=True, keep_results=True, log=None):
        """Serve the requests; with realtime the arrival times are seconds on the
        wall clock, otherwise round numbers.  Returns statistics."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        active = {}
        t0 = time.perf_counter()
        rounds = 0
        updates = 0
        busy = 0.0

        def now():
            return (time.perf_counter() - t0) if realtime else float(rounds)

        while pending or waiting or active:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            while waiting:
                req = waiting[0]
                h = self.pool.admit(req.net, req.key)
                if h is None:
                    if not active:
                        raise RuntimeError("request %d does not fit into an empty pool" % req.rid)
                    break
                waiting.popleft()
                req.handle, req.admitted = h, now()
                req.results = []
                req._chain = ChainStats(req.net.n_ra) if req.mode == 1 else None
                active[h] = req
            if not active:
                if realtime and pending:
                    time.sleep(max(0.0, pending[0].arrival - now()))
                elif not realtime:
                    rounds += 1
                continue
            tr0 = time.perf_counter()
            res = self.pool.run({h: r.trial_kw(len(r.results)) for h, r in active.items()})
            busy += time.perf_counter() - tr0
            rounds += 1
            steps = int(round(self.pool.params.trial_ms / self.pool.params.dt)) - 1
            for h, r in list(active.items()):
                r.results.append(res[h] if keep_results else None)
                if r._chain is not None:
                    r._chain.add_result(res[h])
                updates += (r.net.n_ra + r.net.n_i) * steps
                if len(r.results) == r.n_trials:
                    if r.mode == 0:
                        r.synapses = self.pool.download_synapses(h)
                    else:
                        r.stats = r._chain.result()
                        r._chain.close()
                        r._chain = None
                    self.pool.release(h)
                    r.finished = now()
                    del active[h]
            if log:
                log("round %d: %d resident, %d waiting, %d pending" % (rounds, len(active), len(waiting),
                                                                       len(pending)))
        wall = time.perf_counter() - t0
        lat = np.array([r.finished - r.arrival for r in requests])
        return {
            "requests": len(requests),
            "rounds": rounds,
            "wall_s": wall,
            "busy_s": busy,
            "neuron_updates": updates,
            "neuron_updates_per_s": updates / wall if wall > 0 else 0.0,
            "latency_p50": float(np.percentile(lat, 50)),
            "latency_p99": float(np.percentile(lat, 99)),
            "latency_max": float(lat.max()),
        }


class AsyncServer:
    """Requests served through hvc_submit / hvc_poll.

    An arriving request's network is admitted as soon as it fits (first come,
    first served) and all of its trials are submitted at once; the pool's
    serving thread runs rounds of whatever is queued (one trial per network
    and round, a network's trials in order), so a request that arrives while a
    round runs has its first trial in the next round, and networks of finished
    requests leave at the next round boundary.  The caller's thread only admits,
    submits, polls and releases."""

    def __init__(self, pool: Pool):
        self.pool = pool

    def run(self, requests, keep_results=True, poll_s=0.0005, log=None, timeout_s=None):
        """Serve the requests (arrival times in seconds on the wall clock);
        returns statistics.  timeout_s bounds the whole run (TimeoutError)."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        active = {}            # handle -> (request, tickets)
        t0 = time.perf_counter()
        updates = 0
        last_round = -1
        steps = int(round(self.pool.params.trial_ms / self.pool.params.dt)) - 1

        def now():
            return time.perf_counter() - t0

        while pending or waiting or active:
            t = now()
            if timeout_s is not None and t > timeout_s:
                raise TimeoutError("%d requests unfinished after %g s" % (len(pending) + len(waiting) + len(active),
                                                                         timeout_s))
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            while waiting:
                req = waiting[0]
                h = self.pool.admit(req.net, req.key)
                if h is None:
                    if not active:
                        raise RuntimeError("request %d does not fit into an empty pool" % req.rid)
                    break
                waiting.popleft()
                req.handle, req.admitted = h, now()
                req.results = []
                req._chain = ChainStats(req.net.n_ra) if req.mode == 1 else None
                tickets = deque(self.pool.submit(h, **req.trial_kw(k)) for k in range(req.n_trials))
                active[h] = (req, tickets)
            progressed = False
            for h, (r, tickets) in list(active.items()):
                while tickets:
                    got = self.pool.poll(tickets[0])
                    if got is None:
                        break
                    tickets.popleft()
                    progressed = True
                    rnd, res = got
                    last_round = max(last_round, rnd)
                    r.results.append(res if keep_results else None)
                    if r._chain is not None:
                        r._chain.add_result(res)
                    updates += (r.net.n_ra + r.net.n_i) * steps
                if not tickets:
                    if r.mode == 0:
                        r.synapses = self.pool.download_synapses(h)
                    else:
                        r.stats = r._chain.result()
                        r._chain.close()
                        r._chain = None
                    self.pool.release(h)
                    r.finished = now()
                    del active[h]
                    if log:
                        log("request %d done: %d resident, %d waiting" % (r.rid, len(active), len(waiting)))
            if not progressed:
                nxt = pending[0].arrival - now() if (pending and not active and not waiting) else poll_s
                time.sleep(max(poll_s, min(nxt, 0.05)))
        wall = now()
        lat = np.array([r.finished - r.arrival for r in requests])
        return {
            "requests": len(requests),
            "rounds": last_round + 1,
            "wall_s": wall,
            "neuron_updates": updates,
            "neuron_updates_per_s": updates / wall if wall > 0 else 0.0,
            "latency_p50": float(np.percentile(lat, 50)),
            "latency_p99": float(np.percentile(lat, 99)),
            "latency_max": float(lat.max()),
        }
