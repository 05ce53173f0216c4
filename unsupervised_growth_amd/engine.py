"""Python handle on the HIP engine (libhvc_engine.so, C ABI in include/hvc_engine.h).

``Engine`` owns one ``hvc_ctx``: one network resident in device memory, one HIP
stream.  Every call goes through the C ABI; there is no CPU fallback.
"""
import ctypes
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from . import _abi
from .network import NETWORK_UPDATE_FREQUENCY, TIMESTEP, TRIAL_DURATION, Network, NoiseParams, SynapticParams

RA_MODEL_DEBRABAND, RA_MODEL_RK4 = 0, 1
I_MODEL_DP8, I_MODEL_RK4 = 0, 1
RA_STATE_NAMES = ("Vs", "Vd", "n", "l", "h", "r", "c", "Ca", "Ge_s", "Gi_s", "Ge_d", "Gi_d")
I_STATE_NAMES = ("V", "n", "m", "h", "w", "Ge", "Gi")


class EngineError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (_abi.STATUS_NAMES.get(status, status), message))
        self.status = status


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class SpikeTrains(Sequence):
    """Spike times neuron by neuron, as the engine returns them (a count per
    neuron and the times concatenated); item i is neuron i's array (a view).
    Offsets are computed on first access, so a result that is only counted or
    passed on costs no per-neuron work."""

    def __init__(self, counts, times):
        self.counts = counts
        self.times = times
        self._off = None

    def _offsets(self):
        if self._off is None:
            self._off = np.zeros(len(self.counts) + 1, np.int64)
            np.cumsum(self.counts, out=self._off[1:])
        return self._off

    def __len__(self):
        return len(self.counts)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        off = self._offsets()
        return self.times[off[i]:off[i + 1]]

    def __iter__(self):
        off = self._offsets()
        for i in range(len(self.counts)):
            yield self.times[off[i]:off[i + 1]]

    @property
    def total(self):
        return len(self.times)


@dataclass
class TrialResult:
    soma: Sequence          # per HVC-RA neuron: array of somatic spike times (ms) (SpikeTrains)
    dend: Sequence          # per HVC-RA neuron: dendritic spike times
    interneuron: Sequence   # per HVC-I neuron: spike times
    bad_values: int
    warnings: tuple
    ra_state: np.ndarray = None
    i_state: np.ndarray = None


@dataclass
class SynapseState:
    w: np.ndarray
    act_cnt: np.ndarray
    act_ids: np.ndarray
    sup_cnt: np.ndarray
    sup_ids: np.ndarray
    remodeled: np.ndarray

    def active_lists(self):
        return np.split(self.act_ids, np.cumsum(self.act_cnt)[:-1])

    def super_lists(self):
        return np.split(self.sup_ids, np.cumsum(self.sup_cnt)[:-1])


def make_params(*, ra_model=RA_MODEL_DEBRABAND, i_model=I_MODEL_DP8, precision=0, dt=TIMESTEP,
                trial_ms=TRIAL_DURATION, sync_ms=NETWORK_UPDATE_FREQUENCY, noise=None, ra_poisson=0,
                i_poisson=1, seed=(1, 0), syn=None, spike_capacity=0):
    noise = noise or NoiseParams()
    syn = syn or SynapticParams()
    p = _abi.Params()
    p.ra_model, p.i_model, p.precision = ra_model, i_model, precision
    p.dt, p.trial_ms, p.sync_ms = dt, trial_ms, sync_ms
    p.mu_s, p.sigma_s = noise.white_noise_mean_soma, noise.white_noise_std_soma
    p.mu_d, p.sigma_d = noise.white_noise_mean_dend, noise.white_noise_std_dend
    p.ra_poisson, p.i_poisson = int(ra_poisson), int(i_poisson)
    p.seed_lo, p.seed_hi = int(seed[0]) & 0xffffffff, int(seed[1]) & 0xffffffff
    for k in SynapticParams.__dataclass_fields__:
        setattr(p.syn, k, getattr(syn, k))
    p.spike_capacity = spike_capacity
    return p


class Engine:
    """One network on one device."""

    def __init__(self, params=None, device=0, **kw):
        self.lib = _abi.load()
        self.params = params if params is not None else make_params(**kw)
        self.ctx = ctypes.c_void_p()
        st = self.lib.hvc_create(ctypes.byref(self.params), device, ctypes.byref(self.ctx))
        if st != _abi.HVC_OK:
            msg = self.lib.hvc_last_error(self.ctx).decode() if self.ctx else ""
            if self.ctx:
                self.lib.hvc_destroy(self.ctx)
            self.ctx = None
            raise EngineError(st, msg)
        self.net = None
        self.nets = []

    def _check(self, st):
        if st != _abi.HVC_OK:
            raise EngineError(st, self.lib.hvc_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.hvc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------
    @staticmethod
    def _desc(net: Network):
        keep = {k: (_i32(v) if v.dtype.kind in "iu" else _f64(v)) for k, v in net.__dict__.items()
                if isinstance(v, np.ndarray)}
        d = _abi.NetworkDesc()
        d.n_ra, d.n_i = net.n_ra, net.n_i
        ip, dp = _abi.c_int_p, _abi.c_double_p
        for name in ("ra2i_ptr", "ra2i_tgt", "i2ra_ptr", "i2ra_tgt", "rr_ptr", "rr_tgt", "act_cnt",
                     "act_ids", "sup_cnt", "sup_ids", "remodeled", "train_ids"):
            setattr(d, name, _p(keep[name], ip))
        for name in ("ra2i_w", "ra2i_d", "i2ra_w", "i2ra_d", "rr_d", "rr_w", "erest", "gca", "spread"):
            setattr(d, name, _p(keep[name], dp))
        d.n_train = len(net.train_ids)
        return d, keep

    def upload(self, net: Network):
        d, keep = self._desc(net)
        self._check(self.lib.hvc_upload_network(self.ctx, ctypes.byref(d)))
        self.net = net
        self.nets = [net]
        return self

    def upload_batch(self, nets, keys):
        """Several independent networks in this context (the batched form).
        keys: one (seed_lo, seed_hi) pair per network."""
        descs, keeps = [], []
        for n in nets:
            d, k = self._desc(n)
            descs.append(d)
            keeps.append(k)
        arr = (_abi.NetworkDesc * len(nets))(*descs)
        kk = np.ascontiguousarray(np.asarray(keys, np.uint32).reshape(len(nets), 2))
        self._check(self.lib.hvc_upload_batch(self.ctx, len(nets), arr,
                                              kk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        self.nets = list(nets)
        self.net = self.nets[0] if len(self.nets) == 1 else None
        return self

    def run_trial(self, mode=0, training=True, trial_index=0, first_trial=False, mean_injection_time=100.0,
                  want_state=False):
        """mode 0: growth trial (STDP); mode 1: chain-test trial (no STDP)."""
        if self.net is None:
            raise EngineError(_abi.HVC_E_STATE, "upload a network first")
        t = _abi.Trial(mode, int(bool(training)), int(trial_index) & 0xffffffff, int(bool(first_trial)),
                       float(mean_injection_time))
        n_ra, n_i = self.net.n_ra, self.net.n_i
        counts = [np.zeros(n_ra, np.int32), np.zeros(n_ra, np.int32), np.zeros(max(n_i, 1), np.int32)]
        ras = np.zeros((n_ra, 12)) if want_state else None
        ins = np.zeros((max(n_i, 1), 7)) if want_state else None
        o = _abi.TrialOut()
        ip, dp = _abi.c_int_p, _abi.c_double_p
        o.soma_counts, o.dend_counts, o.i_counts = (_p(c, ip) for c in counts)
        o.ra_state, o.i_state = _p(ras, dp), _p(ins, dp)
        st = self.lib.hvc_run_trial(self.ctx, ctypes.byref(t), ctypes.byref(o))
        if st not in (_abi.HVC_OK, _abi.HVC_E_NUMERIC):
            self._check(st)
        lists = []
        for kind, total in enumerate((o.soma_total, o.dend_total, o.i_total)):
            c = np.zeros(n_i if kind == 2 else n_ra, np.int32)
            times = np.zeros(max(total, 1))
            tot = ctypes.c_int()
            self._check(self.lib.hvc_trial_spikes(self.ctx, kind, _p(c, ip) if len(c) else None,
                                                  _p(times, dp), len(times), ctypes.byref(tot)))
            lists.append(SpikeTrains(c, times[:tot.value]) if len(c) else [])
        return TrialResult(lists[0], lists[1], lists[2], o.bad_values, tuple(o.warnings), ras,
                           ins[:n_i] if ins is not None else None)

    def _spikes(self, net_index, kind, n):
        c = np.zeros(n, np.int32)
        tot = ctypes.c_int()
        ip, dp = _abi.c_int_p, _abi.c_double_p
        self._check(self.lib.hvc_batch_spikes(self.ctx, net_index, kind, _p(c, ip) if n else None, None, 0,
                                              ctypes.byref(tot)))
        times = np.zeros(max(tot.value, 1))
        self._check(self.lib.hvc_batch_spikes(self.ctx, net_index, kind, _p(c, ip) if n else None,
                                              _p(times, dp), len(times), ctypes.byref(tot)))
        return SpikeTrains(c, times[:tot.value]) if n else []

    def run_batch(self, trials):
        """One trial of every network of a batch; trials: list of dicts with the
        keyword arguments of run_trial (mode must agree).  Returns TrialResults."""
        if not getattr(self, "nets", None):
            raise EngineError(_abi.HVC_E_STATE, "upload networks first")
        B = len(self.nets)
        if len(trials) != B:
            raise ValueError("one trial per network")
        ts = (_abi.Trial * B)()
        for b, kw in enumerate(trials):
            ts[b] = _abi.Trial(int(kw.get("mode", 0)), int(bool(kw.get("training", True))),
                               int(kw.get("trial_index", 0)) & 0xffffffff, int(bool(kw.get("first_trial", False))),
                               float(kw.get("mean_injection_time", 100.0)))
        outs = (_abi.TrialOut * B)()   # no buffers: only the spike totals, bad_values and warnings come back
        st = self.lib.hvc_run_batch(self.ctx, ts, outs)
        if st not in (_abi.HVC_OK, _abi.HVC_E_NUMERIC):
            self._check(st)
        out = []
        for b, n in enumerate(self.nets):
            out.append(TrialResult(self._spikes(b, 0, n.n_ra), self._spikes(b, 1, n.n_ra),
                                   self._spikes(b, 2, n.n_i), outs[b].bad_values, tuple(outs[b].warnings)))
        return out

    def download_synapses(self, net_index=None) -> SynapseState:
        b = 0 if net_index is None else net_index
        net = self.nets[b]
        E, n = net.n_candidates, net.n_ra
        w = np.zeros(max(E, 1))
        ac, sc, rem = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        ai, si = np.zeros(max(E, 1), np.int32), np.zeros(max(n * max(1, self.params.syn.Nss), 1), np.int32)
        s = _abi.Synapses(_p(w, _abi.c_double_p), _p(ac, _abi.c_int_p), _p(ai, _abi.c_int_p),
                          _p(sc, _abi.c_int_p), _p(si, _abi.c_int_p), _p(rem, _abi.c_int_p))
        if net_index is None:
            self._check(self.lib.hvc_download_synapses(self.ctx, ctypes.byref(s)))
        else:
            self._check(self.lib.hvc_batch_download_synapses(self.ctx, b, ctypes.byref(s)))
        return SynapseState(w[:E], ac, ai[:int(ac.sum())].copy(), sc, si[:int(sc.sum())].copy(), rem)

    def upload_synapses(self, syn: SynapseState, net_index=None):
        w, ac, ai = _f64(syn.w), _i32(syn.act_cnt), _i32(syn.act_ids)
        sc, si, rem = _i32(syn.sup_cnt), _i32(syn.sup_ids), _i32(syn.remodeled)
        s = _abi.Synapses(_p(w, _abi.c_double_p), _p(ac, _abi.c_int_p), _p(ai, _abi.c_int_p),
                          _p(sc, _abi.c_int_p), _p(si, _abi.c_int_p), _p(rem, _abi.c_int_p))
        if net_index is None:
            self._check(self.lib.hvc_upload_synapses(self.ctx, ctypes.byref(s)))
        else:
            self._check(self.lib.hvc_batch_upload_synapses(self.ctx, net_index, ctypes.byref(s)))

    def patch_neurons(self, ids, erest, gca):
        ids, erest, gca = _i32(ids), _f64(erest), _f64(gca)
        self._check(self.lib.hvc_patch_neurons(self.ctx, _p(ids, _abi.c_int_p), len(ids),
                                               _p(erest, _abi.c_double_p), _p(gca, _abi.c_double_p)))

    # ------------------------------------------------------------------
    def trace_ra(self, erest, gca, kick, exc_w, inh_g, trial=0):
        n_steps, n = kick.shape
        out = np.zeros((n_steps + 1, n, 12))
        fired = np.zeros((n_steps, n), np.uint8)
        a = [_f64(x) for x in (erest, gca, kick, exc_w, inh_g)]
        dp = _abi.c_double_p
        self._check(self.lib.hvc_trace_ra(self.ctx, n, n_steps, _p(a[0], dp), _p(a[1], dp), trial,
                                          _p(a[2], dp), _p(a[3], dp), _p(a[4], dp), _p(out, dp),
                                          _p(fired, _abi.c_u8_p)))
        return out, fired

    def trace_i(self, exc_g, inh_g, trial=0, first_trial=False):
        n_steps, n = exc_g.shape
        out = np.zeros((n_steps + 1, n, 7))
        fired = np.zeros((n_steps, n), np.uint8)
        a = [_f64(x) for x in (exc_g, inh_g)]
        dp = _abi.c_double_p
        self._check(self.lib.hvc_trace_i(self.ctx, n, n_steps, int(bool(first_trial)), trial, _p(a[0], dp),
                                         _p(a[1], dp), _p(out, dp), _p(fired, _abi.c_u8_p)))
        return out, fired

    def eval_math(self, fn, a, b=None):
        """Device exp (fn 0), log (fn 1) or constant-divisor division a / b (fn 2)."""
        a = _f64(a)
        b = _f64(b) if b is not None else None
        out = np.zeros_like(a)
        dp = _abi.c_double_p
        self._check(self.lib.hvc_eval_math(self.ctx, fn, len(a), _p(a, dp), _p(b, dp), _p(out, dp)))
        return out

    def tuning(self):
        t = _abi.Tuning()
        self._check(self.lib.hvc_get_tuning(self.ctx, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def set_tuning(self, **kw):
        t = _abi.Tuning(**{**self.tuning(), **kw})
        self._check(self.lib.hvc_set_tuning(self.ctx, ctypes.byref(t)))

    def autotune(self, windows=200):
        """Time the launch-shape candidates on the resident networks and keep the
        fastest; returns (tuning, per-window ms of integrate, deliver, sync)."""
        t = _abi.Tuning()
        ms = np.zeros(3)
        self._check(self.lib.hvc_autotune(self.ctx, windows, ctypes.byref(t), _p(ms, _abi.c_double_p)))
        return {f: getattr(t, f) for f, _ in t._fields_}, ms

    def set_timing(self, on=True):
        self._check(self.lib.hvc_set_timing(self.ctx, int(bool(on))))

    def set_branch_counting(self, on=True):
        """Count the synapse-state machine's branches from now on (cleared)."""
        self._check(self.lib.hvc_set_branch_counting(self.ctx, int(bool(on))))

    def branch_counts(self):
        """{branch name: passes} since set_branch_counting(True)."""
        c = np.zeros(len(_abi.BRANCH_COUNTERS), np.int32)
        self._check(self.lib.hvc_branch_counts(self.ctx, _p(c, _abi.c_int_p)))
        return dict(zip(_abi.BRANCH_COUNTERS, (int(x) for x in c)))

    def kernel_timing(self, name):
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.hvc_kernel_timing(self.ctx, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def pool_caps(max_nets, max_ra, max_i, max_rr, max_ra2i, max_i2ra, max_in_degree, spikes_per_neuron=0):
    return _abi.PoolCaps(max_nets, max_ra, max_i, max_rr, max_ra2i, max_i2ra, max_in_degree, spikes_per_neuron)


def network_sizes(net: Network):
    """(n_ra, n_i, candidates, RA->I, I->RA) of a network."""
    return net.n_ra, net.n_i, int(net.rr_ptr[-1]), int(net.ra2i_ptr[-1]), int(net.i2ra_ptr[-1])


class Pool(Engine):
    """A serving pool (hvc_pool_*): networks are admitted and released between
    trials and every run advances all resident networks by one trial."""

    def __init__(self, caps, params=None, device=0, **kw):
        self.lib = _abi.load()
        self.params = params if params is not None else make_params(**kw)
        self.caps = caps
        self.ctx = ctypes.c_void_p()
        st = self.lib.hvc_pool_create(ctypes.byref(self.params), device, ctypes.byref(caps), ctypes.byref(self.ctx))
        if st != _abi.HVC_OK:
            msg = self.lib.hvc_last_error(self.ctx).decode() if self.ctx else ""
            if self.ctx:
                self.lib.hvc_destroy(self.ctx)
            self.ctx = None
            raise EngineError(st, msg)
        self.net = None
        self.nets = []
        self.resident = {}      # handle -> Network
        self._tickets = {}      # submitted ticket -> handle

    def footprint(self, net: Network):
        """Device bytes the network occupies in this pool."""
        b = ctypes.c_size_t()
        self._check(self.lib.hvc_pool_footprint(ctypes.byref(self.params), ctypes.byref(self.caps),
                                                *network_sizes(net), ctypes.byref(b), None))
        return b.value

    def pool_bytes(self):
        b = ctypes.c_size_t()
        self._check(self.lib.hvc_pool_footprint(ctypes.byref(self.params), ctypes.byref(self.caps), 0, 0, 0, 0, 0,
                                                None, ctypes.byref(b)))
        return b.value

    def admit(self, net: Network, key):
        """Handle of the admitted network, or None when it does not fit now."""
        d, keep = self._desc(net)
        h = ctypes.c_int()
        st = self.lib.hvc_pool_admit(self.ctx, ctypes.byref(d), int(key[0]), int(key[1]), ctypes.byref(h))
        if st == _abi.HVC_E_CAPACITY:
            return None
        self._check(st)
        self.resident[h.value] = net
        return h.value

    def release(self, handle):
        self._check(self.lib.hvc_pool_release(self.ctx, handle))
        self.resident.pop(handle, None)

    def _pool_spikes(self, handle, kind, n):
        c = np.zeros(n, np.int32)
        tot = ctypes.c_int()
        ip, dp = _abi.c_int_p, _abi.c_double_p
        self._check(self.lib.hvc_pool_spikes(self.ctx, handle, kind, _p(c, ip) if n else None, None, 0,
                                             ctypes.byref(tot)))
        times = np.zeros(max(tot.value, 1))
        self._check(self.lib.hvc_pool_spikes(self.ctx, handle, kind, _p(c, ip) if n else None, _p(times, dp),
                                             len(times), ctypes.byref(tot)))
        return SpikeTrains(c, times[:tot.value]) if n else []

    @staticmethod
    def _trial(kw):
        return _abi.Trial(int(kw.get("mode", 1)), int(bool(kw.get("training", True))),
                          int(kw.get("trial_index", 0)) & 0xffffffff, int(bool(kw.get("first_trial", False))),
                          float(kw.get("mean_injection_time", 100.0)))

    def run(self, trials):
        """trials: {handle: run_trial keyword dict} for every resident network;
        returns {handle: TrialResult}."""
        hs = list(trials)
        n = len(hs)
        handles = np.asarray(hs, np.int32)
        ts = (_abi.Trial * max(n, 1))()
        outs = (_abi.TrialOut * max(n, 1))()
        for k, h in enumerate(hs):
            ts[k] = self._trial(trials[h])
        st = self.lib.hvc_pool_run(self.ctx, n, _p(handles, _abi.c_int_p), ts, outs)
        if st not in (_abi.HVC_OK, _abi.HVC_E_NUMERIC):
            self._check(st)
        res = {}
        for k, h in enumerate(hs):
            net = self.resident[h]
            res[h] = TrialResult(self._pool_spikes(h, 0, net.n_ra), self._pool_spikes(h, 1, net.n_ra),
                                 self._pool_spikes(h, 2, net.n_i), outs[k].bad_values, tuple(outs[k].warnings))
        return res

    def submit(self, handle, **kw):
        """Queue one trial of resident network `handle` (hvc_submit); returns the
        ticket at once.  The pool's serving thread runs it in a later round."""
        t = ctypes.c_uint64()
        tr = self._trial(kw)
        self._check(self.lib.hvc_submit(self.ctx, int(handle), ctypes.byref(tr), ctypes.byref(t)))
        self._tickets[t.value] = handle
        return t.value

    def poll(self, ticket):
        """hvc_poll: None while the ticket's trial is queued or running, else
        (round, TrialResult); the ticket is then released.  Raises EngineError
        when the trial failed (HVC_E_NUMERIC gives the result with bad_values)."""
        net = self.resident[self._tickets[ticket]]
        sizes = (net.n_ra, net.n_ra, net.n_i)
        counts = [np.zeros(max(n, 1), np.int32) for n in sizes]
        times = [np.zeros(1) for _ in sizes]
        rnd = ctypes.c_longlong()
        for attempt in range(2):
            o = _abi.TrialOut()
            o.soma_counts, o.dend_counts, o.i_counts = (_p(c, _abi.c_int_p) for c in counts)
            o.soma_times, o.dend_times, o.i_times = (_p(t, _abi.c_double_p) for t in times)
            o.soma_cap, o.dend_cap, o.i_cap = (len(t) if attempt else 0 for t in times)
            st = self.lib.hvc_poll(self.ctx, ticket, ctypes.byref(o), ctypes.byref(rnd))
            if st == _abi.HVC_PENDING:
                return None
            if st != _abi.HVC_E_ARG or attempt:
                break
            # spike-time buffers sized from the totals, then the same ticket again
            times = [np.zeros(max(tot, 1)) for tot in (o.soma_total, o.dend_total, o.i_total)]
        if st not in (_abi.HVC_OK, _abi.HVC_E_NUMERIC):
            self._tickets.pop(ticket, None)
            raise EngineError(st, "hvc_poll")
        self._tickets.pop(ticket, None)
        tots = (o.soma_total, o.dend_total, o.i_total)
        trains = [SpikeTrains(counts[k][:sizes[k]], times[k][:tots[k]]) if sizes[k] else [] for k in range(3)]
        return rnd.value, TrialResult(trains[0], trains[1], trains[2], o.bad_values, tuple(o.warnings))

    def download_synapses(self, handle) -> SynapseState:
        net = self.resident[handle]
        E, n = net.n_candidates, net.n_ra
        w = np.zeros(max(E, 1))
        ac, sc, rem = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        ai, si = np.zeros(max(E, 1), np.int32), np.zeros(max(n * max(1, self.params.syn.Nss), 1), np.int32)
        s = _abi.Synapses(_p(w, _abi.c_double_p), _p(ac, _abi.c_int_p), _p(ai, _abi.c_int_p),
                          _p(sc, _abi.c_int_p), _p(si, _abi.c_int_p), _p(rem, _abi.c_int_p))
        self._check(self.lib.hvc_pool_download_synapses(self.ctx, handle, ctypes.byref(s)))
        return SynapseState(w[:E], ac, ai[:int(ac.sum())].copy(), sc, si[:int(sc.sum())].copy(), rem)
