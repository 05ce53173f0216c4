"""The host growth driver (C ABI include/hvc_driver.h) from Python.

``GrowthDriver`` runs the reference's growth loop chain_growth
(c++/src/NetworkGrowthSimulator.cpp:5731-5900, "NGS") around the engine: one
growth trial, check_bad_values, the periodic state files, then the bookkeeping
(activity history, driven / silent neurons, maturation, replacement).  All of
it runs in C++; this module only marshals arrays.

``activity_update`` and ``maturation_update`` expose the two bookkeeping steps
on host arrays (no device needed).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from .engine import Engine, EngineError, _f64, _i32, _p, make_params
from .network import MaturationParams, Network
from .topology import TopologyParams


ACTIVITY_WINDOW = 1000   # trials check_neuron_activity reads (NetworkGrowthSimulator.cpp:9586)


def _mat_c(m: MaturationParams, activity_window=ACTIVITY_WINDOW):
    c = _abi.MaturationParamsC()
    for k in ("E_REST_MATURE", "E_REST_IMMATURE", "GCA_MATURE", "GCA_IMMATURE", "DEATH_RATE_THRESHOLD"):
        setattr(c, k, float(getattr(m, k)))
    c.RATE_WINDOW_LONG = int(m.RATE_WINDOW_LONG)
    c.ACTIVITY_WINDOW = int(activity_window)
    return c


@dataclass
class DriverReport:
    trial: int
    mean_injection_time: float
    soma_spikes: int
    dend_spikes: int
    i_spikes: int
    bad_values: int
    became_driven: int
    matured: int
    replaced: int
    saved: int


class GrowthDriver:
    """One network growing on one device.

    net: the Network (its erest / gca are the current neuron properties);
    ra_xyz, i_xyz: positions (N x 3); model_interneuron_distance: the distance
    unit of the positions; topology: resampling parameters; resume: optional
    dict with mature, maturation_scale, trials_after_replacement, history
    (N x RATE_WINDOW_LONG, newest first) and trial_number; activity_window:
    the trials the activity check reads and the trials after a replacement
    before the neuron may be replaced again (the reference's 1000; at most
    RATE_WINDOW_LONG).
    """

    def __init__(self, net: Network, ra_xyz, i_xyz, model_interneuron_distance, topology: TopologyParams,
                 maturation: MaturationParams = None, resume=None, seed=15000, params=None, device=0,
                 activity_window=ACTIVITY_WINDOW, **kw):
        self.lib = _abi.load()
        self.params = params if params is not None else make_params(**kw)
        self.maturation = maturation or MaturationParams()
        self.net = net
        self.n_ra, self.n_i = net.n_ra, net.n_i
        desc, self._keep = Engine._desc(net)
        self._xyz = (_f64(np.asarray(ra_xyz).reshape(-1)), _f64(np.asarray(i_xyz).reshape(-1)))
        geo = _abi.Geometry(_p(self._xyz[0], _abi.c_double_p), _p(self._xyz[1], _abi.c_double_p),
                            float(model_interneuron_distance))
        top = topology.c()
        self.activity_window = int(activity_window)
        mat = _mat_c(self.maturation, self.activity_window)
        res = None
        if resume is not None:
            self._res = {k: _i32(np.asarray(resume[k]).reshape(-1)) if resume.get(k) is not None else None
                         for k in ("mature", "maturation_scale", "trials_after_replacement", "history")}
            res = _abi.DriverResume(*(_p(self._res[k], _abi.c_int_p) if self._res[k] is not None else None
                                      for k in ("mature", "maturation_scale", "trials_after_replacement",
                                                "history")),
                                    int(resume.get("trial_number", 0)))
        self.d = ctypes.c_void_p()
        st = self.lib.hvc_driver_create(ctypes.byref(self.params), device, ctypes.byref(desc), ctypes.byref(geo),
                                        ctypes.byref(top), ctypes.byref(mat), ctypes.byref(res) if res else None,
                                        int(seed) & 0xffffffff, ctypes.byref(self.d))
        if st != _abi.HVC_OK:
            msg = self.lib.hvc_driver_last_error(self.d).decode() if self.d else ""
            if self.d:
                self.lib.hvc_driver_destroy(self.d)
            self.d = None
            raise EngineError(st, msg)

    def _check(self, st):
        if st != _abi.HVC_OK:
            raise EngineError(st, self.lib.hvc_driver_last_error(self.d).decode())

    def close(self):
        if self.d:
            self.lib.hvc_driver_destroy(self.d)
            self.d = None

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
    def set_output(self, directory, save_freq_short=0, save_freq_long=0):
        self._check(self.lib.hvc_driver_set_output(self.d, directory.encode() if directory else None,
                                                   save_freq_short, save_freq_long))

    def trial(self, training=True) -> DriverReport:
        """One pass of the growth loop.  Raises EngineError(HVC_E_NUMERIC) when
        the trial fails check_bad_values (the loop then refuses to continue)."""
        r = _abi.DriverReport()
        st = self.lib.hvc_driver_trial(self.d, int(bool(training)), ctypes.byref(r))
        self._check(st)
        return DriverReport(**{k: getattr(r, k) for k, _ in _abi.DriverReport._fields_})

    def run(self, n_trials, training=True, log=None):
        reports = []
        for _ in range(n_trials):
            reports.append(self.trial(training))
            if log:
                log(reports[-1])
        return reports

    def write_state(self, directory, file_ending="", what="full"):
        code = {"graph": 0, "full": 1, "initial": 2}[what]
        self._check(self.lib.hvc_driver_write_state(self.d, directory.encode(), file_ending.encode(), code))

    @property
    def trial_number(self):
        return self.lib.hvc_driver_trial_number(self.d)

    def maturation_state(self):
        n = self.n_ra
        m, s, t = (np.zeros(n, np.int32) for _ in range(3))
        e, g = np.zeros(n), np.zeros(n)
        self._check(self.lib.hvc_driver_maturation(self.d, _p(m, _abi.c_int_p), _p(s, _abi.c_int_p),
                                                   _p(e, _abi.c_double_p), _p(g, _abi.c_double_p),
                                                   _p(t, _abi.c_int_p)))
        return dict(mature=m, maturation_scale=s, erest=e, gca=g, trials_after_replacement=t)

    def history(self):
        h = np.zeros((self.n_ra, self.maturation.RATE_WINDOW_LONG), np.int32)
        self._check(self.lib.hvc_driver_history(self.d, _p(h, _abi.c_int_p)))
        return h

    def coordinates(self):
        a, b = np.zeros((self.n_ra, 3)), np.zeros((self.n_i, 3))
        self._check(self.lib.hvc_driver_coordinates(self.d, _p(a, _abi.c_double_p), _p(b, _abi.c_double_p)))
        return a, b

    def replaced(self):
        n = ctypes.c_int()
        self._check(self.lib.hvc_driver_replaced(self.d, None, None, 0, ctypes.byref(n)))
        ids, trials = np.zeros(max(n.value, 1), np.int32), np.zeros(max(n.value, 1), np.int32)
        self._check(self.lib.hvc_driver_replaced(self.d, _p(ids, _abi.c_int_p), _p(trials, _abi.c_int_p),
                                                 len(ids), ctypes.byref(n)))
        return ids[:n.value], trials[:n.value]

    def connections(self):
        na, nb = ctypes.c_longlong(), ctypes.c_longlong()
        z = [None] * 8
        self._check(self.lib.hvc_driver_connections(self.d, ctypes.byref(na), ctypes.byref(nb), *z))
        ap, bp = np.zeros(self.n_ra + 1, np.int32), np.zeros(self.n_i + 1, np.int32)
        at, bt = np.zeros(max(na.value, 1), np.int32), np.zeros(max(nb.value, 1), np.int32)
        aw, ad = np.zeros(max(na.value, 1)), np.zeros(max(na.value, 1))
        bw, bd = np.zeros(max(nb.value, 1)), np.zeros(max(nb.value, 1))
        ip, dp = _abi.c_int_p, _abi.c_double_p
        self._check(self.lib.hvc_driver_connections(self.d, ctypes.byref(na), ctypes.byref(nb), _p(ap, ip),
                                                    _p(at, ip), _p(aw, dp), _p(ad, dp), _p(bp, ip), _p(bt, ip),
                                                    _p(bw, dp), _p(bd, dp)))
        return dict(ra2i_ptr=ap, ra2i_tgt=at[:na.value], ra2i_w=aw[:na.value], ra2i_d=ad[:na.value],
                    i2ra_ptr=bp, i2ra_tgt=bt[:nb.value], i2ra_w=bw[:nb.value], i2ra_d=bd[:nb.value])

    def rr_delays(self):
        d = np.zeros(self.net.n_candidates)
        self._check(self.lib.hvc_driver_rr_delays(self.d, _p(d, _abi.c_double_p)))
        return d

    def chain_test(self, num_trials, directory=None):
        """test_chain (NGS:1313-1482) on the current network: returns the jitter
        statistics as a dict of per-neuron arrays (formats.JITTER_FIELDS)."""
        from .formats import JITTER_FIELDS
        out = np.zeros((self.n_ra, 8))
        self._check(self.lib.hvc_driver_chain_test(self.d, num_trials, directory.encode() if directory else None,
                                                   _p(out, _abi.c_double_p)))
        d = {k: out[:, i].copy() for i, k in enumerate(JITTER_FIELDS)}
        d["num_test_trials"] = num_trials
        return d

    def set_tuning(self, integrate_block, i_lanes, deliver_block, sync_block, ra_lanes=0, deliver_lists=1):
        """Launch shapes of the driver's engine (kept across rebuilds)."""
        t = _abi.Tuning(integrate_block, i_lanes, deliver_block, sync_block, ra_lanes, deliver_lists)
        st = self.lib.hvc_set_tuning(self.engine_ctx(), ctypes.byref(t))
        if st != _abi.HVC_OK:
            raise EngineError(st, "hvc_set_tuning")

    def autotune(self, windows=400):
        """Time the launch-shape candidates on the driver's network and keep the
        fastest (kept across rebuilds); returns the shape."""
        t = _abi.Tuning()
        ms = (ctypes.c_double * 3)()
        st = self.lib.hvc_autotune(self.engine_ctx(), windows, ctypes.byref(t), ms)
        if st != _abi.HVC_OK:
            raise EngineError(st, "hvc_autotune")
        return {f: getattr(t, f) for f, _ in t._fields_}

    def engine_ctx(self):
        """The current engine context (valid until the next trial that replaces neurons)."""
        return self.lib.hvc_driver_engine(self.d)


def activity_update(history, head, soma_counts, trials_after_replacement, maturation_scale,
                    death_rate_threshold, recent=None, activity_window=ACTIVITY_WINDOW):
    """check_neuron_activity (NGS:9573-9627) on host arrays; history is the ring
    [n][window] with logical entry 0 at `head`; recent (optional, int32 per
    neuron) the running count of spiking trials among the newest
    activity_window.  Updates history, recent and maturation_scale in place;
    returns (new head, neurons to replace, driven count)."""
    lib = _abi.load()
    n, window = history.shape
    assert history.dtype == np.int32 and history.flags.c_contiguous
    assert maturation_scale.dtype == np.int32 and maturation_scale.flags.c_contiguous
    assert recent is None or (recent.dtype == np.int32 and recent.flags.c_contiguous)
    h = ctypes.c_int(head)
    to_rep = np.zeros(max(n, 1), np.int32)
    nr, nd = ctypes.c_int(), ctypes.c_int()
    st = lib.hvc_activity_update(n, window, int(activity_window), _p(history, _abi.c_int_p), ctypes.byref(h),
                                 _p(recent, _abi.c_int_p),
                                 _p(_i32(soma_counts), _abi.c_int_p), _p(_i32(trials_after_replacement), _abi.c_int_p),
                                 _p(maturation_scale, _abi.c_int_p), float(death_rate_threshold),
                                 _p(to_rep, _abi.c_int_p), ctypes.byref(nr), ctypes.byref(nd))
    if st != _abi.HVC_OK:
        raise ValueError("hvc_activity_update: %s" % _abi.STATUS_NAMES.get(st, st))
    return h.value, to_rep[:nr.value].copy(), nd.value


def maturation_update(maturation: MaturationParams, maturation_scale, mature, erest, gca):
    """update_neuron_properties_sameDendrite_diffMaturationRate (NGS:2152-2183) in
    place on host arrays; returns the number of neurons that became mature."""
    lib = _abi.load()
    for a, t in ((mature, np.int32), (erest, np.float64), (gca, np.float64)):
        assert a.dtype == t and a.flags.c_contiguous
    m = _mat_c(maturation)
    k = ctypes.c_int()
    st = lib.hvc_maturation_update(len(mature), ctypes.byref(m), _p(_i32(maturation_scale), _abi.c_int_p),
                                   _p(mature, _abi.c_int_p), _p(erest, _abi.c_double_p), _p(gca, _abi.c_double_p),
                                   ctypes.byref(k))
    if st != _abi.HVC_OK:
        raise ValueError("hvc_maturation_update: %s" % _abi.STATUS_NAMES.get(st, st))
    return k.value
