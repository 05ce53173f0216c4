"""Test-side bindings of the oracle (oracle/liboracle.so) and of the reference
harness libraries in oracle/_ref (built from the reference sources when the
checkout is present).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

ip = ctypes.POINTER(ctypes.c_int)
dp = ctypes.POINTER(ctypes.c_double)
u8p = ctypes.POINTER(ctypes.c_uint8)
u32p = ctypes.POINTER(ctypes.c_uint32)


def P(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class OraSynParams(ctypes.Structure):
    _fields_ = [("Nss", ctypes.c_int), ("A_P", ctypes.c_double), ("A_D", ctypes.c_double),
                ("T_0", ctypes.c_double), ("T_P", ctypes.c_double), ("T_D", ctypes.c_double),
                ("TAU_P", ctypes.c_double), ("TAU_D", ctypes.c_double), ("BETA", ctypes.c_double),
                ("BETA_SUPERSYNAPSE", ctypes.c_double), ("ACTIVATION_THRESHOLD", ctypes.c_double),
                ("SUPERSYNAPSE_THRESHOLD", ctypes.c_double), ("WEIGHT_MAX", ctypes.c_double)]


class OraNetDesc(ctypes.Structure):
    _fields_ = [("n_ra", ctypes.c_int), ("n_i", ctypes.c_int), ("ra_model", ctypes.c_int),
                ("i_model", ctypes.c_int), ("dt", ctypes.c_double), ("trial_ms", ctypes.c_double),
                ("sync_ms", ctypes.c_double), ("mu_s", ctypes.c_double), ("sigma_s", ctypes.c_double),
                ("mu_d", ctypes.c_double), ("sigma_d", ctypes.c_double), ("ra_poisson", ctypes.c_int),
                ("i_poisson", ctypes.c_int), ("seed_lo", ctypes.c_uint32), ("seed_hi", ctypes.c_uint32),
                ("ra2i_ptr", ip), ("ra2i_tgt", ip), ("ra2i_w", dp), ("ra2i_d", dp),
                ("i2ra_ptr", ip), ("i2ra_tgt", ip), ("i2ra_w", dp), ("i2ra_d", dp),
                ("rr_ptr", ip), ("rr_tgt", ip), ("rr_d", dp),
                ("erest", dp), ("gca", dp), ("n_train", ctypes.c_int), ("train_ids", ip), ("spread", dp),
                ("syn", OraSynParams)]


_ora = None


def oracle():
    global _ora
    if _ora is None:
        lib = ctypes.CDLL(ORACLE)
        lib.ora_exp.restype = ctypes.c_double
        lib.ora_exp.argtypes = [ctypes.c_double]
        lib.ora_log.restype = ctypes.c_double
        lib.ora_log.argtypes = [ctypes.c_double]
        lib.ora_philox4x32.argtypes = [u32p, u32p, u32p]
        lib.ora_create.restype = ctypes.c_void_p
        lib.ora_create.argtypes = [ctypes.POINTER(OraNetDesc)]
        lib.ora_destroy.argtypes = [ctypes.c_void_p]
        lib.ora_destroy.restype = None
        lib.ora_set_synapses.argtypes = [ctypes.c_void_p, dp, ip, ip, ip, ip, ip]
        lib.ora_get_synapses.argtypes = [ctypes.c_void_p, dp, ip, ip, ip, ip, ip]
        lib.ora_set_neuron_params.argtypes = [ctypes.c_void_p, dp, dp]
        lib.ora_run_trial.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                      ctypes.c_double]
        lib.ora_spike_counts.argtypes = [ctypes.c_void_p, ctypes.c_int, ip]
        lib.ora_spike_times.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        lib.ora_warnings.argtypes = [ctypes.c_void_p, ip]
        lib.ora_final_state.argtypes = [ctypes.c_void_p, dp, dp]
        lib.ora_ltp.restype = ctypes.c_double
        lib.ora_ltp.argtypes = [ctypes.POINTER(OraSynParams), ctypes.c_double, ctypes.c_double, ip]
        lib.ora_ltd.restype = ctypes.c_double
        lib.ora_ltd.argtypes = [ctypes.POINTER(OraSynParams), ctypes.c_double, ctypes.c_double, ip]
        lib.ora_clock.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, ip]
        lib.ora_num_steps.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.ora_poisson_u32.restype = ctypes.c_uint32
        lib.ora_poisson_u32.argtypes = [ctypes.c_uint32] * 6
        lib.ora_set_probe.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        lib.ora_update_synapse.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        lib.ora_axon_remodeling.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.ora_threads.restype = ctypes.c_int
        lib.ora_threads.argtypes = []
        lib.ora_set_threads.restype = None
        lib.ora_set_threads.argtypes = [ctypes.c_int]
        _ora = lib
    return _ora


def ref_built(variant):
    """True when oracle/_ref holds the harness library of this variant (checks
    the file only: nothing is loaded, so collecting tests loads no reference code)."""
    return os.path.exists(os.path.join(REF_DIR, "libhvcref_%s.so" % variant))


def ref(variant):
    """variant 'ora' (oracle exp/log interposed), 'glibc' or 'noise'; None when not built."""
    if not ref_built(variant):
        return None
    return ctypes.CDLL(os.path.join(REF_DIR, "libhvcref_%s.so" % variant))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    oracle().ora_philox4x32(c, k, o)
    return list(o)


# ---------------- single-neuron traces (same protocol for oracle and ref) ----------------
def ra_inputs(n, steps, seed=5, kick_step=100, exc_scale=20.0, inh_scale=0.5):
    rng = np.random.default_rng(seed)
    erest = np.where(np.arange(n) % 2 == 0, -80.0, -65.0)
    gca = np.where(np.arange(n) % 2 == 0, 55.0, 15.0)
    kick = np.zeros((steps, n))
    exc = np.zeros((steps, n))
    inh = np.zeros((steps, n))
    if kick_step is not None and kick_step < steps:
        kick[kick_step, :] = 3.0
    for s in range(12, steps, 12):
        m = rng.random(n) < 0.3
        exc[s, m] = rng.uniform(0, exc_scale, m.sum())
        m = rng.random(n) < 0.2
        inh[s, m] = rng.uniform(0, inh_scale, m.sum())
    return erest, gca, kick, exc, inh


def i_inputs(n, steps, seed=6, exc_scale=1.0, inh_scale=0.5):
    rng = np.random.default_rng(seed)
    exc = np.zeros((steps, n))
    inh = np.zeros((steps, n))
    for s in range(12, steps, 12):
        m = rng.random(n) < 0.3
        exc[s, m] = rng.uniform(0, exc_scale, m.sum())
        m = rng.random(n) < 0.2
        inh[s, m] = rng.uniform(0, inh_scale, m.sum())
    return exc, inh


def run_ra_trace(lib, prefix, model, inputs, mu_s=0.0, sigma_s=0.1, mu_d=0.0, sigma_d=0.2, poisson=0,
                 seed=(7, 0), trial=0, dt=0.02):
    erest, gca, kick, exc, inh = (np.ascontiguousarray(x, np.float64) for x in inputs)
    steps, n = kick.shape
    st = np.zeros((steps + 1, n, 12))
    fi = np.zeros((steps, n), np.uint8)
    f = getattr(lib, prefix + "_ra_trace")
    f.restype = ctypes.c_int
    r = f(model, n, steps, P(erest, dp), P(gca, dp), ctypes.c_double(mu_s), ctypes.c_double(sigma_s),
          ctypes.c_double(mu_d), ctypes.c_double(sigma_d), poisson, ctypes.c_uint32(seed[0]),
          ctypes.c_uint32(seed[1]), ctypes.c_uint32(trial), ctypes.c_double(dt), P(kick, dp), P(exc, dp),
          P(inh, dp), P(st, dp), P(fi, u8p))
    assert r == 0
    return st, fi


def run_i_trace(lib, prefix, model, inputs, poisson=1, first_trial=0, seed=(7, 0), trial=3, dt=0.02):
    exc, inh = (np.ascontiguousarray(x, np.float64) for x in inputs)
    steps, n = exc.shape
    st = np.zeros((steps + 1, n, 7))
    fi = np.zeros((steps, n), np.uint8)
    f = getattr(lib, prefix + "_i_trace")
    f.restype = ctypes.c_int
    r = f(model, n, steps, poisson, first_trial, ctypes.c_uint32(seed[0]), ctypes.c_uint32(seed[1]),
          ctypes.c_uint32(trial), ctypes.c_double(dt), P(exc, dp), P(inh, dp), P(st, dp), P(fi, u8p))
    assert r == 0
    return st, fi


# ---------------- network trials on the oracle ----------------
class OracleSim:
    """The oracle's trial engine on a unsupervised_growth_amd.network.Network."""

    def __init__(self, net, *, ra_model=0, i_model=0, dt=0.02, trial_ms=500.0, sync_ms=0.25, noise=None,
                 ra_poisson=0, i_poisson=1, seed=(1, 0), syn=None):
        from unsupervised_growth_amd.network import NoiseParams, SynapticParams
        noise = noise or NoiseParams()
        syn = syn or SynapticParams()
        self.net = net
        self.keep = {k: (np.ascontiguousarray(v, np.int32) if v.dtype.kind in "iu" else
                         np.ascontiguousarray(v, np.float64))
                     for k, v in net.__dict__.items() if isinstance(v, np.ndarray)}
        d = OraNetDesc()
        d.n_ra, d.n_i, d.ra_model, d.i_model = net.n_ra, net.n_i, ra_model, i_model
        d.dt, d.trial_ms, d.sync_ms = dt, trial_ms, sync_ms
        d.mu_s, d.sigma_s = noise.white_noise_mean_soma, noise.white_noise_std_soma
        d.mu_d, d.sigma_d = noise.white_noise_mean_dend, noise.white_noise_std_dend
        d.ra_poisson, d.i_poisson = ra_poisson, i_poisson
        d.seed_lo, d.seed_hi = seed
        for name in ("ra2i_ptr", "ra2i_tgt", "i2ra_ptr", "i2ra_tgt", "rr_ptr", "rr_tgt", "train_ids"):
            setattr(d, name, P(self.keep[name], ip))
        for name in ("ra2i_w", "ra2i_d", "i2ra_w", "i2ra_d", "rr_d", "erest", "gca", "spread"):
            setattr(d, name, P(self.keep[name], dp))
        d.n_train = len(net.train_ids)
        for k in syn.__dataclass_fields__:
            setattr(d.syn, k, getattr(syn, k))
        self.syn = d.syn
        self.lib = oracle()
        self.h = self.lib.ora_create(ctypes.byref(d))
        assert self.h, "ora_create failed"
        k = self.keep
        r = self.lib.ora_set_synapses(self.h, P(k["rr_w"], dp), P(k["act_cnt"], ip), P(k["act_ids"], ip),
                                      P(k["sup_cnt"], ip), P(k["sup_ids"], ip), P(k["remodeled"], ip))
        assert r == 0, "ora_set_synapses failed"

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_neuron_params(self, erest, gca):
        erest = np.ascontiguousarray(erest, np.float64)
        gca = np.ascontiguousarray(gca, np.float64)
        self.lib.ora_set_neuron_params(self.h, P(erest, dp), P(gca, dp))

    def set_synapses(self, w, act_cnt, act_ids, sup_cnt, sup_ids, remodeled):
        a = [np.ascontiguousarray(w, np.float64)] + [np.ascontiguousarray(x, np.int32) for x in
                                                      (act_cnt, act_ids, sup_cnt, sup_ids, remodeled)]
        r = self.lib.ora_set_synapses(self.h, P(a[0], dp), *[P(x, ip) for x in a[1:]])
        assert r == 0

    def run_trial(self, mode=0, training=True, trial_index=0, first_trial=False, mean_injection_time=100.0):
        rc = self.lib.ora_run_trial(self.h, mode, int(bool(training)), trial_index, int(bool(first_trial)),
                                    ctypes.c_double(mean_injection_time))
        out = {"rc": rc}
        for kind, name in ((0, "soma"), (1, "dend"), (2, "interneuron")):
            n = self.net.n_i if kind == 2 else self.net.n_ra
            c = np.zeros(max(n, 1), np.int32)
            tot = self.lib.ora_spike_counts(self.h, kind, P(c, ip))
            t = np.zeros(max(tot, 1))
            self.lib.ora_spike_times(self.h, kind, P(t, dp))
            out[name] = np.split(t[:tot], np.cumsum(c[:n])[:-1]) if n else []
        w3 = np.zeros(3, np.int32)
        self.lib.ora_warnings(self.h, P(w3, ip))
        out["warnings"] = tuple(int(x) for x in w3)
        ras = np.zeros((self.net.n_ra, 12))
        ins = np.zeros((max(self.net.n_i, 1), 7))
        self.lib.ora_final_state(self.h, P(ras, dp), P(ins, dp))
        out["ra_state"], out["i_state"] = ras, ins[:self.net.n_i]
        return out

    def synapses(self):
        E, n = self.net.n_candidates, self.net.n_ra
        w = np.zeros(max(E, 1))
        ac, sc, rem = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        ai, si = np.zeros(max(E, 1), np.int32), np.zeros(max(E, 1), np.int32)
        self.lib.ora_get_synapses(self.h, P(w, dp), P(ac, ip), P(ai, ip), P(sc, ip), P(si, ip), P(rem, ip))
        return {"w": w[:E], "act_cnt": ac, "act_ids": ai[:ac.sum()], "sup_cnt": sc, "sup_ids": si[:sc.sum()],
                "remodeled": rem}
