"""Chain-test statistics (C ABI hvc_chain_stats_* in include/hvc_driver.h).

calculate_and_write_jitter (c++/src/NetworkGrowthSimulator.cpp:1641-1782): per
HVC-RA neuron over the chain-test trials, the probability of a somatic spike,
the mean number of somatic spikes per trial, mean and standard deviation of the
first somatic spike time, and the same four for dendritic spikes; -1 where a
mean or deviation is undefined.  The columns follow jitter.bin
(formats.JITTER_FIELDS).
"""
import ctypes

import numpy as np

from . import _abi
from .engine import SpikeTrains
from .formats import JITTER_FIELDS


def _flat(trains):
    if isinstance(trains, SpikeTrains):
        return (np.ascontiguousarray(trains.counts, np.int32),
                np.ascontiguousarray(trains.times if trains.total else np.zeros(1), np.float64))
    counts = np.array([len(t) for t in trains], np.int32)
    times = np.concatenate([np.asarray(t, np.float64) for t in trains]) if len(trains) else np.zeros(0)
    return counts, np.ascontiguousarray(times if len(times) else np.zeros(1))


class ChainStats:
    def __init__(self, n_ra):
        self.lib = _abi.load()
        self.n_ra = n_ra
        self.h = ctypes.c_void_p()
        st = self.lib.hvc_chain_stats_create(n_ra, ctypes.byref(self.h))
        if st != _abi.HVC_OK:
            raise ValueError("hvc_chain_stats_create: %s" % _abi.STATUS_NAMES.get(st, st))

    def add(self, soma, dend):
        """One trial: per-neuron spike-time arrays (somatic, dendritic)."""
        sc, st_ = _flat(soma)
        dc, dt = _flat(dend)
        ip, dp = _abi.c_int_p, _abi.c_double_p
        st = self.lib.hvc_chain_stats_add(self.h, sc.ctypes.data_as(ip), st_.ctypes.data_as(dp),
                                          dc.ctypes.data_as(ip), dt.ctypes.data_as(dp))
        if st != _abi.HVC_OK:
            raise ValueError("hvc_chain_stats_add: %s" % _abi.STATUS_NAMES.get(st, st))

    def add_result(self, res):
        self.add(res.soma, res.dend)

    def array(self):
        """(num_trials, n_ra x 8 array in jitter.bin order)."""
        out = np.zeros((self.n_ra, 8))
        n = ctypes.c_int()
        st = self.lib.hvc_chain_stats_result(self.h, ctypes.byref(n), out.ctypes.data_as(_abi.c_double_p))
        if st != _abi.HVC_OK:
            raise ValueError("hvc_chain_stats_result: %s" % _abi.STATUS_NAMES.get(st, st))
        return n.value, out

    def result(self):
        n, a = self.array()
        d = {k: a[:, i].copy() for i, k in enumerate(JITTER_FIELDS)}
        d["num_test_trials"] = n
        return d

    def write(self, path):
        st = self.lib.hvc_chain_stats_write(self.h, str(path).encode())
        if st != _abi.HVC_OK:
            raise ValueError("hvc_chain_stats_write: %s" % _abi.STATUS_NAMES.get(st, st))

    def close(self):
        if self.h:
            self.lib.hvc_chain_stats_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
