"""hvc_trace_ra_noalloc (no device allocation) against hvc_trace_ra, through the C ABI.

The entry point takes device arrays and keeps its trace tables in the
library's fixed per-device buffer; its results must equal hvc_trace_ra's bit
for bit (FP64 and FP32 forms), also on a second call with a smaller input,
which reuses the buffer the first call filled.  The test's device arrays come
from the HIP runtime the engine library itself loaded (hipMalloc / hipMemcpy
through ctypes).
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import ra_inputs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from unsupervised_growth_amd import _abi
    _abi.load()
    h = ctypes.CDLL("libamdhip64.so.7")   # the runtime the engine library loaded (same SONAME)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipSetDevice):
        f.restype = ctypes.c_int
    return h


class DevArrays:
    """Device copies of host arrays, freed on exit."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(nbytes, 1)) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *a):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _noalloc(eng, hip, erest, gca, kick, exc, inh, trial):
    n_steps, n = kick.shape
    with DevArrays(hip) as d:
        args = [d.put(np.asarray(x, np.float64)) for x in (erest, gca, kick, exc, inh)]
        out = d.alloc(8 * (n_steps + 1) * n * 12)
        fired = d.alloc(n_steps * n)
        st = eng.lib.hvc_trace_ra_noalloc(eng.ctx, n, n_steps, args[0], args[1], trial, args[2], args[3], args[4],
                                          out, fired)
        eng._check(st)
        return d.get(out, (n_steps + 1, n, 12), np.float64), d.get(fired, (n_steps, n), np.uint8)


@pytest.mark.parametrize("precision,model", [(0, 0), (1, 0), (0, 1)])
def test_noalloc_trace_equals_trace(precision, model):
    from unsupervised_growth_amd.engine import Engine
    print("noalloc trace precision=%d model=%d" % (precision, model), flush=True)
    with Engine(precision=precision, ra_model=model, seed=(7, 0)) as eng:
        hip = _hip()
        for n, steps, trial in ((96, 2500, 2), (40, 900, 5)):
            inp = ra_inputs(n, steps, seed=n)
            ref_state, ref_fired = eng.trace_ra(*inp, trial=trial)
            st, fi = _noalloc(eng, hip, *inp, trial=trial)
            print("  n=%d steps=%d spikes=%d" % (n, steps, int((ref_fired & 1).sum())), flush=True)
            assert (ref_fired & 1).sum() > 0
            np.testing.assert_array_equal(fi, ref_fired)
            np.testing.assert_array_equal(st, ref_state)


def test_noalloc_rejects_host_arrays_and_long_traces():
    from unsupervised_growth_amd._abi import HVC_E_ARG
    from unsupervised_growth_amd.engine import Engine
    with Engine() as eng:
        hip = _hip()
        n, steps = 8, 16
        host = np.zeros(steps * n * 12 + 12 * n)
        hp = ctypes.c_void_p(host.ctypes.data)
        with DevArrays(hip) as d:
            dev = d.alloc(host.nbytes)
            dfi = d.alloc(steps * n)
            st = eng.lib.hvc_trace_ra_noalloc(eng.ctx, n, steps, hp, hp, 0, hp, hp, hp, dev, dfi)
            assert st == HVC_E_ARG
            assert b"device arrays" in eng.lib.hvc_last_error(eng.ctx)
            st = eng.lib.hvc_trace_ra_noalloc(eng.ctx, n, 16385, dev, dev, 0, dev, dev, dev, dev, dfi)
            assert st == HVC_E_ARG
            assert b"16384" in eng.lib.hvc_last_error(eng.ctx)
