"""Write the golden fixtures of tests/golden from the reference neuron classes.

Runs in the build container only (needs oracle/_ref, compiled from the
reference sources by oracle/Makefile).  Each fixture is data: the inputs are
regenerated from the seeds in tests/oracle_lib.py, the outputs are what the
reference classes HH2_test / HH2_buffer / HHI_buffer produced:
  ref_ra_m{model}_p{poisson}.npz : fired flags [steps][n], states every 100 steps
  ref_i_m0_f{first}.npz          : same for HVC-I (DP8)
  ref_poisson_noise.npz          : draws of the real poisson_noise.cpp (random,
                                   get_spike_time) and the raw mt19937 stream
                                   with the same seed.
  ref_host_rng.npz               : Poisson_noise::random(G) for varying G and
                                   sample_integer(0, hi), seed 15000.
Variant "ora" links the oracle's exp/log in place of libm's (bit-exact
targets); variant "glibc" keeps libm (tolerance targets).
    python tests/golden/make_golden.py [--host-rng]
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import i_inputs, ra_inputs, ref, run_i_trace, run_ra_trace  # noqa: E402

N, STEPS, EVERY = 8, 2000, 100


def main():
    for variant in ("ora", "glibc"):
        lib = ref(variant)
        if lib is None:
            raise SystemExit("oracle/_ref not built; run make -C oracle")
        for model, pois in ((0, 0), (1, 0), (1, 1)):
            st, fi = run_ra_trace(lib, "ref", model, ra_inputs(N, STEPS), poisson=pois, seed=(7, 0), trial=2)
            np.savez_compressed(os.path.join(HERE, "ref_%s_ra_m%d_p%d.npz" % (variant, model, pois)),
                                fired=fi, states=st[::EVERY])
        for first in (0, 1):
            st, fi = run_i_trace(lib, "ref", 0, i_inputs(N, STEPS, exc_scale=0.3), first_trial=first,
                                 seed=(7, 0), trial=3)
            np.savez_compressed(os.path.join(HERE, "ref_%s_i_m0_f%d.npz" % (variant, first)),
                                fired=fi, states=st[::EVERY])
    lib = ref("noise")
    out = {}
    for kind, arg in ((0, 0.45), (1, 250.0)):
        d = np.zeros(1024)
        lib.refnoise_draws(ctypes.c_uint(12345), kind, ctypes.c_double(arg), 1024,
                           d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        out["draws%d" % kind] = d
    raw = np.zeros(1024, np.uint32)
    lib.refnoise_mt_raw(ctypes.c_uint(12345), 1024, raw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    out["mt_raw"] = raw
    np.savez_compressed(os.path.join(HERE, "ref_poisson_noise.npz"), **out)
    host_rng(lib)


def host_rng(lib):
    # the host generator of the growth loop: random(G) with varying G and
    # sample_integer(0, hi), seed 15000 (the reference's rank-0 seed)
    rng = np.random.default_rng(3)
    G = np.concatenate([[1.0, 2.0 * np.pi, 0.0, 300.0], rng.uniform(0, 10, 508)])
    r = np.zeros(len(G))
    lib.refnoise_random_seq(ctypes.c_uint(15000), len(G), G.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    hi = np.concatenate([[0, 1, 2, 2999, 2147483646], rng.integers(0, 100000, 507)]).astype(np.int32)
    si = np.zeros(len(hi), np.int32)
    lib.refnoise_sample_integer(ctypes.c_uint(15000), len(hi), hi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                si.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    np.savez_compressed(os.path.join(HERE, "ref_host_rng.npz"), G=G, random=r, hi=hi, sample_integer=si)


if __name__ == "__main__":
    if "--host-rng" in sys.argv:      # only ref_host_rng.npz
        host_rng(ref("noise"))
    else:
        main()
