"""Network topology generation through the C ABI of include/hvc_topology.h.

``generate`` is the reference's generate_network_topology
(c++/src/NetworkGrowthSimulator.cpp:110-155, "NGS"): HVC-I on a jittered
lattice, HVC-RA uniform, Gaussian distance rules for RA->I and I->RA, and the
clustered and random training sets.  ``prepare`` is what new_chain_growth does
to a read topology before growth (NGS:1882-1885): delays from lengths, weights
U(0,1) * G_max.  ``growth_network`` assembles the engine's Network for a
growth run the way new_chain_growth initialises it (NGS:1887-1899): dense
RA->RA candidates with distance delays, zero weights, empty synapse lists,
training neurons mature, every other neuron immature.

The generator is the reference's (std::mt19937, seed 15000 on rank 0,
NGS:2443), so the defaults give the reference's own network for the same sizes
and parameters.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from .network import MaturationParams, Network

REFERENCE_SEED = 15000          # NGS:2443, rank 0
ARRANGEMENTS = {"sphere": 0, "square": 1}


@dataclass
class TopologyParams:
    """TopologyParameters (c++/include/ConfigurationNetworkTopology.h:11-22) and
    ConnectionParameters (c++/include/ConfigurationNetworkGrowth.h:97-105)."""
    arrangement: str = "sphere"
    A_RA2I: float = 0.0
    SIGMA_RA2I: float = 0.0
    B_I2RA: float = 0.0
    SIGMA_I2RA: float = 0.0
    Gei_max: float = 0.25
    Gie_max: float = 0.0
    delay_constant: float = 0.0

    def c(self):
        p = _abi.TopologyParams()
        p.arrangement = ARRANGEMENTS[self.arrangement]
        for k in ("A_RA2I", "SIGMA_RA2I", "B_I2RA", "SIGMA_I2RA", "Gei_max", "Gie_max", "delay_constant"):
            setattr(p, k, float(getattr(self, k)))
        return p


@dataclass
class Topology:
    n_ra: int
    n_i: int
    dim: int
    model_interneuron_distance: float
    ra_xyz: np.ndarray          # (n_ra, 3)
    i_xyz: np.ndarray           # (n_i, 3)
    ra2i_ptr: np.ndarray
    ra2i_tgt: np.ndarray
    ra2i_w: np.ndarray
    ra2i_len: np.ndarray        # microns
    ra2i_d: np.ndarray          # ms
    i2ra_ptr: np.ndarray
    i2ra_tgt: np.ndarray
    i2ra_w: np.ndarray
    i2ra_len: np.ndarray
    i2ra_d: np.ndarray
    training_clustered: np.ndarray
    training_random: np.ndarray


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _from_c(t):
    n_ra, n_i, n_tr = t.n_ra, t.n_i, t.n_tr
    a = _arr(t.ra2i_ptr, n_ra + 1, np.int32)
    b = _arr(t.i2ra_ptr, n_i + 1, np.int32)
    ea, eb = int(a[-1]), int(b[-1])
    return Topology(n_ra=n_ra, n_i=n_i, dim=t.dim, model_interneuron_distance=t.model_interneuron_distance,
                    ra_xyz=_arr(t.ra_xyz, 3 * n_ra, np.float64).reshape(n_ra, 3),
                    i_xyz=_arr(t.i_xyz, 3 * n_i, np.float64).reshape(n_i, 3),
                    ra2i_ptr=a, ra2i_tgt=_arr(t.ra2i_tgt, ea, np.int32), ra2i_w=_arr(t.ra2i_w, ea, np.float64),
                    ra2i_len=_arr(t.ra2i_len, ea, np.float64), ra2i_d=_arr(t.ra2i_d, ea, np.float64),
                    i2ra_ptr=b, i2ra_tgt=_arr(t.i2ra_tgt, eb, np.int32), i2ra_w=_arr(t.i2ra_w, eb, np.float64),
                    i2ra_len=_arr(t.i2ra_len, eb, np.float64), i2ra_d=_arr(t.i2ra_d, eb, np.float64),
                    training_clustered=_arr(t.training_clustered, n_tr, np.int32),
                    training_random=_arr(t.training_random, n_tr, np.int32))


def _check(st, what):
    if st != _abi.HVC_OK:
        raise ValueError("%s failed: %s" % (what, _abi.STATUS_NAMES.get(st, st)))


def generate(n_ra, n_i, n_tr, params: TopologyParams, seed=REFERENCE_SEED) -> Topology:
    """generate_network_topology (NGS:110-155) with a generator seeded `seed`."""
    lib = _abi.load()
    out = ctypes.POINTER(_abi.Topology)()
    p = params.c()
    _check(lib.hvc_topology_generate(n_ra, n_i, n_tr, ctypes.byref(p), seed & 0xffffffff, ctypes.byref(out)),
           "hvc_topology_generate")
    try:
        return _from_c(out.contents)
    finally:
        lib.hvc_topology_free(out)


def _to_c(t: Topology, keep):
    c = _abi.Topology()
    c.n_ra, c.n_i, c.n_tr, c.dim = t.n_ra, t.n_i, len(t.training_clustered), t.dim
    c.model_interneuron_distance = t.model_interneuron_distance
    for name in ("ra_xyz", "i_xyz", "ra2i_w", "ra2i_len", "ra2i_d", "i2ra_w", "i2ra_len", "i2ra_d"):
        a = np.ascontiguousarray(getattr(t, name), np.float64).reshape(-1)
        keep[name] = a
        setattr(c, name, a.ctypes.data_as(_abi.c_double_p))
    for name in ("ra2i_ptr", "ra2i_tgt", "i2ra_ptr", "i2ra_tgt", "training_clustered", "training_random"):
        a = np.ascontiguousarray(getattr(t, name), np.int32)
        keep[name] = a
        setattr(c, name, a.ctypes.data_as(_abi.c_int_p))
    return c


def prepare(t: Topology, params: TopologyParams, seed=REFERENCE_SEED, rr_delays=True):
    """sample_axonal_delays + resample_weights (NGS:466-526, 721-737) with a
    fresh generator seeded `seed`.  Returns (topology, RA->RA delay matrix or None)."""
    lib = _abi.load()
    keep = {}
    c = _to_c(t, keep)
    rr = np.zeros((t.n_ra, t.n_ra)) if rr_delays else None
    p = params.c()
    _check(lib.hvc_topology_prepare(ctypes.byref(c), ctypes.byref(p), seed & 0xffffffff,
                                    rr.ctypes.data_as(_abi.c_double_p) if rr is not None else None),
           "hvc_topology_prepare")
    out = Topology(**{**t.__dict__})
    for name in ("ra2i_w", "ra2i_d", "i2ra_w", "i2ra_d"):
        setattr(out, name, keep[name].copy())
    return out, rr


def growth_network(t: Topology, rr_delay, training, maturation: MaturationParams = None, spread=None) -> Network:
    """The network new_chain_growth starts from (NGS:1887-1899): dense RA->RA
    candidates (delays rr_delay, weights 0), no active or super synapses,
    training neurons mature (Erest, GCa mature), the rest immature."""
    mat = maturation or MaturationParams()
    n = t.n_ra
    training = np.sort(np.asarray(training, np.int32))
    erest = np.full(n, mat.E_REST_IMMATURE)
    gca = np.full(n, mat.GCA_IMMATURE)
    erest[training] = mat.E_REST_MATURE
    gca[training] = mat.GCA_MATURE
    rr_ptr = (np.arange(n + 1, dtype=np.int64) * n).astype(np.int32)
    rr_tgt = np.tile(np.arange(n, dtype=np.int32), n)
    return Network(n_ra=n, n_i=t.n_i,
                   ra2i_ptr=t.ra2i_ptr, ra2i_tgt=t.ra2i_tgt, ra2i_w=t.ra2i_w, ra2i_d=t.ra2i_d,
                   i2ra_ptr=t.i2ra_ptr, i2ra_tgt=t.i2ra_tgt, i2ra_w=t.i2ra_w, i2ra_d=t.i2ra_d,
                   rr_ptr=rr_ptr, rr_tgt=rr_tgt, rr_d=np.ascontiguousarray(rr_delay, np.float64).reshape(-1),
                   rr_w=np.zeros(n * n),
                   act_cnt=np.zeros(n, np.int32), act_ids=np.zeros(0, np.int32),
                   sup_cnt=np.zeros(n, np.int32), sup_ids=np.zeros(0, np.int32),
                   remodeled=np.zeros(n, np.int32), erest=erest, gca=gca,
                   train_ids=training, spread=np.zeros(len(training)) if spread is None else np.asarray(spread, float))


def sphere_model_interneuron_distance(n_i):
    """MODEL_INTERNEURON_DISTANCE of the sphere arrangement (NGS:296-312): the
    distance on the sphere between the first two points of the HVC-I lattice."""
    import math
    dphi = math.pi * (3.0 - math.sqrt(5.0))
    dz = 2 / float(n_i)
    z = 1 - dz / 2.0
    x1, y1 = math.sqrt(1.0 - z * z), 0.0
    x2 = math.cos(dphi) * math.sqrt(1.0 - (z - dz) * (z - dz))
    y2 = math.sin(dphi) * math.sqrt(1.0 - (z - dz) * (z - dz))
    c = x1 * x2 + y1 * y2 + z * (z - dz)
    return 0.0 if c > 1.0 else math.acos(c)


def mt19937(seed, n, skip=0):
    """n raw words of std::mt19937 seeded `seed` after `skip` draws."""
    lib = _abi.load()
    out = np.zeros(n, np.uint32)
    _check(lib.hvc_mt19937(seed & 0xffffffff, skip, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
           "hvc_mt19937")
    return out
