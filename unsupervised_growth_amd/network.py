"""Network description and synthetic networks for the HVC growth trial.

A ``Network`` carries the arrays the engine's C ABI takes (``hvc_network`` in
include/hvc_engine.h):

* RA->I and I->RA connections as CSR by presynaptic neuron -- the reference's
  ``syn_ID_RA_I``, ``weights_RA_I``, ``axonal_delays_RA_I`` and the I->RA
  counterparts (c++/include/NetworkGrowthSimulator.h:142-160);
* the RA->RA candidate set as CSR by presynaptic neuron with ascending targets,
  its delays and weights (the reference keeps dense ``N_RA x N_RA`` arrays,
  NetworkGrowthSimulator.h:142-172; ``rr="dense"`` builds exactly that set);
* the synapse state: active and super lists in insertion order and the
  remodeled flags (``active_synapses``, ``supersynapses``, ``remodeled``);
* per-RA maturation properties (``rest_potential``, ``GCa``) and the training
  neurons with their spread times.
"""
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

TIMESTEP = 0.02              # NetworkGrowthSimulator.cpp:15
NETWORK_UPDATE_FREQUENCY = 0.25  # NetworkGrowthSimulator.cpp:16
TRIAL_DURATION = 500.0       # NetworkGrowthSimulator.cpp:17
WAITING_TIME = 100.0         # NetworkGrowthSimulator.cpp:14
G_TRAINING_KICK = 3.0        # NetworkGrowthSimulator.cpp:22


@dataclass
class SynapticParams:
    """SynapticParameters (c++/include/ConfigurationNetworkGrowth.h:11-34)."""
    Nss: int = 2
    A_P: float = 2.0
    A_D: float = 0.2
    T_0: float = 1.0
    T_P: float = 5.0
    T_D: float = 5.0
    TAU_P: float = 20.0
    TAU_D: float = 20.0
    BETA: float = 0.05
    BETA_SUPERSYNAPSE: float = 0.02
    ACTIVATION_THRESHOLD: float = 0.5
    SUPERSYNAPSE_THRESHOLD: float = 2.0
    WEIGHT_MAX: float = 10.0


@dataclass
class NoiseParams:
    """NoiseParameters (ConfigurationNetworkGrowth.h:77-89)."""
    white_noise_mean_soma: float = 0.0
    white_noise_std_soma: float = 0.1
    white_noise_mean_dend: float = 0.0
    white_noise_std_dend: float = 0.2


@dataclass
class MaturationParams:
    """MaturationParameters (ConfigurationNetworkGrowth.h:37-75) with the
    reference's defaults.  The live growth path reads E_REST_*, GCA_*,
    DEATH_RATE_THRESHOLD and RATE_WINDOW_LONG; the other properties are kept so
    that configuration files round-trip."""
    E_REST_MATURE: float = -80.0
    E_REST_IMMATURE: float = -80.0
    GCA_MATURE: float = 55.0
    GCA_IMMATURE: float = 0.0
    DEATH_RATE_THRESHOLD: float = 0.01
    RATE_WINDOW_LONG: int = 2000
    E_GABA_MATURE: float = -80.0
    E_GABA_IMMATURE: float = -55.0
    AD_MATURE: float = 10000.0
    AD_IMMATURE: float = 1000.0
    GK_MATURE: float = 8.0
    GK_IMMATURE: float = 8.0
    GNA_MATURE: float = 60.0
    GNA_IMMATURE: float = 60.0
    GCAK_MATURE: float = 150.0
    GCAK_IMMATURE: float = 0.0
    GSL_MATURE: float = 0.1
    GSL_IMMATURE: float = 0.1
    GDL_MATURE: float = 0.1
    GDL_IMMATURE: float = 0.1
    RC_MATURE: float = 55.0
    RC_IMMATURE: float = 5.5


@dataclass
class Network:
    n_ra: int
    n_i: int
    ra2i_ptr: np.ndarray
    ra2i_tgt: np.ndarray
    ra2i_w: np.ndarray
    ra2i_d: np.ndarray
    i2ra_ptr: np.ndarray
    i2ra_tgt: np.ndarray
    i2ra_w: np.ndarray
    i2ra_d: np.ndarray
    rr_ptr: np.ndarray
    rr_tgt: np.ndarray
    rr_d: np.ndarray
    rr_w: np.ndarray
    act_cnt: np.ndarray
    act_ids: np.ndarray
    sup_cnt: np.ndarray
    sup_ids: np.ndarray
    remodeled: np.ndarray
    erest: np.ndarray
    gca: np.ndarray
    train_ids: np.ndarray
    spread: np.ndarray

    @property
    def n_candidates(self):
        return int(self.rr_ptr[-1])

    def copy(self):
        return replace(self, **{k: (v.copy() if isinstance(v, np.ndarray) else v)
                                for k, v in self.__dict__.items()})


def _csr(rows, n_pre):
    ptr = np.zeros(n_pre + 1, np.int32)
    for i, r in enumerate(rows):
        ptr[i + 1] = ptr[i] + len(r)
    return ptr


def _delays(rng, n, kind, dt):
    kind, *args = kind if isinstance(kind, tuple) else (kind,)
    if kind == "quantised":          # k * dt, k uniform in {0..kmax}
        kmax = args[0]
        return rng.integers(0, kmax + 1, n).astype(np.float64) * dt
    if kind == "uniform":            # U(lo, hi)
        lo, hi = args
        return rng.uniform(lo, hi, n)
    if kind == "lognormal":          # median, sigma, clip
        med, sig, clip = args
        return np.minimum(np.exp(np.log(med) + sig * rng.standard_normal(n)), clip)
    raise ValueError(kind)


def random_network(n_ra, n_i, *, n_train=10, p_ra2i=0.1, p_i2ra=0.2, rr="dense", gei_max=0.25,
                   gie_max=0.5, delays=("quantised", 250), rr_delays=None, w_rr=0.0, seed=1,
                   dt=TIMESTEP, positions=None, spread_max=0.0, immature_frac=0.0,
                   maturation: Optional[MaturationParams] = None):
    """Synthetic network.

    * RA->I / I->RA: each pair connected with probability p (constant, or a
      callable of the distance); weights ``U(0,1) * G_max`` as sample_G
      (NetworkGrowthSimulator.cpp:739-742).
    * RA->RA candidates: ``"dense"`` = all targets (the reference's dense set),
      or an int k = k nearest other neurons.
    * delays: ("quantised", kmax) -> k*dt, ("uniform", lo, hi), ("lognormal", median, sigma, clip).
    """
    rng = np.random.default_rng(seed)
    mat = maturation or MaturationParams()
    if positions is None:
        pos_ra = rng.uniform(0, 1, (n_ra, 3))
        pos_i = rng.uniform(0, 1, (n_i, 3))
    else:
        pos_ra, pos_i = positions

    def connect(pa, pb, p, gmax):
        rows_t, rows_w, rows_d = [], [], []
        for a in range(len(pa)):
            if len(pb) == 0:
                rows_t.append(np.zeros(0, np.int32))
                rows_w.append(np.zeros(0))
                continue
            if callable(p):
                d = np.linalg.norm(pb - pa[a], axis=1)
                prob = p(d)
            else:
                prob = np.full(len(pb), p)
            tg = np.nonzero(rng.random(len(pb)) < prob)[0].astype(np.int32)
            rows_t.append(tg)
            rows_w.append(rng.random(len(tg)) * gmax)
        ptr = _csr(rows_t, len(pa))
        tgt = np.concatenate(rows_t).astype(np.int32) if rows_t else np.zeros(0, np.int32)
        w = np.concatenate(rows_w) if rows_w else np.zeros(0)
        d = _delays(rng, len(tgt), delays, dt)
        return ptr, tgt, w, d

    ra2i_ptr, ra2i_tgt, ra2i_w, ra2i_d = connect(pos_ra, pos_i, p_ra2i, gei_max)
    i2ra_ptr, i2ra_tgt, i2ra_w, i2ra_d = connect(pos_i, pos_ra, p_i2ra, gie_max)

    if rr == "dense":
        rr_ptr = np.arange(n_ra + 1, dtype=np.int32) * n_ra
        rr_tgt = np.tile(np.arange(n_ra, dtype=np.int32), n_ra)
    else:
        k = int(rr)
        rows = []
        for i in range(n_ra):
            d = np.linalg.norm(pos_ra - pos_ra[i], axis=1)
            d[i] = np.inf
            nn = np.sort(np.argsort(d)[:k]).astype(np.int32)
            rows.append(nn)
        rr_ptr = _csr(rows, n_ra)
        rr_tgt = np.concatenate(rows).astype(np.int32)
    E = int(rr_ptr[-1])
    rr_d = _delays(rng, E, rr_delays or delays, dt)
    rr_w = np.full(E, float(w_rr)) if np.isscalar(w_rr) else np.asarray(w_rr, np.float64)

    train_ids = np.sort(rng.choice(n_ra, size=min(n_train, n_ra), replace=False)).astype(np.int32)
    spread = rng.uniform(0, spread_max, len(train_ids)) if spread_max > 0 else np.zeros(len(train_ids))
    erest = np.full(n_ra, mat.E_REST_MATURE)
    gca = np.full(n_ra, mat.GCA_MATURE)
    if immature_frac > 0:
        imm = rng.random(n_ra) < immature_frac
        imm[train_ids] = False
        erest[imm] = mat.E_REST_IMMATURE
        gca[imm] = mat.GCA_IMMATURE
    return Network(
        n_ra=n_ra, n_i=n_i,
        ra2i_ptr=ra2i_ptr, ra2i_tgt=ra2i_tgt, ra2i_w=ra2i_w, ra2i_d=ra2i_d,
        i2ra_ptr=i2ra_ptr, i2ra_tgt=i2ra_tgt, i2ra_w=i2ra_w, i2ra_d=i2ra_d,
        rr_ptr=rr_ptr, rr_tgt=rr_tgt, rr_d=rr_d, rr_w=rr_w,
        act_cnt=np.zeros(n_ra, np.int32), act_ids=np.zeros(0, np.int32),
        sup_cnt=np.zeros(n_ra, np.int32), sup_ids=np.zeros(0, np.int32),
        remodeled=np.zeros(n_ra, np.int32), erest=erest, gca=gca,
        train_ids=train_ids, spread=spread)


def sphere_positions(rng, n):
    """Uniform points on the unit sphere (the reference's RA arrangement, NGS:358-401)."""
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def gaussian_probability(peak, width):
    return lambda d: peak * np.exp(-d * d / (2 * width * width))
