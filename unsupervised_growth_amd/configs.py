"""The benchmark and parity workloads, built as synthetic networks.

C1  300 RA + 30 I, positions uniform in the unit box, RA->I p = 0.1, I->RA
    p = 0.2, RA->RA candidates dense and empty (w = 0), delays k*dt with k
    uniform in {0..250} (0-5 ms), 10 training neurons, white noise sigma 0.1 /
    0.2, 500 ms trials (the reference's CPU-runnable case).
C2  2000 RA + 550 I on the unit sphere, Gaussian distance rule (peak 0.3,
    width 0.1), lognormal delays (median 2 ms, sigma 0.5, clip 15 ms), dense
    RA->RA candidates, 1000 ms trials, noise seed 7.
C3  20000 RA + 5500 I on the unit sphere, the same distance rule, 100 RA->RA
    candidates per RA (nearest), delays U(0, 10) ms.
C4  batched form: 64 independent networks, N_RA = round(exp(U(ln 500, ln 8000))),
    N_I = round(0.275 N_RA), each built like C3 with delays U(0, 8) ms and seed
    100 + b (also its noise key), 1000 ms trials.
Weights follow sample_G: U(0,1) * G_max (NetworkGrowthSimulator.cpp:739-742).
Each builder returns (network, engine keyword arguments).
"""
import numpy as np

from .network import (MaturationParams, Network, NoiseParams, SynapticParams, TIMESTEP, _csr, _delays,
                      sphere_positions)


def _gaussian_csr(rng, pa, pb, peak, width, gmax, delays, dt, chunk=2048):
    rows_t, rows_w = [], []
    for a0 in range(0, len(pa), chunk):
        blk = pa[a0:a0 + chunk]
        if len(pb) == 0:
            rows_t += [np.zeros(0, np.int32)] * len(blk)
            rows_w += [np.zeros(0)] * len(blk)
            continue
        d2 = ((blk[:, None, :] - pb[None, :, :]) ** 2).sum(-1)
        prob = peak * np.exp(-d2 / (2 * width * width))
        hit = rng.random(prob.shape) < prob
        for r in range(len(blk)):
            tg = np.nonzero(hit[r])[0].astype(np.int32)
            rows_t.append(tg)
            rows_w.append(rng.random(len(tg)) * gmax)
    ptr = _csr(rows_t, len(pa))
    tgt = np.concatenate(rows_t).astype(np.int32) if rows_t else np.zeros(0, np.int32)
    w = np.concatenate(rows_w) if rows_w else np.zeros(0)
    return ptr, tgt, w, _delays(rng, len(tgt), delays, dt)


def _nearest_candidates(pos, k):
    from scipy.spatial import cKDTree
    tree = cKDTree(pos)
    _, idx = tree.query(pos, k=k + 1)
    rows = []
    for i in range(len(pos)):
        nn = idx[i][idx[i] != i][:k]
        rows.append(np.sort(nn).astype(np.int32))
    return _csr(rows, len(pos)), np.concatenate(rows).astype(np.int32)


def _finish(rng, n_ra, n_i, ra2i, i2ra, rr_ptr, rr_tgt, rr_d, n_train, spread_max=0.0):
    mat = MaturationParams()
    train_ids = np.sort(rng.choice(n_ra, size=n_train, replace=False)).astype(np.int32)
    spread = rng.uniform(0, spread_max, n_train) if spread_max > 0 else np.zeros(n_train)
    E = int(rr_ptr[-1])
    return Network(n_ra=n_ra, n_i=n_i,
                   ra2i_ptr=ra2i[0], ra2i_tgt=ra2i[1], ra2i_w=ra2i[2], ra2i_d=ra2i[3],
                   i2ra_ptr=i2ra[0], i2ra_tgt=i2ra[1], i2ra_w=i2ra[2], i2ra_d=i2ra[3],
                   rr_ptr=rr_ptr, rr_tgt=rr_tgt, rr_d=rr_d, rr_w=np.zeros(E),
                   act_cnt=np.zeros(n_ra, np.int32), act_ids=np.zeros(0, np.int32),
                   sup_cnt=np.zeros(n_ra, np.int32), sup_ids=np.zeros(0, np.int32),
                   remodeled=np.zeros(n_ra, np.int32),
                   erest=np.full(n_ra, mat.E_REST_MATURE), gca=np.full(n_ra, mat.GCA_MATURE),
                   train_ids=train_ids, spread=spread)


def c1(seed=1):
    from .network import random_network
    net = random_network(300, 30, n_train=10, p_ra2i=0.1, p_i2ra=0.2, rr="dense",
                         delays=("quantised", 250), seed=seed)
    return net, dict(trial_ms=500.0, seed=(seed, 0), noise=NoiseParams(0.0, 0.1, 0.0, 0.2))


def c2(seed=7, trial_ms=1000.0):
    rng = np.random.default_rng(seed)
    n_ra, n_i = 2000, 550
    pr, pi = sphere_positions(rng, n_ra), sphere_positions(rng, n_i)
    dl = ("lognormal", 2.0, 0.5, 15.0)
    ra2i = _gaussian_csr(rng, pr, pi, 0.3, 0.1, 0.25, dl, TIMESTEP)
    i2ra = _gaussian_csr(rng, pi, pr, 0.3, 0.1, 0.5, dl, TIMESTEP)
    rr_ptr = np.arange(n_ra + 1, dtype=np.int32) * n_ra
    rr_tgt = np.tile(np.arange(n_ra, dtype=np.int32), n_ra)
    rr_d = _delays(rng, len(rr_tgt), dl, TIMESTEP)
    net = _finish(rng, n_ra, n_i, ra2i, i2ra, rr_ptr, rr_tgt, rr_d, 10)
    return net, dict(trial_ms=trial_ms, seed=(seed, 0))


def sphere_network(rng, n_ra, n_i, delay_max, k_rr=100, n_train=None, spread_max=20.0):
    """RA and I uniform on the unit sphere, Gaussian distance rule (peak 0.3,
    width 0.1) for RA->I and I->RA, the k_rr nearest RA as RA->RA candidates,
    delays U(0, delay_max) ms."""
    pr, pi = sphere_positions(rng, n_ra), sphere_positions(rng, n_i)
    dl = ("uniform", 0.0, delay_max)
    ra2i = _gaussian_csr(rng, pr, pi, 0.3, 0.1, 0.25, dl, TIMESTEP)
    i2ra = _gaussian_csr(rng, pi, pr, 0.3, 0.1, 0.5, dl, TIMESTEP)
    rr_ptr, rr_tgt = _nearest_candidates(pr, min(k_rr, n_ra - 1))
    rr_d = _delays(rng, len(rr_tgt), dl, TIMESTEP)
    if n_train is None:
        n_train = max(10, n_ra // 200)
    return _finish(rng, n_ra, n_i, ra2i, i2ra, rr_ptr, rr_tgt, rr_d, n_train, spread_max=spread_max)


def c4(n_nets=64, seed0=100):
    """The batched workload: (networks, noise keys, engine keyword arguments)."""
    nets, keys = [], []
    for b in range(n_nets):
        rng = np.random.default_rng(seed0 + b)
        n_ra = int(round(np.exp(rng.uniform(np.log(500.0), np.log(8000.0)))))
        n_i = int(round(0.275 * n_ra))
        nets.append(sphere_network(rng, n_ra, n_i, 8.0))
        keys.append((seed0 + b, 0))
    return nets, keys, dict(trial_ms=1000.0)


def c3(seed=11, k_rr=100):
    rng = np.random.default_rng(seed)
    n_ra, n_i = 20000, 5500
    pr, pi = sphere_positions(rng, n_ra), sphere_positions(rng, n_i)
    dl = ("uniform", 0.0, 10.0)
    ra2i = _gaussian_csr(rng, pr, pi, 0.3, 0.1, 0.25, dl, TIMESTEP)
    i2ra = _gaussian_csr(rng, pi, pr, 0.3, 0.1, 0.5, dl, TIMESTEP)
    rr_ptr, rr_tgt = _nearest_candidates(pr, k_rr)
    rr_d = _delays(rng, len(rr_tgt), dl, TIMESTEP)
    net = _finish(rng, n_ra, n_i, ra2i, i2ra, rr_ptr, rr_tgt, rr_d, 100, spread_max=20.0)
    return net, dict(trial_ms=500.0, seed=(seed, 0))


CONFIGS = {"c1": c1, "c2": c2, "c3": c3}
