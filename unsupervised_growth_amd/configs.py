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
C5  served form: chain-test requests (50 trials, kick onset 100 ms + N(0, 1) ms
    per trial) on grown networks of 1000 (25 %), 4000 (50 %) and 20000 (25 %)
    HVC-RA, N_I = 0.275 N_RA, Poisson arrivals at 20 requests/s.  A grown
    network is synthetic: its active RA->RA synapses form a feed-forward chain
    of groups (each neuron drives `fan` neurons of the next group), the state
    the growth rule converges to; the training neurons are the first group.
Weights follow sample_G: U(0,1) * G_max (NetworkGrowthSimulator.cpp:739-742).
Each builder returns (network, engine keyword arguments).
"""
import numpy as np

from .network import (MaturationParams, Network, NoiseParams, SynapticParams, TIMESTEP, _csr, _delays,
                      sphere_positions)


def _gaussian_csr(rng, pa, pb, peak, width, gmax, delays, dt, chunk=2048):
    """Pairs (a, b) connected with probability peak * exp(-|pa - pb|^2 / (2 width^2)):
    one uniform per pair of a chunk of rows (row-major), then the weights of the
    chunk's connections in row order, then the delays of all connections."""
    tg_chunks, w_chunks, counts = [], [], []
    den = 2 * width * width
    for a0 in range(0, len(pa), chunk):
        blk = pa[a0:a0 + chunk]
        if len(pb) == 0:
            counts.append(np.zeros(len(blk), np.int64))
            continue
        dx = blk[:, 0:1] - pb[None, :, 0]
        d2 = dx * dx
        dy = blk[:, 1:2] - pb[None, :, 1]
        d2 += dy * dy
        dz = blk[:, 2:3] - pb[None, :, 2]
        d2 += dz * dz
        hit = rng.random(d2.shape) < peak * np.exp(-d2 / den)
        rows, cols = np.nonzero(hit)
        counts.append(np.bincount(rows, minlength=len(blk)))
        tg_chunks.append(cols.astype(np.int32))
        w_chunks.append(rng.random(len(cols)) * gmax)
    ptr = np.zeros(len(pa) + 1, np.int32)
    if counts:
        np.cumsum(np.concatenate(counts), out=ptr[1:])
    tgt = np.concatenate(tg_chunks).astype(np.int32) if tg_chunks else np.zeros(0, np.int32)
    w = np.concatenate(w_chunks) if w_chunks else np.zeros(0)
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


def c2_geometry(seed=7):
    """Positions of the c2 network (its first draws) and the topology
    parameters that resample a replaced neuron by the c2 rule (Gaussian
    distance rule, peak 0.3, width 0.1 on the sphere), in the reference's units
    (distance on the sphere x 40 / MODEL_INTERNEURON_DISTANCE microns):
    (ra_xyz, i_xyz, model_interneuron_distance, TopologyParams).  Replaced
    neurons' new connections take distance-proportional delays."""
    from .topology import TopologyParams, sphere_model_interneuron_distance
    rng = np.random.default_rng(seed)
    pr, pi = sphere_positions(rng, 2000), sphere_positions(rng, 550)
    mid = sphere_model_interneuron_distance(550)
    width = 0.1 * 40.0 / mid
    return pr, pi, mid, TopologyParams(arrangement="sphere", A_RA2I=0.3, SIGMA_RA2I=width, B_I2RA=0.3,
                                       SIGMA_I2RA=width, Gei_max=0.25, Gie_max=0.5, delay_constant=1.0)


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


def c3_geometry(seed=11):
    """Positions of the c3 network (its first draws) and the topology
    parameters that resample a replaced neuron by the same rule, in the
    reference's units (distance on the sphere x 40 / MODEL_INTERNEURON_DISTANCE
    microns): (ra_xyz, i_xyz, model_interneuron_distance, TopologyParams)."""
    from .topology import TopologyParams, sphere_model_interneuron_distance
    rng = np.random.default_rng(seed)
    pr, pi = sphere_positions(rng, 20000), sphere_positions(rng, 5500)
    mid = sphere_model_interneuron_distance(5500)
    width = 0.1 * 40.0 / mid
    return pr, pi, mid, TopologyParams(arrangement="sphere", A_RA2I=0.3, SIGMA_RA2I=width, B_I2RA=0.3,
                                       SIGMA_I2RA=width, Gei_max=0.25, Gie_max=0.5, delay_constant=1.0)


def grown_chain_network(rng, n_ra, n_i, group=20, fan=10, w_chain=3000.0, delay=(1.0, 2.0), spread_max=1.0):
    """A network whose active RA->RA synapses form a chain of groups of `group`
    neurons: each neuron has `fan` candidates in the next group, all active with
    weight w_chain (raiseExcWeight adds w / Ad, Ad = 10000, HH2_test.h:89), delays
    U(delay).  RA->I and I->RA follow the sphere distance rule of C3."""
    pr, pi = sphere_positions(rng, n_ra), sphere_positions(rng, n_i)
    dl = ("uniform", 0.0, 10.0)
    ra2i = _gaussian_csr(rng, pr, pi, 0.3, 0.1, 0.25, dl, TIMESTEP)
    i2ra = _gaussian_csr(rng, pi, pr, 0.3, 0.1, 0.5, dl, TIMESTEP)
    order = rng.permutation(n_ra)
    n_groups = n_ra // group
    rows = [np.zeros(0, np.int32) for _ in range(n_ra)]
    for g in range(n_groups - 1):
        nxt = order[(g + 1) * group:(g + 2) * group]
        for i in order[g * group:(g + 1) * group]:
            rows[i] = np.sort(rng.choice(nxt, size=min(fan, len(nxt)), replace=False)).astype(np.int32)
    rr_ptr = _csr(rows, n_ra)
    rr_tgt = np.concatenate(rows).astype(np.int32)
    E = len(rr_tgt)
    rr_d = rng.uniform(delay[0], delay[1], E)
    net = _finish(rng, n_ra, n_i, ra2i, i2ra, rr_ptr, rr_tgt, rr_d, group, spread_max=spread_max)
    net.train_ids = np.sort(order[:group]).astype(np.int32)
    net.rr_w = np.full(E, w_chain)
    net.act_cnt = np.diff(rr_ptr).astype(np.int32)
    net.act_ids = rr_tgt.copy()
    return net


def c5_networks(seed=500, variants=2, sizes=(1000, 4000, 20000)):
    """Grown chain networks of the served workload: `variants` per size."""
    out = {}
    for k, n_ra in enumerate(sizes):
        out[n_ra] = [grown_chain_network(np.random.default_rng(seed + 10 * k + v), n_ra, int(round(0.275 * n_ra)))
                     for v in range(variants)]
    return out


def c5_requests(nets_by_size, n_requests=24, rate=20.0, trials=50, seed=7,
                mix=((1000, 0.25), (4000, 0.5), (20000, 0.25))):
    """(arrival s, network, key, kick onsets) of the served workload."""
    rng = np.random.default_rng(seed)
    t = 0.0
    sizes = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix])
    out = []
    for r in range(n_requests):
        t += rng.exponential(1.0 / rate)
        n_ra = sizes[rng.choice(len(sizes), p=probs / probs.sum())]
        var = nets_by_size[n_ra]
        net = var[rng.integers(len(var))]
        onsets = 100.0 + rng.normal(0.0, 1.0, trials)
        out.append((t, net, (seed * 1000 + r, 1), onsets))
    return out


CONFIGS = {"c1": c1, "c2": c2, "c3": c3}
