"""ora_growth.py -- CPU restatement of the host side of the growth loop.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the driver and the
topology generator (include/hvc_driver.h, include/hvc_topology.h): tests/ may
import it; the product (unsupervised_growth_amd) never imports it.  Plain
Python loops, written from the reference's algorithm, not from the product's
C++ ("NGS" = c++/src/NetworkGrowthSimulator.cpp of the reference).

Parity pins (tests/test_growth_host.py):
  Mt19937         : the C++ standard's known answer (10 000th word of the
                    default-seeded engine is 4123659995), the raw stream of the
                    reference build's std::mt19937 (tests/golden/ref_poisson_noise.npz)
                    and numpy's legacy-seeded MT19937.
  random, sample_integer : the reference's own Poisson_noise compiled into
                    oracle/_ref (tests/golden/ref_host_rng.npz).
  distances       : closed forms (orthogonal, equal, antipodal points).
  topology        : invariants (minimum distance, connection rule statistics)
                    and the reference seed's draw order; the product's C++ must
                    equal it bit for bit.
  activity_update : closed forms (driven at exactly 500 of the last 1000 trials,
                    silent below the death threshold only after 1000 trials).
  maturation_update : the closed form x_k = x_m + (x_0 - x_m) e^(-k/s) within
                    rounding, and the step count to maturity.
  replace_neurons : invariants (no edge touches a replaced neuron except the
                    resampled ones, lists keep their order) and the product's
                    files after a replacement, bit for bit.
"""
import math
from collections import deque

# ---------------------------------------------------------------- generator
class Mt19937:
    """std::mt19937 (Matsumoto and Nishimura 1998; C++ [rand.eng.mers],
    [rand.predef]): 624 words, init by x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i."""
    N, M = 624, 397

    def __init__(self, seed=5489):
        mt = [seed & 0xffffffff]
        for i in range(1, self.N):
            mt.append((1812433253 * (mt[-1] ^ (mt[-1] >> 30)) + i) & 0xffffffff)
        self.mt = mt
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % self.N] & 0x7fffffff)
            mt[i] = mt[(i + self.M) % self.N] ^ (y >> 1) ^ (0x9908b0df if y & 1 else 0)
        self.idx = 0

    def __call__(self):
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9d2c5680
        y ^= (y << 15) & 0xefc60000
        y ^= y >> 18
        return y & 0xffffffff

    # Poisson_noise::random (c++/src/poisson_noise.cpp:17-20): (double)G * x / max
    def random(self, G):
        return float(G) * float(self()) / 4294967295.0

    # Poisson_noise::sample_integer (c++/include/poisson_noise.h:26): libstdc++'s
    # uniform_int_distribution for a 32-bit engine, Lemire's nearly divisionless
    # method (D. Lemire, ACM TOMACS 29(1), 2019)
    def sample_integer(self, lo, hi):
        r = hi - lo + 1
        prod = self() * r
        low = prod & 0xffffffff
        if low < r:
            threshold = ((1 << 32) - r) % r
            while low < threshold:
                prod = self() * r
                low = prod & 0xffffffff
        return lo + (prod >> 32)


# ---------------------------------------------------------------- geometry
def distance_on_sphere(a, b):
    """NGS:36-44: acos of the dot product, 0 when it rounds above 1."""
    c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    return 0.0 if c > 1.0 else math.acos(c)


def distance2d(a, b):
    """NGS:31-34."""
    return math.sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]))


EXPERIMENTAL_INTERNEURON_DISTANCE = 40.0   # NGS:19
MIN_INTERSOMATIC_DISTANCE = 10.0           # NGS:20
AXONAL_DELAY_PER_MICRON = 0.01             # NGS:25
PI = 3.14159265358979323846
SPONTANEOUS, DRIVEN = 100000, 1000         # NGS:23-24
WAITING_TIME = 100.0                       # NGS:14


def gaussian_p(peak, sigma, d):
    """p_RA2I / p_I2RA (NGS:707-719)."""
    sigma_square = sigma * sigma
    return peak * math.exp(-d * d / (2 * sigma_square))


# ---------------------------------------------------------------- topology
def generate_topology(n_ra, n_i, n_tr, p, seed):
    """generate_network_topology (NGS:110-155).  p: dict with arrangement
    ('sphere' / 'square'), A_RA2I, SIGMA_RA2I, B_I2RA, SIGMA_I2RA.  Returns a
    dict of plain lists."""
    g = Mt19937(seed)
    ns = 0.15
    ra = [[0.0, 0.0, 0.0] for _ in range(n_ra)]
    ii = [[0.0, 0.0, 0.0] for _ in range(n_i)]
    sphere = p["arrangement"] == "sphere"
    if sphere:                                      # NGS:294-401
        dphi = PI * (3.0 - math.sqrt(5.0))
        dz = 2 / float(n_i)
        phi = 0.0
        z = 1 - dz / 2.0
        x1, y1 = math.cos(phi) * math.sqrt(1.0 - z * z), math.sin(phi) * math.sqrt(1.0 - z * z)
        x2 = math.cos(phi + dphi) * math.sqrt(1.0 - (z - dz) * (z - dz))
        y2 = math.sin(phi + dphi) * math.sqrt(1.0 - (z - dz) * (z - dz))
        mid = distance_on_sphere((x1, y1, z), (x2, y2, z - dz))
        close_d = mid * MIN_INTERSOMATIC_DISTANCE / EXPERIMENTAL_INTERNEURON_DISTANCE
        for i in range(n_i):
            while True:
                theta = math.acos(z)
                tn = -ns * mid + g.random(2 * ns * mid)
                pn = (-ns * mid + g.random(2 * ns * mid)) / math.sin(theta)
                c = (math.cos(phi + pn) * math.sin(theta + tn), math.sin(phi + pn) * math.sin(theta + tn),
                     math.cos(theta + tn))
                if not any(distance_on_sphere(c, ii[j]) < close_d for j in range(i)):
                    break
            ii[i] = list(c)
            z = z - dz
            phi = phi + dphi
        for i in range(n_ra):
            while True:
                z = 2 * g.random(1.0) - 1
                phi = g.random(2.0 * PI)
                c = (math.sin(math.acos(z)) * math.cos(phi), math.sin(math.acos(z)) * math.sin(phi), z)
                if any(distance_on_sphere(c, ii[j]) < close_d for j in range(n_i)):
                    continue
                if any(distance_on_sphere(c, ra[j]) < close_d for j in range(i)):
                    continue
                break
            ra[i] = list(c)
        dist = distance_on_sphere
    else:                                           # NGS:228-292
        mid = 1 / (math.sqrt(n_i) + 1)
        close_d = mid * MIN_INTERSOMATIC_DISTANCE / EXPERIMENTAL_INTERNEURON_DISTANCE
        side = int(math.sqrt(n_i))
        for i in range(side):
            for k in range(side):
                ii[i * side + k][0] = float(i + 1) / (math.sqrt(n_i) + 1) - ns * mid + g.random(2 * ns * mid)
                ii[i * side + k][1] = float(k + 1) / (math.sqrt(n_i) + 1) - ns * mid + g.random(2 * ns * mid)
        for i in range(n_ra):
            while True:
                x = 0.5 / (math.sqrt(n_i) + 1) + g.random(1.0 - 1.0 / (math.sqrt(n_i) + 1))
                y = 0.5 / (math.sqrt(n_i) + 1) + g.random(1.0 - 1.0 / (math.sqrt(n_i) + 1))
                c = (x, y)
                if any(distance2d(c, ii[j]) < close_d for j in range(n_i)):
                    continue
                if any(distance2d(c, ra[j]) < close_d for j in range(i)):
                    continue
                break
            ra[i][0], ra[i][1] = x, y
        dist = distance2d
    # sample_connections (NGS:620-705): weights sample_G(0.0), delays 0
    a = {"tgt": [[] for _ in range(n_ra)], "len": [[] for _ in range(n_ra)]}
    for i in range(n_ra):
        for j in range(n_i):
            d = dist(ra[i], ii[j]) * EXPERIMENTAL_INTERNEURON_DISTANCE / mid
            if g.random(1) < gaussian_p(p["A_RA2I"], p["SIGMA_RA2I"], d):
                g.random(0.0)
                a["tgt"][i].append(j)
                a["len"][i].append(d)
    b = {"tgt": [[] for _ in range(n_i)], "len": [[] for _ in range(n_i)]}
    for i in range(n_i):
        for j in range(n_ra):
            d = dist(ii[i], ra[j]) * EXPERIMENTAL_INTERNEURON_DISTANCE / mid
            if g.random(1) < gaussian_p(p["B_I2RA"], p["SIGMA_I2RA"], d):
                g.random(0.0)
                b["tgt"][i].append(j)
                b["len"][i].append(d)
    # write_different_training_arrangements (NGS:406-455)
    clustered, rnd = [], []
    if n_tr > 0:
        dd = [dist(ra[0], ra[i]) for i in range(1, n_ra)]
        order = sorted(range(len(dd)), key=lambda k: dd[k])
        clustered = [0] + [order[k] + 1 for k in range(n_tr - 1)]
        pool = list(range(n_ra))
        for i in range(n_tr):
            r = g.sample_integer(0, n_ra - i - 1)
            rnd.append(pool[r])
            pool[r] = pool[-1]
            pool.pop()
    return dict(mid=mid, ra=ra, i=ii, ra2i=a, i2ra=b, clustered=clustered, random=rnd)


# ---------------------------------------------------------------- bookkeeping
def activity_update(history, counts, trials_after, scale, death_rate_threshold, window=1000):
    """check_neuron_activity (NGS:9573-9627) with its window of 1000 trials
    (NGS:9586) as a parameter.  history: list of deques (newest first, maxlen
    RATE_WINDOW_LONG, kept full as boost::circular_buffer is); push_front drops
    the oldest.  Returns (to_replace, newly driven)."""
    to_replace, driven = [], 0
    for i in range(len(counts)):
        history[i].appendleft(int(counts[i]))
        robustness = sum(1.0 for j in range(window) if history[i][j] > 0) / float(window)
        if robustness >= 0.5 and scale[i] != DRIVEN:
            scale[i] = DRIVEN
            driven += 1
        if robustness < death_rate_threshold and trials_after[i] > window:
            to_replace.append(i)
    return to_replace, driven


def maturation_update(m, scale, mature, erest, gca):
    """update_neuron_properties_sameDendrite_diffMaturationRate (NGS:2152-2183)."""
    k = 0
    for i in range(len(mature)):
        if mature[i] == 1:
            continue
        erest[i] = erest[i] + (erest[i] - m["E_REST_MATURE"]) * (math.exp(-1.0 / scale[i]) - 1)
        gca[i] = gca[i] + (gca[i] - m["GCA_MATURE"]) * (math.exp(-1.0 / scale[i]) - 1)
        if gca[i] >= m["GCA_MATURE"] - 0.5 or erest[i] <= m["E_REST_MATURE"] - 0.5:
            gca[i] = m["GCA_MATURE"]
            erest[i] = m["E_REST_MATURE"]
            mature[i] = 1
            k += 1
    return k


def replace_neurons(s, R, g, topo, mat):
    """replace_neurons (NGS:15858-15870) on a state s with dense RA->RA arrays:
    s['w'][i][j], s['rr_d'][i][j], s['act'][i], s['sup'][i], s['remodeled'],
    s['a'] / s['b'] = {'tgt','w','d'} lists per RA / per I, s['ra'], s['i'],
    s['mid'], maturation lists and s['history'] (deques).  g: the generator;
    topo: dict of A_RA2I ... delay_constant."""
    n_ra, n_i = len(s["ra"]), len(s["i"])
    # remove_neurons_from_network (NGS:15757-15847)
    for i in range(n_ra):
        for r in R:
            if r in s["act"][i]:
                s["act"][i].remove(r)
            if r in s["sup"][i]:
                s["sup"][i].remove(r)
                s["remodeled"][i] = 0
            s["w"][i][r] = 0.0
    for r in R:
        s["act"][r] = []
        s["sup"][r] = []
        s["w"][r] = [0.0] * n_ra
    for i in range(n_i):
        for r in R:
            if r in s["b"]["tgt"][i]:
                k = s["b"]["tgt"][i].index(r)
                for key in ("tgt", "w", "d"):
                    del s["b"][key][i][k]
    for r in R:
        for key in ("tgt", "w", "d"):
            s["a"][key][r] = []
    # sample_coordinates_for_replaced (NGS:157-220)
    fixed = [i for i in range(n_ra) if i not in set(R)]
    close_d = s["mid"] * MIN_INTERSOMATIC_DISTANCE / EXPERIMENTAL_INTERNEURON_DISTANCE
    for r in R:
        while True:
            z = 2 * g.random(1.0) - 1
            phi = g.random(2.0 * PI)
            c = (math.sin(math.acos(z)) * math.cos(phi), math.sin(math.acos(z)) * math.sin(phi), z)
            if any(distance_on_sphere(c, s["i"][j]) < close_d for j in range(n_i)):
                continue
            if any(distance_on_sphere(c, s["ra"][j]) < close_d for j in fixed):
                continue
            break
        s["ra"][r] = list(c)
        fixed.append(r)
    # sample_connectionsAndDelays_for_replaced (NGS:528-618)
    for r in R:
        for j in range(n_i):
            d = distance_on_sphere(s["ra"][r], s["i"][j]) * EXPERIMENTAL_INTERNEURON_DISTANCE / s["mid"]
            if g.random(1) < gaussian_p(topo["A_RA2I"], topo["SIGMA_RA2I"], d):
                s["a"]["w"][r].append(g.random(topo["Gei_max"]))
                s["a"]["tgt"][r].append(j)
                s["a"]["d"][r].append(d * AXONAL_DELAY_PER_MICRON * topo["delay_constant"])
    for i in range(n_i):
        for r in R:
            d = distance_on_sphere(s["i"][i], s["ra"][r]) * EXPERIMENTAL_INTERNEURON_DISTANCE / s["mid"]
            if g.random(1) < gaussian_p(topo["B_I2RA"], topo["SIGMA_I2RA"], d):
                s["b"]["w"][i].append(g.random(topo["Gie_max"]))
                s["b"]["tgt"][i].append(r)
                s["b"]["d"][i].append(d * AXONAL_DELAY_PER_MICRON * topo["delay_constant"])
    for i in range(n_ra):
        for r in R:
            ln = distance_on_sphere(s["ra"][i], s["ra"][r]) * EXPERIMENTAL_INTERNEURON_DISTANCE / s["mid"] \
                if r != i else 0.0
            s["rr_d"][i][r] = ln * AXONAL_DELAY_PER_MICRON * topo["delay_constant"]
    for r in R:
        for j in range(n_ra):
            ln = distance_on_sphere(s["ra"][r], s["ra"][j]) * EXPERIMENTAL_INTERNEURON_DISTANCE / s["mid"] \
                if j != r else 0.0
            s["rr_d"][r][j] = ln * AXONAL_DELAY_PER_MICRON * topo["delay_constant"]
    # update_replaced_neurons (NGS:2109-2133)
    for r in R:
        s["erest"][r] = mat["E_REST_IMMATURE"]
        s["gca"][r] = mat["GCA_IMMATURE"]
        s["mature"][r] = 0
        s["scale"][r] = SPONTANEOUS
        s["remodeled"][r] = 0
        s["trials_after"][r] = 0
        for _ in range(len(s["history"][r])):
            s["history"][r].append(0)


# ---------------------------------------------------------------- chain test
def chain_statistics(trials):
    """calculate_and_write_jitter (NGS:1641-1782) over trials given as (soma,
    dend) lists of per-neuron spike-time lists, in trial order.  Returns a list
    of 8-tuples per neuron in jitter.bin's order (write_jitter, NGS:17818-17857)."""
    n = len(trials[0][0])
    T = float(len(trials))
    out = []
    for i in range(n):
        row = []
        for kind in (0, 1):
            first = [tr[kind][i][0] for tr in trials if len(tr[kind][i]) >= 1]
            counts = [len(tr[kind][i]) for tr in trials]
            total = 0.0
            for c in counts:
                total += c
            mean = -1
            if first:
                s = 0.0
                for x in first:
                    s += x
                mean = s / float(len(first))
            acc = 0
            for x in first:
                acc += (x - mean) * (x - mean)
            std = math.sqrt(acc / (float(len(first)) - 1)) if len(first) > 1 else -1
            row += [float(len(first)) / T, total / T, float(mean), float(std)]
        out.append(tuple(row))
    return out
