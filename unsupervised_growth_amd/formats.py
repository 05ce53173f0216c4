"""The reference's on-disk formats: readers and writers.

Every file is little-endian int32 / float64, laid out as the reference's writers
lay it out (c++/src/NetworkGrowthSimulator.cpp, "NGS") and as its Python
readers read it (python_scripts/reading.py).  The readers keep reading.py's
names and return tuples; values come back as float64 (reading.py narrows some
to float32).  The host growth driver (include/hvc_driver.h) writes the same
bytes from C++; these writers let Python code produce resumable states and
test the driver's files.

  connections   RA_I / I_RA_connections*.bin: N, per neuron {id, n, n x {tgt i32, w, length, delay}}
                (NGS:18312-18365; reading.py:22-72)
  coordinates   RA_xy*.bin / I_xy.bin: N, dim, MODEL_INTERNEURON_DISTANCE, N x dim f64
                (NGS:18431-18500; reading.py:74-117)
  weights       weights*.bin: N, trial, N x N f64 (NGS:17760-17779; reading.py:134-158)
  synapses      RA_RA_{active,super}_connections*.bin: N, trial, per neuron {n, ids}
                (NGS:17531-17578; reading.py:822-844)
  delays        axonal_delays_*.bin: N, trial, per neuron {n, f64...} (NGS:17733-17758; reading.py:269-296)
  spike times   spike_times_*.bin: trial, TRIAL_DURATION f64, N, per neuron {n, f64...}
                (NGS:17916-18033; reading.py:628-656)
  maturation    maturation_properties*.bin: N, trial, per neuron {mature i32, scale i32, Erest, GCa}
                (NGS:17650-17668; reading.py:160-185)
  indicators    mature*.bin, remodeled_indicators*.bin, replacement_history*.bin: N, trial, N i32
                (NGS:17632-17687, 18035-18049; reading.py:187-245)
  history       activity_history*.bin: N, RATE_WINDOW_LONG, trial, N x W i32 newest first
                (NGS:17711-17731; reading.py:247-267)
  statistics    weight_statistics.bin {trial, mean, std}*, num_synapses.bin {trial, active, super}*
                (NGS:17580-17630, 17885-17914; reading.py:723-748, 796-820)
  training      training_neurons*.bin: N_TR, ids; training_spread.bin: N_TR, f64
                (NGS:18502-18538; reading.py:1021-1053); num_neurons.bin: N_RA, N_I (NGS:18540-18550)
  replaced      replaced_neurons.bin {trial, n, ids}* (NGS:17859-17883; reading.py:750-781)
  jitter        jitter.bin: N, num_test_trials, per neuron 8 f64 (NGS:1641-1782; reading.py:298-342)
"""
import os
from dataclasses import dataclass

import numpy as np

I32 = np.dtype("<i4")
F64 = np.dtype("<f8")


class _Reader:
    def __init__(self, filename):
        with open(filename, "rb") as f:
            self.data = f.read()
        self.pos = 0

    def i32(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.data, I32, k, self.pos)
        self.pos += 4 * k
        return int(a[0]) if n is None else a.astype(np.int32)

    def f64(self, n=None):
        k = 1 if n is None else n
        a = np.frombuffer(self.data, F64, k, self.pos)
        self.pos += 8 * k
        return float(a[0]) if n is None else a.astype(np.float64)

    def left(self):
        return len(self.data) - self.pos


class _Writer:
    def __init__(self, filename, append=False):
        self.f = open(filename, "ab" if append else "wb")

    def i32(self, *v):
        self.f.write(np.asarray(v, I32).tobytes())

    def f64(self, *v):
        self.f.write(np.asarray(v, F64).tobytes())

    def ai32(self, a):
        self.f.write(np.ascontiguousarray(a, I32).tobytes())

    def af64(self, a):
        self.f.write(np.ascontiguousarray(a, F64).tobytes())

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---------------------------------------------------------------------- readers
def read_connections(filename):
    """(N, targets_ID, weights, syn_lengths, axonal_delays), lists per neuron."""
    r = _Reader(filename)
    n = r.i32()
    tg, w, ln, dl = [], [], [], []
    rec = np.dtype([("t", I32), ("w", F64), ("l", F64), ("d", F64)])
    for _ in range(n):
        r.i32()  # neuron id
        k = r.i32()
        a = np.frombuffer(r.data, rec, k, r.pos)
        r.pos += rec.itemsize * k
        tg.append(a["t"].astype(np.int32).tolist())
        w.append(a["w"].astype(np.float64).tolist())
        ln.append(a["l"].astype(np.float64).tolist())
        dl.append(a["d"].astype(np.float64).tolist())
    return n, tg, w, ln, dl


def read_coordinates_full(filename):
    """(N, dim, model_interneuron_distance, coordinates (N, dim))."""
    r = _Reader(filename)
    n, dim = r.i32(), r.i32()
    mid = r.f64()
    return n, dim, mid, r.f64(n * dim).reshape(n, dim)


def read_coordinates(filename):
    """Coordinates (N, dim), as reading.read_coordinates."""
    return read_coordinates_full(filename)[3]


def read_weights(filename):
    r = _Reader(filename)
    n, trial = r.i32(), r.i32()
    return n, trial, r.f64(n * n).reshape(n, n)


def read_synapses(filename):
    """(N, trial_number, targets) -- active or super synapse lists."""
    r = _Reader(filename)
    n, trial = r.i32(), r.i32()
    out = []
    for _ in range(n):
        k = r.i32()
        out.append(r.i32(k).tolist())
    return n, trial, out


def read_axonal_delays(filename):
    r = _Reader(filename)
    n, trial = r.i32(), r.i32()
    out = []
    for _ in range(n):
        k = r.i32()
        out.append(r.f64(k).tolist())
    return n, trial, out


def read_spike_times(filename):
    """(trial_number, trial_duration, per-neuron spike-time arrays for all N neurons)."""
    r = _Reader(filename)
    trial, duration, n = r.i32(), r.f64(), r.i32()
    out = []
    for _ in range(n):
        k = r.i32()
        out.append(r.f64(k))
    return trial, duration, out


def read_time_info(filename):
    """(trial_number, simulation_time, spike_times, neuron_fired) of the neurons that
    fired, as reading.read_time_info."""
    trial, duration, trains = read_spike_times(filename)
    spikes = [t.tolist() for t in trains if len(t)]
    fired = [[i] * len(t) for i, t in enumerate(trains) if len(t)]
    return trial, duration, spikes, fired


def read_maturation_properties(filename):
    """(N, trial_number, mature_indicators, maturation_rate, Erest, GCa)."""
    r = _Reader(filename)
    n, trial = r.i32(), r.i32()
    rec = np.dtype([("m", I32), ("s", I32), ("e", F64), ("g", F64)])
    a = np.frombuffer(r.data, rec, n, r.pos)
    return (n, trial, a["m"].astype(np.int32), a["s"].astype(np.int32), a["e"].astype(np.float64),
            a["g"].astype(np.float64))


def _read_indicators(filename):
    r = _Reader(filename)
    n, trial = r.i32(), r.i32()
    return n, trial, r.i32(n)


read_mature_indicators = _read_indicators
read_remodeled_indicators = _read_indicators
read_replacement_history = _read_indicators


def read_activity_history(filename):
    """(N, trial_number, activity_history (N, RATE_WINDOW_LONG)), newest trial first."""
    r = _Reader(filename)
    n, w, trial = r.i32(), r.i32(), r.i32()
    return n, trial, r.i32(n * w).reshape(n, w)


def read_weight_statistics(filename):
    r = _Reader(filename)
    rec = np.dtype([("t", I32), ("m", F64), ("s", F64)])
    a = np.frombuffer(r.data, rec, r.left() // rec.itemsize)
    return a["t"].tolist(), a["m"].tolist(), a["s"].tolist()


def read_num_synapses(filename):
    r = _Reader(filename)
    a = r.i32(r.left() // 4).reshape(-1, 3)
    return a[:, 0].tolist(), a[:, 1].tolist(), a[:, 2].tolist()


def read_num_neurons(filename):
    r = _Reader(filename)
    return r.i32(), r.i32()


def read_training_neurons(filename):
    r = _Reader(filename)
    return r.i32(r.i32())


def read_training_spread(filename):
    r = _Reader(filename)
    return r.f64(r.i32())


def read_replaced_neurons(filename):
    r = _Reader(filename)
    out = {}
    while r.left() > 0:
        trial, k = r.i32(), r.i32()
        out[trial] = r.i32(k).tolist()
    return out


JITTER_FIELDS = ("probability_soma_spike", "average_num_soma_spikes_in_trial", "mean_first_soma_spike_time",
                 "std_first_soma_spike_time", "probability_dend_spike", "average_num_dend_spikes_in_trial",
                 "mean_first_dend_spike_time", "std_first_dend_spike_time")


def read_jitter(filename):
    """(N, num_test_trials, then the 8 per-neuron arrays of JITTER_FIELDS)."""
    r = _Reader(filename)
    n, trials = r.i32(), r.i32()
    a = r.f64(8 * n).reshape(n, 8)
    return (n, trials) + tuple(a[:, k].copy() for k in range(8))


# ---------------------------------------------------------------------- writers
def write_connections(filename, targets, weights, lengths, delays):
    with _Writer(filename) as w:
        w.i32(len(targets))
        rec = np.dtype([("t", I32), ("w", F64), ("l", F64), ("d", F64)])
        for i, t in enumerate(targets):
            w.i32(i, len(t))
            a = np.zeros(len(t), rec)
            a["t"], a["w"], a["l"], a["d"] = t, weights[i], lengths[i], delays[i]
            w.f.write(a.tobytes())


def write_coordinates(filename, xyz, model_interneuron_distance, dim=3):
    xyz = np.asarray(xyz, np.float64)
    with _Writer(filename) as w:
        w.i32(len(xyz), dim)
        w.f64(model_interneuron_distance)
        w.af64(xyz[:, :dim])


def write_weights(filename, weights, trial):
    weights = np.asarray(weights, np.float64)
    with _Writer(filename) as w:
        w.i32(len(weights), trial)
        w.af64(weights)


def write_synapses(filename, lists, trial):
    with _Writer(filename) as w:
        w.i32(len(lists), trial)
        for t in lists:
            w.i32(len(t))
            w.ai32(t)


def write_axonal_delays(filename, rows, trial):
    with _Writer(filename) as w:
        w.i32(len(rows), trial)
        for r in rows:
            w.i32(len(r))
            w.af64(r)


def write_spike_times(filename, trains, trial, trial_duration):
    with _Writer(filename) as w:
        w.i32(trial)
        w.f64(trial_duration)
        w.i32(len(trains))
        for t in trains:
            w.i32(len(t))
            w.af64(t)


def write_maturation_properties(filename, mature, scale, erest, gca, trial):
    rec = np.dtype([("m", I32), ("s", I32), ("e", F64), ("g", F64)])
    a = np.zeros(len(mature), rec)
    a["m"], a["s"], a["e"], a["g"] = mature, scale, erest, gca
    with _Writer(filename) as w:
        w.i32(len(mature), trial)
        w.f.write(a.tobytes())


def write_indicators(filename, values, trial):
    with _Writer(filename) as w:
        w.i32(len(values), trial)
        w.ai32(values)


def write_activity_history(filename, history, trial):
    history = np.asarray(history, np.int32)
    with _Writer(filename) as w:
        w.i32(history.shape[0], history.shape[1], trial)
        w.ai32(history)


def weight_statistics(weights):
    """Mean and standard deviation of all N x N weights as write_weight_statistics
    accumulates them (NGS:17580-17630): row sums added in turn, squared
    deviations pair by pair, divided by N^2 - 1."""
    weights = np.asarray(weights, np.float64)
    n = len(weights)
    mean = 0.0
    for row in weights:
        s = 0.0
        for x in row:
            s += float(x)
        mean += s
    mean /= float(n) * float(n)
    sq = 0.0
    for row in weights:
        for x in row:
            sq += (float(x) - mean) * (float(x) - mean)
    if n * n > 1:
        sq = float(np.sqrt(sq / (float(n) * float(n) - 1)))
    return mean, sq


def append_weight_statistics(filename, trial, mean, std):
    with _Writer(filename, append=trial != 0) as w:
        w.i32(trial)
        w.f64(mean, std)


def append_num_synapses(filename, trial, active, super_):
    with _Writer(filename, append=trial != 0) as w:
        w.i32(trial, active, super_)


def write_training_neurons(filename, ids):
    with _Writer(filename) as w:
        w.i32(len(ids))
        w.ai32(ids)


def write_training_spread(filename, spread):
    with _Writer(filename) as w:
        w.i32(len(spread))
        w.af64(spread)


def write_num_neurons(filename, n_ra, n_i):
    with _Writer(filename) as w:
        w.i32(n_ra, n_i)


def append_replaced_neurons(filename, trial, ids):
    with _Writer(filename, append=trial != 0) as w:
        w.i32(trial, len(ids))
        w.ai32(ids)


def write_jitter(filename, num_test_trials, columns):
    """columns: the 8 per-neuron arrays of JITTER_FIELDS."""
    a = np.stack([np.asarray(c, np.float64) for c in columns], axis=1)
    with _Writer(filename) as w:
        w.i32(len(a), num_test_trials)
        w.af64(a)


# ---------------------------------------------------------------------- states
@dataclass
class GrowthState:
    """What read_network_state reads to resume growth (NGS:1943-2030)."""
    trial: int
    n_ra: int
    n_i: int
    model_interneuron_distance: float
    ra_xyz: np.ndarray
    i_xyz: np.ndarray
    ra2i: tuple            # (targets, weights, lengths, delays) lists per RA
    i2ra: tuple            # per I
    weights: np.ndarray    # (N, N)
    rr_delays: np.ndarray  # (N, N)
    active: list
    super: list
    mature: np.ndarray
    maturation_scale: np.ndarray
    erest: np.ndarray
    gca: np.ndarray
    history: np.ndarray    # (N, W) newest first
    trials_after_replacement: np.ndarray
    remodeled: np.ndarray
    training: np.ndarray
    spread: np.ndarray


def _name(directory, base, ending=""):
    return os.path.join(directory, base + ending + ".bin")


def read_network_state(directory, trial, file_ending=None):
    """The full state written at trial `trial` (file ending "_<trial>", none for
    trial 0 as in NGS:1947-1950, or `file_ending`), with the initial files of
    the run (NGS:1943-2030)."""
    e = file_ending if file_ending is not None else ("_%d" % trial if trial > 0 else "")
    n_ra, n_i = read_num_neurons(_name(directory, "num_neurons"))
    _, dim, mid, ra = read_coordinates_full(_name(directory, "RA_xy", e))
    _, _, _, ii = read_coordinates_full(_name(directory, "I_xy"))
    pad = lambda a: np.pad(a, ((0, 0), (0, 3 - a.shape[1])))  # noqa: E731
    a = read_connections(_name(directory, "RA_I_connections", e))
    b = read_connections(_name(directory, "I_RA_connections", e))
    _, _, w = read_weights(_name(directory, "weights", e))
    _, _, rrd = read_axonal_delays(_name(directory, "axonal_delays_RA2RA", e))
    _, _, act = read_synapses(_name(directory, "RA_RA_active_connections", e))
    _, _, sup = read_synapses(_name(directory, "RA_RA_super_connections", e))
    _, _, mat, sc, er, gc = read_maturation_properties(_name(directory, "maturation_properties", e))
    _, _, hist = read_activity_history(_name(directory, "activity_history", e))
    _, _, rep = read_replacement_history(_name(directory, "replacement_history", e))
    _, _, rem = read_remodeled_indicators(_name(directory, "remodeled_indicators", e))
    return GrowthState(trial=trial, n_ra=n_ra, n_i=n_i, model_interneuron_distance=mid, ra_xyz=pad(ra),
                       i_xyz=pad(ii), ra2i=a[1:], i2ra=b[1:], weights=w, rr_delays=np.array(rrd, np.float64),
                       active=act, super=sup, mature=mat, maturation_scale=sc, erest=er, gca=gc, history=hist,
                       trials_after_replacement=rep, remodeled=rem,
                       training=np.sort(read_training_neurons(_name(directory, "training_neurons"))),
                       spread=read_training_spread(_name(directory, "training_spread")))


def _csr(lists, dtype):
    ptr = np.zeros(len(lists) + 1, np.int32)
    ptr[1:] = np.cumsum([len(x) for x in lists])
    flat = np.concatenate([np.asarray(x, dtype) for x in lists]) if lists else np.zeros(0, dtype)
    return ptr, flat.astype(dtype) if len(flat) else np.zeros(0, dtype)


def state_to_network(s: GrowthState):
    """Engine network (dense RA->RA candidates) of a read state."""
    from .network import Network
    n = s.n_ra
    ra2i_ptr, ra2i_tgt = _csr(s.ra2i[0], np.int32)
    i2ra_ptr, i2ra_tgt = _csr(s.i2ra[0], np.int32)
    act_cnt, act_ids = _csr(s.active, np.int32)
    sup_cnt, sup_ids = _csr(s.super, np.int32)
    return Network(n_ra=n, n_i=s.n_i,
                   ra2i_ptr=ra2i_ptr, ra2i_tgt=ra2i_tgt, ra2i_w=_csr(s.ra2i[1], np.float64)[1],
                   ra2i_d=_csr(s.ra2i[3], np.float64)[1],
                   i2ra_ptr=i2ra_ptr, i2ra_tgt=i2ra_tgt, i2ra_w=_csr(s.i2ra[1], np.float64)[1],
                   i2ra_d=_csr(s.i2ra[3], np.float64)[1],
                   rr_ptr=(np.arange(n + 1, dtype=np.int64) * n).astype(np.int32),
                   rr_tgt=np.tile(np.arange(n, dtype=np.int32), n),
                   rr_d=np.ascontiguousarray(s.rr_delays, np.float64).reshape(-1),
                   rr_w=np.ascontiguousarray(s.weights, np.float64).reshape(-1),
                   act_cnt=np.diff(act_cnt).astype(np.int32), act_ids=act_ids,
                   sup_cnt=np.diff(sup_cnt).astype(np.int32), sup_ids=sup_ids,
                   remodeled=s.remodeled.astype(np.int32), erest=s.erest, gca=s.gca,
                   train_ids=s.training.astype(np.int32), spread=s.spread)
