"""Growth runs as the reference sets them up, on the C++ driver.

* ``write_network_topology`` / ``read_network_topology``: the files of
  generate_network_topology (c++/src/NetworkGrowthSimulator.cpp:110-155, "NGS")
  and read_network_topology (NGS:2628-2650).
* ``new_growth``: new_chain_growth (NGS:1865-1930) -- a generated topology gets
  its delays and weights, the training neurons are made mature, the initial
  files and both configuration files are written, and a GrowthDriver is
  returned ready to run chain_growth's loop (``driver.run``).
* ``continue_growth``: continue_chain_growth (NGS:2035-2058) -- resume from the
  state files written at a trial.
* ``test_chain``: test_chain (NGS:1313-1482) on a saved state.

The driver draws from its own std::mt19937 (seed ``seed``); the reference's
single generator also feeds the neuron noise, so its later draws differ.
"""
import os

import numpy as np

from . import formats as F
from .config import (ConnectionParams, GrowthConfig, read_growth_configuration, read_topology_configuration,
                     write_growth_configuration, write_topology_configuration)
from .driver import GrowthDriver
from .engine import make_params
from .network import TRIAL_DURATION
from .topology import REFERENCE_SEED, Topology, TopologyParams, growth_network, prepare


def _p(directory, name):
    return os.path.join(directory, name)


def write_pajek(path, t: Topology, training, arrangement="sphere"):
    """write_pajek_topology (NGS:18367-18429): vertices (training neurons Yellow,
    other HVC-RA Green, HVC-I Red) and weighted arcs, numbers as std::ostream
    prints them (%g)."""
    tr = set(int(x) for x in training)
    dim = 3 if arrangement == "sphere" else 2
    lines = ["*Vertices %d" % (t.n_ra + t.n_i)]
    for i in range(t.n_ra):
        xyz = " ".join("%g" % v for v in t.ra_xyz[i, :dim])
        lines.append('%d "%d" %s ic %s' % (i + 1, i, xyz, "Yellow" if i in tr else "Green"))
    for i in range(t.n_i):
        xyz = " ".join("%g" % v for v in t.i_xyz[i, :dim])
        lines.append('%d "%d" %s ic Red' % (i + t.n_ra + 1, i + t.n_ra, xyz))
    lines.append("*Arcs")
    for i in range(t.n_ra):
        for e in range(t.ra2i_ptr[i], t.ra2i_ptr[i + 1]):
            lines.append("%d %d %g c Green" % (i + 1, t.ra2i_tgt[e] + t.n_ra + 1, t.ra2i_w[e]))
    for i in range(t.n_i):
        for e in range(t.i2ra_ptr[i], t.i2ra_ptr[i + 1]):
            lines.append("%d %d %g c Red" % (i + t.n_ra + 1, t.i2ra_tgt[e] + 1, t.i2ra_w[e]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _lists(ptr, *arrays):
    return [[a[ptr[i]:ptr[i + 1]].tolist() for i in range(len(ptr) - 1)] for a in arrays]


def write_network_topology(directory, t: Topology, params: TopologyParams):
    """The files generate_network_topology writes (NGS:141-154), plus the
    network_parameters.cfg read_network_topology expects beside them (NGS:2638)."""
    dim = t.dim
    F.write_training_neurons(_p(directory, "training_neurons_clustered.bin"), t.training_clustered)
    F.write_training_neurons(_p(directory, "training_neurons_random.bin"), t.training_random)
    F.write_coordinates(_p(directory, "I_xy.bin"), t.i_xyz, t.model_interneuron_distance, dim)
    F.write_coordinates(_p(directory, "RA_xy.bin"), t.ra_xyz, t.model_interneuron_distance, dim)
    F.write_connections(_p(directory, "RA_I_connections.bin"), *_lists(t.ra2i_ptr, t.ra2i_tgt, t.ra2i_w, t.ra2i_len,
                                                                       t.ra2i_d))
    F.write_connections(_p(directory, "I_RA_connections.bin"), *_lists(t.i2ra_ptr, t.i2ra_tgt, t.i2ra_w, t.i2ra_len,
                                                                       t.i2ra_d))
    F.write_num_neurons(_p(directory, "num_neurons.bin"), t.n_ra, t.n_i)
    write_pajek(_p(directory, "network_topology.net"), t, t.training_random, params.arrangement)
    write_topology_configuration(_p(directory, "network_parameters.cfg"), params)


def _csr(lists, dtype):
    ptr = np.zeros(len(lists) + 1, np.int32)
    ptr[1:] = np.cumsum([len(x) for x in lists])
    flat = [x for row in lists for x in row]
    return ptr, np.array(flat, dtype)


def read_network_topology(directory, connection: ConnectionParams = None):
    """read_network_topology (NGS:2628-2650) -> (Topology, TopologyParams)."""
    params = read_topology_configuration(_p(directory, "network_parameters.cfg"), connection)
    n_ra, n_i = F.read_num_neurons(_p(directory, "num_neurons.bin"))
    _, dim, mid, ra = F.read_coordinates_full(_p(directory, "RA_xy.bin"))
    _, _, _, ii = F.read_coordinates_full(_p(directory, "I_xy.bin"))
    pad = lambda a: np.pad(a, ((0, 0), (0, 3 - a.shape[1])))  # noqa: E731
    _, at, aw, al, ad = F.read_connections(_p(directory, "RA_I_connections.bin"))
    _, bt, bw, bl, bd = F.read_connections(_p(directory, "I_RA_connections.bin"))
    ap, atg = _csr(at, np.int32)
    bp, btg = _csr(bt, np.int32)

    def flat(x):
        return np.array([v for row in x for v in row], np.float64)

    tr_c = _p(directory, "training_neurons_clustered.bin")
    tr_r = _p(directory, "training_neurons_random.bin")
    t = Topology(n_ra=n_ra, n_i=n_i, dim=dim, model_interneuron_distance=mid, ra_xyz=pad(ra), i_xyz=pad(ii),
                 ra2i_ptr=ap, ra2i_tgt=atg, ra2i_w=flat(aw), ra2i_len=flat(al), ra2i_d=flat(ad),
                 i2ra_ptr=bp, i2ra_tgt=btg, i2ra_w=flat(bw), i2ra_len=flat(bl), i2ra_d=flat(bd),
                 training_clustered=F.read_training_neurons(tr_c) if os.path.exists(tr_c) else np.zeros(0, np.int32),
                 training_random=F.read_training_neurons(tr_r) if os.path.exists(tr_r) else np.zeros(0, np.int32))
    return t, params


def engine_params(cfg: GrowthConfig, trial_ms=TRIAL_DURATION, **kw):
    return make_params(noise=cfg.noise, syn=cfg.synaptic, trial_ms=trial_ms, **kw)


def new_growth(network_dir, training_file, output_dir, cfg: GrowthConfig, trial_ms=TRIAL_DURATION,
               seed=REFERENCE_SEED, save_freq_short=0, save_freq_long=0, device=0, **engine_kw):
    """new_chain_growth (NGS:1865-1930) on the driver; returns the GrowthDriver."""
    t, params = read_network_topology(network_dir, cfg.connection)
    write_topology_configuration(_p(output_dir, "network_parameters.cfg"), params)
    t, rr = prepare(t, params, seed=seed)
    training = np.sort(F.read_training_neurons(training_file))
    net = growth_network(t, rr, training, cfg.maturation, spread=np.zeros(len(training)))
    drv = GrowthDriver(net, t.ra_xyz, t.i_xyz, t.model_interneuron_distance, params, cfg.maturation, seed=seed,
                       params=engine_params(cfg, trial_ms, **engine_kw), device=device)
    drv.write_state(output_dir, what="initial")
    write_growth_configuration(_p(output_dir, "growth_parameters.cfg"), cfg)
    drv.set_output(output_dir, save_freq_short, save_freq_long)
    return drv


def _resume(data_dir, starting_trial, seed, trial_ms, device, engine_kw):
    cfg = read_growth_configuration(_p(data_dir, "growth_parameters.cfg"))
    params = read_topology_configuration(_p(data_dir, "network_parameters.cfg"), cfg.connection)
    st = F.read_network_state(data_dir, starting_trial)
    net = F.state_to_network(st)
    resume = dict(mature=st.mature, maturation_scale=st.maturation_scale,
                  trials_after_replacement=st.trials_after_replacement, history=st.history,
                  trial_number=starting_trial)
    if st.history.shape[1] != cfg.maturation.RATE_WINDOW_LONG:
        raise ValueError("activity history window %d differs from RATE_WINDOW_LONG %d"
                         % (st.history.shape[1], cfg.maturation.RATE_WINDOW_LONG))
    drv = GrowthDriver(net, st.ra_xyz, st.i_xyz, st.model_interneuron_distance, params, cfg.maturation,
                       resume=resume, seed=seed, params=engine_params(cfg, trial_ms, **engine_kw), device=device)
    return drv, cfg


def continue_growth(data_dir, starting_trial, output_dir, trial_ms=TRIAL_DURATION, seed=REFERENCE_SEED,
                    save_freq_short=0, save_freq_long=0, device=0, **engine_kw):
    """continue_chain_growth (NGS:2035-2058): the driver resumes from the state
    files of `starting_trial` and repeats that trial number first, as the
    reference does."""
    drv, _ = _resume(data_dir, starting_trial, seed, trial_ms, device, engine_kw)
    drv.set_output(output_dir, save_freq_short, save_freq_long)
    return drv


def test_chain(data_dir, starting_trial, num_trials, output_dir, trial_ms=TRIAL_DURATION, seed=REFERENCE_SEED,
               device=0, **engine_kw):
    """test_chain (NGS:1313-1482): chain-test trials on a saved state; writes
    test_spike_times_*_<k>.bin and jitter.bin and returns the statistics."""
    drv, _ = _resume(data_dir, starting_trial, seed, trial_ms, device, engine_kw)
    with drv:
        return drv.chain_test(num_trials, output_dir)
