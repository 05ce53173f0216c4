"""The host growth driver on the device: one growth trial, then the reference's
bookkeeping (check_neuron_activity, maturation, replace_neurons) in C++.

The driver's state files at the trial (written where chain_growth writes them,
before the bookkeeping) are the oracle's input; the oracle
(oracle/ora_growth.py) applies the bookkeeping and the replacement with the
same generator draws, and the driver's state after the trial must equal it bit
for bit.  The next trial, run on the driver's rebuilt (or patched) engine, must
equal a fresh engine loaded from those files.
"""
import os
import sys
from collections import deque

import numpy as np
import pytest

from oracle_lib import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ora_growth as og  # noqa: E402

pytestmark = pytest.mark.gpu

TOPO = dict(arrangement="sphere", A_RA2I=0.35, SIGMA_RA2I=60.0, B_I2RA=0.3, SIGMA_I2RA=70.0, Gei_max=0.25,
            Gie_max=0.5, delay_constant=2.0)
TRIAL_MS = 300.0
SEED = 15000


def _setup(n_ra=100, n_i=30, n_tr=8):
    from unsupervised_growth_amd.network import MaturationParams
    from unsupervised_growth_amd.topology import TopologyParams, generate, growth_network, prepare
    p = TopologyParams(**TOPO)
    t = generate(n_ra, n_i, n_tr, p, seed=SEED)
    t, rr = prepare(t, p, seed=SEED)
    mat = MaturationParams(E_REST_IMMATURE=-60.0, GCA_IMMATURE=10.0)
    net = growth_network(t, rr, t.training_random, mat)
    # some synapses already active / super, so that removal has work to do
    rng = np.random.default_rng(5)
    n = n_ra
    act = [[] for _ in range(n)]
    sup = [[] for _ in range(n)]
    w = net.rr_w.reshape(n, n)
    for i in range(n):
        for j in rng.choice(n, 3, replace=False):
            if j == i:
                continue
            if rng.random() < 0.3 and len(sup[i]) < 2:
                w[i, j] = 2.5
                act[i].append(int(j))
                sup[i].append(int(j))
            else:
                w[i, j] = 0.8
                act[i].append(int(j))
    net.act_cnt = np.array([len(a) for a in act], np.int32)
    net.act_ids = np.array([x for a in act for x in a], np.int32)
    net.sup_cnt = np.array([len(s) for s in sup], np.int32)
    net.sup_ids = np.array([x for s in sup for x in s], np.int32)
    return t, net, p, mat


def _make(net, t, p, mat, resume):
    from unsupervised_growth_amd.driver import GrowthDriver
    from unsupervised_growth_amd.engine import make_params
    params = make_params(precision=0, trial_ms=TRIAL_MS, seed=(7, 0))
    return GrowthDriver(net, t.ra_xyz, t.i_xyz, t.model_interneuron_distance, p, mat, resume=resume, seed=SEED,
                        params=params)


def _oracle_state(d, ending):
    """Files written by the driver -> the oracle's dense state."""
    from unsupervised_growth_amd import formats as F
    _, _, mid, ra = F.read_coordinates_full(os.path.join(d, "RA_xy%s.bin" % ending))
    _, _, _, ii = F.read_coordinates_full(os.path.join(d, "I_xy.bin"))
    _, at, aw, _, ad = F.read_connections(os.path.join(d, "RA_I_connections%s.bin" % ending))
    _, bt, bw, _, bd = F.read_connections(os.path.join(d, "I_RA_connections%s.bin" % ending))
    _, _, w = F.read_weights(os.path.join(d, "weights%s.bin" % ending))
    _, _, rrd = F.read_axonal_delays(os.path.join(d, "axonal_delays_RA2RA%s.bin" % ending))
    _, _, act = F.read_synapses(os.path.join(d, "RA_RA_active_connections%s.bin" % ending))
    _, _, sup = F.read_synapses(os.path.join(d, "RA_RA_super_connections%s.bin" % ending))
    _, _, m, sc, er, gc = F.read_maturation_properties(os.path.join(d, "maturation_properties%s.bin" % ending))
    _, _, h = F.read_activity_history(os.path.join(d, "activity_history%s.bin" % ending))
    _, _, rep = F.read_replacement_history(os.path.join(d, "replacement_history%s.bin" % ending))
    _, _, rem = F.read_remodeled_indicators(os.path.join(d, "remodeled_indicators%s.bin" % ending))
    return dict(mid=mid, ra=ra.tolist(), i=ii.tolist(), a=dict(tgt=at, w=aw, d=ad), b=dict(tgt=bt, w=bw, d=bd),
                w=w.tolist(), rr_d=rrd, act=act, sup=sup, mature=m.tolist(), scale=sc.tolist(), erest=er.tolist(),
                gca=gc.tolist(), history=[deque(x.tolist(), maxlen=h.shape[1]) for x in h],
                trials_after=rep.tolist(), remodeled=rem.tolist())


def _assert_state_equal(d, ending, s):
    got = _oracle_state(d, ending)
    for k in ("mid", "ra", "i", "w", "rr_d", "act", "sup", "mature", "scale", "erest", "gca", "trials_after",
              "remodeled"):
        assert got[k] == s[k], k
    for side in ("a", "b"):
        for k in ("tgt", "w", "d"):
            assert got[side][k] == s[side][k], (side, k)
    assert [list(x) for x in got["history"]] == [list(x) for x in s["history"]]


def _mat_dict(mat):
    return dict(E_REST_MATURE=mat.E_REST_MATURE, E_REST_IMMATURE=mat.E_REST_IMMATURE, GCA_MATURE=mat.GCA_MATURE,
                GCA_IMMATURE=mat.GCA_IMMATURE)


def test_replacement_equals_oracle_and_rebuilt_engine(tmp_path):
    from unsupervised_growth_amd import formats as F
    from unsupervised_growth_amd.engine import Engine, make_params
    t, net, p, mat = _setup()
    n = net.n_ra
    ids = np.arange(n)
    hist = np.zeros((n, mat.RATE_WINDOW_LONG), np.int32)
    hist[ids % 2 == 0, 3:1000:50] = 1                        # 20 of the last 1000 trials: not silent
    after = np.where(ids % 3 == 0, 900, 1500).astype(np.int32)  # 901 after this trial: too young to die
    scale = np.where(ids % 5 == 0, 1000, 100000).astype(np.int32)
    resume = dict(mature=np.isin(ids, net.train_ids).astype(np.int32), maturation_scale=scale,
                  trials_after_replacement=after, history=hist, trial_number=0)
    d = str(tmp_path)
    print("driver replacement: n_ra=%d n_i=%d" % (n, net.n_i), flush=True)
    with _make(net, t, p, mat, resume) as drv:
        drv.write_state(d, what="initial")
        drv.set_output(d, save_freq_short=1, save_freq_long=1)
        r0 = drv.trial(training=True)
        drv.write_state(d, "_after", "full")
        r1 = drv.trial(training=True)
        replaced_ids, replaced_trials = drv.replaced()
    g = og.Mt19937(SEED)
    assert r0.mean_injection_time == og.WAITING_TIME + g.random(TRIAL_MS - 2 * og.WAITING_TIME)
    assert r0.saved == 3 and r0.soma_spikes > 0
    # the oracle's bookkeeping from the files written at trial 0
    s = _oracle_state(d, "_0")
    _, _, soma = F.read_spike_times(os.path.join(d, "spike_times_soma_0.bin"))
    counts = [len(x) for x in soma]
    s["trials_after"] = [a + 1 for a in s["trials_after"]]
    R, driven = og.activity_update(s["history"], counts, s["trials_after"], s["scale"], mat.DEATH_RATE_THRESHOLD)
    matured = og.maturation_update(_mat_dict(mat), s["scale"], s["mature"], s["erest"], s["gca"])
    og.replace_neurons(s, R, g, TOPO, _mat_dict(mat))
    assert r0.replaced == len(R) > 10 and r0.became_driven == driven and r0.matured == matured
    assert sorted(R) == R and all(i % 2 == 1 and i % 3 != 0 for i in R)
    assert replaced_ids[:len(R)].tolist() == R and (replaced_trials[:len(R)] == 0).all()
    _assert_state_equal(d, "_after", s)
    # the graph state and the statistics files
    assert F.read_num_synapses(os.path.join(d, "num_synapses.bin"))[0] == [0, 1]
    # appended by the full state at trial 0, the explicit write and trial 1
    assert F.read_weight_statistics(os.path.join(d, "weight_statistics.bin"))[0] == [0, 1, 1]
    _, _, w0 = F.read_weights(os.path.join(d, "weights_0.bin"))
    assert F.read_weight_statistics(os.path.join(d, "weight_statistics.bin"))[1][0] == F.weight_statistics(w0)[0]
    # trial 1 ran on the rebuilt engine: a fresh engine loaded from the files gives the same spikes
    st = F.read_network_state(d, 1, file_ending="_after")
    with Engine(make_params(precision=0, trial_ms=TRIAL_MS, seed=(7, 0))) as e:
        e.upload(F.state_to_network(st))
        res = e.run_trial(mode=0, training=True, trial_index=1, first_trial=False,
                          mean_injection_time=r1.mean_injection_time)
    for kind, fname in (("soma", "soma"), ("dend", "dend"), ("interneuron", "interneuron")):
        _, _, trains = F.read_spike_times(os.path.join(d, "spike_times_%s_1.bin" % fname))
        for a, b in zip(getattr(res, kind), trains):
            np.testing.assert_array_equal(a, b)


def test_maturation_patch_and_resume(tmp_path):
    """Without replacement the driver patches Erest / GCa into the engine; the
    next trial equals a fresh engine with the updated properties.  A driver
    resumed from the files writes the same files back."""
    from unsupervised_growth_amd import formats as F
    from unsupervised_growth_amd.engine import Engine, make_params
    t, net, p, mat = _setup(n_ra=80, n_i=24)
    d = str(tmp_path)
    print("driver maturation: n_ra=%d n_i=%d" % (net.n_ra, net.n_i), flush=True)
    with _make(net, t, p, mat, None) as drv:
        drv.write_state(d, what="initial")
        drv.set_output(d, 0, 1)
        r = drv.run(2)
        assert drv.trial_number == 2 and all(x.replaced == 0 for x in r)
    s0 = _oracle_state(d, "_0")
    _, _, soma = F.read_spike_times(os.path.join(d, "spike_times_soma_0.bin"))
    s0["trials_after"] = [a + 1 for a in s0["trials_after"]]
    og.activity_update(s0["history"], [len(x) for x in soma], s0["trials_after"], s0["scale"], 0.01)
    og.maturation_update(_mat_dict(mat), s0["scale"], s0["mature"], s0["erest"], s0["gca"])
    s1 = _oracle_state(d, "_1")
    for k in ("mature", "scale", "erest", "gca", "trials_after"):
        assert s1[k] == s0[k], k
    assert [list(x) for x in s1["history"]] == [list(x) for x in s0["history"]]
    assert s1["erest"] != _oracle_state(d, "_0")["erest"]          # immature neurons moved
    # trial 1 ran with the patched properties: synapses after trial 0, properties from trial 1's file
    st = F.read_network_state(d, 0, file_ending="_0")
    _, _, m1, sc1, er1, gc1 = F.read_maturation_properties(os.path.join(d, "maturation_properties_1.bin"))
    st.erest, st.gca = er1, gc1
    with Engine(make_params(precision=0, trial_ms=TRIAL_MS, seed=(7, 0))) as e:
        e.upload(F.state_to_network(st))
        res = e.run_trial(mode=0, training=True, trial_index=1, first_trial=False,
                          mean_injection_time=r[1].mean_injection_time)
    _, _, trains = F.read_spike_times(os.path.join(d, "spike_times_soma_1.bin"))
    for a, b in zip(res.soma, trains):
        np.testing.assert_array_equal(a, b)
    # resume from trial 1's files: the driver writes the same state back
    st1 = F.read_network_state(d, 1)
    net1 = F.state_to_network(st1)
    resume = dict(mature=st1.mature, maturation_scale=st1.maturation_scale,
                  trials_after_replacement=st1.trials_after_replacement, history=st1.history, trial_number=1)
    out = tmp_path / "re"
    out.mkdir()
    with _make(net1, t, p, mat, resume) as drv:
        drv.write_state(str(out), "_1", "full")
    for name in ("RA_xy", "RA_I_connections", "I_RA_connections", "weights", "axonal_delays_RA2RA",
                 "axonal_delays_RA2I", "axonal_delays_I2RA", "RA_RA_active_connections", "RA_RA_super_connections",
                 "mature", "maturation_properties", "activity_history", "replacement_history",
                 "remodeled_indicators"):
        a = open(os.path.join(d, name + "_1.bin"), "rb").read()
        b = open(os.path.join(str(out), name + "_1.bin"), "rb").read()
        assert a == b, name


def test_chain_test_statistics(tmp_path):
    """test_chain on the driver's network: its trial files equal chain-test
    trials of a fresh engine, and jitter.bin equals the statistics of those."""
    from unsupervised_growth_amd import formats as F
    from unsupervised_growth_amd.chain import ChainStats
    from unsupervised_growth_amd.engine import Engine, make_params
    t, net, p, mat = _setup(n_ra=90, n_i=26)
    d = str(tmp_path)
    print("driver chain test: n_ra=%d n_i=%d" % (net.n_ra, net.n_i), flush=True)
    with _make(net, t, p, mat, None) as drv:
        stats = drv.chain_test(4, d)
        assert drv.trial_number == 0
    cs = ChainStats(net.n_ra)
    with Engine(make_params(precision=0, trial_ms=TRIAL_MS, seed=(7, 0))) as e:
        e.upload(net)
        for k in range(4):
            res = e.run_trial(mode=1, training=True, trial_index=0x80000000 + k, first_trial=(k == 0),
                              mean_injection_time=100.0)
            cs.add_result(res)
            for kind in ("soma", "dend", "interneuron"):
                tr, dur, trains = F.read_spike_times(os.path.join(d, "test_spike_times_%s_%d.bin" % (kind, k)))
                assert tr == 0 and dur == TRIAL_MS
                for a, b in zip(getattr(res, kind), trains):
                    np.testing.assert_array_equal(a, b)
    n, want = cs.array()
    got = F.read_jitter(os.path.join(d, "jitter.bin"))
    assert got[:2] == (net.n_ra, 4)
    for k, name in enumerate(F.JITTER_FIELDS):
        np.testing.assert_array_equal(got[2 + k], want[:, k])
        np.testing.assert_array_equal(stats[name], want[:, k])
    assert (want[:, 0] > 0).any()


def test_growth_runs_end_to_end(tmp_path):
    """The reference's run sequence on the driver: topology files
    (generate_network_topology), new_chain_growth with periodic saves,
    continue_chain_growth from a saved trial, test_chain on a saved state.
    The resumed run reads what the first run wrote and repeats the saved trial
    number first, as the reference does."""
    from unsupervised_growth_amd import formats as F
    from unsupervised_growth_amd.config import GrowthConfig
    from unsupervised_growth_amd.growth import continue_growth, new_growth, test_chain, write_network_topology
    from unsupervised_growth_amd.topology import TopologyParams, generate
    p = TopologyParams(**TOPO)
    t = generate(70, 20, 6, p, seed=SEED)
    net_dir, out, out2, out3 = (tmp_path / k for k in ("net", "out", "out2", "chain"))
    for x in (net_dir, out, out2, out3):
        x.mkdir()
    write_network_topology(str(net_dir), t, p)
    cfg = GrowthConfig()
    cfg.connection.Gei_max, cfg.connection.Gie_max, cfg.connection.delay_constant = 0.25, 0.5, 2.0
    cfg.maturation.GCA_IMMATURE, cfg.maturation.E_REST_IMMATURE = 10.0, -60.0
    print("growth end to end: n_ra=%d n_i=%d" % (t.n_ra, t.n_i), flush=True)
    with new_growth(str(net_dir), str(net_dir / "training_neurons_random.bin"), str(out), cfg, trial_ms=TRIAL_MS,
                    save_freq_short=1, save_freq_long=2) as drv:
        reps = drv.run(3)
    assert [r.saved for r in reps] == [3, 1, 3]
    assert F.read_num_synapses(str(out / "num_synapses.bin"))[0] == [0, 1, 2]
    st = F.read_network_state(str(out), 2)
    assert st.trial == 2 and st.n_ra == 70 and sorted(st.training.tolist()) == sorted(t.training_random.tolist())
    with continue_growth(str(out), 2, str(out2), trial_ms=TRIAL_MS, save_freq_long=1) as drv:
        r = drv.trial()
        assert r.trial == 2 and drv.trial_number == 3
    _, _, w_saved = F.read_weights(str(out / "weights_2.bin"))
    _, _, w_again = F.read_weights(str(out2 / "weights_2.bin"))
    assert w_saved.shape == w_again.shape
    stats = test_chain(str(out), 2, 3, str(out3), trial_ms=TRIAL_MS)
    assert stats["num_test_trials"] == 3
    got = F.read_jitter(str(out3 / "jitter.bin"))
    np.testing.assert_array_equal(got[2], stats["probability_soma_spike"])
    assert os.path.exists(str(out3 / "test_spike_times_soma_2.bin"))


def test_replacement_every_25_trials_equals_oracle(tmp_path):
    """The cadence of the 2000-neuron growth workload: with an activity window
    of 25 trials (check_neuron_activity's 1000, NetworkGrowthSimulator.cpp:9586,
    as a parameter), no neuron is replaced before its 26th trial after the
    start, silent neurons are replaced at that trial, and the driver's state
    after it equals the oracle's bookkeeping (oracle/ora_growth.py with the same
    window and generator draws) applied to the files written at that trial."""
    from unsupervised_growth_amd import formats as F
    from unsupervised_growth_amd.driver import GrowthDriver
    from unsupervised_growth_amd.engine import make_params
    W = 25
    t, net, p, mat = _setup()
    d = str(tmp_path)
    params = make_params(precision=0, trial_ms=TRIAL_MS, seed=(7, 0))
    print("activity window %d: n_ra=%d n_i=%d" % (W, net.n_ra, net.n_i), flush=True)
    reports = []
    with GrowthDriver(net, t.ra_xyz, t.i_xyz, t.model_interneuron_distance, p, mat, seed=SEED, params=params,
                      activity_window=W) as drv:
        drv.write_state(d, what="initial")
        for k in range(W + 1):
            if k == W:
                drv.set_output(d, save_freq_short=1, save_freq_long=1)
            reports.append(drv.trial(training=True))
        drv.write_state(d, "_after", "full")
        replaced_ids, replaced_trials = drv.replaced()
    assert all(r.replaced == 0 for r in reports[:W]), [r.replaced for r in reports]
    g = og.Mt19937(SEED)
    for r in reports:
        assert r.mean_injection_time == og.WAITING_TIME + g.random(TRIAL_MS - 2 * og.WAITING_TIME)
    s = _oracle_state(d, "_%d" % W)
    _, _, soma = F.read_spike_times(os.path.join(d, "spike_times_soma_%d.bin" % W))
    counts = [len(x) for x in soma]
    s["trials_after"] = [a + 1 for a in s["trials_after"]]
    R, driven = og.activity_update(s["history"], counts, s["trials_after"], s["scale"], mat.DEATH_RATE_THRESHOLD,
                                   window=W)
    matured = og.maturation_update(_mat_dict(mat), s["scale"], s["mature"], s["erest"], s["gca"])
    og.replace_neurons(s, R, g, TOPO, _mat_dict(mat))
    last = reports[W]
    print("trial %d: replaced %d, driven %d, matured %d" % (W, last.replaced, last.became_driven, last.matured),
          flush=True)
    assert last.replaced == len(R) > 0 and last.became_driven == driven and last.matured == matured
    assert replaced_ids[:len(R)].tolist() == R and (replaced_trials[:len(R)] == W).all()
    _assert_state_equal(d, "_after", s)


def test_autotuned_shape_survives_replacement(tmp_path):
    """The driver's engine autotuned once keeps its launch shape through the
    engine rebuild of a replacement, and (FP64: every shape gives the same
    bits) grows exactly as a driver on the default shape: same reports, same
    state files after the replacement."""
    from unsupervised_growth_amd import _abi
    from unsupervised_growth_amd.driver import GrowthDriver
    from unsupervised_growth_amd.engine import make_params
    import ctypes
    W = 3
    t, net, p, mat = _setup()
    params = make_params(precision=0, trial_ms=TRIAL_MS, seed=(7, 0))
    reports, dirs = [], []
    for tuned in (False, True):
        d = str(tmp_path / ("tuned" if tuned else "default"))
        os.makedirs(d)
        with GrowthDriver(net, t.ra_xyz, t.i_xyz, t.model_interneuron_distance, p, mat, seed=SEED, params=params,
                          activity_window=W) as drv:
            shape = drv.autotune(windows=40) if tuned else None
            print("tuned=%s shape %s" % (tuned, shape), flush=True)
            rs = [drv.trial(training=True) for _ in range(W + 2)]
            assert any(r.replaced > 0 for r in rs), "the window of 3 trials must replace neurons"
            if tuned:
                tu = _abi.Tuning()
                assert drv.lib.hvc_get_tuning(drv.engine_ctx(), ctypes.byref(tu)) == _abi.HVC_OK
                assert {f: getattr(tu, f) for f, _ in tu._fields_} == shape
            drv.write_state(d, "_end", "full")
        reports.append(rs)
        dirs.append(d)
    for a, b in zip(*reports):
        assert a == b
    for name in sorted(os.listdir(dirs[0])):
        with open(os.path.join(dirs[0], name), "rb") as fa, open(os.path.join(dirs[1], name), "rb") as fb:
            assert fa.read() == fb.read(), name
