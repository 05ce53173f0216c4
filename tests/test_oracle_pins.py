"""Known answers of the oracle's trial logic, derived by hand from the reference.

Each test builds a network of one to six neurons, runs the oracle's trial, and
checks what NetworkGrowthSimulator.cpp ("NGS") fixes for that case, written out
here from the NGS lines -- not by a second simulator:

* the clock is internal_time += 0.02 with a synchronisation when
  internal_time > network_time (NGS:8925, 9291, 9563);
* the training kick adds G_TRAINING_KICK = 3 to Ge_d before the step of the
  first step whose clock exceeds spread + mean_injection_time (NGS:8933-8940);
* a spike at clock T is delivered through a delay d at the first step whose
  clock is >= T + d (NGS:9039-9062, 9069-9076, 9260-9284), and the delivered
  weight reaches the target at the next synchronisation: raiseExcWeight adds
  w / 10000 to Ge_d, raiseI adds w to Gi_d, HVC-I raiseE adds w to Ge
  (NGS:9309-9366; HH2_test.h:89, HH2_test.cpp:444-447, HHI_buffer.cpp:693-696);
* a delivered pre event is stored at its exact time T + d, a post event at
  the clock of the synchronisation that follows the spike (NGS:9093-9102,
  9386); LTD on delivery uses dt = post - (T + d), the synchronisation's STDP
  dt = clock - pre (NGS:9111-9137, 9403-9542), with the LTP / LTD branches of
  NGS:16615-16652;
* update_synapse's four rules and axon_remodeling (NGS:16546-16592,
  16363-16415), and potentiation_decay (NGS:16346-16361).

Spike times come from the oracle's own output: they are the neuron seam,
pinned bit for bit against the reference neuron classes (test_oracle_ref.py).
Exponentials of the hand-derived weights use Python's exp, which may differ
from the oracle's by one ulp, hence the 1e-14 relative tolerance there.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle_lib import OracleSim, dp, oracle

DT, SYNC = 0.02, 0.25


def _clock(trial_ms):
    """internal_time after each step and the synchronisation steps (NGS:8906-8925, 9291, 9563)."""
    n = int(trial_ms / DT)
    clock = np.zeros(n)
    sync = np.zeros(n, bool)
    t, net = 0.0, SYNC
    for k in range(1, n):
        t += DT
        clock[k] = t
        if t > net or k == n - 1:
            sync[k] = True
            net += SYNC
    return clock, sync


def _net(n_ra, n_i, rr=(), ra2i=(), i2ra=(), train=(), spread=(), active=(), supers=(), remodeled=None):
    """rr / ra2i / i2ra: (pre, post, delay, weight) tuples; rows sorted by target."""
    from unsupervised_growth_amd.network import Network

    def csr(edges, n_pre):
        rows = [sorted((e for e in edges if e[0] == i), key=lambda e: e[1]) for i in range(n_pre)]
        ptr = np.zeros(n_pre + 1, np.int32)
        for i, r in enumerate(rows):
            ptr[i + 1] = ptr[i] + len(r)
        flat = [e for r in rows for e in r]
        return (ptr, np.array([e[1] for e in flat], np.int32), np.array([e[3] for e in flat], float),
                np.array([e[2] for e in flat], float))

    a, b, r = csr(ra2i, n_ra), csr(i2ra, n_i), csr(rr, n_ra)
    act = [[j for (i2, j) in active if i2 == i] for i in range(n_ra)]
    sup = [[j for (i2, j) in supers if i2 == i] for i in range(n_ra)]
    return Network(n_ra=n_ra, n_i=n_i, ra2i_ptr=a[0], ra2i_tgt=a[1], ra2i_w=a[2], ra2i_d=a[3],
                   i2ra_ptr=b[0], i2ra_tgt=b[1], i2ra_w=b[2], i2ra_d=b[3],
                   rr_ptr=r[0], rr_tgt=r[1], rr_d=r[3], rr_w=r[2],
                   act_cnt=np.array([len(x) for x in act], np.int32),
                   act_ids=np.array([j for x in act for j in x], np.int32),
                   sup_cnt=np.array([len(x) for x in sup], np.int32),
                   sup_ids=np.array([j for x in sup for j in x], np.int32),
                   remodeled=np.zeros(n_ra, np.int32) if remodeled is None else np.asarray(remodeled, np.int32),
                   erest=np.full(n_ra, -80.0), gca=np.full(n_ra, 55.0),
                   train_ids=np.array(train, np.int32), spread=np.array(spread, float))


class Probe:
    """One state variable of one neuron after every step of the next trials."""

    def __init__(self, sim, kind, neuron, var, trial_ms):
        self.out = np.zeros(int(trial_ms / DT))
        assert sim.lib.ora_set_probe(sim.h, kind, neuron, var, self.out.ctypes.data_as(dp)) == 0


GE_D, GI_D, I_GE = 10, 11, 5


def _first_step_at_or_after(clock, T):
    return int(np.nonzero(clock >= T)[0][0])


def _sync_at_or_after(sync, t):
    return int(np.nonzero(sync[t:])[0][0]) + t


def _first_jump(x):
    """First step whose value is not the previous one decayed (the previous value is 0 here)."""
    nz = np.nonzero(x)[0]
    return int(nz[0]) if len(nz) else None


def test_training_kick_step_and_size():
    trial_ms, spread, mean = 40.0, 3.3, 10.0
    net = _net(1, 0, train=[0], spread=[spread])
    clock, _ = _clock(trial_ms)
    sim = OracleSim(net, trial_ms=trial_ms, i_poisson=0)
    p = Probe(sim, 0, 0, GE_D, trial_ms)
    sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=mean)
    t_kick = int(np.nonzero(clock > spread + mean)[0][0])       # strict >, NGS:8935
    assert _first_jump(p.out) == t_kick
    # raiseE before the step, then the step's decay exp(-((c + dt) - c) / tExc), tExc = 5
    # (HH2_test.cpp:287, 861-864)
    c = clock[t_kick - 1]
    assert p.out[t_kick] == pytest.approx(3.0 * math.exp(-((c + DT) - c) / 5.0), rel=1e-14)
    # growth trial without training: no kick
    p2 = Probe(sim, 0, 0, GE_D, trial_ms)
    sim.run_trial(mode=0, training=False, trial_index=1, mean_injection_time=mean)
    assert not p2.out.any()


@pytest.mark.parametrize("delay", [0.0, 1.37, 2 * DT])
def test_ra_to_ra_arrival_and_sync(delay):
    """Active RA->RA synapse 0 -> 1: neuron 1's Ge_d is 0 until the
    synchronisation at or after the first step with clock >= T0 + d, where it
    equals w / 10000 (the step's decay acts on 0 before the sync adds it).
    With d = 0 the spike is delivered in its own step (zero delay is immediate
    delivery)."""
    trial_ms, w = 40.0, 0.7
    net = _net(2, 0, rr=[(0, 1, delay, w), (1, 0, 1.0, 0.0)], train=[0], spread=[0.0], active=[(0, 1)])
    clock, sync = _clock(trial_ms)
    sim = OracleSim(net, trial_ms=trial_ms, i_poisson=0)
    p = Probe(sim, 0, 1, GE_D, trial_ms)
    r = sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=5.0)
    T0 = r["soma"][0][0]
    t0 = int(np.nonzero(clock == T0)[0][0])
    t_arr = _first_step_at_or_after(clock, T0 + delay)
    if delay == 0.0:
        assert t_arr == t0
    S = _sync_at_or_after(sync, t_arr)
    assert _first_jump(p.out) == S
    assert p.out[S] == w / 10000.0
    # the growth trial's synapse is unchanged by its first delivery (no post event of 1 yet)
    assert len(r["soma"][1]) == 0


def test_ra_to_i_and_i_to_ra_arrival():
    """RA 0 -> HVC-I 0 (weight wa, delay da) and HVC-I 0 -> RA 1 (wi, di):
    the HVC-I Ge jumps by wa at the sync after T0 + da; RA 1's Gi_d by wi at the
    sync after T_I + di, T_I the interneuron's first spike."""
    trial_ms, wa, da, wi, di = 40.0, 5.0, 0.73, 0.9, 1.11
    net = _net(2, 1, rr=[(0, 1, 1.0, 0.0), (1, 0, 1.0, 0.0)], ra2i=[(0, 0, da, wa)], i2ra=[(0, 1, di, wi)],
               train=[0], spread=[0.0])
    clock, sync = _clock(trial_ms)
    sim = OracleSim(net, trial_ms=trial_ms, i_poisson=0)
    pi = Probe(sim, 1, 0, I_GE, trial_ms)
    r = sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=5.0)
    T0 = r["soma"][0][0]
    S = _sync_at_or_after(sync, _first_step_at_or_after(clock, T0 + da))
    assert _first_jump(pi.out) == S and pi.out[S] == wa
    assert len(r["interneuron"][0]) > 0, "the RA input must drive the interneuron"
    TI = r["interneuron"][0][0]
    pr = Probe(sim, 0, 1, GI_D, trial_ms)
    r2 = sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=5.0)
    assert r2["interneuron"][0][0] == TI
    S2 = _sync_at_or_after(sync, _first_step_at_or_after(clock, TI + di))
    assert _first_jump(pr.out) == S2 and pr.out[S2] == wi


def _syn(**kw):
    from unsupervised_growth_amd.network import SynapticParams
    return SynapticParams(**kw)


def _ltp_exp(p, w, dt):
    return min(p.WEIGHT_MAX, w + p.A_P * math.exp(-(dt - p.T_0 - p.T_P) / p.TAU_P))


def _check_one_post_event(spikes, clock, sync):
    """Every spike of a burst after the first lies in a later window and within
    30 ms of the first: the burst makes one post event, at the sync after its
    first spike (NGS:8956-8967, 9386)."""
    t = [int(np.nonzero(clock == x)[0][0]) for x in spikes]
    S = _sync_at_or_after(sync, t[0])
    assert all(k > S for k in t[1:]) and spikes[-1] < spikes[0] + 30.0
    return clock[S]


def test_stdp_exponential_branches_known_answer():
    """Two trained neurons, 0 kicked 20 ms before 1.  Edge 0 -> 1 (w = 0):
    pre event P01 = T0 + d01; at neuron 1's post event (sync clock S1) LTP with
    dt = S1 - P01 > T_0 + T_P: w += A_P exp(-(dt - T_0 - T_P) / TAU_P), which
    activates the synapse.  Edge 1 -> 0 (w = 1, active): its first delivery
    P10 = T1 + d10 meets neuron 0's post event S0: LTD with dt = S0 - P10 <
    T_0 - T_D: w -= w A_D exp((dt - T_0 + T_D) / TAU_D).  End of trial: both
    below SUPERSYNAPSE_THRESHOLD, so w -= BETA; both stay active."""
    p = _syn()
    trial_ms, d01, d10 = 80.0, 1.5, 2.5
    net = _net(2, 0, rr=[(0, 1, d01, 0.0), (1, 0, d10, 1.0)], train=[0, 1], spread=[0.0, 20.0], active=[(1, 0)])
    clock, sync = _clock(trial_ms)
    sim = OracleSim(net, trial_ms=trial_ms, i_poisson=0, syn=p)
    r = sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=10.0)
    s0, s1 = r["soma"][0], r["soma"][1]
    S0 = _check_one_post_event(s0, clock, sync)
    S1 = _check_one_post_event(s1, clock, sync)
    P01, P10 = s0[0] + d01, s1[0] + d10
    assert all(x < P01 + 30.0 for x in s0 + d01)      # one pre event per edge
    assert all(x < P10 + 30.0 for x in s1 + d10)
    dt01 = S1 - P01
    dt10 = S0 - P10
    assert dt01 > p.T_0 + p.T_P and dt10 < p.T_0 - p.T_D
    w01 = _ltp_exp(p, 0.0, dt01) - p.BETA
    w10 = max(0.0, 1.0 - 1.0 * p.A_D * math.exp((dt10 - p.T_0 + p.T_D) / p.TAU_D)) - p.BETA
    syn = sim.synapses()
    assert syn["w"][0] == pytest.approx(w01, rel=1e-14)
    assert syn["w"][1] == pytest.approx(w10, rel=1e-14)
    assert list(syn["act_ids"]) == [1, 0] and list(syn["act_cnt"]) == [1, 1]
    assert list(syn["sup_cnt"]) == [0, 0]
    assert r["warnings"] == (0, 0, 0)


def test_stdp_linear_branches_known_answer():
    """Neuron 1 kicked 3 ms after neuron 0, short delays: LTP at neuron 1's post
    event falls in the linear branch (T_0 < dt <= T_0 + T_P):
    w += A_P (dt - T_0) / T_P; the first delivery 1 -> 0 meets neuron 0's post
    event with T_0 - T_D <= dt <= T_0: w += w A_D (dt - T_0) / T_D."""
    p = _syn()
    trial_ms, d01, d10 = 60.0, 0.3, 0.2
    net = _net(2, 0, rr=[(0, 1, d01, 0.0), (1, 0, d10, 1.0)], train=[0, 1], spread=[0.0, 3.0], active=[(1, 0)])
    clock, sync = _clock(trial_ms)
    sim = OracleSim(net, trial_ms=trial_ms, i_poisson=0, syn=p)
    r = sim.run_trial(mode=0, training=True, trial_index=0, first_trial=True, mean_injection_time=10.0)
    s0, s1 = r["soma"][0], r["soma"][1]
    S0 = _check_one_post_event(s0, clock, sync)
    S1 = _check_one_post_event(s1, clock, sync)
    P01, P10 = s0[0] + d01, s1[0] + d10
    dt01 = S1 - P01
    dt10 = S0 - P10
    assert p.T_0 < dt01 <= p.T_0 + p.T_P, dt01
    assert p.T_0 - p.T_D <= dt10 <= p.T_0, dt10
    # every delivery of neuron 0's burst precedes neuron 1's post event, and
    # later deliveries are within 30 ms of the first (no new pre event)
    assert all(x < P01 + 30.0 for x in s0 + d01) and all(x < P10 + 30.0 for x in s1 + d10)
    w01 = (0.0 + p.A_P * (dt01 - p.T_0) / p.T_P) - p.BETA
    w10 = (1.0 + 1.0 * p.A_D * (dt10 - p.T_0) / p.T_D) - p.BETA
    syn = sim.synapses()
    assert syn["w"][0] == pytest.approx(w01, rel=1e-14, abs=0)
    assert syn["w"][1] == pytest.approx(w10, rel=1e-14, abs=0)
    assert list(syn["act_ids"]) == [1, 0]


def _row_sim(w, active, supers, remodeled=0, Nss=2):
    """Row 0 with candidates 1..5 (weights w[j-1]); no neurons ever run."""
    net = _net(6, 0, rr=[(0, j, 1.0, w[j - 1]) for j in range(1, 6)], active=[(0, j) for j in active],
               supers=[(0, j) for j in supers], remodeled=[remodeled] + [0] * 5)
    return OracleSim(net, trial_ms=1.0, i_poisson=0, syn=_syn(Nss=Nss))


def _row_state(sim):
    s = sim.synapses()
    a = list(s["act_ids"][:s["act_cnt"][0]])
    u = list(s["sup_ids"][:s["sup_cnt"][0]])
    return a, u, int(s["remodeled"][0])


# update_synapse(0, 3): ACTIVATION_THRESHOLD 0.5, SUPERSYNAPSE_THRESHOLD 2.0, Nss 2.
# (w of target 3, active list, super list, remodeled) -> (active list, super list, remodeled, missing-entry warnings)
UPDATE_SYNAPSE_TABLE = [
    (0.3, [1], [], 0, [1], [], 0, 0),               # below both thresholds, inactive: nothing
    (0.7, [1], [], 0, [1, 3], [], 0, 0),            # rule 1: activate, appended
    (0.7, [1], [], 1, [1], [], 1, 0),               # rule 1 blocked by remodeled
    (2.5, [], [], 0, [3], [3], 0, 0),               # rules 1 and 2: active and super
    (2.5, [1, 2, 3], [1, 2], 0, [1, 2, 3], [1, 2], 0, 0),   # rule 2 blocked: saturated (|super| = Nss)
    (2.5, [1, 3], [1], 0, [1, 3], [1, 3], 0, 0),    # rule 2: super, appended
    (1.0, [3, 1], [3, 1], 1, [3, 1], [1], 0, 0),    # rule 3: un-super, remodeled cleared, active kept
    (0.2, [2, 3, 4], [3], 0, [2, 4], [], 0, 0),     # rules 3 and 4: out of both lists, order kept
    (0.2, [3], [], 0, [], [], 0, 0),                # rule 4: deactivate
    (0.2, [], [], 0, [], [], 0, 0),                 # below, inactive: nothing
    (0.7, [3], [], 0, [3], [], 0, 0),               # active and above: nothing
    (2.5, [3, 1], [3, 1], 1, [3, 1], [3, 1], 1, 0),  # remodeled supersynapse above: nothing
    (0.5, [], [], 0, [3], [], 0, 0),                # w == ACTIVATION_THRESHOLD activates (>=)
    (2.0, [3], [], 0, [3], [3], 0, 0),              # w == SUPERSYNAPSE_THRESHOLD is super (>=)
]


@pytest.mark.parametrize("w3,act,sup,rem,act2,sup2,rem2,miss", UPDATE_SYNAPSE_TABLE)
def test_update_synapse_truth_table(w3, act, sup, rem, act2, sup2, rem2, miss):
    w = [0.1, 0.1, w3, 0.1, 0.1]
    for j in act:
        if j != 3:
            w[j - 1] = 2.5 if j in sup else 0.7
    sim = _row_sim(w, act, sup, rem)
    assert sim.lib.ora_update_synapse(sim.h, 0, 3) == 0
    assert _row_state(sim) == (act2, sup2, rem2)
    warn = np.zeros(3, np.int32)
    sim.lib.ora_warnings(sim.h, warn.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert warn[2] == miss


def test_axon_remodeling_known_answer():
    """Saturated row, active list (5, 2, 4), super list (4, 2): remodeling makes
    the active set the super set, in the super list's order, and sets remodeled."""
    sim = _row_sim([0.1, 2.5, 0.1, 2.5, 0.7], [5, 2, 4], [4, 2])
    assert sim.lib.ora_axon_remodeling(sim.h, 0) == 0
    assert _row_state(sim) == ([4, 2], [4, 2], 1)
    # the active indicator of 5 is cleared: a weight above threshold does not re-list it (remodeled)
    assert sim.lib.ora_update_synapse(sim.h, 0, 5) == 0
    assert _row_state(sim) == ([4, 2], [4, 2], 1)


def test_potentiation_decay_closed_form():
    """A growth trial without a spike (no kick, no noise) changes the synapses
    only by potentiation_decay and update_all_synapses: every candidate
    w -> max(0, w - (w >= SUPERSYNAPSE_THRESHOLD ? BETA_SUPERSYNAPSE : BETA)),
    then the four rules on the new weights, targets in ascending order."""
    from unsupervised_growth_amd.network import NoiseParams
    p = _syn()
    w = [2.01, 0.52, 0.55, 0.03, 2.1]                        # targets 1..5 of row 0
    net = _net(6, 0, rr=[(0, j, 1.0, w[j - 1]) for j in range(1, 6)] + [(1, 0, 1.0, 0.3)],
               active=[(0, 1), (0, 2), (0, 5)], supers=[(0, 1), (0, 5)])
    sim = OracleSim(net, trial_ms=5.0, i_poisson=0, syn=p, noise=NoiseParams(0.0, 0.0, 0.0, 0.0))
    r = sim.run_trial(mode=0, training=False, trial_index=0, first_trial=True)
    assert sum(map(len, r["soma"])) == 0
    syn = sim.synapses()
    want = [max(0.0, x - (p.BETA_SUPERSYNAPSE if x >= p.SUPERSYNAPSE_THRESHOLD else p.BETA)) for x in w + [0.3]]
    assert list(syn["w"]) == want                               # exact: one subtraction each
    # 2.01 -> 1.99: un-super (remodeled stays 0); 0.52 -> 0.47: deactivated;
    # 0.55 -> 0.5 (0.55 - 0.05 >= 0.5 in binary64): activated, appended;
    # 0.03 -> 0; 2.1 -> 2.08: stays super and active
    assert (0.55 - 0.05) >= 0.5
    assert _row_state(sim) == ([1, 5, 3], [5], 0)
