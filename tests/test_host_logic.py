"""Host-side logic without a device: synthetic networks, configs, op counts."""
import numpy as np

from unsupervised_growth_amd import configs, opcount
from unsupervised_growth_amd.network import random_network


def _check_csr(ptr, tgt, n_post, ascending=False):
    assert ptr[0] == 0 and np.all(np.diff(ptr) >= 0)
    assert len(tgt) == ptr[-1]
    if len(tgt):
        assert tgt.min() >= 0 and tgt.max() < n_post
    for i in range(len(ptr) - 1):
        row = tgt[ptr[i]:ptr[i + 1]]
        assert len(set(row.tolist())) == len(row)
        if ascending:
            assert np.all(np.diff(row) > 0)


def test_random_network_shapes():
    net = random_network(30, 7, n_train=5, rr=4, seed=2)
    _check_csr(net.ra2i_ptr, net.ra2i_tgt, 7)
    _check_csr(net.i2ra_ptr, net.i2ra_tgt, 30)
    _check_csr(net.rr_ptr, net.rr_tgt, 30, ascending=True)
    assert net.n_candidates == 30 * 4
    assert np.all(net.rr_tgt != np.repeat(np.arange(30), 4))
    assert len(net.train_ids) == 5 and len(set(net.train_ids.tolist())) == 5


def test_c1_config_matches_description():
    net, kw = configs.c1()
    assert (net.n_ra, net.n_i) == (300, 30)
    assert net.n_candidates == 300 * 300 and np.all(net.rr_w == 0)
    # quantised delays k * dt, k in 0..250
    k = net.ra2i_d / 0.02
    assert np.all(np.abs(k - np.round(k)) < 1e-9) and k.max() <= 250
    assert kw["trial_ms"] == 500.0 and len(net.train_ids) == 10


def test_opcounts_are_stable():
    # the roofline uses these; a change of the stepper expression trees changes them
    assert opcount.count_ra_step(0).n == {"add": 290, "mul": 224, "div": 120, "exp": 32}
    assert opcount.count_ra_step(1).n == {"add": 250, "mul": 170, "div": 105, "exp": 28}
    assert opcount.count_i_step(0).n == {"add": 650, "mul": 661, "div": 392, "exp": 84}
