"""CPU checks of host-side logic around the kernels: the ensemble members'
parameter draws, the longest-first cost estimate, the autotuner's cache
keys and the relaxed-variant accuracy classification used by the tests."""
import math

import numpy as np

from pip_amd import autotune, problems
from pip_amd.ensemble import cost_estimate
from tests.util import order_one


def test_ensemble_member_draws_cover_the_stated_ranges():
    p1 = problems.ensemble_members(256, seed=7)
    nx = np.array([p["nx"] for p in p1])
    assert nx.min() >= 256 and nx.max() <= 4096 and len(set(nx)) > 200
    assert all(0.1 <= p["xi_n"] <= 0.99 for p in p1)
    assert all(1.0 <= p["col"] <= 1.0e4 for p in p1)
    assert all(2.0 <= 1.0 / p["p_right"] <= 100.0 for p in p1)
    assert p1 == problems.ensemble_members(256, seed=7)  # seeded, reproducible
    p2 = problems.ensemble_members_2d(64, seed=9)
    n = np.array([p["n"] for p in p2])
    assert n.min() >= 32 and n.max() <= 256
    assert all(0.1 <= p["xi_n"] <= 0.99 and 1.0 <= p["col"] <= 1.0e4 for p in p2)


def test_cost_estimate_scales_with_size_and_survives_nonfinite_states():
    p = problems.ensemble_members(256, seed=7)[0]
    cfg, uh, um, _ = problems.ensemble_member(dict(p, nx=400))
    c1 = cost_estimate(cfg, uh, um, 0.1)
    cfg2, uh2, um2, _ = problems.ensemble_member(dict(p, nx=800))
    c2 = cost_estimate(cfg2, uh2, um2, 0.1)
    # twice the cells and about twice the steps (dt ~ dx)
    assert 3.5 < c2 / c1 < 4.5
    um_bad = um.copy()
    um_bad[6, 0, 30] = np.nan
    assert cost_estimate(cfg, uh, um_bad, 0.1) == float(cfg.nx)
    assert math.isfinite(cost_estimate(cfg, uh, um, 1.0e30))


def test_autotune_keys_separate_configurations_and_arithmetic():
    kh = problems.kelvin_helmholtz(64)[0]
    ot = problems.orszag_tang(64)[0]
    keys = {autotune.stage_key(c, a) for c in (kh, ot) for a in ("exact", "relaxed")}
    assert len(keys) == 4
    assert autotune.two_pass_ok(kh) and autotune.two_pass_ok(ot)
    assert not autotune.two_pass_ok(kh.replace(flag_col=3))


def test_order_one_classification():
    """The relaxed-variant tests call a variable order one when its mean
    magnitude is >= 5 % of its fluid's mean density."""
    cfg, uh, um, _ = problems.kelvin_helmholtz(32)
    big = order_one(um, cfg)
    assert 0 in big and 4 in big and 5 in big  # rho, E, the guide field B_x
    assert 6 not in big and 8 not in big       # B_y, psi: born from the perturbation
    cfg, uh, um, _ = problems.orszag_tang(32)
    assert set(order_one(um, cfg)) >= {0, 1, 2, 4, 5, 6}


def test_replay_sets_never_hand_a_replay_a_set_in_use():
    """The backward sweep's buffer sets (pip_amd.adjoint.replay_sets): the
    kept newest segments hold their own sets; overlapped, the replay of
    segment i + 1 (running while segment i's adjoint steps read its set)
    writes another set, and only after the adjoint steps of every older
    reader of its set (it waits for the last of them, which is at least two
    segments back); in line, the replays share set 0."""
    import pytest

    from pip_amd.adjoint import replay_sets
    for nseg in range(0, 13):
        for kept in range(0, nseg + 1):
            sets, waits = replay_sets(nseg, kept, False)
            assert sets == [i if i < kept else 0 for i in range(nseg)] and waits == [-1] * nseg
            sets, waits = replay_sets(nseg, kept, True)
            assert sets[:kept] == list(range(kept))
            nsets = len(set(sets))
            assert nsets == (max(kept, min(nseg, 2)) if nseg > kept else kept)
            for i in range(kept, nseg):
                if i >= 1:
                    assert sets[i] != sets[i - 1]  # written while segment i - 1 reads its set
                readers = [j for j in range(i) if sets[j] == sets[i]]
                assert waits[i] == (readers[-1] if readers else -1)
                assert waits[i] == -1 or waits[i] <= i - 2  # enqueued before this replay
                # no kept segment still to be read (newer index = read earlier)
                assert all(j < i for j in range(nseg) if j < kept and sets[j] == sets[i])
    with pytest.raises(ValueError):
        replay_sets(3, 4, True)
