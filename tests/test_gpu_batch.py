"""GPU tests of batched solves of independent isolated sets (agrav_batch_*).

A batch shares one tree build (set index above the Hilbert digits, one domain
cube per set) and one persistent walk launch.  The key properties:
  * each set's result equals a solve of that set alone (bitwise, for the same
    group size and leaf bucket, when the set's tree is the same -- the batch
    keeps fewer key levels per set, which only matters for sets whose own tree
    is deeper than those levels);
  * the accuracy against FP64 direct summation matches the reference's own
    unbinding walk (subfind_loctree.c:564-730, restated in the oracle);
  * the autotuner only changes speed: a solve with the tuned setting equals a
    solve with that setting given by hand.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _sets(sizes, seed):
    from arepo_amd import ics
    g = ics.rng(seed)
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    rel = np.zeros((int(off[-1]), 3), np.float32)
    eps = np.zeros(len(sizes))
    for s, n in enumerate(sizes):
        if n == 0:
            eps[s] = 0.01
            continue
        a = 0.01 * (1 + s % 5)
        r = ics.hernquist_radii(int(n), a, g) if s % 2 else ics._nfw_radii(int(n), 4.0 + s % 11, 30 * a, g)
        rel[off[s]:off[s + 1]] = ics._isotropic(r, g)
        eps[s] = max(float(np.median(r)) / 50.0, 1e-4)
    cen = g.uniform(-5, 5, (len(sizes), 3))
    mass = np.full(int(off[-1]), 1.0 / 1000) * (1 + (np.arange(int(off[-1])) % 3))
    return off, rel, mass, eps, cen


def _params(agrav, precision, walk, opening=0, theta=0.7, leaf_bucket=8):
    # one wavefront per group, as the batched walk uses, so single solves compare bitwise
    return agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01,), precision=precision, walk=walk,
                             opening=opening, err_tol_theta=theta, leaf_bucket=leaf_bucket, group_size=64, group_waves=1)


def _solve_batch(agrav, p, off, rel, mass, eps, cen, tuning=None):
    with agrav.Gravity(p) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        if tuning is not None:
            g.batch_set_tuning(*tuning)
        g.compute_grav_accelerations(first_step=True)
        out = g.get_accelerations()
        out["timings"] = g.timings()
    return out


def _abs_pos(off, rel, cen, s):
    return rel[off[s]:off[s + 1]].astype(np.float64) + cen[s]


def _solve_single(agrav, p, pos, mass, eps):
    p.n_soft_types = 1
    p.softening[0] = eps
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        out = g.get_accelerations()
        tree = g.get_tree()
    return out, int(tree["level"].max()) if len(tree["level"]) else 0


SIZES = [1, 2, 3, 0, 32, 100, 777, 4000, 12000, 64, 65]


@pytest.mark.parametrize("kind", ["fp32_group", "fp64_group", "fp64_plain"])
def test_batch_equals_single_solves(agrav, kind):
    prec = agrav.FP32 if kind == "fp32_group" else agrav.FP64
    walk = agrav.WALK_PLAIN if kind == "fp64_plain" else agrav.WALK_GROUP
    off, rel, mass, eps, cen = _sets(SIZES, 7)
    out = _solve_batch(agrav, _params(agrav, prec, walk), off, rel, mass, eps, cen, tuning=(64, 8))
    bits = int(np.ceil(np.log2(len(SIZES))))
    levels = (63 - bits) // 3
    compared_bitwise = 0
    for s, n in enumerate(SIZES):
        if n == 0:
            continue
        sl = slice(off[s], off[s + 1])
        ref, maxlev = _solve_single(agrav, _params(agrav, prec, walk), _abs_pos(off, rel, cen, s), mass[sl], eps[s])
        if maxlev < levels:
            for f in ("acc", "pot", "nint", "old_acc"):
                assert np.array_equal(out[f][sl], ref[f]), (kind, s, n, f)
            compared_bitwise += 1
        else:
            assert rel_err(out["acc"][sl], ref["acc"]).max() < 1e-3
    assert compared_bitwise >= len([n for n in SIZES if n]) - 2
    # the batched group walk writes the final outputs from its epilogue (no finalize kernel)
    assert (out["timings"]["finalize"] == 0) == (walk == agrav.WALK_GROUP)


def test_many_small_sets_equal_single_solves(agrav):
    # 1000 sets: 10 bits of set index above 17 Hilbert levels per set
    from arepo_amd import ics
    g = ics.rng(5)
    sizes = g.integers(32, 300, 1000)
    off, rel, mass, eps, cen = _sets(list(sizes), 9)
    p = _params(agrav, agrav.FP32, agrav.WALK_GROUP)
    out = _solve_batch(agrav, p, off, rel, mass, eps, cen, tuning=(64, 8))
    for s in [0, 1, 2, 499, 998, 999]:
        sl = slice(off[s], off[s + 1])
        ref, maxlev = _solve_single(agrav, _params(agrav, agrav.FP32, agrav.WALK_GROUP), _abs_pos(off, rel, cen, s),
                                    mass[sl], eps[s])
        assert maxlev < 17
        assert np.array_equal(out["acc"][sl], ref["acc"]) and np.array_equal(out["pot"][sl], ref["pot"]), s


def test_batch_direct_matches_oracle(agrav, oracle):
    # the per-set FP64 direct sum of the batch against the oracle's direct summation
    off, rel, mass, eps, cen = _sets([50, 700, 3000, 1], 11)
    with agrav.Gravity(_params(agrav, agrav.FP64, agrav.WALK_GROUP)) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        acc_d, pot_d = g.gravity_direct()
        sel = np.array([0, 10, 49, 50, 300, 749, 750, 3749, 3750], np.int32)
        ft = g.gravity_forcetest(sel)
    for s in range(4):
        sl = slice(off[s], off[s + 1])
        op = oracle.make_params(periodic=False, box_size=0.0, soft=(eps[s],))
        ao, po = oracle.direct_summation(op, _abs_pos(off, rel, cen, s), mass[sl])
        np.testing.assert_allclose(acc_d[sl], ao, rtol=1e-10, atol=1e-12 * np.abs(ao).max())
        np.testing.assert_allclose(pot_d[sl], po, rtol=1e-10)
    np.testing.assert_array_equal(ft["acc"], acc_d[sel])
    np.testing.assert_array_equal(ft["pot"], pot_d[sel])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_accuracy_matches_reference_unbinding_walk(agrav, oracle, precision):
    # Barnes-Hut potential as in subfind_loctree.c:564-730 (theta 0.7): the batch's
    # error against direct summation is no worse than the reference walk's own
    prec = agrav.FP32 if precision == "fp32" else agrav.FP64
    off, rel, mass, eps, cen = _sets([300, 2500, 9000, 20000], 13)
    out = _solve_batch(agrav, _params(agrav, prec, agrav.WALK_GROUP, theta=0.7), off, rel, mass, eps, cen)
    with agrav.Gravity(_params(agrav, agrav.FP64, agrav.WALK_GROUP)) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        acc_d, pot_d = g.gravity_direct()
    for s in range(4):
        sl = slice(off[s], off[s + 1])
        pos = _abs_pos(off, rel, cen, s)
        pot_o, _ = oracle.group_potential_bh(pos, mass[sl], 2.8 * eps[s], 0.7)
        e_ref = np.abs(pot_o - pot_d[sl]) / np.abs(pot_d[sl])
        e_gpu = np.abs(out["pot"][sl] - pot_d[sl]) / np.abs(pot_d[sl])
        assert np.median(e_gpu) <= np.median(e_ref) * 1.05 + 1e-6, s
        assert np.percentile(e_gpu, 99) <= np.percentile(e_ref, 99) * 1.05 + 1e-6, s
        ea = rel_err(out["acc"][sl], acc_d[sl])
        assert np.median(ea) < 5e-3 and np.percentile(ea, 99) < 5e-2, s


def test_batch_relative_criterion(agrav):
    off, rel, mass, eps, cen = _sets([500, 6000, 15000], 17)
    p = _params(agrav, agrav.FP32, agrav.WALK_GROUP, opening=1)
    out = _solve_batch(agrav, p, off, rel, mass, eps, cen)
    with agrav.Gravity(_params(agrav, agrav.FP64, agrav.WALK_GROUP)) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        acc_d, _ = g.gravity_direct()
    e = rel_err(out["acc"], acc_d)
    assert np.median(e) < 2.5e-3 and np.percentile(e, 99) < 2e-2


def test_autotune_changes_speed_only(agrav):
    off, rel, mass, eps, cen = _sets([40, 90, 200, 1500, 3000, 40000, 60, 5000], 19)
    p = _params(agrav, agrav.FP32, agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        t = g.batch_autotune(sample_particles=30000)
        g.compute_grav_accelerations(first_step=True)
        tuned = g.get_accelerations()
        assert g.batch_get_tuning()["group_size"] == t["group_size"]
    assert t["tuned"] == 1
    assert all(v in (16, 32, 64) for v in t["group_size"])
    assert all(v in (0, 4, 8, 16) for v in t["leaf_bucket"])
    assert t["sample_particles"][3] > 0 and t["best_ms"][3] > 0
    by_hand = _solve_batch(agrav, p, off, rel, mass, eps, cen, tuning=(t["group_size"], t["leaf_bucket"]))
    assert np.array_equal(tuned["acc"], by_hand["acc"]) and np.array_equal(tuned["pot"], by_hand["pot"])


@pytest.mark.parametrize("gs", [16, 32])
def test_batch_small_groups_accurate(agrav, gs):
    off, rel, mass, eps, cen = _sets([33, 1000, 7000], 23)
    out = _solve_batch(agrav, _params(agrav, agrav.FP32, agrav.WALK_GROUP), off, rel, mass, eps, cen, tuning=(gs, 4))
    plain = _solve_batch(agrav, _params(agrav, agrav.FP64, agrav.WALK_PLAIN), off, rel, mass, eps, cen)
    with agrav.Gravity(_params(agrav, agrav.FP64, agrav.WALK_GROUP)) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        acc_d, _ = g.gravity_direct()
    # the group criterion holds for every target of the group, so it is at least
    # as accurate as the per-target Barnes-Hut walk
    e, e_plain = rel_err(out["acc"], acc_d), rel_err(plain["acc"], acc_d)
    assert np.median(e) <= np.median(e_plain) * 1.05 and np.percentile(e, 99) <= np.percentile(e_plain, 99) * 1.05


def test_context_reuse_batch_then_single(agrav):
    off, rel, mass, eps, cen = _sets([100, 2000], 29)
    pos = _abs_pos(off, rel, cen, 1)
    p = _params(agrav, agrav.FP32, agrav.WALK_GROUP)
    ref, _ = _solve_single(agrav, _params(agrav, agrav.FP32, agrav.WALK_GROUP), pos, mass[off[1]:], eps[1])
    p.n_soft_types = 1
    p.softening[0] = eps[1]
    with agrav.Gravity(p) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        g.compute_grav_accelerations()
        g.set_particles(pos, mass[off[1]:])
        g.compute_grav_accelerations()
        again = g.get_accelerations()
    assert np.array_equal(again["acc"], ref["acc"])


def test_batch_rejects_bad_input(agrav):
    off, rel, mass, eps, cen = _sets([10, 20], 31)
    with agrav.Gravity(agrav.make_params(periodic=True, box_size=1.0, soft=(0.01,))) as g:
        with pytest.raises(agrav.AgravError):
            g.batch_set_particles(off, rel, mass, eps)
    with agrav.Gravity(_params(agrav, agrav.FP32, agrav.WALK_GROUP)) as g:
        with pytest.raises(agrav.AgravError):
            g.batch_set_particles(off, rel, mass, np.array([0.01, 0.0]))
        bad = off.copy()
        bad[1] = 35
        with pytest.raises((agrav.AgravError, ValueError)):
            g.batch_set_particles(bad, rel, mass, eps)
        with pytest.raises(agrav.AgravError):
            g.batch_set_tuning(128, 4)
