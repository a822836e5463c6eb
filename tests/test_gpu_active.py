"""GPU tests of active-target walks (agrav_set_active): gravity_tree over TimeBinsGravity.ActiveParticleList
only (gravtree.c:494-520), with the tree built from every particle and the PM force over every particle.

* The FP64 per-target walk of an active subset equals the full evaluation on those particles bitwise (a
  target's walk does not depend on the other targets) and the oracle's walk of the same target list.
* The group walks form their groups from the active targets in Hilbert order; their error against the
  FP64 direct sum stays at most the oracle's on the same targets.
* Inactive particles keep their outputs (zero after set_particles); lists are validated.
"""
import numpy as np
import pytest

from arepo_amd import ics
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "treepm":
        pos, mass, eps = ics.clustered_box(32768, nhalo=16, seed=8)
        kw = dict(periodic=True, box_size=1.0, pmgrid=32, soft=(eps,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
    else:
        pos, mass, eps = ics.halo_disk(32768, seed=9)
        kw = dict(periodic=False, box_size=0.0, pmgrid=0, soft=(eps,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
    return pos, mass, kw


def _subset(n, k, seed):
    return np.sort(np.random.Generator(np.random.Philox(seed)).choice(n, k, replace=False)).astype(np.int32)


@pytest.mark.parametrize("name", ["treepm", "isolated"])
def test_active_plain_walk_equals_full_evaluation(agrav, name):
    pos, mass, kw = _case(name)
    sel = _subset(len(pos), 3000, 5)
    p = agrav.make_params(**kw, opening=1, precision=agrav.FP64, walk=agrav.WALK_PLAIN)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        full = g.get_accelerations()
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.set_active(sel)
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
        t = g.timings()
    assert np.array_equal(r["acc"][sel], full["acc"][sel])
    assert np.array_equal(r["pot"][sel], full["pot"][sel])
    assert np.array_equal(r["nint"][sel], full["nint"][sel])
    assert t["interactions"] == int(full["nint"][sel].sum())
    rest = np.setdiff1d(np.arange(len(pos)), sel)
    assert not r["acc"][rest].any() and not r["nint"][rest].any()


def test_active_plain_walk_equals_oracle_target_walk(agrav, oracle):
    # relative criterion with a given OldAcc, tree only: the oracle walks the same target list
    pos, mass, kw = _case("isolated")
    sel = _subset(len(pos), 2500, 6)
    oa = (np.random.Generator(np.random.Philox(7)).uniform(0.5, 2.0, len(pos)) * 10).astype(np.float32)
    with agrav.Gravity(agrav.make_params(**kw, opening=1, precision=agrav.FP64, walk=agrav.WALK_PLAIN)) as g:
        g.set_particles(pos, mass)
        g.set_old_acc(oa)
        g.set_active(sel)
        g.construct_forcetree()
        g.gravity_tree(use_bh=False)
        g.gravity_force_finalize()
        r = g.get_accelerations()
    tree = oracle.Tree(oracle.make_params(**kw, opening=1), pos, mass)
    acc, pot, nint, _ = tree.walk(False, old_acc=oa, targets=sel)
    e = rel_err(r["acc"][sel], acc)
    assert np.percentile(e, 99.9) < 1e-9, float(e.max())
    assert np.mean(r["nint"][sel] == nint) > 0.999


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("name", ["treepm", "isolated"])
def test_active_group_walk_error_vs_direct_sum(agrav, oracle, name, prec):
    # groups of active targets span more space than groups of all particles, so they open at least as
    # many nodes: the error on the active targets is no worse than the same path's full evaluation
    pos, mass, kw = _case(name)
    sel = _subset(len(pos), 3000, 11)
    P = agrav.FP32 if prec == "fp32" else agrav.FP64
    p = agrav.make_params(**kw, opening=1, precision=P, walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        full = g.get_accelerations()
        ft = g.gravity_forcetest(sel)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.set_active(sel)
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
    o = oracle.compute_grav_accelerations(oracle.make_params(**kw, opening=1), pos, mass, first_step=True)
    e_act = rel_err(r["acc"][sel], ft["acc"])
    e_full = rel_err(full["acc"][sel], ft["acc"])
    e_or = rel_err(o["acc"][sel], ft["acc"])
    print("\n%s %s: active median %.3g p99 %.3g | full %.3g %.3g | oracle %.3g %.3g"
          % (name, prec, np.median(e_act), np.percentile(e_act, 99), np.median(e_full), np.percentile(e_full, 99),
             np.median(e_or), np.percentile(e_or, 99)))
    assert np.median(e_act) <= 1.05 * np.median(e_full)
    assert np.percentile(e_act, 99) <= 1.05 * np.percentile(e_full, 99)
    assert np.median(e_act) <= 1.05 * np.median(e_or)


def test_inactive_particles_keep_their_outputs(agrav):
    pos, mass, kw = _case("treepm")
    a = _subset(len(pos), 4000, 21)
    b = _subset(len(pos), 4000, 22)
    p = agrav.make_params(**kw, opening=1, precision=agrav.FP32, walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r0 = g.get_accelerations()
        g.set_active(a)
        g.compute_grav_accelerations(first_step=False)
        r1 = g.get_accelerations()
        g.set_active(b)
        g.compute_grav_accelerations(first_step=False)
        r2 = g.get_accelerations()
        g.set_active([])  # an empty list evaluates nobody
        g.compute_grav_accelerations(first_step=False)
        r3 = g.get_accelerations()
        g.set_active(None)  # every particle again
        g.compute_grav_accelerations(first_step=False)
        r4 = g.get_accelerations()
    only_a = np.setdiff1d(a, b)
    rest = np.setdiff1d(np.arange(len(pos)), np.union1d(a, b))
    assert np.array_equal(r2["acc"][only_a], r1["acc"][only_a])
    assert np.array_equal(r2["acc"][rest], r0["acc"][rest])
    assert np.array_equal(r3["acc"], r2["acc"])
    # active targets were evaluated afresh: close to the full evaluation, not equal to stale values
    assert np.median(rel_err(r2["acc"][b], r0["acc"][b])) < 5e-3
    assert np.median(rel_err(r4["acc"], r0["acc"])) < 5e-3


def test_bad_active_lists(agrav):
    pos, mass, kw = _case("isolated")
    with agrav.Gravity(agrav.make_params(**kw, precision=agrav.FP32, walk=agrav.WALK_GROUP)) as g:
        with pytest.raises(agrav.AgravError):
            g.set_active([1, 2, 3])  # before set_particles
        g.set_particles(pos, mass)
        for bad in ([0, 5, 5], [len(pos)], [-1]):
            g.set_active(bad)
            with pytest.raises(agrav.AgravError) as ei:
                g.compute_grav_accelerations(first_step=True)
            assert ei.value.code == agrav.ERR_BAD_PARAM
        g.set_active([0, 1, 2])
        g.compute_grav_accelerations(first_step=True)
        assert np.isfinite(g.get_accelerations()["acc"][:3]).all()
