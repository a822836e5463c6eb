"""GPU tests of the partial tree of HIERARCHICAL_GRAVITY (agrav_set_partial_tree): builds insert only the
particles of the active list (forcetree.c:91-95, :166-171) in the domain cube of every particle, and the
finalize may keep OldAcc (accel.c:231-232).

* The tree of a sorted active subset is the tree of that subset alone whenever the domain cube is the same
  (a periodic box, TreePM or tree-only with Ewald, with one or two softening types; or an isolated set whose
  subset holds the extreme particles of every axis): the walk's
  tree forces, potentials and interaction counts must equal those of a context holding only the subset,
  bit for bit, for the FP64 per-target walk and the FP32 group walk.  The FP64 walk also equals the
  oracle's walk of the subset's own tree.
* The long-range PM force still covers every particle: PM results computed on a full tree before a partial
  build are carried over unchanged, and a PM computed while the tree is partial equals the full one.
* OldAcc is kept on request and updated otherwise; inactive particles keep their outputs; lists are checked
  before the build indexes with them.
"""
import numpy as np
import pytest

from arepo_amd import ics
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _sel(n, k, seed):
    return np.sort(np.random.Generator(np.random.Philox(seed)).choice(n, k, replace=False)).astype(np.int32)


def _oldacc(n, seed):
    return (np.random.Generator(np.random.Philox(seed)).uniform(0.5, 2.0, n) * 10).astype(np.float32)


def _case(name):
    if name == "ewald":
        # periodic tree-only: Ewald correction per interaction, domain = the box
        pos, mass, eps = ics.clustered_box(16384, nhalo=8, seed=14)
        kw = dict(periodic=True, box_size=1.0, pmgrid=0, soft=(eps,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
        return pos, mass, kw, _sel(len(pos), 3000, 15)
    if name == "treepm":
        pos, mass, eps = ics.clustered_box(32768, nhalo=16, seed=8)
        kw = dict(periodic=True, box_size=1.0, pmgrid=32, soft=(eps,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
        sel = _sel(len(pos), 6000, 3)
    else:
        pos, mass, eps = ics.halo_disk(32768, seed=9)
        kw = dict(periodic=False, box_size=0.0, pmgrid=0, soft=(eps,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
        # the extreme particles of every axis give the subset the domain cube of all particles
        ext = np.concatenate([pos.argmin(0), pos.argmax(0)])
        sel = np.union1d(_sel(len(pos), 5000, 4), ext).astype(np.int32)
    return pos, mass, kw, sel


WALKS = [("plain_fp64", "WALK_PLAIN", "FP64"), ("group_fp32", "WALK_GROUP", "FP32")]


@pytest.mark.parametrize("walk", WALKS, ids=[w[0] for w in WALKS])
@pytest.mark.parametrize("name", ["treepm", "isolated", "ewald", "treepm_2soft"])
def test_partial_tree_equals_tree_of_the_subset(agrav, name, walk):
    pos, mass, kw, sel = _case(name.replace("_2soft", ""))
    stype = None
    if name.endswith("_2soft"):
        # two softening types (forcetree_walk.c:294-320 opens nodes with a larger softening nearby)
        kw = dict(kw, soft=(kw["soft"][0], 3 * kw["soft"][0]))
        stype = (np.random.Generator(np.random.Philox(16)).random(len(pos)) < 0.3).astype(np.uint8)
    oa = _oldacc(len(pos), 12)
    p = agrav.make_params(**kw, opening=1, precision=getattr(agrav, walk[2]), walk=getattr(agrav, walk[1]))
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass, softtype=stype)
        g.set_old_acc(oa)
        g.set_partial_tree(True)
        g.set_active(sel)
        g.compute_grav_accelerations(first_step=False)
        part = g.get_components(tree=True, pm=kw["pmgrid"] > 0)
        out = g.get_accelerations()
        perm, _ = g.get_order()
    with agrav.Gravity(p) as g:
        g.set_particles(pos[sel], mass[sel], softtype=None if stype is None else stype[sel])
        g.set_old_acc(oa[sel])
        g.compute_grav_accelerations(first_step=False)
        alone = g.get_components(tree=True)
        out_alone = g.get_accelerations()
        perm_alone, _ = g.get_order()
    assert np.array_equal(perm, sel[perm_alone])  # the same Hilbert order of the same particles
    assert np.array_equal(part["acc_tree"][sel], alone["acc_tree"])
    assert np.array_equal(part["pot_tree"][sel], alone["pot_tree"])
    assert np.array_equal(out["nint"][sel], out_alone["nint"])
    rest = np.setdiff1d(np.arange(len(pos)), sel)
    assert not out["acc"][rest].any() and not out["nint"][rest].any()
    if kw["pmgrid"] > 0:
        # final = G a_tree + a_PM, with the PM of every particle (G = 1 here)
        e = rel_err(out["acc"][sel], part["acc_tree"][sel] + part["acc_pm"][sel])
        assert e.max() < 1e-12


def test_partial_tree_fp64_equals_oracle_subset_tree(agrav, oracle):
    pos, mass, kw, sel = _case("isolated")
    oa = _oldacc(len(pos), 13)
    with agrav.Gravity(agrav.make_params(**kw, opening=1, precision=agrav.FP64, walk=agrav.WALK_PLAIN)) as g:
        g.set_particles(pos, mass)
        g.set_old_acc(oa)
        g.set_partial_tree(True)
        g.set_active(sel)
        g.construct_forcetree()
        g.gravity_tree(use_bh=False)
        g.gravity_force_finalize()
        r = g.get_accelerations()
    tree = oracle.Tree(oracle.make_params(**kw, opening=1), pos[sel], mass[sel])
    acc, pot, nint, _ = tree.walk(False, old_acc=oa[sel])
    e = rel_err(r["acc"][sel], acc.reshape(-1, 3))
    assert np.percentile(e, 99.9) < 1e-9, float(e.max())
    assert np.mean(r["nint"][sel] == nint) > 0.999


def test_pm_of_every_particle_survives_a_partial_build(agrav):
    pos, mass, kw, sel = _case("treepm")
    p = agrav.make_params(**kw, opening=0, precision=agrav.FP64, walk=agrav.WALK_PLAIN)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.construct_forcetree()
        g.long_range_force()  # on the full tree: kept in its sorted order
        pm_full = g.get_components(tree=False, pm=True)
        g.set_partial_tree(True)
        g.set_active(sel)
        g.construct_forcetree()  # replaces the order: the PM results go back to caller order
        g.gravity_tree(use_bh=True)
        g.gravity_force_finalize()
        out = g.get_accelerations()
        comp = g.get_components(tree=True, pm=True)
        g.long_range_force()  # a PM while the tree is partial deposits every particle from caller order
        pm_again = g.get_components(tree=False, pm=True)
    assert np.array_equal(comp["acc_pm"], pm_full["acc_pm"]) and np.array_equal(comp["pot_pm"], pm_full["pot_pm"])
    assert rel_err(out["acc"][sel], comp["acc_tree"][sel] + pm_full["acc_pm"][sel]).max() < 1e-12
    assert rel_err(pm_again["acc_pm"], pm_full["acc_pm"]).max() < 1e-12
    assert np.abs(pm_again["pot_pm"] - pm_full["pot_pm"]).max() <= 1e-12 * np.abs(pm_full["pot_pm"]).max()


def test_oldacc_kept_or_updated(agrav):
    pos, mass, kw, sel = _case("treepm")
    p = agrav.make_params(**kw, opening=1, precision=agrav.FP32, walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r0 = g.get_accelerations()
        g.set_partial_tree(True, set_oldacc=False)  # a timebin below the highest occupied one
        g.set_active(sel)
        g.compute_grav_accelerations(first_step=False)
        r1 = g.get_accelerations()
        g.set_partial_tree(True, set_oldacc=True)  # the highest occupied timebin
        g.compute_grav_accelerations(first_step=False)
        r2 = g.get_accelerations()
        comp = g.get_components(tree=True, pm=True)
        g.set_partial_tree(False)  # back to the full tree over every particle
        g.set_active(None)
        g.compute_grav_accelerations(first_step=False)
        r3 = g.get_accelerations()
    rest = np.setdiff1d(np.arange(len(pos)), sel)
    assert np.array_equal(r1["old_acc"], r0["old_acc"])
    assert np.array_equal(r1["acc"][rest], r0["acc"][rest])
    # forces among the active particles only: far from the full ones
    assert np.median(rel_err(r1["acc"][sel], r0["acc"][sel])) > 1e-2
    assert np.array_equal(r2["old_acc"][rest], r0["old_acc"][rest])
    want = np.linalg.norm(comp["acc_tree"][sel] + comp["acc_pm"][sel], axis=1).astype(np.float32)
    assert np.allclose(r2["old_acc"][sel], want, rtol=1e-6)
    assert np.median(rel_err(r3["acc"], r0["acc"])) < 5e-3


def test_partial_tree_lists_are_checked(agrav):
    pos, mass, kw, sel = _case("isolated")
    with agrav.Gravity(agrav.make_params(**kw, precision=agrav.FP32, walk=agrav.WALK_GROUP)) as g:
        g.set_particles(pos, mass)
        g.set_partial_tree(True)
        for bad in ([0, 5, 5], [len(pos)], [-1]):
            g.set_active(bad)
            with pytest.raises(agrav.AgravError) as ei:
                g.compute_grav_accelerations(first_step=True)
            assert ei.value.code == agrav.ERR_BAD_PARAM
        g.set_active(sel)
        g.construct_forcetree()
        g.set_active(sel[:100])  # a new list after the partial build
        with pytest.raises(agrav.AgravError) as ei:
            g.gravity_tree(use_bh=True)
        assert ei.value.code == agrav.ERR_STATE
        g.set_active([])  # an empty list: an empty tree, nobody evaluated
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
        assert not r["acc"].any()
        g.set_active(sel[:2000])
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
        assert np.isfinite(r["acc"][sel[:2000]]).all() and (r["nint"][sel[:2000]] > 0).all()


@pytest.mark.parametrize("walk", WALKS, ids=[w[0] for w in WALKS])
def test_partial_tree_with_deep_tie_runs(agrav, walk):
    # uniform particles in the unit periodic box plus clumps inside single level-21 cells (first key words
    # equal, order from the second words): the active half of the clumps forms tie runs of ~1200 (global
    # bitonic sort), ~150 (LDS) and 2 (per-thread); the partial build is the context's first build
    g = np.random.Generator(np.random.Philox(31))
    uni = g.random((6000, 3))
    ln = 1.00001
    corner = 0.5 - 0.5 * ln
    cell = ln / 2 ** 21
    parts = [uni]
    for j, (nc, off) in enumerate(((2400, (2 ** 20, 2 ** 20 - 7, 2 ** 20 + 5)), (300, (2 ** 19 + 3, 2 ** 18, 2 ** 20 - 1)))):
        origin = corner + np.array(off) * cell
        parts.append(origin + (0.4 + 0.2 * g.random((nc, 3))) * cell)
    pos = np.concatenate(parts)
    pos[10:20] = pos[20:30]  # coincident pairs
    mass = g.uniform(0.5, 1.5, len(pos)) / len(pos)
    sel = np.arange(0, len(pos), 2, dtype=np.int32)
    kw = dict(periodic=True, box_size=1.0, pmgrid=0, soft=(1e-9,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
    p = agrav.make_params(**kw, opening=0, precision=getattr(agrav, walk[2]), walk=getattr(agrav, walk[1]))
    with agrav.Gravity(p) as gr:
        gr.set_particles(pos, mass)
        gr.set_partial_tree(True)
        gr.set_active(sel)
        gr.compute_grav_accelerations(first_step=True)
        part = gr.get_components(tree=True)
        perm, keys = gr.get_order()
        nodes = gr.get_tree()
    with agrav.Gravity(p) as gr:
        gr.set_particles(pos[sel], mass[sel])
        gr.compute_grav_accelerations(first_step=True)
        alone = gr.get_components(tree=True)
        perm_alone, keys_alone = gr.get_order()
        nodes_alone = gr.get_tree()
    assert nodes["level"].max() > 21  # cells split below the first key word
    assert np.array_equal(keys, keys_alone) and np.array_equal(perm, sel[perm_alone])
    assert np.array_equal(nodes["len"], nodes_alone["len"]) and np.array_equal(nodes["npart"], nodes_alone["npart"])
    assert np.array_equal(part["acc_tree"][sel], alone["acc_tree"])
    assert np.array_equal(part["pot_tree"][sel], alone["pot_tree"])
