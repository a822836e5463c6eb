"""GPU tests at the full sizes of the single-set workloads.

* 128^3 = 2,097,152 clustered particles (half uniform, half in 256 Hernquist
  spheres) in the periodic unit box, TreePM with PMGRID = 128, ASMTH 1.25,
  RCUT 4.5, ErrTolForceAcc 0.0025, softening 1/40 of the mean spacing (the
  benchmark's workload).  OldAcc comes from a Barnes-Hut seeding pass.  The
  FP64 plain walk + PM equals the oracle (its tree walk + PM) on 16,384
  seeded targets; the FP32 group walk's error against the FP64 Ewald direct
  sum (FORCETEST) is no worse than the oracle's, in total and in its tree
  (short-range) part; the PM part of the error is printed beside it.

* 32^3 = 32,768 uniform particles in the periodic unit box, tree-only with the
  Ewald correction, relative criterion (ErrTolForceAcc 0.0025) seeded by a
  Barnes-Hut pass (theta 0.7), softening 1/40 of the mean spacing, one solve
  from scratch.  The FP64 plain walk equals the oracle over all particles;
  the FP32 group walk's error against the Ewald direct sum is no worse than
  the oracle's.
* 1,048,576-particle isolated Hernquist halo + exponential disk, softening
  0.01, ErrTolForceAcc 0.001 / 0.0025 / 0.01: FP32-offset + compensated group
  walk and FP64 walks against the FP64 direct sum on a seeded 16,384-target
  subset, next to the oracle walk on the same targets with the same OldAcc.

* 4096 independent isolated groups, 100,311,619 particles (32 ... 199,900 per
  group, NFW- and Hernquist-like; the batched unbinding workload), one batched
  FP32 solve of potentials and accelerations with Barnes-Hut theta 0.7 after
  the autotuner: on 512 targets in each of 32 groups, the potential error
  against the FP64 direct sum over the group is no worse than that of the
  oracle's restatement of the reference's unbinding walk
  (subfind_loctree.c:564-730) on the same targets; no group falls back.

Each test prints its error table (run pytest with -s to see it).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _stats(e):
    return float(np.median(e)), float(np.percentile(e, 99))


def test_clustered_128cubed_treepm(agrav, oracle):
    from arepo_amd import ics
    pos, mass, eps = ics.clustered_box(128 ** 3, seed=2)
    n = len(pos)
    sel = np.sort(np.random.Generator(np.random.Philox(99)).choice(n, 16384, replace=False)).astype(np.int32)
    kw = dict(periodic=True, box_size=1.0, pmgrid=128, asmth=1.25, rcut=4.5, soft=(eps,), err_tol_theta=0.7,
              err_tol_force_acc=0.0025, opening=1)
    # OldAcc from the Barnes-Hut seeding pass (accel.c:62-72) of the FP64 per-target walk
    with agrav.Gravity(agrav.make_params(**kw, precision=agrav.FP64, walk=agrav.WALK_PLAIN)) as g:
        g.set_particles(pos, mass)
        g.construct_forcetree()
        g.long_range_force()
        g.gravity_tree(use_bh=True)
        g.gravity_force_finalize()
        oa = g.get_accelerations()["old_acc"]
        g.set_old_acc(oa)
        g.compute_grav_accelerations(first_step=False)
        r64 = g.get_accelerations()
        c64 = g.get_components(tree=True, pm=True)
        ft = g.gravity_forcetest(sel)
    with agrav.Gravity(agrav.make_params(**kw, precision=agrav.FP32, walk=agrav.WALK_GROUP, leaf_bucket=4,
                                         eval_potential=False)) as g:
        g.set_particles(pos, mass)
        g.set_old_acc(oa)
        g.compute_grav_accelerations(first_step=False)
        r32 = g.get_accelerations()
        c32 = g.get_components(tree=True, pm=True)
    # the oracle: the reference's PM and its relative walk of the selected targets with the same OldAcc
    op = oracle.make_params(**kw)
    acc_pm_o, _ = oracle.pm_periodic(op, pos, mass)
    acc_t_o, _, nint_o, _ = oracle.Tree(op, pos, mass).walk(False, old_acc=oa, targets=sel, nthreads=16)
    acc_o = acc_t_o + acc_pm_o[sel]
    e = rel_err(r64["acc"][sel], acc_o)
    assert np.percentile(e, 99.9) < 1e-8, (np.percentile(e, 99.9), e.max())
    assert np.mean(r64["nint"][sel] == nint_o) > 0.999
    assert np.percentile(rel_err(c64["acc_pm"][sel], acc_pm_o[sel]), 99.9) < 1e-8
    # accuracy against the FP64 Ewald direct sum, in total and split: the tree part against the exact
    # short-range sum, the PM part against the long-range rest (gravtree_forcetest.c:629-638), both
    # relative to the total acceleration
    a_norm = np.linalg.norm(ft["acc"], axis=1)

    def split(acc_tree, acc_pm):
        e_tree = np.linalg.norm(acc_tree - ft["acc_short"], axis=1) / a_norm
        e_pm = np.linalg.norm(acc_pm - (ft["acc"] - ft["acc_short"]), axis=1) / a_norm
        return _stats(e_tree), _stats(e_pm)

    e_or, e_32 = rel_err(acc_o, ft["acc"]), rel_err(r32["acc"][sel], ft["acc"])
    (t_or, pm_or), (t_32, pm_32) = split(acc_t_o, acc_pm_o[sel]), split(c32["acc_tree"][sel], c32["acc_pm"][sel])
    print("\n128^3 TreePM: oracle median %.3g p99 %.3g (tree %.3g/%.3g, pm %.3g/%.3g) | FP32 group median %.3g p99 "
          "%.3g (tree %.3g/%.3g, pm %.3g/%.3g)" % (*_stats(e_or), *t_or, *pm_or, *_stats(e_32), *t_32, *pm_32))
    assert np.median(e_32) <= np.median(e_or) * 1.05 and np.percentile(e_32, 99) <= np.percentile(e_or, 99) * 1.05
    assert t_32[0] <= t_or[0] * 1.05 and t_32[1] <= t_or[1] * 1.05
    # the PM part is the reference's mesh force at FP32 on the mesh: it matches the oracle's
    assert pm_32[1] <= pm_or[1] * 1.05


def test_uniform_32cubed_periodic_ewald(agrav, oracle):
    from arepo_amd import ics
    pos, mass, eps = ics.uniform_box(32768, seed=1)
    kw = dict(periodic=True, box_size=1.0, pmgrid=0, soft=(eps,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
    with agrav.Gravity(agrav.make_params(**kw, opening=1, precision=agrav.FP64, walk=agrav.WALK_PLAIN)) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r64 = g.get_accelerations()
        sel = np.sort(np.random.Generator(np.random.Philox(77)).choice(len(pos), 4096, replace=False))
        sel = sel.astype(np.int32)
        ft = g.gravity_forcetest(sel)
    with agrav.Gravity(agrav.make_params(**kw, opening=1, precision=agrav.FP32, walk=agrav.WALK_GROUP)) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r32 = g.get_accelerations()
    o = oracle.compute_grav_accelerations(oracle.make_params(**kw, opening=1), pos, mass, first_step=True)
    e = rel_err(r64["acc"], o["acc"])
    assert np.percentile(e, 99.9) < 1e-8, e.max()
    assert np.mean(r64["nint"] == o["nint"]) > 0.999
    e_or = rel_err(o["acc"][sel], ft["acc"])
    e_32 = rel_err(r32["acc"][sel], ft["acc"])
    print("\n32^3 Ewald: oracle median %.3g p99 %.3g | FP32 group median %.3g p99 %.3g"
          % (*_stats(e_or), *_stats(e_32)))
    assert np.median(e_32) <= np.median(e_or) * 1.05 and np.percentile(e_32, 99) <= np.percentile(e_or, 99) * 1.05


@pytest.fixture(scope="module")
def halo_disk():
    from arepo_amd import ics
    pos, mass, eps = ics.halo_disk(1 << 20, seed=3)
    sel = np.sort(np.random.Generator(np.random.Philox(33)).choice(len(pos), 16384, replace=False)).astype(np.int32)
    return pos, mass, eps, sel


@pytest.fixture(scope="module")
def halo_disk_reference(agrav, halo_disk):
    # FP64 direct sum on the subset, and OldAcc from a Barnes-Hut seeding pass (accel.c:62-72)
    pos, mass, eps, sel = halo_disk
    p = agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(eps,), err_tol_theta=0.7,
                          precision=agrav.FP64, walk=agrav.WALK_PLAIN)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        ft = g.gravity_forcetest(sel)
        g.construct_forcetree()
        g.gravity_tree(use_bh=True)
        g.gravity_force_finalize()
        oa = g.get_accelerations()["old_acc"]
    return ft, oa


@pytest.mark.parametrize("tol", [0.001, 0.0025, 0.01])
def test_halo_disk_precision_study(agrav, oracle, halo_disk, halo_disk_reference, tol):
    pos, mass, eps, sel = halo_disk
    ft, oa = halo_disk_reference
    out = {}
    for name, prec, walk in (("fp32_group", agrav.FP32, agrav.WALK_GROUP), ("fp64_group", agrav.FP64, agrav.WALK_GROUP),
                             ("fp64_plain", agrav.FP64, agrav.WALK_PLAIN)):
        p = agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(eps,), err_tol_force_acc=tol, opening=1,
                              precision=prec, walk=walk)
        with agrav.Gravity(p) as g:
            g.set_particles(pos, mass)
            g.set_old_acc(oa)
            g.compute_grav_accelerations(first_step=False)
            out[name] = g.get_accelerations()["acc"][sel]
    op = oracle.make_params(periodic=False, box_size=0.0, soft=(eps,), err_tol_force_acc=tol, opening=1)
    acc_o, _, _, _ = oracle.Tree(op, pos, mass).walk(False, old_acc=oa, targets=sel, nthreads=8)
    # the FP64 plain walk is the reference's walk
    assert np.percentile(rel_err(out["fp64_plain"], acc_o), 99.9) < 1e-8
    e_or = rel_err(acc_o, ft["acc"])
    row = {k: _stats(rel_err(v, ft["acc"])) for k, v in out.items()}
    print("\nhalo+disk tol %g: oracle median %.3g p99 %.3g | %s" % (
        tol, *_stats(e_or), " | ".join("%s median %.3g p99 %.3g" % (k, *v) for k, v in row.items())))
    for k in ("fp32_group", "fp64_group"):
        assert row[k][0] <= np.median(e_or) * 1.05, (k, row[k])
        assert row[k][1] <= np.percentile(e_or, 99) * 1.05, (k, row[k])
    # error follows the tolerance: well below 2 x ErrTolForceAcc at the median
    assert row["fp32_group"][0] < 2 * tol


def test_batched_unbinding_4096_groups_full_size(agrav, oracle):
    from arepo_amd import ics
    off, rel, mass, eps, cen = ics.unbinding_batch(4096, seed=4)
    n = int(off[-1])
    print("batch of %d particles generated" % n, flush=True)
    p = agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01,), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP, err_tol_theta=0.7, opening=0, leaf_bucket=4, eval_potential=True)
    groups = np.arange(0, 4096, 128)
    g0 = np.random.Generator(np.random.Philox(61))
    sel = np.concatenate([off[s] + np.sort(g0.choice(off[s + 1] - off[s], min(512, off[s + 1] - off[s]), replace=False))
                          for s in groups]).astype(np.int32)
    with agrav.Gravity(p) as g:
        g.batch_set_particles(off, rel, mass, eps, centres=cen)
        g.batch_autotune()
        g.compute_grav_accelerations(first_step=True)
        out = g.get_accelerations(fields=("acc", "pot"))
        tim = g.timings()
        ft = g.gravity_forcetest(sel)  # FP64 direct sum over each target's own group
    assert tim["walk_retries"] == 0
    assert np.all(np.isfinite(out["acc"])) and np.all(np.isfinite(out["pot"]))
    ep_gpu = np.abs(out["pot"][sel] - ft["pot"]) / np.abs(ft["pot"])
    ea_gpu = rel_err(out["acc"][sel], ft["acc"])
    ep_ref = np.empty(len(sel))
    k = 0
    for s in groups:
        sl = slice(off[s], off[s + 1])
        pot_o, _ = oracle.group_potential_bh(rel[sl].astype(np.float64) + cen[s], mass[sl], 2.8 * eps[s], 0.7)
        m = min(512, off[s + 1] - off[s])
        idx = sel[k:k + m] - off[s]
        ep_ref[k:k + m] = np.abs(pot_o[idx] - ft["pot"][k:k + m]) / np.abs(ft["pot"][k:k + m])
        k += m
    print("potential error median %.3g p99 %.3g (reference walk %.3g / %.3g); acceleration median %.3g p99 %.3g"
          % (np.median(ep_gpu), np.percentile(ep_gpu, 99), np.median(ep_ref), np.percentile(ep_ref, 99),
             np.median(ea_gpu), np.percentile(ea_gpu, 99)))
    assert np.median(ep_ref) < 2e-3  # the reference walk and the direct sum agree on the convention
    assert np.median(ep_gpu) <= 1.05 * np.median(ep_ref) + 1e-6
    assert np.percentile(ep_gpu, 99) <= 1.05 * np.percentile(ep_ref, 99) + 1e-6
    assert np.median(ea_gpu) < 5e-3 and np.percentile(ea_gpu, 99) < 5e-2
