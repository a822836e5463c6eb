"""GPU tests of how one full step is scheduled, through the C ABI.

compute_grav_accelerations of every particle (accel.c:58-192) runs the periodic
PM on a second stream beside the octree kernels (it needs only the
Hilbert-ordered particles of the build's gather) and reads the build's error
flags back with the walk's counters, after the walk: the walk kernels test the
same flags on the device before they start.  These tests pin that the PM of
such a step is bitwise the PM run alone on the main stream, that a build error
is still reported with the build's own code, and that the context recovers.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _box(n, seed):
    g = np.random.Generator(np.random.Philox(seed))
    pos = np.concatenate([g.random((n // 2, 3)), (0.5 + 0.05 * g.standard_normal((n - n // 2, 3))) % 1.0])
    return pos, np.full(n, 1.0 / n)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_step_pm_equals_pm_on_main_stream(agrav, precision):
    """The PM of a full step (PM stream) is bitwise the PM of long_range_force run on the
    main stream after the same build."""
    pos, mass = _box(30000, 7)
    prec = agrav.FP32 if precision == "fp32" else agrav.FP64
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,), precision=prec, walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        step = g.get_components(tree=False, pm=True)
        acc_step = g.get_accelerations()["acc"]
        g.construct_forcetree()
        g.long_range_force()
        alone = g.get_components(tree=False, pm=True)
        assert g.timings()["treebuild"] > 0
    assert np.array_equal(step["acc_pm"], alone["acc_pm"])
    assert np.array_equal(step["pot_pm"], alone["pot_pm"])
    assert np.all(np.isfinite(acc_step))


@pytest.mark.parametrize("pmgrid", [0, 32])
def test_step_reports_build_error_after_walk(agrav, pmgrid):
    """A particle still outside the periodic box after one wrap: the step returns the
    build's error code (read back after the walk), and the context then solves a valid
    set exactly as a fresh context does."""
    pos, mass = _box(5000, 11)
    bad = pos.copy()
    bad[17] = (2.6, 0.3, 0.3)
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=pmgrid, soft=(0.002,), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(bad, mass)
        for first in (True, False):
            with pytest.raises(agrav.AgravError) as ei:
                g.compute_grav_accelerations(first_step=first)
            assert ei.value.code == agrav.ERR_OUT_OF_BOX
        # no results from the failed step
        with pytest.raises(agrav.AgravError) as ei:
            g.gravity_tree()
        assert ei.value.code == agrav.ERR_STATE
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
    with agrav.Gravity(p) as g2:
        g2.set_particles(pos, mass)
        g2.compute_grav_accelerations(first_step=True)
        r2 = g2.get_accelerations()
    assert np.array_equal(r["acc"], r2["acc"])
    assert np.array_equal(r["nint"], r2["nint"])


def test_build_error_with_ordered_walk_then_valid_step(agrav):
    """The same with enough groups for the costliest-first order (sorted beside the
    build from the last walk's costs): repeated failing steps with the order queued,
    then a valid set equals a fresh context's result."""
    pos, mass = _box(600000, 21)
    bad = pos.copy()
    bad[123] = (2.6, 0.3, 0.3)
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=64, soft=(0.002,), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP, group_size=64, group_waves=1)
    with agrav.Gravity(p) as g:
        g.set_particles(bad, mass)
        for first in (True, False, False):
            with pytest.raises(agrav.AgravError) as ei:
                g.compute_grav_accelerations(first_step=first)
            assert ei.value.code == agrav.ERR_OUT_OF_BOX
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        g.compute_grav_accelerations(first_step=False)
        assert g.timings()["walk_ordered"] == 1
        r = g.get_accelerations()
    with agrav.Gravity(p) as g2:
        g2.set_particles(pos, mass)
        g2.compute_grav_accelerations(first_step=True)
        g2.compute_grav_accelerations(first_step=False)
        r2 = g2.get_accelerations()
    for f in ("acc", "pot", "nint", "old_acc"):
        assert np.array_equal(r[f], r2[f]), f


def test_step_timings_cover_build_pm_and_walk(agrav):
    """The deferred build check still reports the build's duration, and the PM's."""
    pos, mass = _box(20000, 13)
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        for _ in range(2):
            g.compute_grav_accelerations(first_step=False)
            t = g.timings()
            assert t["treebuild"] > 0 and t["pm_grav"] > 0 and t["treewalk"] > 0
            assert t["interactions"] > 0 and t["numnodes"] > 0
        a = g.get_accelerations()["acc"]
    # the same steps on a second context give the same bits (the PM stream adds no
    # dependence on timing)
    with agrav.Gravity(p) as g2:
        g2.set_particles(pos, mass)
        g2.compute_grav_accelerations(first_step=True)
        g2.compute_grav_accelerations(first_step=False)
        g2.compute_grav_accelerations(first_step=False)
        b = g2.get_accelerations()["acc"]
    assert rel_err(a, b).max() == 0.0


@pytest.mark.parametrize("pmgrid,n,gs,waves", [(0, 600000, 64, 1), (64, 600000, 64, 1), (0, 50000, 16, 4)])
def test_costliest_first_walk_equals_hilbert_order_walk(agrav, pmgrid, n, gs, waves):
    """A step whose walk groups start costliest first (by the tile entries each group
    streamed in the last walk of the same grouping, AREPO's GravCost idea) gives bitwise
    the accelerations, potentials and counts of the same step launched in Hilbert order:
    each group's sums do not depend on when it runs.  600,000 particles make 9,375 groups
    of 64; 50,000 particles in groups of 16 shared by 4 wavefronts make 3,125 groups of
    12,500 wavefronts; both are enough for the order to be used."""
    pos, mass = _box(n, 5)
    g0 = np.random.Generator(np.random.Philox(9))
    old = (0.5 + g0.random(n)).astype(np.float32)
    kw = dict(periodic=True, box_size=1.0, pmgrid=pmgrid) if pmgrid else dict(periodic=False, box_size=0.0, pmgrid=0)
    p = agrav.make_params(**kw, soft=(0.002,), precision=agrav.FP32, walk=agrav.WALK_GROUP, group_size=gs,
                          group_waves=waves, eval_potential=True)
    runs = []
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass)
        for _ in range(2):
            g.set_old_acc(old)
            g.compute_grav_accelerations(first_step=False)
            runs.append((g.get_accelerations(), g.timings()))
        # a new set of another size has no counts yet: Hilbert order again
        g.set_particles(pos[: n - 64], mass[: n - 64])
        g.set_old_acc(old[: n - 64])
        g.compute_grav_accelerations(first_step=False)
        assert g.timings()["walk_ordered"] == 0
    (r0, t0), (r1, t1) = runs
    assert t0["walk_ordered"] == 0 and t1["walk_ordered"] == 1
    for k in ("acc", "pot", "nint", "old_acc"):
        assert np.array_equal(r0[k], r1[k]), k
    assert t0["interactions"] == t1["interactions"] and t0["tile_entries"] == t1["tile_entries"]
