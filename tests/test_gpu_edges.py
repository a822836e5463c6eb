"""GPU edge cases of the gravity step, through the C ABI.

* empty particle sets (a rank with no particles is legal in the reference);
* coincident particles: AREPO splits cells down to level 30 and then puts
  coincident points into free slots (forcetree.c:861-877); this library keeps
  a cell still shared at level 30 as one multi-particle leaf.  Accelerations
  and potentials equal the oracle's anyway (all members of such a cell sit at
  one point, so a deep node and its particles contribute the same pair terms);
  only interaction counts may differ;
* a dense clump inside one level-21 cell: its particles tie in the first key
  word and are ordered by the second (runs longer than 1024 are sorted in
  global memory); the octree below level 21 must be the oracle's;
* several softening types (the node softening test, forcetree_walk.c:294-320);
* positions one box outside a periodic box, wrapped once as forcetree.c:404-409;
* device-resident input (agrav_set_particles_device) equal to host input.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _clumped(n, seed, periodic):
    """Uniform particles plus exact duplicates: one clump of 40 copies of a point and
    20 pairs of coincident particles."""
    g = np.random.Generator(np.random.Philox(seed))
    pos = g.random((n, 3)) * (1.0 if periodic else 0.8) + (0.0 if periodic else 0.1)
    pos[1:41] = pos[0]
    for k in range(20):
        pos[100 + 2 * k + 1] = pos[100 + 2 * k]
    mass = g.uniform(0.5, 1.5, n) / n
    return pos, mass


@pytest.mark.parametrize("case", ["isolated", "treepm"])
def test_empty_set(agrav, case):
    kw = dict(periodic=False, box_size=0.0, pmgrid=0) if case == "isolated" else dict(periodic=True, box_size=1.0,
                                                                                    pmgrid=32)
    for prec, walk in ((agrav.FP64, agrav.WALK_PLAIN), (agrav.FP32, agrav.WALK_GROUP)):
        with agrav.Gravity(agrav.make_params(**kw, soft=(0.01,), precision=prec, walk=walk)) as g:
            g.set_particles(np.zeros((0, 3)), np.zeros(0))
            g.compute_grav_accelerations(first_step=True)
            g.compute_grav_accelerations(first_step=False)
            r = g.get_accelerations()
            assert r["acc"].shape == (0, 3) and r["pot"].shape == (0,)
            assert g.timings()["interactions"] == 0
            # the context stays usable for a non-empty set afterwards
            pos = np.array([[0.25, 0.5, 0.5], [0.75, 0.5, 0.5]])
            g.set_particles(pos, np.ones(2))
            g.compute_grav_accelerations(first_step=True)
            assert np.all(np.isfinite(g.get_accelerations()["acc"]))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_coincident_particles_direct_limit(agrav, oracle, precision):
    """With a vanishing opening angle every node is opened and the walk is the direct sum,
    whatever the leaf layout of the duplicates."""
    pos, mass = _clumped(3000, 5, periodic=False)
    prec = agrav.FP64 if precision == "fp64" else agrav.FP32
    ad, pd = oracle.direct_summation(oracle.make_params(periodic=False, box_size=0.0, soft=(0.004,)), pos, mass)
    for walk in (agrav.WALK_PLAIN, agrav.WALK_GROUP) if precision == "fp64" else (agrav.WALK_GROUP,):
        gp = agrav.make_params(periodic=False, box_size=0.0, soft=(0.004,), opening=0, err_tol_theta=1e-4,
                               precision=prec, walk=walk)
        with agrav.Gravity(gp) as g:
            g.set_particles(pos, mass)
            g.compute_grav_accelerations()
            r = g.get_accelerations()
        tol = 1e-10 if precision == "fp64" else 2e-5
        assert np.percentile(rel_err(r["acc"], ad), 99.9) < tol, walk
        np.testing.assert_allclose(r["pot"], pd, rtol=tol)
        # each clump member feels the same force (they sit at one point)
        assert rel_err(r["acc"][1:41], np.repeat(r["acc"][:1], 40, axis=0)).max() < (1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("case", ["isolated_rel", "treepm_rel"])
def test_coincident_particles_match_oracle(agrav, oracle, case):
    periodic = case == "treepm_rel"
    pos, mass = _clumped(6000, 7, periodic)
    kw = dict(periodic=periodic, box_size=1.0 if periodic else 0.0, pmgrid=32 if periodic else 0, soft=(0.003,),
              err_tol_force_acc=0.0025, opening=1)
    gp = agrav.make_params(**kw, precision=agrav.FP64, walk=agrav.WALK_PLAIN)
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
    o = oracle.compute_grav_accelerations(oracle.make_params(**kw), pos, mass, first_step=True)
    e = rel_err(r["acc"], o["acc"])
    assert np.percentile(e, 99.9) < 1e-9, e.max()
    np.testing.assert_allclose(r["pot"], o["pot"], rtol=1e-9, atol=1e-12 * np.abs(o["pot"]).max())


@pytest.mark.parametrize("case", ["isolated_rel", "treepm_rel"])
def test_several_softening_types(agrav, oracle, case):
    periodic = case == "treepm_rel"
    g0 = np.random.Generator(np.random.Philox(41))
    n = 8000
    pos = g0.random((n, 3))
    # a dense clump so that softened pairs and node softening tests occur
    pos[:2000] = 0.5 + 0.02 * (g0.random((2000, 3)) - 0.5)
    mass = np.full(n, 1.0 / n)
    stype = g0.integers(0, 3, n).astype(np.uint8)
    soft = (0.0005, 0.002, 0.006)
    kw = dict(periodic=periodic, box_size=1.0 if periodic else 0.0, pmgrid=32 if periodic else 0, soft=soft,
              err_tol_force_acc=0.0025, opening=1)
    out = {}
    for name, prec, walk in (("fp64_plain", agrav.FP64, agrav.WALK_PLAIN), ("fp32_group", agrav.FP32, agrav.WALK_GROUP),
                             ("fp64_group", agrav.FP64, agrav.WALK_GROUP)):
        with agrav.Gravity(agrav.make_params(**kw, precision=prec, walk=walk)) as g:
            g.set_particles(pos, mass, softtype=stype)
            g.compute_grav_accelerations(first_step=True)
            out[name] = g.get_accelerations()
    op = oracle.make_params(**kw)
    o = oracle.compute_grav_accelerations(op, pos, mass, stype=stype, first_step=True)
    e = rel_err(out["fp64_plain"]["acc"], o["acc"])
    assert np.percentile(e, 99.9) < 1e-9, e.max()
    assert np.mean(out["fp64_plain"]["nint"] == o["nint"]) > 0.999
    # group walks: error against the direct sum no worse than the reference walk's
    if periodic:
        with agrav.Gravity(agrav.make_params(**kw)) as g:
            g.set_particles(pos, mass, softtype=stype)
            sel = np.arange(0, n, 4, dtype=np.int32)
            ref = g.gravity_forcetest(sel)["acc"]
    else:
        sel = np.arange(n)
        ref = oracle.direct_summation(op, pos, mass, stype=stype)[0]
    eo = rel_err(o["acc"][sel], ref)
    for name in ("fp32_group", "fp64_group"):
        eg = rel_err(out[name]["acc"][sel], ref)
        assert np.median(eg) <= 1.05 * np.median(eo) + 2e-6, (name, np.median(eg), np.median(eo))
        assert np.percentile(eg, 99) <= 1.05 * np.percentile(eo, 99) + 2e-6, (name, np.percentile(eg, 99))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_types_with_one_softening_equal_no_types(agrav, prec):
    """With a one-entry softening table, per-particle types (all 0) change nothing: the same
    results, bit for bit, as a set without types (the single-softening walk runs for both)."""
    g0 = np.random.Generator(np.random.Philox(47))
    n = 6000
    pos = g0.random((n, 3))
    pos[:1500] = 0.5 + 0.02 * (g0.random((1500, 3)) - 0.5)
    mass = np.full(n, 1.0 / n)
    kw = dict(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,), err_tol_force_acc=0.0025, opening=1,
              precision=agrav.FP32 if prec == "fp32" else agrav.FP64, walk=agrav.WALK_GROUP)
    out = []
    for st in (None, np.zeros(n, np.uint8)):
        with agrav.Gravity(agrav.make_params(**kw)) as g:
            g.set_particles(pos, mass, softtype=st)
            g.compute_grav_accelerations(first_step=True)
            g.compute_grav_accelerations(first_step=False)
            out.append(g.get_accelerations())
    for k in ("acc", "nint", "old_acc"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_shorter_softening_table_must_cover_types(agrav):
    g0 = np.random.Generator(np.random.Philox(53))
    n = 3000
    pos, mass = g0.random((n, 3)), np.full(n, 1.0 / n)
    stype = (np.arange(n) % 2).astype(np.uint8)
    with agrav.Gravity(agrav.make_params(soft=(0.002, 0.004))) as g:
        g.set_particles(pos, mass, softtype=stype)
        with pytest.raises(agrav.AgravError):
            g.set_softenings([0.002])  # type 1 would have no softening
        g.compute_grav_accelerations(first_step=True)  # the two-entry table still stands
        a2 = g.get_accelerations()["acc"]
        g.set_particles(pos, mass, softtype=np.zeros(n, np.uint8))
        g.set_softenings([0.002])
        g.compute_grav_accelerations(first_step=True)
        a1 = g.get_accelerations()["acc"]
    assert np.all(np.isfinite(a2)) and np.all(np.isfinite(a1))


def test_periodic_positions_wrapped_once(agrav, oracle):
    g0 = np.random.Generator(np.random.Philox(43))
    n = 5000
    pos = g0.random((n, 3))
    mass = np.full(n, 1.0 / n)
    shifted = pos.copy()
    shifted[::3, 0] -= 1.0
    shifted[1::3, 1] += 1.0
    kw = dict(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,), err_tol_force_acc=0.0025, opening=1)
    res = []
    for p in (pos, shifted):
        with agrav.Gravity(agrav.make_params(**kw, precision=agrav.FP64, walk=agrav.WALK_PLAIN)) as g:
            g.set_particles(p, mass)
            g.compute_grav_accelerations(first_step=True)
            res.append(g.get_accelerations())
    o = oracle.compute_grav_accelerations(oracle.make_params(**kw), shifted, mass, first_step=True)
    assert np.percentile(rel_err(res[1]["acc"], o["acc"]), 99.9) < 1e-9
    # the same particles, up to the rounding of one box shift
    assert np.percentile(rel_err(res[1]["acc"], res[0]["acc"]), 99) < 1e-6


class _DeviceBuffers:
    """Device copies of host arrays through the HIP runtime (hipMalloc / hipMemcpy)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        C = self.C
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 1))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p.value

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_device_input_equals_host_input(agrav):
    pos, mass = _clumped(4000, 9, periodic=True)
    stype = np.random.Generator(np.random.Philox(3)).integers(0, 2, len(pos)).astype(np.uint8)
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002, 0.004), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass, softtype=stype)
        g.compute_grav_accelerations(first_step=True)
        host = g.get_accelerations()
    dev_buf = _DeviceBuffers()
    try:
        dpos, dmass, dst = dev_buf.put(pos.astype(np.float64)), dev_buf.put(mass.astype(np.float64)), dev_buf.put(stype)
        bad = stype.copy()
        bad[17] = 2
        dbad = dev_buf.put(bad)
        with agrav.Gravity(p) as g:
            g.set_particles_device(len(pos), dpos, dmass, dst)
            g.compute_grav_accelerations(first_step=True)
            dev = g.get_accelerations()
            assert np.array_equal(dev["acc"], host["acc"]) and np.array_equal(dev["pot"], host["pot"])
            # a softening type beyond the table is rejected on the device copy too
            with pytest.raises(agrav.AgravError) as ei:
                g.set_particles_device(len(pos), dpos, dmass, dbad)
            assert ei.value.code == agrav.ERR_BAD_PARAM
    finally:
        dev_buf.free()


@pytest.mark.parametrize("bad", [2.5, float("nan")])
def test_pm_without_tree_rejects_positions_outside_box(agrav, bad):
    """The PM may run before any tree build (long_range_force right after the particle
    load); positions still outside the box after one wrap are reported, not deposited."""
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=16, soft=(0.01,))
    with agrav.Gravity(p) as g:
        g.set_particles(np.array([[0.5, 0.5, 0.5], [bad, 0.1, 0.1]]), np.ones(2))
        with pytest.raises(agrav.AgravError) as ei:
            g.long_range_force()
        assert ei.value.code == agrav.ERR_OUT_OF_BOX
        with pytest.raises(agrav.AgravError) as ei:
            g.construct_forcetree()
        assert ei.value.code == agrav.ERR_OUT_OF_BOX
        # one box outside is wrapped and accepted
        g.set_particles(np.array([[0.5, 0.5, 0.5], [-0.75, 1.1, 0.1]]), np.ones(2))
        g.long_range_force()
        pm = g.get_components(tree=False, pm=True)["acc_pm"]
        assert np.all(np.isfinite(pm))


@pytest.mark.parametrize("nclump", [300, 1500])
def test_clump_below_level_21_matches_oracle(agrav, oracle, nclump):
    # uniform particles in an isolated domain, plus a clump of nclump particles filling the middle fifth
    # of one level-21 cell of that domain (1500 > 1024: the tie run is sorted in global memory)
    g = np.random.Generator(np.random.Philox(123))
    nu = 4000
    uni = g.random((nu, 3))
    p = oracle.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(1e-9,), err_tol_theta=0.7,
                           err_tol_force_acc=0.0025)
    corner, center, length = oracle.domain_extent(p, uni)
    cell = length / 2 ** 21
    origin = corner + (np.array([2 ** 20, 2 ** 20 - 7, 2 ** 20 + 5]) * cell)
    clump = origin + (0.4 + 0.2 * g.random((nclump, 3))) * cell
    pos = np.concatenate([uni, clump])
    mass = np.full(len(pos), 1.0 / len(pos))
    kw = dict(periodic=False, box_size=0.0, pmgrid=0, soft=(1e-9,), err_tol_theta=0.7, err_tol_force_acc=0.0025)
    o = oracle.compute_grav_accelerations(oracle.make_params(**kw, opening=1), pos, mass, first_step=True)
    with agrav.Gravity(agrav.make_params(**kw, opening=1, precision=agrav.FP64, walk=agrav.WALK_PLAIN)) as gr:
        gr.set_particles(pos, mass)
        gr.compute_grav_accelerations(first_step=True)
        r = gr.get_accelerations()
        tree = gr.get_tree()
    ot = oracle.Tree(oracle.make_params(**kw), pos, mass).export()
    # the same cubes as the oracle's insertion tree, down to its deepest level (above 21 here)
    assert ot["level"].max() > 24
    assert len(tree["level"]) == len(ot["level"]) and tree["level"].max() == ot["level"].max()
    assert np.array_equal(np.sort(tree["len"]), np.sort(ot["len"]))
    e = rel_err(r["acc"], o["acc"])
    assert np.percentile(e, 99.9) < 1e-8, float(e.max())
    assert np.mean(r["nint"] == o["nint"]) > 0.999


def test_rejected_set_particles_state(agrav):
    """A set_particles rejected by its host-side checks leaves the context as it was
    (the old set and its results stay valid); one that fails after the old set was
    given up (a softening type beyond the table found on the device copy) leaves the
    context without particles: later calls report a state error instead of reading
    a mix of old and new buffers.  A good set afterwards works as on a fresh context."""
    pos, mass = _clumped(3000, 21, periodic=True)
    p = agrav.make_params(periodic=True, box_size=1.0, pmgrid=16, soft=(0.002, 0.004), precision=agrav.FP32,
                          walk=agrav.WALK_GROUP)
    bad = np.zeros(len(pos) + 500, dtype=np.uint8)
    bad[7] = 5  # beyond the two softening types
    pos2 = np.random.Generator(np.random.Philox(22)).random((len(pos) + 500, 3))
    m2 = np.full(len(pos2), 1.0 / len(pos2))
    dev_buf = _DeviceBuffers()
    try:
        with agrav.Gravity(p) as g:
            g.set_particles(pos, mass)
            g.compute_grav_accelerations(first_step=True)
            first = g.get_accelerations()
            with pytest.raises(agrav.AgravError) as ei:
                g.set_particles(pos2, m2, softtype=bad)
            assert ei.value.code == agrav.ERR_BAD_PARAM
            kept = g.get_accelerations()
            assert np.array_equal(kept["acc"], first["acc"])
            with pytest.raises(agrav.AgravError) as ei:
                g.set_particles_device(len(pos2), dev_buf.put(pos2), dev_buf.put(m2), dev_buf.put(bad))
            assert ei.value.code == agrav.ERR_BAD_PARAM
            for call in (lambda: g.compute_grav_accelerations(first_step=True), g.construct_forcetree,
                         g.long_range_force, g.get_accelerations):
                with pytest.raises(agrav.AgravError) as ei:
                    call()
                assert ei.value.code == agrav.ERR_STATE
            g.set_particles(pos, mass)
            g.compute_grav_accelerations(first_step=True)
            again = g.get_accelerations()
    finally:
        dev_buf.free()
    assert np.array_equal(again["acc"], first["acc"])


def test_page_locked_host_arrays(agrav):
    """Inputs staged in and results copied to agrav_host_alloc memory give the bits of pageable arrays."""
    g0 = np.random.Generator(np.random.Philox(59))
    n = 20000
    pos, mass = g0.random((n, 3)), g0.random(n) / n
    kw = dict(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,), precision=agrav.FP32, walk=agrav.WALK_GROUP)
    res = []
    for pinned in (False, True):
        with agrav.Gravity(agrav.make_params(**kw)) as g:
            if pinned:
                hp, hm = agrav.host_empty((n, 3)), agrav.host_empty(n)
                hp[:], hm[:] = pos, mass
                out = {"acc": agrav.host_empty((n, 3)), "pot": agrav.host_empty(n)}
                g.set_particles(hp, hm)
            else:
                out = None
                g.set_particles(pos, mass)
            g.compute_grav_accelerations(first_step=True)
            r = g.get_accelerations(fields=("acc", "pot"), out=out)
            res.append({k: np.array(r[k]) for k in ("acc", "pot")})
    assert np.array_equal(res[0]["acc"], res[1]["acc"]) and np.array_equal(res[0]["pot"], res[1]["pot"])
