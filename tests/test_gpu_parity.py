"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle.

Bars: integer/topology data bit-exact; FP64 paths within 1e-9 relative of
the oracle (summation order differs); FP32 paths assessed by their error
against FP64 direct summation, which must not exceed the oracle's own
error by more than the stated margin.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _uniform(n, seed):
    g = np.random.Generator(np.random.Philox(seed))
    return g.random((n, 3)), np.full(n, 1.0 / n)


def _clustered(n, seed, periodic=False):
    from arepo_amd import ics
    g = ics.rng(seed)
    nh = n // 2
    parts = [g.random((n - nh, 3))]
    cent = g.random((8, 3))
    cnt = np.full(8, nh // 8)
    cnt[0] += nh - cnt.sum()
    for c, k in zip(cent, cnt):
        r = ics.hernquist_radii(int(k), 0.02, g)
        parts.append(c + ics._isotropic(r, g))
    pos = np.concatenate(parts)
    if periodic:
        pos %= 1.0
    return pos, np.full(n, 1.0 / n)


def _tree_key(center, ln):
    return np.lexsort((center[:, 2], center[:, 1], center[:, 0], ln))


@pytest.mark.parametrize("periodic", [False, True])
def test_tree_matches_oracle(agrav, oracle, periodic):
    pos, mass = _clustered(6000, 11, periodic)
    box = 1.0 if periodic else 0.0
    gp = agrav.make_params(periodic=periodic, box_size=box, soft=(0.001,))
    op = oracle.make_params(periodic=periodic, box_size=box, soft=(0.001,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        m = g.construct_forcetree()
        tg = g.get_tree()
        perm, keys = g.get_order()
    to = oracle.Tree(op, pos, mass).export()
    assert m == len(to["len"])
    assert np.all(np.diff(keys.astype(np.uint64)) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(len(pos)))
    if periodic:
        # unit box: the domain cube is 1.00001 box around the box centre (domain_box.c:161-168); AREPO's
        # domain key is peano_hilbert_key of the top 42 mantissa bits of (x - corner) / len + 1
        # (domain_toplevel.c:121-126, domain.c:361-370); the first key word holds its top 21 levels
        ln = 1.0 * 1.00001
        corner = 0.5 * (0.0 + 1.0) - 0.5 * ln
        v = (np.mod(pos[perm], 1.0) - corner) * (1.0 / ln) + 1.0
        ic = (v.view(np.uint64) & np.uint64((1 << 52) - 1)) >> np.uint64(10)
        want = np.array([oracle.peano_hilbert_key(int(a), int(b), int(c), 42) >> 63 for a, b, c in ic], dtype=np.uint64)
        assert np.array_equal(keys.astype(np.uint64), want)
    ig = _tree_key(tg["center"], tg["len"])
    io = _tree_key(to["center"], to["len"])
    # geometry is bit-identical (same integer formula); topology identical
    assert np.array_equal(tg["center"][ig], to["center"][io])
    assert np.array_equal(tg["len"][ig], to["len"][io])
    assert np.array_equal(tg["level"][ig], to["level"][io])
    assert np.array_equal(tg["npart"][ig], to["npart"][io])
    np.testing.assert_allclose(tg["mass"][ig], to["mass"][io], rtol=1e-13)
    np.testing.assert_allclose(tg["com"][ig], to["com"][io], rtol=1e-12, atol=1e-15)
    assert tg["npart"][0] == len(pos) and tg["level"][0] == 0


def _walk_both(agrav, oracle, pos, mass, periodic, pmgrid, use_bh, old_acc=None, walk=None, precision=None, eps=0.002,
               errtol=0.0025, theta=0.7):
    box = 1.0 if periodic else 0.0
    kw = dict(periodic=periodic, box_size=box, pmgrid=pmgrid, soft=(eps,), err_tol_force_acc=errtol, err_tol_theta=theta)
    gp = agrav.make_params(**kw, walk=agrav.WALK_PLAIN if walk is None else walk,
                           precision=agrav.FP64 if precision is None else precision)
    op = oracle.make_params(**kw)
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        if old_acc is not None:
            g.set_old_acc(old_acc)
        g.construct_forcetree()
        g.gravity_tree(use_bh=use_bh)
        comp = g.get_components(tree=True)
        tim = g.timings()
    t = oracle.Tree(op, pos, mass)
    ewt = oracle.ewald_table_tree(box) if (periodic and pmgrid == 0) else None
    acc, pot, nint, tot = t.walk(use_bh, old_acc=old_acc, ewald_tab=ewt)
    return comp, acc, pot, nint, tot, tim


@pytest.mark.parametrize("case", ["iso_bh", "iso_rel", "treepm_rel", "ewald_bh"])
def test_plain_walk_matches_oracle(agrav, oracle, case):
    if case == "iso_bh":
        pos, mass = _clustered(5000, 3)
        args = dict(periodic=False, pmgrid=0, use_bh=True)
    elif case == "iso_rel":
        pos, mass = _clustered(5000, 4)
        oa = np.random.Generator(np.random.Philox(9)).uniform(0.5, 2.0, len(pos)).astype(np.float32)
        args = dict(periodic=False, pmgrid=0, use_bh=False, old_acc=oa)
    elif case == "treepm_rel":
        pos, mass = _clustered(8000, 5, periodic=True)
        oa = np.random.Generator(np.random.Philox(8)).uniform(0.5, 2.0, len(pos)).astype(np.float32)
        args = dict(periodic=True, pmgrid=32, use_bh=False, old_acc=oa)
    else:
        pos, mass = _uniform(2000, 6)
        args = dict(periodic=True, pmgrid=0, use_bh=True)
    comp, acc, pot, nint, tot, tim = _walk_both(agrav, oracle, pos, mass, **args)
    e = rel_err(comp["acc_tree"], acc)
    assert np.percentile(e, 99.9) < 1e-9, np.max(e)
    np.testing.assert_allclose(comp["pot_tree"], pot, rtol=1e-9, atol=1e-12 * np.abs(pot).max())
    assert tim["interactions"] == tot or abs(tim["interactions"] - tot) <= 1e-4 * tot


def _direct(oracle, pos, mass, periodic, pmgrid, eps):
    box = 1.0 if periodic else 0.0
    op = oracle.make_params(periodic=periodic, box_size=box, pmgrid=pmgrid, soft=(eps,))
    return oracle.direct_summation(op, pos, mass)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["iso_rel", "treepm_rel", "ewald_rel"])
def test_group_walk_accuracy(agrav, oracle, precision, case):
    """Group walk: its error against direct summation (same pair kernel) is at most the
    per-particle walk's (AREPO's) error, up to FP32 rounding, and it never does fewer interactions."""
    prec = agrav.FP64 if precision == "fp64" else agrav.FP32
    if case == "iso_rel":
        pos, mass = _clustered(6000, 21)
        periodic, pmgrid = False, 0
    elif case == "treepm_rel":
        pos, mass = _clustered(8000, 22, periodic=True)
        periodic, pmgrid = True, 32
    else:
        # tree-only periodic: sources anywhere in the box, pair offsets up to half a box
        pos, mass = _uniform(6000, 23)
        periodic, pmgrid = True, 0
    if case == "ewald_rel":
        # the periodic (Ewald) direct sum of FORCETEST, pinned against the oracle elsewhere
        with agrav.Gravity(agrav.make_params(periodic=True, box_size=1.0, pmgrid=0, soft=(0.002,))) as g:
            g.set_particles(pos, mass)
            ad = g.gravity_forcetest(np.arange(len(pos), dtype=np.int32))["acc"]
    else:
        ad, _ = _direct(oracle, pos, mass, periodic, pmgrid, 0.002)
    oa = np.linalg.norm(ad, axis=1).astype(np.float32)
    comp, acc, pot, nint, tot, tim = _walk_both(agrav, oracle, pos, mass, periodic, pmgrid, False, old_acc=oa,
                                                walk=agrav.WALK_GROUP, precision=prec)
    eg = rel_err(comp["acc_tree"], ad)
    eo = rel_err(acc, ad)
    slack = 1.0 if precision == "fp64" else 1.05
    floor = 1e-12 if precision == "fp64" else 2e-6
    assert np.median(eg) <= slack * np.median(eo) + floor, (np.median(eg), np.median(eo))
    assert np.percentile(eg, 99) <= slack * np.percentile(eo, 99) + floor, (np.percentile(eg, 99), np.percentile(eo, 99))
    assert tim["interactions"] >= tot


def test_pm_matches_oracle(agrav, oracle):
    pos, mass = _clustered(20000, 31, periodic=True)
    gp = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,))
    op = oracle.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        g.long_range_force()
        c = g.get_components(tree=False, pm=True)
    acc, pot = oracle.pm_periodic(op, pos, mass)
    e = rel_err(c["acc_pm"], acc)
    assert np.max(e) < 1e-9
    np.testing.assert_allclose(c["pot_pm"], pot, rtol=1e-9, atol=1e-12)


def test_pm_fp32_close(agrav, oracle):
    pos, mass = _clustered(20000, 32, periodic=True)
    gp = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,), precision=agrav.FP32)
    op = oracle.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        g.long_range_force()
        c = g.get_components(tree=False, pm=True)
    acc, pot = oracle.pm_periodic(op, pos, mass)
    e = rel_err(c["acc_pm"], acc)
    assert np.percentile(e, 99) < 1e-4


@pytest.mark.parametrize("case", ["treepm", "iso", "ewald"])
def test_compute_matches_oracle(agrav, oracle, case):
    if case == "treepm":
        pos, mass = _clustered(8000, 41, periodic=True)
        kw = dict(periodic=True, box_size=1.0, pmgrid=32)
    elif case == "iso":
        pos, mass = _clustered(6000, 42)
        kw = dict(periodic=False, box_size=0.0, pmgrid=0)
    else:
        pos, mass = _uniform(2000, 43)
        kw = dict(periodic=True, box_size=1.0, pmgrid=0)
    gp = agrav.make_params(**kw, soft=(0.002,), walk=agrav.WALK_PLAIN)
    op = oracle.make_params(**kw, soft=(0.002,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations(first_step=True)
        r = g.get_accelerations()
    o = oracle.compute_grav_accelerations(op, pos, mass, first_step=True)
    e = rel_err(r["acc"], o["acc"])
    assert np.percentile(e, 99.9) < 1e-8, np.max(e)
    np.testing.assert_allclose(r["pot"], o["pot"], rtol=1e-8, atol=1e-10 * np.abs(o["pot"]).max())
    assert np.mean(r["nint"] == o["nint"]) > 0.999
    np.testing.assert_allclose(r["old_acc"], o["old_acc"], rtol=1e-5)


def test_forcetest_matches_oracle(agrav, oracle):
    pos, mass = _clustered(3000, 51, periodic=True)
    sel = np.arange(0, 3000, 97, dtype=np.int32)
    gp = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,))
    op = oracle.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        r = g.gravity_forcetest(sel)
        tabs = g.get_tables(ewald_forcetest=True)
    # the GPU's 129^3 table equals the exact series at its grid points
    tab = tabs["ewald_forcetest"].reshape(4, 129, 129, 129)
    gsel = np.random.Generator(np.random.Philox(5)).integers(0, 129, (40, 3))
    for i, j, k in gsel:
        ex = oracle.forcetest_ewald_exact(1.0, 0.5 * i / 128, 0.5 * j / 128, 0.5 * k / 128)
        np.testing.assert_allclose(tab[:, i, j, k], ex, rtol=1e-10, atol=1e-10)
    # full direct sum with the exact series (tiny target set)
    acc, accs, pot, pots = oracle.forcetest(op, pos, mass, sel, ewald=2)
    e = rel_err(r["acc"], acc)
    assert np.max(e) < 2e-4  # trilinear table vs exact series
    e2 = rel_err(r["acc_short"], accs)
    assert np.max(e2) < 1e-10
    np.testing.assert_allclose(r["pot_short"], pots, rtol=1e-10)


def test_ewald_tree_table(agrav, oracle):
    gp = agrav.make_params(periodic=True, box_size=2.0, pmgrid=0, soft=(0.01,))
    with agrav.Gravity(gp) as g:
        tabs = g.get_tables(ewald_tree=True)
    tab = tabs["ewald_tree"].reshape(4, 65, 65, 65)
    idx = np.random.Generator(np.random.Philox(6)).integers(0, 65, (30, 3))
    for i, j, k in idx:
        f = oracle.ewald_force(0.5 * i / 64, 0.5 * j / 64, 0.5 * k / 64)
        p = oracle.ewald_psi(0.5 * i / 64, 0.5 * j / 64, 0.5 * k / 64)
        np.testing.assert_allclose(tab[:3, i, j, k], f / 4.0, rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(tab[3, i, j, k], p / 2.0, rtol=1e-11, atol=1e-13)
    tf, tp = oracle.short_range_tables(4.5)
    assert np.array_equal(tabs["tab_force"], tf) and np.array_equal(tabs["tab_pot"], tp)


def test_direct_matches_oracle(agrav, oracle):
    pos, mass = _clustered(1500, 61, periodic=True)
    gp = agrav.make_params(periodic=True, box_size=1.0, pmgrid=16, soft=(0.002,))
    op = oracle.make_params(periodic=True, box_size=1.0, pmgrid=16, soft=(0.002,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        acc, pot = g.gravity_direct()
    ao, po = oracle.direct_summation(op, pos, mass)
    assert np.max(rel_err(acc, ao)) < 1e-11
    np.testing.assert_allclose(pot, po, rtol=1e-11)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_deterministic(agrav, precision):
    pos, mass = _clustered(30000, 71, periodic=True)
    prec = agrav.FP64 if precision == "fp64" else agrav.FP32
    gp = agrav.make_params(periodic=True, box_size=1.0, pmgrid=32, soft=(0.001,), precision=prec)
    out = []
    for _ in range(2):
        with agrav.Gravity(gp) as g:
            g.set_particles(pos, mass)
            g.compute_grav_accelerations(first_step=True)
            out.append(g.get_accelerations())
    assert np.array_equal(out[0]["acc"], out[1]["acc"])
    assert np.array_equal(out[0]["pot"], out[1]["pot"])


def test_error_paths(agrav):
    with pytest.raises(agrav.AgravError) as ei:
        agrav.Gravity(agrav.make_params(soft=(0.0,)))
    assert ei.value.code == agrav.ERR_BAD_PARAM
    with pytest.raises(agrav.AgravError) as ei:
        agrav.Gravity(agrav.make_params(pmgrid=48))
    assert ei.value.code == agrav.ERR_BAD_PARAM
    g = agrav.Gravity(agrav.make_params(periodic=True, pmgrid=0, soft=(0.01,)))
    with pytest.raises(agrav.AgravError) as ei:
        g.construct_forcetree()
    assert ei.value.code == agrav.ERR_STATE
    pos = np.array([[0.5, 0.5, 0.5], [2.7, 0.1, 0.1]])
    g.set_particles(pos, np.ones(2))
    with pytest.raises(agrav.AgravError) as ei:
        g.construct_forcetree()
    assert ei.value.code == agrav.ERR_OUT_OF_BOX
    with pytest.raises(agrav.AgravError) as ei:
        g.gravity_tree()
    assert ei.value.code == agrav.ERR_STATE
    g.close()


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_tiny_sets(agrav, oracle, n):
    g0 = np.random.Generator(np.random.Philox(100 + n))
    pos = g0.random((n, 3))
    mass = g0.uniform(0.5, 1.5, n)
    # Barnes-Hut with a tiny opening angle opens every node: the walk is the direct sum
    gp = agrav.make_params(periodic=False, box_size=0.0, soft=(0.01,), walk=agrav.WALK_GROUP, opening=0,
                           err_tol_theta=1e-4)
    op = oracle.make_params(periodic=False, box_size=0.0, soft=(0.01,))
    with agrav.Gravity(gp) as g:
        g.set_particles(pos, mass)
        g.compute_grav_accelerations()
        r = g.get_accelerations()
    ad, pd = oracle.direct_summation(op, pos, mass)
    np.testing.assert_allclose(r["acc"], ad, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r["pot"], pd, rtol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_group_walk_wavefront_split(agrav, periodic):
    """A group's traversal shared by 2 or 4 wavefronts evaluates the same entries as one
    wavefront (same interaction count); only the summation order differs, and every
    setting is deterministic."""
    pos, mass = _clustered(20000, 61, periodic=periodic)
    kw = dict(periodic=periodic, box_size=1.0 if periodic else 0.0, pmgrid=32 if periodic else 0, soft=(0.002,),
              precision=agrav.FP32, walk=agrav.WALK_GROUP)
    res = {}
    for waves in (1, 2, 4):
        outs = []
        for rep in range(2):
            with agrav.Gravity(agrav.make_params(**kw, group_waves=waves)) as g:
                g.set_particles(pos, mass)
                g.compute_grav_accelerations(first_step=True)
                outs.append((g.get_accelerations(), g.timings()["interactions"]))
        assert np.array_equal(outs[0][0]["acc"], outs[1][0]["acc"]), waves
        res[waves] = outs[0]
    for waves in (2, 4):
        assert res[waves][1] == res[1][1]
        np.testing.assert_array_equal(res[waves][0]["nint"], res[1][0]["nint"])
        assert rel_err(res[waves][0]["acc"], res[1][0]["acc"]).max() < 1e-4
