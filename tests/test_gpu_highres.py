"""GPU tests of the high-resolution PM region (PLACEHIGHRESREGION) through the C ABI.

A periodic TreePM box (PMGRID = 32) holds uniform low-resolution particles (type 2) and a clump of
high-resolution particles (type 1, mask 2) near the centre: the zoom set-up of the reference's
cosmo_zoom_gravity_only_3d example at a small size.  Against the oracle: the region's geometry,
the PM force of both grids, the FP64 per-target walk with Asmth[1] / Rcut[1] for the region's
targets (forces to 1e-8, equal interaction counts), FORCETEST and the direct sum with the per-target
split scale, re-sizing when a high-resolution particle leaves the region, and the FP32 group walk's
accuracy against the Ewald direct sum no worse than the oracle's.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

KW = dict(periodic=True, box_size=1.0, pmgrid=32, asmth=1.25, rcut=4.5, err_tol_theta=0.7, err_tol_force_acc=0.0025)


def _zoom(n_lo=24000, n_hi=12000, seed=5):
    g = np.random.Generator(np.random.Philox(seed))
    lo = g.random((n_lo, 3))
    hi = 0.5 + 0.03 * g.standard_normal((n_hi, 3))
    hi = np.clip(hi, 0.42, 0.58)
    pos = np.concatenate([lo, hi])
    mass = np.concatenate([np.full(n_lo, 0.8 / n_lo), np.full(n_hi, 0.2 / n_hi)])
    ptype = np.concatenate([np.full(n_lo, 2), np.full(n_hi, 1)]).astype(np.uint8)
    stype = np.concatenate([np.zeros(n_lo), np.ones(n_hi)]).astype(np.uint8)
    return pos, mass, ptype, stype


def _gpu(agrav, pos, mass, ptype, stype, precision, walk, first_step=True, old_acc=None):
    p = agrav.make_params(**KW, soft=(0.002, 0.0005), precision=precision, walk=walk, highres_mask=2,
                          enlarge_region=1.1)
    g = agrav.Gravity(p)
    g.set_particles(pos, mass, softtype=stype)
    g.set_particle_types(ptype)
    if old_acc is not None:
        g.set_old_acc(old_acc)
    g.compute_grav_accelerations(first_step=first_step)
    return g


def test_highres_step_matches_oracle(agrav, oracle):
    pos, mass, ptype, stype = _zoom()
    g = _gpu(agrav, pos, mass, ptype, stype, agrav.FP64, agrav.WALK_PLAIN)
    r = g.get_accelerations()
    reg = g.highres_region()
    comp = g.get_components(tree=True, pm=True)
    sel = np.sort(np.random.Generator(np.random.Philox(9)).choice(len(pos), 600, replace=False)).astype(np.int32)
    ft = g.gravity_forcetest(sel)
    g.close()
    op = oracle.make_params(**KW, soft=(0.002, 0.0005))
    hr = oracle.make_highres(2, 1.1)
    o = oracle.compute_grav_accelerations(op, pos, mass, stype=stype, first_step=True, highres=hr, ptype=ptype, nthreads=16)
    # the region (pm_init_regionsize)
    for k in ("xmin", "xmax", "corner", "upper"):
        np.testing.assert_allclose(reg[k], np.array(getattr(hr, k)), rtol=0, atol=1e-14)
    assert reg["meshsize"] == pytest.approx(hr.meshsize, rel=1e-14)
    assert reg["asmth"] == pytest.approx(hr.asmth, rel=1e-14) and reg["rcut"] == pytest.approx(hr.rcut, rel=1e-14)
    flags = oracle.highres_flags(op, hr, pos, ptype)
    assert flags.sum() > 12000 and flags.sum() < len(pos)
    # both grids' PM force
    acc_p, _ = oracle.pm_periodic(op, pos, mass)
    acc_h, _ = oracle.pm_highres(op, hr, pos, mass, ptype)
    assert np.percentile(rel_err(comp["acc_pm"], acc_p + acc_h), 99.9) < 1e-9
    # the step: forces, potentials, interaction counts
    e = rel_err(r["acc"], o["acc"])
    assert np.percentile(e, 99.9) < 1e-8, (np.percentile(e, 99.9), e.max())
    assert np.mean(r["nint"] == o["nint"]) > 0.999
    np.testing.assert_allclose(r["pot"], o["pot"], rtol=1e-8, atol=1e-8 * np.abs(o["pot"]).max())
    # FORCETEST with the per-target split scale (gravtree_forcetest.c:480-483)
    # (the short-range parts do not depend on the periodic images, so the oracle skips them here)
    fo, fso, po, pso = oracle.forcetest(op, pos, mass, sel, stype=stype, ewald=0, highres=hr, hrflag=flags, nthreads=16)
    assert np.max(rel_err(ft["acc_short"], fso)) < 1e-6
    np.testing.assert_allclose(ft["pot_short"], pso, rtol=1e-9)


def test_highres_group_walk_accuracy(agrav, oracle):
    pos, mass, ptype, stype = _zoom(seed=6)
    g = _gpu(agrav, pos, mass, ptype, stype, agrav.FP64, agrav.WALK_PLAIN)
    r64 = g.get_accelerations()
    sel = np.sort(np.random.Generator(np.random.Philox(10)).choice(len(pos), 2000, replace=False)).astype(np.int32)
    ft = g.gravity_forcetest(sel)
    g.close()
    e_ref = rel_err(r64["acc"][sel], ft["acc"])
    for prec in (agrav.FP32, agrav.FP64):
        g = _gpu(agrav, pos, mass, ptype, stype, prec, agrav.WALK_GROUP)
        rg = g.get_accelerations()
        g.close()
        e = rel_err(rg["acc"][sel], ft["acc"])
        print("\nhigh-res zoom: reference walk median %.3g p99 %.3g | group (%s) median %.3g p99 %.3g"
              % (np.median(e_ref), np.percentile(e_ref, 99), "fp32" if prec == agrav.FP32 else "fp64", np.median(e),
                 np.percentile(e, 99)))
        assert np.median(e) <= np.median(e_ref) * 1.05 and np.percentile(e, 99) <= np.percentile(e_ref, 99) * 1.05


def test_highres_direct_matches_oracle(agrav, oracle):
    pos, mass, ptype, stype = _zoom(n_lo=1500, n_hi=700, seed=7)
    g = _gpu(agrav, pos, mass, ptype, stype, agrav.FP64, agrav.WALK_PLAIN)
    acc, pot = g.gravity_direct()
    g.close()
    op = oracle.make_params(**KW, soft=(0.002, 0.0005))
    hr = oracle.make_highres(2, 1.1)
    oracle.highres_regionsize(op, hr, pos, ptype)
    flags = oracle.highres_flags(op, hr, pos, ptype)
    ad, pd = oracle.direct_summation(op, pos, mass, stype=stype, highres=hr, hrflag=flags)
    assert np.max(rel_err(acc, ad)) < 1e-10
    np.testing.assert_allclose(pot, pd, rtol=1e-10)


def test_highres_region_resized_when_left(agrav, oracle):
    """The region is kept while its particles stay inside and sized anew when one has left it
    (longrange.c:129-140); the forces then match an oracle that does the same."""
    pos, mass, ptype, stype = _zoom(n_lo=8000, n_hi=4000, seed=8)
    p = agrav.make_params(**KW, soft=(0.002, 0.0005), precision=agrav.FP64, walk=agrav.WALK_PLAIN, highres_mask=2,
                          enlarge_region=1.1)
    op = oracle.make_params(**KW, soft=(0.002, 0.0005))
    hr = oracle.make_highres(2, 1.1)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass, softtype=stype)
        g.set_particle_types(ptype)
        g.compute_grav_accelerations(first_step=True)
        o = oracle.compute_grav_accelerations(op, pos, mass, stype=stype, first_step=True, highres=hr, ptype=ptype)
        reg0 = g.highres_region()
        # a small move keeps the region; moving one high-resolution particle out re-sizes it
        pos2 = pos.copy()
        pos2[-1] = reg0["xmax"] + 0.02
        oa = g.get_accelerations()["old_acc"]
        g.set_particles(pos2, mass, softtype=stype)
        g.set_particle_types(ptype)
        g.set_old_acc(oa)
        g.compute_grav_accelerations(first_step=False)
        r = g.get_accelerations()
        reg1 = g.highres_region()
    assert reg1["xmax"][0] > reg0["xmax"][0]
    o2 = oracle.compute_grav_accelerations(op, pos2, mass, stype=stype, first_step=False, old_acc=o["old_acc"], highres=hr,
                                           ptype=ptype)
    assert np.array(hr.xmax)[0] == pytest.approx(reg1["xmax"][0], abs=1e-14)
    assert np.percentile(rel_err(r["acc"], o2["acc"]), 99.9) < 1e-7


def test_highres_errors(agrav):
    pos, mass, ptype, stype = _zoom(n_lo=3000, n_hi=500, seed=9)
    p = agrav.make_params(**KW, soft=(0.002, 0.0005), highres_mask=2)
    with agrav.Gravity(p) as g:
        g.set_particles(pos, mass, softtype=stype)
        with pytest.raises(agrav.AgravError) as ei:  # the types are missing
            g.compute_grav_accelerations()
        assert ei.value.code == agrav.ERR_STATE
        g.set_particle_types(np.full(len(pos), 2, np.uint8))  # no particle of the high-resolution types
        with pytest.raises(agrav.AgravError) as ei:
            g.compute_grav_accelerations()
        assert ei.value.code == agrav.ERR_BAD_PARAM
    with pytest.raises(agrav.AgravError):  # a high-resolution region needs a periodic PM box
        agrav.Gravity(agrav.make_params(periodic=False, box_size=0.0, soft=(0.01,), highres_mask=2))
