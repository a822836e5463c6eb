"""CPU tests of the oracle's high-resolution PM region (PLACEHIGHRESREGION).

The reference places a second, non-periodic PM grid of GRIDBOOST x PMGRID cells over the particles of
the high-resolution types (pm_nonperiodic.c).  A target in that region takes the short-range force with
the finer split scale Asmth[1] (forcetree_walk.c:143-150), the band between the two split scales comes
from the high-resolution grid, and the rest from the periodic grid.  The oracle is pinned by:

* the region geometry of pm_init_regionsize (pm_nonperiodic.c:135-245) recomputed here;
* a numpy restatement of pmforce_nonperiodic(1) with its kernel (zero-padded convolution);
* the physics: for a pair in the region, periodic PM + high-resolution PM equals FORCETEST's long-range
  part at the fine split scale, and the whole step equals the exact Ewald force, to the meshes' error;
* the potential constant of accel.c:269-274 with the mass of each PM region (forcetree.c:440-456).
"""
import numpy as np
import pytest

GRIDBOOST = 2


def _region_numpy(pmgrid, asmth_cells, rcut_cells, box, pos_hr, enlarge):
    """pm_init_regionsize (pm_nonperiodic.c:135-245), high-resolution grid, from the particles of the mask."""
    grid = GRIDBOOST * pmgrid
    lo, hi = pos_hr.min(0), pos_hr.max(0)
    size = (hi - lo).max() * (enlarge if enlarge > 0 else 1.0)
    xmin = (lo + hi) / 2 - size / 2
    xmax = xmin + size
    inner = size
    tms = size * 2.001 * grid / (grid - 10)
    corner = xmin - 2.0005 * tms / grid
    rc0 = rcut_cells * asmth_cells * box / pmgrid
    if 2 * tms / grid < rc0:
        tms = 2 * (inner + 2 * rc0) * grid / (grid - 2)
        corner = xmin - 1.0001 * rc0
        if 2 * tms / grid > rc0:
            tms = 2 * (inner + 2 * rc0) * grid / (grid - 10)
            corner = xmin - 1.0001 * (rc0 + 2 * tms / grid)
    upper = corner + (grid / 2 - 1) * tms / grid
    return dict(xmin=xmin, xmax=xmax, corner=corner, upper=upper, meshsize=tms, asmth=asmth_cells * tms / grid)


def _pm_highres_numpy(pos, mass, hrflag, pmgrid, reg, asmth_cells, asmth0):
    """pmforce_nonperiodic(1) (pm_nonperiodic.c:1508-1775) with pm_setup_nonperiodic_kernel (:1778-1987),
    G = 1: CIC deposit of the particles in [Corner, UpperCorner) onto the zero-padded grid; the kernel
    -(erfc(u Asmth[1] / Asmth[0]) - erfc(u)) / r, u = r / (2 ASMTH / GRID) with r in units of the mesh size
    and wrapped to [-1/2, 1/2) (-(1 - Asmth[1] / Asmth[0]) / (sqrt(pi) ASMTH / GRID) at r = 0), forward FFT,
    CIC deconvolution [sinc(pi k_x / GRID) sinc(pi k_y / GRID) sinc(pi k_z / GRID)]^-4 (k above GRID / 2 taken
    negative), product with the density's transform, inverse FFT; potential G / L^4 (L / GRID)^3 times it,
    force that / (2 L / GRID) times the 4-point difference on the cells 2 .. GRID/2 - 3; CIC readout for
    the high-resolution particles."""
    from scipy.special import erfc
    G = GRIDBOOST * pmgrid
    L = reg["meshsize"]
    fs = G / L
    rho = np.zeros((G, G, G))
    x = pos - reg["corner"]
    inside = np.all((pos >= reg["corner"]) & (pos < reg["upper"]), axis=1)
    s = np.floor(fs * x).astype(int)
    d = fs * x - s
    corners = []
    for c in range(8):
        b = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        w = np.prod(np.where(b, d, 1 - d), axis=1)
        idx = tuple((s + b).T)
        corners.append((idx, w))
        np.add.at(rho, tuple(a[inside] for a in idx), (mass * w)[inside])
    q = np.arange(G) / G
    q = np.where(q >= 0.5, q - 1.0, q)
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    ratio = reg["asmth"] / asmth0
    u = 0.5 * r / (asmth_cells / G)
    with np.errstate(divide="ignore", invalid="ignore"):
        kern = np.where(r > 0, -(erfc(u * ratio) - erfc(u)) / r, -(1 - ratio) / (np.sqrt(np.pi) * asmth_cells / G))
    kk = np.arange(G)
    kk = np.where(kk > G // 2, kk - G, kk)

    def sinc(k):
        a = np.pi * k / G
        return np.where(k == 0, 1.0, np.sin(a) / np.where(k == 0, 1.0, a))

    KX, KY, KZ = np.meshgrid(kk, kk, kk, indexing="ij")
    dec = np.where(KX * KX + KY * KY + KZ * KZ > 0, 1.0 / (sinc(KX) * sinc(KY) * sinc(KZ)) ** 4, 1.0)
    phi = np.fft.ifftn(np.fft.fftn(rho) * np.fft.fftn(kern) * dec).real * G ** 3
    cell = L / G
    fac_pot = 1.0 / L ** 4 * cell ** 3
    H = G // 2
    forces = []
    for ax in range(3):
        f = np.zeros_like(phi)
        core = (slice(2, H - 2),) * 3
        fd = (4.0 / 3) * (np.roll(phi, 1, ax) - np.roll(phi, -1, ax)) - (1.0 / 6) * (np.roll(phi, 2, ax) - np.roll(phi, -2, ax))
        f[core] = fac_pot / (2 * cell) * fd[core]
        forces.append(f)
    acc = np.zeros((len(pos), 3))
    pot = np.zeros(len(pos))
    for idx, w in corners:
        ok = hrflag.astype(bool)
        sel = tuple(a[ok] for a in idx)
        pot[ok] += w[ok] * fac_pot * phi[sel]
        for ax in range(3):
            acc[ok, ax] += w[ok] * forces[ax][sel]
    return acc, pot


def _markers(lo, width):
    return lo + width * np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], float)


@pytest.mark.parametrize("pmgrid,width,enlarge", [(16, 0.1, 0.0), (16, 0.6, 1.1), (32, 0.05, 1.1), (8, 0.5, 0.0)])
def test_region_geometry(oracle, pmgrid, width, enlarge):
    g = np.random.Generator(np.random.Philox(pmgrid))
    pos = g.random((500, 3))
    ty = np.zeros(500, np.uint8)
    pos[:40] = 0.3 + width * g.random((40, 3))
    ty[:40] = 1
    p = oracle.make_params(periodic=True, box_size=1.0, pmgrid=pmgrid, soft=(1e-3,))
    r = oracle.highres_regionsize(p, oracle.make_highres(2, enlarge), pos, ty)
    want = _region_numpy(pmgrid, 1.25, 4.5, 1.0, pos[:40], enlarge)
    for k in ("xmin", "xmax", "corner", "upper"):
        np.testing.assert_allclose(np.array(getattr(r, k)), want[k], rtol=0, atol=1e-14)
    assert r.meshsize == pytest.approx(want["meshsize"], rel=1e-14)
    assert r.asmth == pytest.approx(want["asmth"], rel=1e-14) and r.rcut == pytest.approx(4.5 * want["asmth"], rel=1e-14)
    # every high-resolution particle lies in the region, and the region's particles in the padded grid's
    # inner half, two cells from its edges
    flags = oracle.highres_flags(p, r, pos, ty)
    assert flags[:40].all()
    inside = np.all((pos >= np.array(r.xmin)) & (pos <= np.array(r.xmax)), axis=1)
    np.testing.assert_array_equal(flags.astype(bool), inside | (ty == 1))
    cells = (pos[flags.astype(bool)] - np.array(r.corner)) * (GRIDBOOST * pmgrid) / r.meshsize
    assert cells.min() >= 2 and cells.max() < GRIDBOOST * pmgrid / 2 - 2


def test_pm_highres_equals_numpy_restatement(oracle):
    pmgrid = 8
    g = np.random.Generator(np.random.Philox(12))
    n = 120
    pos = g.random((n, 3))
    pos[:30] = 0.4 + 0.2 * g.random((30, 3))
    ty = np.zeros(n, np.uint8)
    ty[:30] = 1
    mass = g.uniform(0.5, 1.5, n)
    p = oracle.make_params(periodic=True, box_size=1.0, pmgrid=pmgrid, soft=(1e-3,))
    r = oracle.highres_regionsize(p, oracle.make_highres(2, 1.1), pos, ty)
    flags = oracle.highres_flags(p, r, pos, ty)
    acc, pot = oracle.pm_highres(p, r, pos, mass, ty)
    reg = dict(corner=np.array(r.corner), upper=np.array(r.upper), meshsize=r.meshsize, asmth=r.asmth)
    acc_n, pot_n = _pm_highres_numpy(pos, mass, flags, pmgrid, reg, 1.25, 1.25 / pmgrid)
    assert flags.sum() >= 30
    assert np.abs(acc - acc_n).max() < 1e-12 * np.abs(acc_n).max()
    assert np.abs(pot - pot_n).max() < 1e-12 * np.abs(pot_n).max()
    assert np.all(acc[flags == 0] == 0) and np.all(pot[flags == 0] == 0)
    # a high-resolution particle outside the region is reported (the reference re-sizes the region)
    moved = pos.copy()
    moved[0] = [0.9, 0.9, 0.9]
    assert oracle.pm_highres(p, r, moved, mass, ty) is None


def test_highres_step_against_ewald(oracle):
    """A test particle and a unit mass, the region set by eight massless markers of the high-resolution
    type around the box centre (PMGRID = 32: Asmth[0] = 0.039, Asmth[1] = 0.018).  Over 60 pairs at
    0.5 - 8 Asmth[1]:
    * periodic + high-resolution PM = FORCETEST's long-range force at the fine split scale (its
      short-range part taken with Asmth[1], gravtree_forcetest.c:480-483); mean error per separation bin
      measured 0.13 %, 0.46 %, 0.96 %, 0.92 % of the total force;
    * the whole step (tree with Asmth[1], Rcut[1], both meshes) = the exact Ewald force, errors alike;
    * the step's potential - the Ewald potential = M pi / (a^2 V) - 4 pi Asmth[1]^2 M / V with a the split
      scale of the source's region (accel.c:269-274 with forcetree.c:440-456), to the mesh's 0.5."""
    G = 32
    a0 = 1.25 / G
    p = oracle.make_params(periodic=True, box_size=1.0, pmgrid=G, soft=(1e-5,), opening=0, err_tol_theta=1e-4)
    mk = _markers(0.45, 0.1)
    r = oracle.highres_regionsize(p, oracle.make_highres(2), mk, np.ones(8, np.uint8))
    a1 = r.asmth
    assert 0.4 * a0 < a1 < 0.5 * a0
    g = np.random.Generator(np.random.Philox(3))
    rows = []
    for _ in range(60):
        sep = a1 * np.exp(g.uniform(np.log(0.5), np.log(8)))
        u = g.normal(size=3)
        u /= np.linalg.norm(u)
        p0 = 0.46 + 0.08 * g.random(3)
        pos = np.vstack([[p0, p0 + sep * u], mk])
        mass = np.r_[1e-12, 1.0, np.full(8, 1e-12)]
        ty = np.r_[0, 0, np.ones(8)].astype(np.uint8)
        fl = oracle.highres_flags(p, r, pos, ty)
        assert fl[0] == 1
        acc_p, _ = oracle.pm_periodic(p, pos, mass)
        acc_h, _ = oracle.pm_highres(p, r, pos, mass, ty)
        ft, fts, fp, _ = oracle.forcetest(p, pos, mass, np.array([0], np.int32), ewald=2, highres=r, hrflag=fl)
        res = oracle.compute_grav_accelerations(p, pos, mass, first_step=False, highres=r, ptype=ty)
        an = np.linalg.norm(ft[0])
        a_src = a1 if fl[1] else a0
        rows.append((sep / a1, np.linalg.norm(acc_p[0] + acc_h[0] - (ft[0] - fts[0])) / an,
                     np.linalg.norm(res["acc"][0] - ft[0]) / an,
                     res["pot"][0] - fp[0] - (np.pi / a_src ** 2 - 4 * np.pi * a1 ** 2), fl[1]))
    rows = np.array(rows)
    for (lo, hi), bound in zip(((0.5, 1), (1, 2), (2, 4), (4, 8)), (0.002, 0.007, 0.014, 0.013)):
        m = (rows[:, 0] >= lo) & (rows[:, 0] < hi)
        assert m.sum() >= 8
        assert rows[m, 1].mean() < bound, (lo, hi, rows[m, 1].mean())
        assert rows[m, 2].mean() < bound, (lo, hi, rows[m, 2].mean())
    assert rows[:, 2].max() < 0.03
    assert np.abs(rows[:, 3]).max() < 0.6
    assert 0 < rows[:, 4].sum() < len(rows)  # sources inside and outside the region both occur
