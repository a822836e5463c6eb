"""CPU tests that pin the oracle's SUBFIND unbinding loop (subfind_serial.c:429-637).

The reference ships no unbinding fixtures, so the loop is pinned by closed
forms a plausible mistake would break: the binding energy of a member at rest
(potential plus the self-energy added back), the kinetic factor 1/2 and its
sign on both sides of escape speed, the comoving terms (1/a on velocities, the
Hubble flow, G/a on the potential), the thermal energy of gas, and the pass
counts the quarter-per-pass removal rule and the phase switch imply.
"""
import numpy as np
import pytest


def _up(oracle, **kw):
    kw.setdefault("des_link_ngb", 0)
    return oracle.make_unbind_params(**kw)


def test_loctree_extent_closed_form(oracle):
    g = np.random.default_rng(3)
    pos = g.normal(size=(500, 3)) * [1.0, 2.0, 0.5] + [3.0, -1.0, 7.0]
    cube = oracle.loctree_extent(pos)
    lo, hi = pos.min(0), pos.max(0)
    assert np.array_equal(cube[:3], 0.5 * (lo + hi))
    assert cube[3] == pytest.approx(1.001 * (hi - lo).max(), rel=1e-15)


def test_member_at_rest_has_minus_partner_potential(oracle):
    # two members at rest: E_1 = G pot_1 + G m_1/eps with pot_1 = -m_2/d - m_1/eps (self term of the
    # spline kernel at r = 0, -2.8 m/h), so E_1 = -G m_2 / d exactly
    d, m1, m2, eps = 0.7, 0.3, 1.1, 0.01
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    for G in (1.0, 2.5):
        r = oracle.subfind_unbind(_up(oracle, G=G), pos, np.zeros((2, 3)), np.array([m1, m2]), eps)
        assert r["len"] == 2 and r["bound"].all() and r["iterations"] == 1
        assert r["energy"][0] == pytest.approx(-G * m2 / d, rel=1e-12)
        assert r["energy"][1] == pytest.approx(-G * m1 / d, rel=1e-12)
        assert r["potential"][0] == pytest.approx(G * (-m2 / d - m1 / eps), rel=1e-12)


def _satellites(speed2, d=1.0, M=1.0, m=1e-9):
    """A heavy member at the origin and light members at distance d on +-x, +-y with tangential speeds^2."""
    pos = np.array([[0, 0, 0], [d, 0, 0], [-d, 0, 0], [0, d, 0], [0, -d, 0]], dtype=float)
    vel = np.zeros((5, 3))
    tang = np.array([[0, 1, 0], [0, -1, 0], [-1, 0, 0], [1, 0, 0]], dtype=float)
    for k, v2 in enumerate(speed2):
        vel[k + 1] = tang[k] * np.sqrt(v2)
    mass = np.array([M, m, m, m, m])
    return pos, vel, mass


def test_kinetic_half_and_escape_speed(oracle):
    # E = -M/d + v^2/2: 0.9 of escape speed squared stays, 1.1 goes; a factor 1 instead of 1/2
    # (or a flipped sign) would remove the first as well (or keep the second)
    pos, vel, mass = _satellites([2 * 0.9, 2 * 1.1, 0.0, 0.0])
    r = oracle.subfind_unbind(_up(oracle), pos, vel, mass, 1e-3)
    assert list(r["bound"]) == [True, True, False, True, True]
    assert r["len"] == 4
    assert r["energy"][1] == pytest.approx(-1 + 0.9, rel=1e-6)
    assert r["energy"][2] == pytest.approx(-1 + 1.1, rel=1e-6)
    # pass 1 removes the fast member, pass 2 (energies only) nobody, pass 3 (full) nobody
    assert r["iterations"] == 3


def test_comoving_terms(oracle):
    # physical peculiar velocity w: Vel = a w - H a^2 (x - s) makes dv = w; the potential is G/a pot
    a, H, G = 0.5, 2.0, 1.0
    pos, w, mass = _satellites([2 * 0.9 * G / a, 2 * 1.1 * G / a, 0.0, 0.0])
    s = (mass[:, None] * pos).sum(0) / mass.sum()
    vel = a * w - H * a * a * (pos - s)
    r = oracle.subfind_unbind(_up(oracle, G=G, time_a=a, hubble_a=H, comoving=True), pos, vel, mass, 1e-3)
    assert list(r["bound"]) == [True, True, False, True, True]
    assert r["energy"][1] == pytest.approx(G / a * (-1 + 0.9), rel=1e-6)
    assert r["energy"][3] == pytest.approx(-G / a, rel=1e-6)
    # the same velocities read as non-comoving: no 1/a, no Hubble flow, G instead of G/a, so a satellite
    # at rest in the comoving frame moves at H a^2 d
    r2 = oracle.subfind_unbind(_up(oracle, G=G), pos, vel, mass, 1e-3)
    assert r2["energy"][3] == pytest.approx(-G + 0.5 * (H * a * a) ** 2, rel=1e-6)


def test_gas_thermal_energy_and_len_non_gas(oracle):
    # member 3 is gas with u = 2 M/d (E = +M/d), member 4 gas with u = 0 (bound)
    pos, vel, mass = _satellites([2 * 0.9, 2 * 1.1, 0.0, 0.0])
    ptype = np.array([1, 1, 1, 0, 0], dtype=np.uint8)
    ut = np.array([0, 0, 0, 2.0, 0.0])
    r = oracle.subfind_unbind(_up(oracle), pos, vel, mass, 1e-3, utherm=ut, ptype=ptype)
    assert list(r["bound"]) == [True, True, False, False, True]
    assert r["len"] == 3 and r["len_non_gas"] == 2
    # two unbound members but at most a quarter per pass: pass 1 takes the gas member (E = 1), pass 2
    # (energies only) the fast one, pass 3 nobody (switch to a full pass), pass 4 nobody
    assert r["iterations"] == 4


def _cluster_and_fast(nb, nf, seed=5):
    g = np.random.default_rng(seed)
    cpos = g.normal(size=(nb, 3)) * 0.05
    fpos = g.normal(size=(nf, 3)) * 0.05
    fdir = g.normal(size=(nf, 3))
    fdir /= np.linalg.norm(fdir, axis=1)[:, None]
    fvel = fdir * np.linspace(100.0, 200.0, nf)[:, None]  # distinct, far above escape speed
    pos = np.concatenate([cpos, fpos])
    vel = np.concatenate([np.zeros((nb, 3)), fvel])
    mass = np.concatenate([np.full(nb, 1.0 / nb), np.full(nf, 1e-9)])
    return pos, vel, mass


@pytest.mark.parametrize("nf, passes", [(60, 4), (30, 3), (100, 5)])
def test_quarter_per_pass_rule(oracle, nf, passes):
    # 100 cold members + nf light fast ones.  With len members, one pass removes the unbound ones above
    # the int(0.25 len)-th largest energy:
    #   nf=60:  160 -> 120 (40 go) -> 100 (20, energies only) -> nobody -> full pass: nobody   = 4 passes
    #   nf=30:  130 -> 100 (all 30, k = 32 falls on a bound member) -> nobody -> nobody         = 3 passes
    #   nf=100: 200 -> 150 (50) -> 113 (37, energies only) -> 100 (13, energies only) -> nobody -> nobody
    #           = 5 passes
    pos, vel, mass = _cluster_and_fast(100, nf)
    r = oracle.subfind_unbind(_up(oracle), pos, vel, mass, 1e-3)
    assert r["len"] == 100
    assert r["bound"][:100].all() and not r["bound"][100:].any()
    assert r["iterations"] == passes


def test_des_link_ngb_stops_the_loop(oracle):
    pos, vel, mass = _cluster_and_fast(100, 60)
    # after pass 1 there are 120 members; 130 stops the loop there, before any pass is counted
    r = oracle.subfind_unbind(_up(oracle, des_link_ngb=130), pos, vel, mass, 1e-3)
    assert r["len"] == 120 and r["iterations"] == 0


def test_max_iter_is_reported(oracle):
    pos, vel, mass = _cluster_and_fast(100, 60)
    r = oracle.subfind_unbind(_up(oracle, max_iter=1), pos, vel, mass, 1e-3)
    assert r["len"] == -1


def test_energy_only_pass_rescales_weakly_bound_potential(oracle):
    # In an energy-only pass the reference multiplies the kept potential of the members at or above the
    # weakly bound limit by G/a once more (subfind_serial.c:491-497).  Two satellites above escape speed,
    # E_a = +0.05 G and E_b = +0.15 G, two at rest.
    #   G = 1 (the scaling is a no-op): pass 1 removes b, pass 2 a, pass 3 nobody, pass 4 (full) nobody.
    #   G = 2: pass 1 removes b with weakly bound limit -2 (the 3rd largest energy, 2P - 1 = 3); pass 2
    #   doubles every potential again, so a is bound (-4 + 2.1) and nobody goes; pass 3 (full, fresh
    #   potentials) removes a; passes 4 and 5 remove nobody.
    for G, passes in ((1.0, 4), (2.0, 5)):
        pos, vel, mass = _satellites([2 * G * 1.05, 2 * G * 1.15, 0.0, 0.0])
        r = oracle.subfind_unbind(_up(oracle, G=G), pos, vel, mass, 1e-3)
        assert list(r["bound"]) == [True, False, False, True, True]
        assert r["iterations"] == passes, G
        # the last pass was a full one: fresh potentials G x walk over the members left (the heavy one,
        # the opposite satellite at 2d, and the self term)
        assert r["potential"][3] == pytest.approx(G * (-1.0 - 1e-9 / 2.0 - 1e-9 / 1e-3), rel=1e-12)
