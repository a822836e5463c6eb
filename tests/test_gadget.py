"""CPU tests of the Gadget format-1 reader (arepo_amd.gadget, restating read_ic.c's ICFormat 1 load) on the
reference's own example initial conditions, kept as fixtures under tests/golden/."""
import os
import struct

import numpy as np
import pytest

from arepo_amd import gadget as G

HERE = os.path.dirname(os.path.abspath(__file__))
COSMO = os.path.join(HERE, "golden", "L50n32_ics")
GALAXY = os.path.join(HERE, "golden", "galaxy_ics_every12")


def test_cosmo_box_header_and_mass_closed_form():
    s = G.read_gadget1(COSMO)
    h = s.header
    assert h.npart == [0, 32768, 0, 0, 0, 0] and h.num_files == 1
    assert h.time == 1.0 / 128 and h.redshift == 127.0 and h.box_size == 50000.0
    assert s.ptype.tolist() == [1] * 32768
    # the particle mass is the mean matter density times the box volume over N:
    # Omega0 3 H0^2 / (8 pi G) L^3 / N with H0 = 0.1 h km/s/kpc and G of the code units (begrun.c:273)
    Gc = G.gravity_constant(3.085678e21, 1.989e43, 1e5)
    rho_m = h.omega0 * 3 * 0.1 ** 2 / (8 * np.pi * Gc)
    assert s.mass[0] == pytest.approx(rho_m * h.box_size ** 3 / 32768, rel=2e-3)
    assert np.all(s.mass == h.mass[1])
    assert s.pos.min() >= 0 and s.pos.max() < h.box_size
    assert np.array_equal(np.sort(s.ids), np.arange(1, 32769))


def test_gravity_constant_and_softening():
    assert G.gravity_constant(3.085678e21, 1.989e43, 1e5) == pytest.approx(43018.7, rel=1e-5)
    # set_softenings: comoving 40 at a = 1/128 is 0.31 physical, below the 20 cap
    assert G.softening_comoving(40.0, 20.0, 1 / 128, True) == 40.0
    assert G.softening_comoving(40.0, 20.0, 1.0, True) == 20.0
    assert G.softening_comoving(40.0, 20.0, 1.0, False) == 40.0


@pytest.mark.parametrize("path", [COSMO, GALAXY])
def test_rewrite_is_byte_identical(path, tmp_path):
    s = G.read_gadget1(path)
    out = tmp_path / "rt"
    G.write_gadget1(str(out), s)
    assert out.read_bytes() == open(path, "rb").read()


def test_swapped_byte_order_reads_the_same(tmp_path):
    s = G.read_gadget1(COSMO)
    out = tmp_path / "be"
    G.write_gadget1(str(out), s, byteorder=">")
    assert struct.unpack(">i", out.read_bytes()[:4])[0] == 256
    b = G.read_gadget1(str(out))
    assert b.header.npart == s.header.npart and b.header.box_size == s.header.box_size
    assert np.array_equal(b.pos, s.pos) and np.array_equal(b.ids, s.ids) and np.array_equal(b.mass, s.mass)


def test_broken_files_are_rejected(tmp_path):
    data = open(COSMO, "rb").read()
    (tmp_path / "trunc").write_bytes(data[:100000])
    with pytest.raises(ValueError):
        G.read_gadget1(str(tmp_path / "trunc"))
    (tmp_path / "nohdr").write_bytes(struct.pack("<i", 12) + bytes(12) + struct.pack("<i", 12))
    with pytest.raises(ValueError):
        G.read_gadget1(str(tmp_path / "nohdr"))


def test_galaxy_subsample_components():
    s = G.read_gadget1(GALAXY)
    assert s.header.npart == [0, 8334, 8334, 8334, 0, 0]
    # a per-particle mass block (zero mass table), one mass per component
    assert s.header.mass == [0.0] * 6
    for t in (1, 2, 3):
        m = s.mass[s.ptype == t]
        assert np.all(m == m[0])
    # the disk is flat: its vertical extent is far below its radial one
    d = s.pos[s.ptype == 2]
    assert np.std(d[:, 2]) < 0.2 * np.std(d[:, 0])


def test_internal_velocities_comoving():
    v = np.array([[1.0, 2.0, 3.0]])
    assert np.allclose(G.internal_velocities(v, 0.25, True), v * 0.5 * 0.25)
    assert np.array_equal(G.internal_velocities(v, 0.25, False), v)
