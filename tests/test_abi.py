"""CPU tests of the C ABI boundary: the library loads, exports exactly what
include/agrav.h declares, host-side helpers behave, and calls without a GPU
fail with an error code instead of falling back to the CPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available


def _declared():
    src = open(os.path.join(ROOT, "include", "agrav.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(agrav_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols(agrav):
    names = _declared()
    assert len(names) >= 20
    lib = C.CDLL(agrav.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(agrav.EXPORTED) == set(names)


def test_status_strings(agrav):
    L = agrav.lib()
    for code in range(0, -9, -1):
        s = L.agrav_status_string(code)
        assert s and s != b"unknown"


def test_default_params_match_reference_defaults(agrav):
    p = agrav.Params()
    agrav.lib().agrav_default_params(C.byref(p))
    assert p.asmth == 1.25 and p.rcut == 4.5  # allvars.h:461, :469
    assert p.type_of_opening_criterion == 1 and p.G == 1.0


def test_create_without_gpu_fails_loudly(agrav):
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(agrav.AgravError) as ei:
        agrav.Gravity(agrav.make_params())
    assert ei.value.code in (agrav.ERR_NO_DEVICE, agrav.ERR_HIP)


@pytest.mark.parametrize("bad", [
    dict(periodic=False, box_size=0.0, pmgrid=32),  # isolated TreePM: the reference leaves Asmth/Rcut unset
    dict(periodic=True, box_size=1.0, pmgrid=48),   # not a power of two
    dict(periodic=True, box_size=1.0, pmgrid=32, rcut=0.0),
    dict(periodic=True, box_size=0.0),
])
def test_invalid_params_rejected_before_device(agrav, bad):
    # checked before the device is touched, so the same code comes back with or without a GPU
    with pytest.raises(agrav.AgravError) as ei:
        agrav.Gravity(agrav.make_params(**bad))
    assert ei.value.code == agrav.ERR_BAD_PARAM


def _cells(level):
    """All cells of a 2^level grid, keyed by the top `level` bits of the 21-bit key coordinates."""
    n = 1 << level
    idx = np.arange(n)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


@pytest.mark.parametrize("level", [1, 2, 3])
def test_hilbert_key_is_a_hilbert_curve(agrav, level):
    shift = 21 - level
    cells = _cells(level)
    keys = np.array([agrav.hilbert_key21(*(c << shift)) >> (3 * shift) for c in cells])
    # bijection onto 0 .. 8^level - 1
    assert np.array_equal(np.sort(keys), np.arange(8 ** level))
    order = cells[np.argsort(keys)]
    # consecutive cells along the curve are face neighbours (the defining Hilbert property)
    steps = np.abs(np.diff(order, axis=0)).sum(1)
    assert np.all(steps == 1)


def test_hilbert_key_prefix_property(agrav):
    # the key's top 3L bits depend only on the top L bits of each coordinate: every octree cube is a
    # contiguous key range (what the octree construction relies on)
    g = np.random.Generator(np.random.Philox(1))
    for _ in range(200):
        L = int(g.integers(1, 21))
        base = g.integers(0, 1 << 21, 3)
        mask = ~np.int64((1 << (21 - L)) - 1)
        a = base & mask
        b = a + g.integers(0, 1 << (21 - L), 3)
        ka = agrav.hilbert_key21(*a) >> (3 * (21 - L))
        kb = agrav.hilbert_key21(*b) >> (3 * (21 - L))
        assert ka == kb


def test_hilbert_key_equals_peano_hilbert_key(agrav, oracle):
    # the library's 24-state table (derived from the curve's quadrant description) gives the reference's
    # peano_hilbert_key (peano.c:374-390), restated in the oracle from its orientation/digit tables
    g = np.random.Generator(np.random.Philox(3))
    pts = g.integers(0, 1 << 21, (10000, 3))
    m = (1 << 21) - 1
    pts[:8] = [[0, 0, 0], [m, m, m], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1 << 20, 0, 0], [0, 1 << 20, 0], [0, 0, 1 << 20]]
    for x, y, z in pts:
        assert agrav.hilbert_key21(x, y, z) == oracle.peano_hilbert_key(x, y, z, 21)
    # the generator's packed table for the 30 levels of the octree keys, and the first 30 of the 42 levels
    # of AREPO's domain keys (BITS_PER_DIMENSION, allvars.h:370-372)
    from arepo_amd.gen_hilbert_table import derive, table_key
    tab = derive()
    for x, y, z in g.integers(0, 1 << 42, (500, 3), dtype=np.int64):
        k42 = oracle.peano_hilbert_key(x, y, z, 42)
        assert table_key(tab, int(x) >> 12, int(y) >> 12, int(z) >> 12, 30) == k42 >> 36
        assert table_key(tab, int(x), int(y), int(z), 42) == k42


def test_hilbert_key_fits_63_bits(agrav):
    m = (1 << 21) - 1
    for c in [(0, 0, 0), (m, m, m), (m, 0, 0), (0, m, 0), (0, 0, m)]:
        assert agrav.hilbert_key21(*c) < (1 << 63)


def test_host_alloc_edge_cases_need_no_device(agrav):
    import ctypes as C
    L = agrav.lib()
    p = C.c_void_p(1)
    assert L.agrav_host_alloc(0, C.byref(p)) == 0 and p.value is None  # nothing to allocate
    assert L.agrav_host_alloc(16, None) == -1                           # no place for the pointer
    assert L.agrav_host_free(None) == 0
