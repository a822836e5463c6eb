"""GPU test of agrav_gravity_direct_noalloc, the entry point that allocates
nothing (its partial sums live in the library's fixed device array).

Its results must equal agrav_gravity_direct's, bit for bit, on two calls in a
row (the second smaller, so stale partial sums of the first would show), in a
periodic TreePM box, a periodic tree-only (Ewald) box and an isolated set with
two softening types; the device's free memory must not change across a call,
and sizes beyond the maximum are rejected.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Dev:
    """Caller-owned device buffers through the HIP runtime."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 1))) == 0
        self.ptrs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(C.c_void_p(p), C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(p), C.c_size_t(out.nbytes), 2) == 0
        return out

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        assert self.hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


CASES = {
    "treepm": dict(periodic=True, box_size=1.0, pmgrid=32, soft=(0.002,)),
    "ewald": dict(periodic=True, box_size=1.0, pmgrid=0, soft=(0.002,)),
    "isolated2": dict(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01, 0.03)),
}


def _set(n, seed, periodic, ntypes):
    g = np.random.Generator(np.random.Philox(seed))
    pos = g.random((n, 3)) if periodic else g.normal(size=(n, 3)) * 0.3
    pos[: n // 4] = 0.5 + 0.01 * g.normal(size=(n // 4, 3)) if periodic else 0.01 * g.normal(size=(n // 4, 3))
    if periodic:
        pos %= 1.0
    mass = g.uniform(0.5, 1.5, n) / n
    st = g.integers(0, ntypes, n).astype(np.uint8) if ntypes > 1 else None
    return pos, mass, st


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_equals_gravity_direct(agrav, case):
    kw = CASES[case]
    p = agrav.make_params(**kw, precision=agrav.FP64, walk=agrav.WALK_PLAIN)
    ntypes = len(kw["soft"])
    dev = _Dev()
    try:
        with agrav.Gravity(p) as g:
            for call, (n, seed) in enumerate([(5000, 11), (1500, 12)]):
                pos, mass, st = _set(n, seed, kw["periodic"], ntypes)
                print("noalloc case=%s call=%d n=%d" % (case, call, n), flush=True)
                g.set_particles(pos, mass, softtype=st)
                ref_acc, ref_pot = g.gravity_direct()
                d_pos, d_mass = dev.put(pos), dev.put(mass)
                d_st = dev.put(st) if st is not None else None
                d_acc, d_pot = dev.alloc(24 * n), dev.alloc(8 * n)
                dev.sync()
                before = dev.free_bytes()
                g.gravity_direct_noalloc(n, d_pos, d_mass, d_st, d_acc, d_pot)
                dev.sync()
                assert dev.free_bytes() == before  # nothing allocated
                acc = dev.get(d_acc, (n, 3))
                pot = dev.get(d_pot, n)
                assert np.array_equal(acc, ref_acc), "acc differs: max %g" % np.abs(acc - ref_acc).max()
                assert np.array_equal(pot, ref_pot)
    finally:
        dev.free()


def test_noalloc_rejects_too_many_particles(agrav):
    p = agrav.make_params(periodic=False, box_size=0.0, pmgrid=0, soft=(0.01,))
    dev = _Dev()
    try:
        n = agrav.DIRECT_NOALLOC_MAX_N + 1
        d = dev.alloc(24 * n)
        with agrav.Gravity(p) as g:
            with pytest.raises(agrav.AgravError) as ei:
                g.gravity_direct_noalloc(n, d, d, None, d, None)
            assert ei.value.code == agrav.ERR_CAPACITY
            # the largest size is accepted, without a potential output
            n = agrav.DIRECT_NOALLOC_MAX_N
            pos, mass, _ = _set(n, 5, False, 1)
            d_pos, d_mass, d_acc = dev.put(pos), dev.put(mass), dev.alloc(24 * n)
            g.gravity_direct_noalloc(n, d_pos, d_mass, None, d_acc, None)
            dev.sync()
            g.set_particles(pos, mass)
            ref_acc, _ = g.gravity_direct()
            assert np.array_equal(dev.get(d_acc, (n, 3)), ref_acc)
    finally:
        dev.free()
