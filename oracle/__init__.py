"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only the
tests, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module.  Functions restate AREPO's gravity step; each wrapper
names the C function whose header cites the reference file:line.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NTAB = 127


class Params(C.Structure):
    _fields_ = [
        ("periodic", C.c_int),
        ("box_size", C.c_double),
        ("pmgrid", C.c_int),
        ("asmth", C.c_double),
        ("rcut", C.c_double),
        ("G", C.c_double),
        ("err_tol_theta", C.c_double),
        ("err_tol_force_acc", C.c_double),
        ("opening", C.c_int),
        ("nsoft", C.c_int),
        ("soft", C.c_double * 8),
    ]


def make_params(periodic=True, box_size=1.0, pmgrid=0, asmth=1.25, rcut=4.5, G=1.0, err_tol_theta=0.7,
                err_tol_force_acc=0.0025, opening=1, soft=(0.01,)):
    p = Params()
    p.periodic = int(periodic)
    p.box_size = box_size
    p.pmgrid = pmgrid
    p.asmth = asmth
    p.rcut = rcut
    p.G = G
    p.err_tol_theta = err_tol_theta
    p.err_tol_force_acc = err_tol_force_acc
    p.opening = opening
    p.nsoft = len(soft)
    for i, s in enumerate(soft):
        p.soft[i] = s
    return p


class HighRes(C.Structure):
    """orc_highres: the PLACEHIGHRESREGION region (mask, enlarge in; the rest filled by the sizing)."""
    _fields_ = [("mask", C.c_int), ("enlarge", C.c_double), ("valid", C.c_int), ("xmin", C.c_double * 3),
                ("xmax", C.c_double * 3), ("corner", C.c_double * 3), ("upper", C.c_double * 3),
                ("meshsize", C.c_double), ("asmth", C.c_double), ("rcut", C.c_double)]


def make_highres(mask, enlarge=0.0):
    r = HighRes()
    r.mask = int(mask)
    r.enlarge = float(enlarge)
    return r


class UnbindParams(C.Structure):
    _fields_ = [
        ("G", C.c_double),
        ("time_a", C.c_double),
        ("hubble_a", C.c_double),
        ("comoving", C.c_int),
        ("theta", C.c_double),
        ("des_link_ngb", C.c_int),
        ("max_iter", C.c_int),
    ]


class TimestepParams(C.Structure):
    _fields_ = [
        ("err_tol_int_accuracy", C.c_double),
        ("max_size_timestep", C.c_double),
        ("min_size_timestep", C.c_double),
        ("time_a", C.c_double),
        ("hubble_a", C.c_double),
        ("comoving", C.c_int),
    ]


def gravity_timesteps(p, acc, mass, acc_pm=None, stype=None, eta=0.025, max_dt=0.01, min_dt=0.0, time_a=1.0,
                      hubble_a=0.0, comoving=False):
    """get_timestep_gravity's dt per particle and the PM displacement constraint (timestep.c:284-344, :525-565)."""
    tp = TimestepParams(eta, max_dt, min_dt, time_a, hubble_a, int(comoving))
    acc = np.ascontiguousarray(acc, dtype=np.float64).reshape(-1, 3)
    n = len(acc)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    ap = None if acc_pm is None else np.ascontiguousarray(acc_pm, dtype=np.float64).reshape(-1, 3)
    st = None if stype is None else np.ascontiguousarray(stype, dtype=np.uint8)
    dt = np.zeros(max(n, 1))
    dd = C.c_double(0)
    below = lib().orc_gravity_timesteps(C.byref(tp), C.byref(p), n, acc, None if ap is None else ap.ctypes.data, mass,
                                        None if st is None else st.ctypes.data, dt, C.byref(dd))
    return dt[:n], dd.value, below


def make_unbind_params(G=1.0, time_a=1.0, hubble_a=0.0, comoving=False, theta=0.7, des_link_ngb=20,
                       max_iter=300000):
    return UnbindParams(G, time_a, hubble_a, int(comoving), theta, des_link_ngb, max_iter)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def _declare(L):
    P = C.POINTER(Params)
    L.orc_set_softenings.argtypes = [P, _dp]
    L.orc_short_range_tables.argtypes = [C.c_double, _fp, _fp]
    L.orc_domain_extent.argtypes = [P, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_double)]
    L.orc_tree_build.argtypes = [P, C.c_int, _dp, _dp, C.c_void_p, C.c_double, C.POINTER(C.c_void_p)]
    L.orc_tree_build.restype = C.c_int
    L.orc_tree_free.argtypes = [C.c_void_p]
    L.orc_tree_numnodes.argtypes = [C.c_void_p]
    L.orc_tree_export.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip, _ip]
    L.orc_tree_walk.argtypes = [P, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, _dp, _dp, _ip,
                                C.c_int]
    L.orc_tree_walk.restype = C.c_longlong
    L.orc_pm_periodic.argtypes = [P, C.c_int, _dp, _dp, _dp, _dp]
    L.orc_ewald_psi.argtypes = [C.c_double] * 3
    L.orc_ewald_psi.restype = C.c_double
    L.orc_ewald_force.argtypes = [C.c_double] * 3 + [_dp]
    L.orc_ewald_table_tree.argtypes = [C.c_int, C.c_double, _dp, C.c_int]
    L.orc_ewald_corr.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.orc_ewald_pot_corr.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_ewald_pot_corr.restype = C.c_double
    L.orc_forcetest_ewald_exact.argtypes = [C.c_double] * 4 + [_dp]
    L.orc_forcetest_ewald_table.argtypes = [C.c_int, C.c_double, _dp, C.c_int]
    L.orc_forcetest_ewald_lookup.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.orc_forcetest.argtypes = [P, C.c_int, _dp, _dp, C.c_void_p, C.c_int, _ip, C.c_int, C.c_void_p, C.c_int,
                                _dp, _dp, _dp, _dp, C.c_int]
    L.orc_compute_grav_accelerations.argtypes = [P, C.c_int, _dp, _dp, C.c_void_p, C.c_int, _fp, _dp, _dp, _ip,
                                                 C.c_int, _dp]
    L.orc_compute_grav_accelerations.restype = C.c_longlong
    L.orc_direct_summation.argtypes = [P, C.c_int, _dp, _dp, C.c_void_p, _dp, _dp, C.c_int]
    L.orc_group_potential_bh.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp]
    L.orc_group_potential_bh.restype = C.c_longlong
    L.orc_fft3d.argtypes = [C.c_int, _dp, _dp, C.c_int]
    L.orc_group_potential_bh_cube.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_void_p, _dp]
    L.orc_group_potential_bh_cube.restype = C.c_longlong
    L.orc_loctree_extent.argtypes = [C.c_int, _dp, _dp]
    L.orc_subfind_unbind.argtypes = [C.POINTER(UnbindParams), C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_void_p, _u8, _dp, _dp, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
    L.orc_subfind_unbind.restype = C.c_int
    L.orc_gravity_timesteps.argtypes = [C.POINTER(TimestepParams), C.POINTER(Params), C.c_int, _dp, C.c_void_p, _dp,
                                        C.c_void_p, _dp, C.POINTER(C.c_double)]
    L.orc_gravity_timesteps.restype = C.c_int
    HR = C.POINTER(HighRes)
    L.orc_highres_regionsize.argtypes = [P, HR, C.c_int, _dp, _u8]
    L.orc_highres_regionsize.restype = C.c_int
    L.orc_highres_flags.argtypes = [P, HR, C.c_int, _dp, _u8, _u8]
    L.orc_pm_highres.argtypes = [P, HR, C.c_int, _dp, _dp, _u8, _dp, _dp]
    L.orc_pm_highres.restype = C.c_int
    L.orc_tree_walk_hr.argtypes = [P, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_void_p, _dp, _dp, _ip, C.c_int]
    L.orc_tree_walk_hr.restype = C.c_longlong
    L.orc_forcetest_hr.argtypes = [P, C.c_void_p, C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, C.c_int, _ip, C.c_int,
                                   C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int]
    L.orc_direct_summation_hr.argtypes = [P, C.c_void_p, C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, _dp, _dp, C.c_int]
    L.orc_compute_grav_accelerations_hr.argtypes = [P, C.c_void_p, C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, C.c_int,
                                                    _fp, _dp, _dp, _ip, C.c_int, _dp]
    L.orc_compute_grav_accelerations_hr.restype = C.c_longlong
    u64 = C.c_uint64
    L.orc_peano_hilbert_key.argtypes = [u64, u64, u64, C.c_int, C.POINTER(u64), C.POINTER(u64)]
    L.orc_peano_hilbert_key_inverse.argtypes = [u64, u64, C.c_int, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]


def _pos(pos):
    return np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)


def _u8ptr(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.uint8).ctypes.data


def force_softening(p):
    out = np.zeros(p.nsoft + 1)
    lib().orc_set_softenings(C.byref(p), out)
    return out


def short_range_tables(rcut=4.5):
    tf = np.zeros(NTAB + 1, np.float32)
    tp = np.zeros(NTAB + 1, np.float32)
    lib().orc_short_range_tables(rcut, tf, tp)
    return tf, tp


def domain_extent(p, pos):
    pos = _pos(pos)
    corner = np.zeros(3)
    center = np.zeros(3)
    ln = C.c_double()
    lib().orc_domain_extent(C.byref(p), len(pos), pos, corner, center, C.byref(ln))
    return corner, center, ln.value


class Tree:
    """Oracle octree (forcetree.c restatement).  Keeps its inputs alive."""

    def __init__(self, p, pos, mass, stype=None, alloc_factor=None):
        self.p = p
        self.pos = _pos(pos)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.stype = None if stype is None else np.ascontiguousarray(stype, dtype=np.uint8)
        n = len(self.pos)
        af = alloc_factor if alloc_factor is not None else (10.0 if n < 1000 else 0.7)
        h = C.c_void_p()
        while True:
            rc = lib().orc_tree_build(C.byref(p), n, self.pos, self.mass, _u8ptr(self.stype), af, C.byref(h))
            if rc >= 0 or alloc_factor is not None:
                break
            af *= 1.15
        self.rc = rc
        self.h = h
        self.n = n

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib().orc_tree_free(self.h)
            self.h = C.c_void_p()

    @property
    def numnodes(self):
        return lib().orc_tree_numnodes(self.h)

    def export(self):
        M = self.numnodes
        center = np.zeros(3 * M)
        ln = np.zeros(M)
        com = np.zeros(3 * M)
        mass = np.zeros(M)
        level = np.zeros(M, np.int32)
        npart = np.zeros(M, np.int32)
        lib().orc_tree_export(self.h, center, ln, com, mass, level, npart)
        return dict(center=center.reshape(-1, 3), len=ln, com=com.reshape(-1, 3), mass=mass, level=level,
                    npart=npart)

    def walk(self, use_bh, old_acc=None, ewald_tab=None, targets=None, nthreads=0, highres=None, hrflag=None):
        """highres/hrflag: the region and the per-particle flags of a high-resolution PM region."""
        nt = self.n if targets is None else len(targets)
        acc = np.zeros(3 * nt)
        pot = np.zeros(nt)
        nint = np.zeros(nt, np.int32)
        oa = None if old_acc is None else np.ascontiguousarray(old_acc, dtype=np.float32)
        tg = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
        tab = None if ewald_tab is None else np.ascontiguousarray(ewald_tab, dtype=np.float64)
        fl = None if hrflag is None else np.ascontiguousarray(hrflag, dtype=np.uint8)
        total = lib().orc_tree_walk_hr(C.byref(self.p), self.h, int(use_bh), None if oa is None else oa.ctypes.data,
                                       None if tab is None else tab.ctypes.data,
                                       None if highres is None else C.addressof(highres),
                                       None if fl is None else fl.ctypes.data, nt,
                                       None if tg is None else tg.ctypes.data, acc, pot, nint, nthreads)
        return acc.reshape(-1, 3), pot, nint, total


def pm_periodic(p, pos, mass):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    acc = np.zeros(3 * len(pos))
    pot = np.zeros(len(pos))
    rc = lib().orc_pm_periodic(C.byref(p), len(pos), pos, mass, acc, pot)
    if rc != 0:
        raise ValueError("orc_pm_periodic failed %d" % rc)
    return acc.reshape(-1, 3), pot


def ewald_psi(x, y, z):
    return lib().orc_ewald_psi(x, y, z)


def ewald_force(x, y, z):
    f = np.zeros(3)
    lib().orc_ewald_force(x, y, z, f)
    return f


_EWALD_CACHE = {}


def ewald_table_tree(box=1.0, en=64, nthreads=0):
    """forcetree_ewald.c table (cached per box size: it takes seconds to compute)."""
    key = (float(box), en)
    if key not in _EWALD_CACHE:
        tab = np.zeros(4 * (en + 1) ** 3)
        lib().orc_ewald_table_tree(en, box, tab, nthreads)
        _EWALD_CACHE[key] = tab
    return _EWALD_CACHE[key]


def ewald_corr(tab, box, d, en=64):
    f = np.zeros(3)
    lib().orc_ewald_corr(tab, en, box, d[0], d[1], d[2], f)
    return f, lib().orc_ewald_pot_corr(tab, en, box, d[0], d[1], d[2])


def forcetest_ewald_exact(box, x, y, z):
    f = np.zeros(4)
    lib().orc_forcetest_ewald_exact(box, x, y, z, f)
    return f


def forcetest_ewald_table(tew_n, box=1.0, nthreads=0):
    tab = np.zeros(4 * (tew_n + 1) ** 3)
    lib().orc_forcetest_ewald_table(tew_n, box, tab, nthreads)
    return tab


def forcetest_ewald_lookup(tab, tew_n, box, d):
    f = np.zeros(4)
    lib().orc_forcetest_ewald_lookup(tab, tew_n, box, d[0], d[1], d[2], f)
    return f


def forcetest(p, pos, mass, sel, stype=None, ewald=0, tab=None, tew_n=0, nthreads=0, highres=None, hrflag=None):
    """FORCETEST direct sum (G=1 units): returns acc, acc_short, pot, pot_short for the selected targets."""
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    sel = np.ascontiguousarray(sel, dtype=np.int32)
    ns = len(sel)
    acc = np.zeros(3 * ns)
    accs = np.zeros(3 * ns)
    pot = np.zeros(ns)
    pots = np.zeros(ns)
    st = None if stype is None else np.ascontiguousarray(stype, dtype=np.uint8)
    tb = None if tab is None else np.ascontiguousarray(tab, dtype=np.float64)
    fl = None if hrflag is None else np.ascontiguousarray(hrflag, dtype=np.uint8)
    lib().orc_forcetest_hr(C.byref(p), None if highres is None else C.addressof(highres),
                           None if fl is None else fl.ctypes.data, len(pos), pos, mass,
                           None if st is None else st.ctypes.data, ns, sel, ewald,
                           None if tb is None else tb.ctypes.data, tew_n, acc, accs, pot, pots, nthreads)
    return acc.reshape(-1, 3), accs.reshape(-1, 3), pot, pots


def compute_grav_accelerations(p, pos, mass, stype=None, first_step=True, old_acc=None, nthreads=0, highres=None,
                               ptype=None):
    """The step (accel.c:58-347); with `highres` (a HighRes, sized when its `valid` is 0) and the particle
    types `ptype`, a high-resolution PM region as PLACEHIGHRESREGION."""
    pos = _pos(pos)
    n = len(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    oa = np.zeros(n, np.float32) if old_acc is None else np.ascontiguousarray(old_acc, dtype=np.float32).copy()
    acc = np.zeros(3 * n)
    pot = np.zeros(n)
    nint = np.zeros(n, np.int32)
    t = np.zeros(3)
    st = None if stype is None else np.ascontiguousarray(stype, dtype=np.uint8)
    ty = None if ptype is None else np.ascontiguousarray(ptype, dtype=np.uint8)
    total = lib().orc_compute_grav_accelerations_hr(C.byref(p), None if highres is None else C.addressof(highres),
                                                    None if ty is None else ty.ctypes.data, n, pos, mass,
                                                    None if st is None else st.ctypes.data, int(first_step), oa, acc,
                                                    pot, nint, nthreads, t)
    if total < 0:
        raise RuntimeError("oracle compute failed %d" % total)
    return dict(acc=acc.reshape(-1, 3), pot=pot, nint=nint, old_acc=oa, total=total,
                t_treebuild=t[0], t_treewalk=t[1], t_pm=t[2])


def direct_summation(p, pos, mass, stype=None, nthreads=0, highres=None, hrflag=None):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    acc = np.zeros(3 * len(pos))
    pot = np.zeros(len(pos))
    st = None if stype is None else np.ascontiguousarray(stype, dtype=np.uint8)
    fl = None if hrflag is None else np.ascontiguousarray(hrflag, dtype=np.uint8)
    lib().orc_direct_summation_hr(C.byref(p), None if highres is None else C.addressof(highres),
                                  None if fl is None else fl.ctypes.data, len(pos), pos, mass,
                                  None if st is None else st.ctypes.data, acc, pot, nthreads)
    return acc.reshape(-1, 3), pot


def highres_regionsize(p, r, pos, ptype):
    """pm_init_regionsize for the high-resolution grid: fills the HighRes r in place."""
    rc = lib().orc_highres_regionsize(C.byref(p), C.byref(r), len(pos), _pos(pos),
                                      np.ascontiguousarray(ptype, dtype=np.uint8))
    if rc != 0:
        raise ValueError("no particle of the high-resolution types")
    return r


def highres_flags(p, r, pos, ptype):
    fl = np.zeros(len(pos), np.uint8)
    lib().orc_highres_flags(C.byref(p), C.byref(r), len(pos), _pos(pos), np.ascontiguousarray(ptype, dtype=np.uint8), fl)
    return fl


def pm_highres(p, r, pos, mass, ptype):
    """The high-resolution PM force (G included) and potential; None when a high-resolution particle is
    outside the region."""
    pos = _pos(pos)
    acc = np.zeros(3 * len(pos))
    pot = np.zeros(len(pos))
    rc = lib().orc_pm_highres(C.byref(p), C.byref(r), len(pos), pos, np.ascontiguousarray(mass, dtype=np.float64),
                              np.ascontiguousarray(ptype, dtype=np.uint8), acc, pot)
    if rc == 1:
        return None
    if rc != 0:
        raise RuntimeError("orc_pm_highres failed %d" % rc)
    return acc.reshape(-1, 3), pot


def group_potential_bh(pos, mass, h, theta):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    pot = np.zeros(len(pos))
    total = lib().orc_group_potential_bh(len(pos), pos, mass, h, theta, pot)
    return pot, total


def group_potential_bh_cube(pos, mass, h, theta, cube=None):
    pos = _pos(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    pot = np.zeros(len(pos))
    cb = None if cube is None else np.ascontiguousarray(cube, dtype=np.float64)
    total = lib().orc_group_potential_bh_cube(len(pos), pos, mass, h, theta,
                                              None if cb is None else cb.ctypes.data, pot)
    return pot, total


def loctree_extent(pos):
    pos = _pos(pos)
    cube = np.zeros(4)
    lib().orc_loctree_extent(len(pos), pos, cube)
    return cube


def subfind_unbind(up, pos, vel, mass, eps, utherm=None, ptype=None, cube=None):
    """subfind_unbind (subfind_serial.c:429-637) of one isolated group.

    Returns dict(bound[n] bool, len, len_non_gas, iterations, potential[n], energy[n]); len is -1 when the
    loop passed max_iter."""
    pos = _pos(pos)
    vel = _pos(vel)
    n = len(pos)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    ut = None if utherm is None else np.ascontiguousarray(utherm, dtype=np.float64)
    ty = None if ptype is None else np.ascontiguousarray(ptype, dtype=np.uint8)
    cb = None if cube is None else np.ascontiguousarray(cube, dtype=np.float64)
    bound = np.zeros(max(n, 1), dtype=np.uint8)
    pot = np.zeros(max(n, 1))
    en = np.zeros(max(n, 1))
    lng = C.c_int(0)
    it = C.c_int(0)
    ln = lib().orc_subfind_unbind(C.byref(up), n, pos, vel, mass, None if ut is None else ut.ctypes.data,
                                  None if ty is None else ty.ctypes.data, eps,
                                  None if cb is None else cb.ctypes.data, bound, pot, en, C.byref(lng), C.byref(it))
    return dict(bound=bound[:n].astype(bool), len=ln, len_non_gas=lng.value, iterations=it.value,
                potential=pot[:n], energy=en[:n])


def fft3d(re, im, sign):
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    n = re.shape[0]
    lib().orc_fft3d(n, re, im, sign)
    return re, im


def peano_hilbert_key(x, y, z, bits):
    """peano_hilbert_key (peano.c:374-390): the key of integer coordinates as a Python int."""
    hi, lo = C.c_uint64(), C.c_uint64()
    lib().orc_peano_hilbert_key(int(x), int(y), int(z), int(bits), C.byref(hi), C.byref(lo))
    return (hi.value << 64) | lo.value


def peano_hilbert_key_inverse(key, bits):
    """peano_hilbert_key_inverse (peano.c:446-496)."""
    x, y, z = C.c_uint64(), C.c_uint64(), C.c_uint64()
    lib().orc_peano_hilbert_key_inverse(int(key) >> 64, int(key) & ((1 << 64) - 1), int(bits), C.byref(x), C.byref(y),
                                        C.byref(z))
    return x.value, y.value, z.value
