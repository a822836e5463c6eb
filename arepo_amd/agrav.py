"""Python mirror of the C ABI in include/agrav.h (ctypes over libagrav.so).

The methods keep the names of the AREPO functions they replace
(construct_forcetree, gravity_tree, long_range_force,
gravity_force_finalize, compute_grav_accelerations, gravity_forcetest, ...).
There is no CPU fallback: if libagrav.so is missing or cannot be loaded this
module raises, and every call goes to the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libagrav.so")

OK, ERR_BAD_PARAM, ERR_CAPACITY, ERR_OUT_OF_BOX, ERR_HIP, ERR_OOM, ERR_STATE, ERR_FFT, ERR_NO_DEVICE = \
    0, -1, -2, -3, -4, -5, -6, -7, -8
FP64, FP32 = 0, 1
WALK_PLAIN, WALK_GROUP = 0, 1
MAX_SOFT = 8
DIRECT_NOALLOC_MAX_N = 16384  # AGRAV_DIRECT_NOALLOC_MAX_N


class AgravError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("agrav error %d: %s" % (code, msg))
        self.code = code


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
        ("type_of_opening_criterion", C.c_int),
        ("n_soft_types", C.c_int),
        ("softening", C.c_double * MAX_SOFT),
        ("precision", C.c_int),
        ("walk", C.c_int),
        ("group_size", C.c_int),
        ("tree_alloc_factor", C.c_double),
        ("leaf_bucket", C.c_int),
        ("eval_potential", C.c_int),
        ("group_waves", C.c_int),
        ("placehighresregion", C.c_int),
        ("enlargeregion", C.c_double),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("treebuild", C.c_double),
        ("treewalk", C.c_double),
        ("pm_grav", C.c_double),
        ("finalize", C.c_double),
        ("walk_kernel", C.c_double),
        ("interactions", C.c_longlong),
        ("numnodes", C.c_int),
        ("walk_retries", C.c_int),
        ("walk_popped", C.c_longlong),
        ("tile_entries", C.c_longlong),
        ("unbind_rounds", C.c_int),
        ("unbind_solves", C.c_int),
        ("unbind_solve_ms", C.c_double),
        ("unbind_pass_ms", C.c_double),
        ("device_bytes", C.c_longlong),
        ("walk_ordered", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libagrav.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libagrav.so not built: run `python -m arepo_amd.build` (%s missing)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


EXPORTED = [
    "agrav_status_string", "agrav_last_error", "agrav_default_params", "agrav_create", "agrav_destroy",
    "agrav_set_softenings", "agrav_set_particles", "agrav_set_particles_device", "agrav_set_old_acc", "agrav_set_active",
    "agrav_set_particle_types", "agrav_highres_region",
    "agrav_set_partial_tree",
    "agrav_construct_forcetree", "agrav_gravity_tree", "agrav_long_range_force", "agrav_gravity_force_finalize",
    "agrav_compute_grav_accelerations", "agrav_get_accelerations", "agrav_get_accelerations_device",
    "agrav_get_components", "agrav_gravity_forcetest", "agrav_gravity_direct", "agrav_gravity_direct_noalloc", "agrav_get_timings", "agrav_get_tree",
    "agrav_get_order", "agrav_get_tables", "agrav_hilbert_key21",
    "agrav_batch_set_particles", "agrav_batch_autotune", "agrav_batch_get_tuning", "agrav_batch_set_tuning",
    "agrav_subfind_unbind", "agrav_gravity_timesteps", "agrav_server_create", "agrav_server_submit", "agrav_server_wait", "agrav_server_poll", "agrav_server_stats",
    "agrav_server_destroy", "agrav_solve_footprint", "agrav_host_alloc", "agrav_host_free",
]


BATCH_CLASSES = 4


class BatchTuning(C.Structure):
    _fields_ = [
        ("group_size", C.c_int * BATCH_CLASSES),
        ("leaf_bucket", C.c_int * BATCH_CLASSES),
        ("class_max_n", C.c_int64 * BATCH_CLASSES),
        ("sample_particles", C.c_int64 * BATCH_CLASSES),
        ("best_ms", C.c_double * BATCH_CLASSES),
        ("tuned", C.c_int),
    ]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k != "tuned" else self.tuned) for k, _ in self._fields_}


class TimestepParams(C.Structure):
    _fields_ = [
        ("err_tol_int_accuracy", C.c_double),
        ("max_size_timestep", C.c_double),
        ("min_size_timestep", C.c_double),
        ("time_a", C.c_double),
        ("hubble_a", C.c_double),
        ("comoving", C.c_int),
    ]


class UnbindParams(C.Structure):
    _fields_ = [
        ("time_a", C.c_double),
        ("hubble_a", C.c_double),
        ("comoving", C.c_int),
        ("des_link_ngb", C.c_int),
        ("max_iter", C.c_int),
    ]


class Request(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("pos", C.c_void_p),
        ("mass", C.c_void_p),
        ("softtype", C.c_void_p),
        ("softening", C.c_double),
        ("acc", C.c_void_p),
        ("pot", C.c_void_p),
    ]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double),
                ("slot", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_longlong), ("failed", C.c_longlong), ("queued", C.c_int), ("running", C.c_int),
                ("peak_running", C.c_int), ("slots", C.c_int), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t),
                ("hbm_budget", C.c_size_t), ("hbm_resident", C.c_size_t), ("evictions", C.c_longlong),
                ("model_exceeded", C.c_longlong)]


def _declare(L):
    vp = C.c_void_p
    L.agrav_status_string.restype = C.c_char_p
    L.agrav_status_string.argtypes = [C.c_int]
    L.agrav_last_error.restype = C.c_char_p
    L.agrav_last_error.argtypes = [vp]
    L.agrav_default_params.argtypes = [C.POINTER(Params)]
    L.agrav_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.agrav_destroy.argtypes = [vp]
    L.agrav_set_softenings.argtypes = [vp, C.c_int, vp]
    L.agrav_set_particles.argtypes = [vp, C.c_int64, vp, vp, vp]
    L.agrav_set_particles_device.argtypes = [vp, C.c_int64, vp, vp, vp]
    L.agrav_set_old_acc.argtypes = [vp, vp]
    L.agrav_set_active.argtypes = [vp, vp, C.c_int64]
    L.agrav_set_partial_tree.argtypes = [vp, C.c_int, C.c_int]
    L.agrav_construct_forcetree.argtypes = [vp, C.POINTER(C.c_int)]
    L.agrav_gravity_tree.argtypes = [vp, C.c_int]
    L.agrav_long_range_force.argtypes = [vp]
    L.agrav_gravity_force_finalize.argtypes = [vp]
    L.agrav_compute_grav_accelerations.argtypes = [vp, C.c_int]
    L.agrav_get_accelerations.argtypes = [vp, vp, vp, vp, vp]
    L.agrav_get_accelerations_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.agrav_get_components.argtypes = [vp, vp, vp, vp, vp]
    L.agrav_gravity_forcetest.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    L.agrav_gravity_direct.argtypes = [vp, vp, vp]
    L.agrav_get_timings.argtypes = [vp, C.POINTER(Timings)]
    L.agrav_get_tree.argtypes = [vp, C.POINTER(C.c_int), vp, vp, vp, vp, vp, vp]
    L.agrav_get_order.argtypes = [vp, vp, vp]
    L.agrav_get_tables.argtypes = [vp, vp, vp, vp, vp]
    L.agrav_batch_set_particles.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    L.agrav_batch_autotune.argtypes = [vp, C.c_int64, C.POINTER(BatchTuning)]
    L.agrav_batch_get_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
    L.agrav_batch_set_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
    L.agrav_gravity_timesteps.argtypes = [vp, C.POINTER(TimestepParams), vp, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int64)]
    L.agrav_subfind_unbind.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(UnbindParams), vp, vp,
                                       vp, vp, vp, vp]
    L.agrav_gravity_direct_noalloc.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp]
    L.agrav_set_particle_types.argtypes = [vp, vp]
    L.agrav_highres_region.argtypes = [vp, C.c_int, vp]
    L.agrav_server_create.argtypes = [C.POINTER(Params), C.c_int, C.c_size_t, C.POINTER(vp)]
    L.agrav_solve_footprint.restype = C.c_size_t
    L.agrav_solve_footprint.argtypes = [C.POINTER(Params), C.c_int64]
    L.agrav_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.agrav_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(RequestResult)]
    L.agrav_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(C.c_int)]
    L.agrav_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    L.agrav_server_destroy.argtypes = [vp]
    L.agrav_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.agrav_host_free.argtypes = [vp]
    L.agrav_hilbert_key21.restype = C.c_uint64
    L.agrav_hilbert_key21.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    for name in EXPORTED:
        if name not in ("agrav_status_string", "agrav_last_error", "agrav_hilbert_key21", "agrav_destroy",
                        "agrav_default_params", "agrav_server_destroy", "agrav_solve_footprint"):
            getattr(L, name).restype = C.c_int
    L.agrav_destroy.restype = None
    L.agrav_server_destroy.restype = None
    L.agrav_default_params.restype = None


def make_params(periodic=True, box_size=1.0, pmgrid=0, asmth=1.25, rcut=4.5, G=1.0, err_tol_theta=0.7,
                err_tol_force_acc=0.0025, opening=1, soft=(0.01,), precision=FP64, walk=WALK_GROUP, group_size=0,
                leaf_bucket=8, eval_potential=True, group_waves=0, highres_mask=0, enlarge_region=0.0):
    p = Params()
    lib().agrav_default_params(C.byref(p))
    p.periodic = int(periodic)
    p.box_size = float(box_size)
    p.pmgrid = int(pmgrid)
    p.asmth = asmth
    p.rcut = rcut
    p.G = G
    p.err_tol_theta = err_tol_theta
    p.err_tol_force_acc = err_tol_force_acc
    p.type_of_opening_criterion = opening
    p.n_soft_types = len(soft)
    for i, s in enumerate(soft):
        p.softening[i] = s
    p.precision = precision
    p.walk = walk
    p.group_size = group_size
    p.leaf_bucket = leaf_bucket
    p.eval_potential = int(eval_potential)
    p.group_waves = int(group_waves)
    p.placehighresregion = int(highres_mask)
    p.enlargeregion = float(enlarge_region)
    return p


def _ptr(a):
    return None if a is None else a.ctypes.data


class _PinnedBlock:
    """Page-locked host memory from agrav_host_alloc, freed when the last array viewing it goes."""

    def __init__(self, nbytes):
        self.L = lib()
        p = C.c_void_p()
        rc = self.L.agrav_host_alloc(max(int(nbytes), 1), C.byref(p))
        if rc != 0:
            raise AgravError(rc, (self.L.agrav_last_error(None) or b"").decode())
        self.ptr = p.value
        self.nbytes = max(int(nbytes), 1)

    def __del__(self):
        if getattr(self, "ptr", None):
            self.L.agrav_host_free(self.ptr)
            self.ptr = None


def host_empty(shape, dtype=np.float64):
    """An uninitialised numpy array in page-locked host memory (agrav_host_alloc): copies between it
    and the device run at the host link's DMA rate."""
    dt = np.dtype(dtype)
    count = int(np.prod(shape)) if np.ndim(shape) else int(shape)
    blk = _PinnedBlock(count * dt.itemsize)
    buf = (C.c_char * blk.nbytes).from_address(blk.ptr)
    buf._agrav_block = blk  # the ctypes buffer keeps the block alive, the array keeps the buffer
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


class Gravity:
    """One solver context (agrav_ctx): device buffers persist between calls."""

    def __init__(self, params: Params):
        self.L = lib()
        h = C.c_void_p()
        self._check(self.L.agrav_create(C.byref(params), C.byref(h)), None)
        self.h = h
        self.params = params
        self.n = 0

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.agrav_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, h=None):
        if rc != OK:
            msg = self.L.agrav_last_error(h if h is not None else (self.h if hasattr(self, "h") else None))
            raise AgravError(rc, msg.decode() if msg else "")

    # -- inputs -----------------------------------------------------------
    def set_particles(self, pos, mass, softtype=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        st = None if softtype is None else np.ascontiguousarray(softtype, dtype=np.uint8)
        self._check(self.L.agrav_set_particles(self.h, len(pos), _ptr(pos), _ptr(mass), _ptr(st)))
        self.n = len(pos)

    def set_particles_device(self, n, d_pos, d_mass, d_softtype=None):
        self._check(self.L.agrav_set_particles_device(self.h, n, d_pos, d_mass, d_softtype))
        self.n = n

    def batch_set_particles(self, offsets, pos_rel, mass, softening, centres=None):
        """Independent isolated sets: set s is particles offsets[s]:offsets[s+1], positions
        as FP32 offsets from its centre (centres: (nsets, 3) or None), one softening per set."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        rel = np.ascontiguousarray(pos_rel, dtype=np.float32).reshape(-1, 3)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        eps = np.ascontiguousarray(softening, dtype=np.float64)
        cen = None if centres is None else np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        nsets = len(off) - 1
        if len(eps) != nsets or off[-1] != len(rel) or len(m) != len(rel) or (cen is not None and len(cen) != nsets):
            raise ValueError("batch arrays do not match the offsets")
        self._check(self.L.agrav_batch_set_particles(self.h, nsets, _ptr(off), _ptr(rel), _ptr(cen), _ptr(m), _ptr(eps)))
        self.n = len(rel)

    def batch_autotune(self, sample_particles=0):
        t = BatchTuning()
        self._check(self.L.agrav_batch_autotune(self.h, int(sample_particles), C.byref(t)))
        return t.as_dict()

    def batch_get_tuning(self):
        t = BatchTuning()
        self._check(self.L.agrav_batch_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def batch_set_tuning(self, group_size, leaf_bucket):
        t = BatchTuning()
        gs = [group_size] * BATCH_CLASSES if np.isscalar(group_size) else list(group_size)
        lb = [leaf_bucket] * BATCH_CLASSES if np.isscalar(leaf_bucket) else list(leaf_bucket)
        for k in range(BATCH_CLASSES):
            t.group_size[k] = int(gs[k])
            t.leaf_bucket[k] = int(lb[k])
        self._check(self.L.agrav_batch_set_tuning(self.h, C.byref(t)))

    def subfind_unbind(self, offsets, pos, vel, mass, softening, utherm=None, ptype=None, cubes=None, time_a=1.0,
                       hubble_a=0.0, comoving=False, des_link_ngb=20, max_iter=0):
        """SUBFIND unbinding (subfind_unbind, subfind_serial.c:429-637) of every group
        offsets[g]:offsets[g+1] at once; theta is the context's err_tol_theta.

        Returns dict(bound[n] bool, len[g], len_non_gas[g], iterations[g], potential[n], energy[n])."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        ng = len(off) - 1
        x = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        eps = np.ascontiguousarray(softening, dtype=np.float64)
        ut = None if utherm is None else np.ascontiguousarray(utherm, dtype=np.float64)
        ty = None if ptype is None else np.ascontiguousarray(ptype, dtype=np.uint8)
        cb = None if cubes is None else np.ascontiguousarray(cubes, dtype=np.float64).reshape(-1, 4)
        n = len(x)
        if ng < 1 or off[-1] != n or len(v) != n or len(m) != n or len(eps) != ng or \
                (cb is not None and len(cb) != ng) or (ut is not None and len(ut) != n) or \
                (ty is not None and len(ty) != n):
            raise ValueError("unbinding arrays do not match the offsets")
        up = UnbindParams(time_a, hubble_a, int(comoving), int(des_link_ngb), int(max_iter))
        bound = np.zeros(max(n, 1), dtype=np.uint8)
        ln = np.zeros(ng, dtype=np.int32)
        lng = np.zeros(ng, dtype=np.int32)
        it = np.zeros(ng, dtype=np.int32)
        pot = np.zeros(max(n, 1))
        en = np.zeros(max(n, 1))
        self._check(self.L.agrav_subfind_unbind(self.h, ng, _ptr(off), _ptr(x), _ptr(v), _ptr(m), _ptr(eps), _ptr(ut),
                                                _ptr(ty), _ptr(cb), C.byref(up), _ptr(bound), _ptr(ln), _ptr(lng),
                                                _ptr(it), _ptr(pot), _ptr(en)))
        return dict(bound=bound[:n].astype(bool), len=ln, len_non_gas=lng, iterations=it, potential=pot[:n],
                    energy=en[:n])

    def set_active(self, active=None):
        """Targets of the next walk/finalize (gravity_tree's TargetList); None: every particle."""
        if active is None:
            self._check(self.L.agrav_set_active(self.h, None, -1))
            return
        a = np.ascontiguousarray(active, dtype=np.int32)
        buf = a if len(a) else np.zeros(1, np.int32)  # an empty list is a list, not "every particle"
        self._check(self.L.agrav_set_active(self.h, _ptr(buf), len(a)))

    def set_partial_tree(self, partial=True, set_oldacc=True):
        """HIERARCHICAL_GRAVITY: builds insert only the active particles; set_oldacc=False keeps OldAcc."""
        self._check(self.L.agrav_set_partial_tree(self.h, int(bool(partial)), int(bool(set_oldacc))))

    def gravity_timesteps(self, eta=0.025, max_dt=0.01, min_dt=0.0, time_a=1.0, hubble_a=0.0, comoving=False):
        """dt per particle (get_timestep_gravity) and the PM displacement constraint after a step."""
        tp = TimestepParams(eta, max_dt, min_dt, time_a, hubble_a, int(comoving))
        dt = np.zeros(max(self.n, 1))
        dd = C.c_double(0)
        nb = C.c_int64(0)
        self._check(self.L.agrav_gravity_timesteps(self.h, C.byref(tp), _ptr(dt), C.byref(dd), C.byref(nb)))
        return dt[:self.n], dd.value, nb.value

    def set_particle_types(self, ptype):
        t = np.ascontiguousarray(ptype, dtype=np.uint8)
        self._check(self.L.agrav_set_particle_types(self.h, _ptr(t)))

    def highres_region(self, reset=False):
        """The high-resolution PM region as a dict (xmin, xmax, corner, upper, meshsize, asmth, rcut)."""
        r = np.zeros(15)
        self._check(self.L.agrav_highres_region(self.h, int(reset), _ptr(r)))
        return dict(xmin=r[0:3], xmax=r[3:6], corner=r[6:9], upper=r[9:12], meshsize=r[12], asmth=r[13], rcut=r[14])

    def set_old_acc(self, old_acc):
        oa = np.ascontiguousarray(old_acc, dtype=np.float32)
        self._check(self.L.agrav_set_old_acc(self.h, _ptr(oa)))

    def set_softenings(self, eps):
        e = np.ascontiguousarray(eps, dtype=np.float64)
        self._check(self.L.agrav_set_softenings(self.h, len(e), _ptr(e)))

    # -- the gravity step ---------------------------------------------------
    def construct_forcetree(self):
        m = C.c_int()
        self._check(self.L.agrav_construct_forcetree(self.h, C.byref(m)))
        return m.value

    def gravity_tree(self, use_bh=False):
        self._check(self.L.agrav_gravity_tree(self.h, int(use_bh)))

    def long_range_force(self):
        self._check(self.L.agrav_long_range_force(self.h))

    def gravity_force_finalize(self):
        self._check(self.L.agrav_gravity_force_finalize(self.h))

    def compute_grav_accelerations(self, first_step=True):
        self._check(self.L.agrav_compute_grav_accelerations(self.h, int(first_step)))

    # -- outputs ------------------------------------------------------------
    def get_accelerations(self, fields=("acc", "pot", "nint", "old_acc"), out=None):
        """The results in caller order; only `fields` are copied back (into the arrays of `out` when given)."""
        n = self.n
        out = dict(out or {})
        make = dict(acc=lambda: np.empty((n, 3)), pot=lambda: np.empty(n), nint=lambda: np.empty(n, np.int32),
                    old_acc=lambda: np.empty(n, np.float32))
        for k in fields:
            if out.get(k) is None:
                out[k] = make[k]()
        self._check(self.L.agrav_get_accelerations(self.h, *(_ptr(out.get(k)) if k in fields else None
                                                             for k in ("acc", "pot", "nint", "old_acc"))))
        return {k: (out[k] if k in fields else None) for k in ("acc", "pot", "nint", "old_acc")}

    def get_components(self, tree=True, pm=False):
        n = self.n
        at = np.zeros((n, 3)) if tree else None
        pt = np.zeros(n) if tree else None
        ap = np.zeros((n, 3)) if pm else None
        pp = np.zeros(n) if pm else None
        self._check(self.L.agrav_get_components(self.h, _ptr(at), _ptr(pt), _ptr(ap), _ptr(pp)))
        return dict(acc_tree=at, pot_tree=pt, acc_pm=ap, pot_pm=pp)

    def gravity_forcetest(self, sel):
        sel = np.ascontiguousarray(sel, dtype=np.int32)
        ns = len(sel)
        acc = np.zeros((ns, 3))
        accs = np.zeros((ns, 3))
        pot = np.zeros(ns)
        pots = np.zeros(ns)
        self._check(self.L.agrav_gravity_forcetest(self.h, ns, _ptr(sel), _ptr(acc), _ptr(accs), _ptr(pot), _ptr(pots)))
        return dict(acc=acc, acc_short=accs, pot=pot, pot_short=pots)

    def gravity_direct(self):
        acc = np.zeros((self.n, 3))
        pot = np.zeros(self.n)
        self._check(self.L.agrav_gravity_direct(self.h, _ptr(acc), _ptr(pot)))
        return acc, pot

    def gravity_direct_noalloc(self, n, d_pos, d_mass, d_softtype, d_acc, d_pot, stream=None):
        """agrav_gravity_direct_noalloc on device pointers (asynchronous on the context's stream by default)."""
        self._check(self.L.agrav_gravity_direct_noalloc(self.h, int(n), d_pos, d_mass, d_softtype, d_acc, d_pot, stream))

    def timings(self):
        t = Timings()
        self._check(self.L.agrav_get_timings(self.h, C.byref(t)))
        return t.as_dict()

    def get_tree(self):
        m = C.c_int()
        self._check(self.L.agrav_get_tree(self.h, C.byref(m), None, None, None, None, None, None))
        M = m.value
        center = np.zeros((M, 3))
        ln = np.zeros(M)
        com = np.zeros((M, 3))
        mass = np.zeros(M)
        level = np.zeros(M, np.int32)
        npart = np.zeros(M, np.int32)
        self._check(self.L.agrav_get_tree(self.h, C.byref(m), _ptr(center), _ptr(ln), _ptr(com), _ptr(mass),
                                          _ptr(level), _ptr(npart)))
        return dict(center=center, len=ln, com=com, mass=mass, level=level, npart=npart)

    def get_order(self):
        """(perm, keys) of the tree's particles (the active ones after a partial tree)."""
        perm = np.full(max(self.n, 1), -1, np.int32)
        keys = np.zeros(max(self.n, 1), np.uint64)
        self._check(self.L.agrav_get_order(self.h, _ptr(perm), _ptr(keys)))
        k = int(np.count_nonzero(perm >= 0))
        return perm[:k], keys[:k]

    def get_tables(self, ewald_tree=False, ewald_forcetest=False):
        tf = np.zeros(128, np.float32)
        tp = np.zeros(128, np.float32)
        et = np.zeros(4 * 65 ** 3) if ewald_tree else None
        ef = np.zeros(4 * 129 ** 3) if ewald_forcetest else None
        self._check(self.L.agrav_get_tables(self.h, _ptr(tf), _ptr(tp), _ptr(et), _ptr(ef)))
        return dict(tab_force=tf, tab_pot=tp, ewald_tree=et, ewald_forcetest=ef)


def solve_footprint(params: Params, n):
    """The serving footprint model: device bytes one solve of n particles adds to a fresh context."""
    return int(lib().agrav_solve_footprint(C.byref(params), int(n)))


def hilbert_key21(x, y, z):
    return int(lib().agrav_hilbert_key21(int(x), int(y), int(z)))


class Server:
    """Continuous-batching server (agrav_server): `slots` resident solver contexts with their own streams."""

    def __init__(self, params: Params, slots=4, hbm_budget_bytes=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.agrav_server_create(C.byref(params), slots, hbm_budget_bytes, C.byref(h))
        if rc != OK:
            raise AgravError(rc, (self.L.agrav_last_error(None) or b"").decode())
        self.h = h
        self._keep = {}

    def submit(self, pos, mass, softening=0.0, softtype=None, potential=True):
        """Queue one solve; returns a ticket.  Arrays are kept alive until wait()."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        st = None if softtype is None else np.ascontiguousarray(softtype, dtype=np.uint8)
        n = len(pos)
        acc = np.zeros((n, 3))
        pot = np.zeros(n) if potential else None
        r = Request(n, _ptr(pos), _ptr(mass), _ptr(st), float(softening), _ptr(acc), _ptr(pot))
        t = C.c_uint64()
        rc = self.L.agrav_server_submit(self.h, C.byref(r), C.byref(t))
        if rc != OK:
            raise AgravError(rc, "submit failed")
        self._keep[t.value] = (pos, mass, st, acc, pot)
        return t.value

    def wait(self, ticket):
        if ticket not in self._keep:
            raise AgravError(ERR_BAD_PARAM, "unknown ticket %r" % (ticket,))
        res = RequestResult()
        rc = self.L.agrav_server_wait(self.h, ticket, C.byref(res))
        pos, mass, st, acc, pot = self._keep.pop(ticket)
        if rc != OK:
            raise AgravError(rc, "request failed")
        return dict(acc=acc, pot=pot, queue_ms=res.queue_ms, run_ms=res.run_ms, total_ms=res.total_ms, slot=res.slot)

    def poll(self, ticket):
        d = C.c_int()
        self.L.agrav_server_poll(self.h, ticket, C.byref(d))
        return bool(d.value)

    def stats(self):
        s = ServerStats()
        self.L.agrav_server_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.agrav_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
