"""ctypes binding of librbh (include/rbh.h).

The library is built in-tree (``rainbow_amd/librbh.so``) by ``build.py``.  There
is no fallback: when the library or a GPU is missing, ``load()`` and
``Context()`` raise, so no step can silently run anywhere but on the device.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbh.so")

RBH_E_ARG, RBH_E_HIP, RBH_E_MEMORY, RBH_E_CAPACITY, RBH_E_COLORS, RBH_E_STATE = -1, -2, -3, -4, -5, -6
SCHEMES = {"gauss_seidel": 0, "parallel_gauss_seidel": 1, "parallel_jacobi": 2,
           "parallel_jacobi_hybrid": 3, "parallel_jacboi_hybrid": 3}
FRICTION = {"isotropic": 0, "anisotropic": 1}
STEP_COUNT_OVERLAPS, STEP_TIMING, STEP_GRAPH, STEP_FORCE_SINGLE, STEP_FORCE_MULTI = 1, 2, 4, 8, 16
TIMING_KEYS = ["update_bvh", "narrow_phase", "contact_determination", "contact_assembly",
               "coloring", "solver_time", "integration", "stepper_time", "contact_point_reduction",
               "post_stabilization_solver_time"]


class RbhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "time_step", "envelope", "absolute_tolerance", "relative_tolerance", "nu_reduce", "nu_increase",
        "too_small_merit_change", "contact_optimization_tolerance", "gap_reduction", "min_gap_value",
        "max_gap_value", "ellipsoid_tolerance", "ellipsoid_expansion")] + \
        [(n, C.c_int) for n in ("max_iterations", "contact_optimization_max_iterations",
                                "ellipsoid_max_iterations", "use_bounce", "use_pre_stabilization",
                                "proximal_solver", "friction_model", "use_post_stabilization")] + \
        [("coloring_seed", C.c_ulonglong)]


class SceneDesc(C.Structure):
    _fields_ = [("n_bodies", C.c_int), ("shape", C.POINTER(C.c_int)), ("fixed", C.POINTER(C.c_int)),
                ("material", C.POINTER(C.c_int)), ("mass", C.POINTER(C.c_double)),
                ("inertia", C.POINTER(C.c_double)), ("gravity", C.POINTER(C.c_double)),
                ("damping", C.POINTER(C.c_double)), ("n_materials", C.c_int),
                ("interactions", C.POINTER(C.c_double)), ("params", Params)]


class StepLog(C.Structure):
    _fields_ = [("step_index", C.c_int), ("contact_points", C.c_int), ("iterations", C.c_int),
                ("exitcode", C.c_int), ("rejects", C.c_int), ("sweeps", C.c_int), ("colors", C.c_int),
                ("solved", C.c_int), ("number_of_overlaps", C.c_ulonglong), ("fw_cell_evals", C.c_ulonglong),
                ("natural_residual", C.c_double), ("max_penetration", C.c_double), ("final_merit", C.c_double),
                ("post_iterations", C.c_int), ("post_exitcode", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("unit_size", "lanes_per_scene", "resident_contacts", "resident_bodies",
                                        "lanes_per_group", "waves_per_group", "color_lds_contacts",
                                        "color_body_bits", "solve_groups_per_cu")]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("compute_units", C.c_int), ("lds_per_block", C.c_int),
                ("hbm_bytes", C.c_ulonglong), ("precision_bytes", C.c_int), ("max_colors", C.c_int)]


_lib = None
_LIVE = weakref.WeakSet()      # open contexts, closed by close_all() at interpreter exit


def close_all():
    """Destroy every open context.  Registered with atexit, so device buffers,
    graphs, events and streams are released while the HIP runtime is still
    up, not from object finalizers during interpreter teardown (which run
    after the runtime's and a profiler's own exit handlers)."""
    for ctx in list(_LIVE):
        ctx.close()


atexit.register(close_all)

# (name, argtypes) of every entry point declared in include/rbh.h and include/rbh_testing.h
_PROTOS = {
    "rbh_version": [],
    "rbh_ctx_create": [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)],
    "rbh_ctx_destroy": [C.c_void_p],
    "rbh_last_error": [C.c_void_p],
    "rbh_device_query": [C.c_void_p, C.POINTER(DeviceInfo)],
    "rbh_memory_usage": [C.c_void_p, C.POINTER(C.c_ulonglong)],
    "rbh_sdf_build": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                      C.c_int, C.c_int, C.c_int, C.c_void_p],
    "rbh_sdf_build_noalloc": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                              C.c_int, C.c_int, C.c_int, C.c_void_p],
    "rbh_shape_upload": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                         C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)],
    "rbh_scene_create": [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_int)],
    "rbh_scene_destroy": [C.c_void_p, C.c_int],
    "rbh_scene_set_params": [C.c_void_p, C.c_int, C.POINTER(Params)],
    "rbh_scene_set_hinges": [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p],
    "rbh_scene_set_state": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "rbh_scene_get_state": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "rbh_scene_set_step_index": [C.c_void_p, C.c_int, C.c_int],
    "rbh_scene_footprint": [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_ulonglong)],
    "rbh_step": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int],
    "rbh_sync": [C.c_void_p],
    "rbh_get_log": [C.c_void_p, C.c_int, C.POINTER(StepLog), C.c_void_p, C.c_int],
    "rbh_get_contacts": [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "rbh_get_timings": [C.c_void_p, C.c_void_p, C.c_int],
    "rbh_get_counters": [C.c_void_p, C.c_void_p, C.c_int],
    "rbh_scene_totals": [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int],
    "rbh_set_tuning": [C.c_void_p, C.c_void_p],
    "rbh_get_tuning": [C.c_void_p, C.c_void_p],
    # include/rbh_testing.h
    "rbh_prefix_sum": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)],
}


def load():
    """Load librbh.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("librbh.so is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "(expected at %s)" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, args in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "rbh_last_error" else C.c_int
    _lib = lib
    return lib


def exported_symbols():
    return list(_PROTOS)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One device context (one HIP stream).  precision: 'fp32' or 'fp64'."""

    def __init__(self, device=0, precision="fp32", hbm_budget=0):
        self.lib = load()
        self.precision = precision
        h = C.c_void_p()
        prec = {"fp32": 4, "fp64": 8}[precision]
        rc = self.lib.rbh_ctx_create(int(device), prec, C.c_ulonglong(int(hbm_budget)), C.byref(h))
        if rc != 0:
            raise RbhError(rc, "rbh_ctx_create failed (code %d): no usable HIP device %d" % (rc, device))
        self.h = h
        self._shape_cache = {}
        _LIVE.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.rbh_ctx_destroy(self.h)
            self.h = None
        _LIVE.discard(self)

    @property
    def closed(self):
        return not getattr(self, "h", None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            msg = self.lib.rbh_last_error(self.h).decode()
            if rc == RBH_E_ARG:
                err = RbhError(rc, msg)
            else:
                err = RbhError(rc, "librbh error %d: %s" % (rc, msg))
            raise err
        return rc

    def device_info(self):
        info = DeviceInfo()
        self._check(self.lib.rbh_device_query(self.h, C.byref(info)))
        return dict(name=info.name.decode(), compute_units=info.compute_units, lds_per_block=info.lds_per_block,
                    hbm_bytes=info.hbm_bytes, precision_bytes=info.precision_bytes, max_colors=info.max_colors)

    def memory_usage(self):
        b = C.c_ulonglong()
        self._check(self.lib.rbh_memory_usage(self.h, C.byref(b)))
        return int(b.value)

    def sdf_build(self, V, T, min_coord, max_coord, I, J, K):
        V = _f64(V)
        T = _i32(T)
        out = np.empty(int(I) * int(J) * int(K), dtype=np.float64)
        self._check(self.lib.rbh_sdf_build(self.h, _ptr(V), len(V), _ptr(T), len(T), _ptr(_f64(min_coord, 3)),
                                           _ptr(_f64(max_coord, 3)), int(I), int(J), int(K), _ptr(out)))
        return out

    def prefix_sum(self, values, cap=None):
        """The step's exclusive prefix sum of a uint8 or int32 array (a test hook,
        include/rbh_testing.h): returns (out[0..n], total), the launch sized for cap."""
        a = np.ascontiguousarray(values)
        if a.dtype not in (np.uint8, np.int32):
            raise TypeError("prefix_sum takes uint8 or int32 values")
        n = int(a.size)
        cap = max(n, 1) if cap is None else int(cap)
        out = np.empty(n + 1, dtype=np.int32)
        tot = C.c_int()
        self._check(self.lib.rbh_prefix_sum(self.h, a.ctypes.data_as(C.c_void_p), a.dtype.itemsize, n, cap,
                                            out.ctypes.data_as(C.c_void_p), C.byref(tot)))
        return out, tot.value

    def sdf_build_noalloc(self, V, T, min_coord, max_coord, I, J, K):
        """sdf_build staged in the library's fixed per-GPU array (no device allocation)."""
        V = _f64(V)
        T = _i32(T)
        out = np.empty(int(I) * int(J) * int(K), dtype=np.float64)
        self._check(self.lib.rbh_sdf_build_noalloc(self.h, _ptr(V), len(V), _ptr(T), len(T),
                                                   _ptr(_f64(min_coord, 3)), _ptr(_f64(max_coord, 3)),
                                                   int(I), int(J), int(K), _ptr(out)))
        return out

    def shape_upload(self, V, T, values, I, J, K, min_coord, max_coord):
        V, T, values = _f64(V), _i32(T), _f64(values)
        key = (V.tobytes(), T.tobytes(), values.tobytes(), int(I), int(J), int(K),
               _f64(min_coord).tobytes(), _f64(max_coord).tobytes())
        hk = hash(key)
        if hk in self._shape_cache and self._shape_cache[hk][0] == key:
            return self._shape_cache[hk][1]
        sid = C.c_int()
        self._check(self.lib.rbh_shape_upload(self.h, _ptr(V), len(V), _ptr(T), len(T), _ptr(values), int(I),
                                              int(J), int(K), _ptr(_f64(min_coord, 3)), _ptr(_f64(max_coord, 3)),
                                              C.byref(sid)))
        self._shape_cache[hk] = (key, sid.value)
        return sid.value

    @staticmethod
    def make_desc(shape, fixed, material, mass, inertia, gravity, damping, interactions, params):
        """Build an rbh_scene_desc; returns (desc, keepalive arrays)."""
        arrs = dict(shape=_i32(shape), fixed=_i32(fixed), material=_i32(material), mass=_f64(mass),
                    inertia=_f64(inertia), gravity=_f64(gravity), damping=_f64(damping),
                    interactions=_f64(interactions))
        n = len(arrs["shape"])
        nmat = int(round(np.sqrt(arrs["interactions"].size / 4)))
        d = SceneDesc()
        d.n_bodies = n
        d.shape = arrs["shape"].ctypes.data_as(C.POINTER(C.c_int))
        d.fixed = arrs["fixed"].ctypes.data_as(C.POINTER(C.c_int))
        d.material = arrs["material"].ctypes.data_as(C.POINTER(C.c_int))
        d.mass = arrs["mass"].ctypes.data_as(C.POINTER(C.c_double))
        d.inertia = arrs["inertia"].ctypes.data_as(C.POINTER(C.c_double))
        d.gravity = arrs["gravity"].ctypes.data_as(C.POINTER(C.c_double))
        d.damping = arrs["damping"].ctypes.data_as(C.POINTER(C.c_double))
        d.n_materials = nmat
        d.interactions = arrs["interactions"].ctypes.data_as(C.POINTER(C.c_double))
        d.params = params
        return d, arrs

    def scene_create(self, desc):
        sid = C.c_int()
        self._check(self.lib.rbh_scene_create(self.h, C.byref(desc), C.byref(sid)))
        return sid.value

    def scene_footprint(self, desc):
        b = C.c_ulonglong()
        self._check(self.lib.rbh_scene_footprint(self.h, C.byref(desc), C.byref(b)))
        return int(b.value)

    def scene_destroy(self, sid):
        self._check(self.lib.rbh_scene_destroy(self.h, int(sid)))

    def set_params(self, sid, params):
        self._check(self.lib.rbh_scene_set_params(self.h, int(sid), C.byref(params)))

    def set_hinges(self, sid, parent, child, arm_p, arm_c, axis_p, axis_c):
        """Hinges of the scene (scene body indices; body-frame arms and axes, (n, 3))."""
        parent, child = _i32(parent), _i32(child)
        arrs = [_f64(np.asarray(a, dtype=np.float64).reshape(-1, 3)) for a in (arm_p, arm_c, axis_p, axis_c)]
        self._check(self.lib.rbh_scene_set_hinges(self.h, int(sid), len(parent), _ptr(parent), _ptr(child),
                                                  *[_ptr(a) for a in arrs]))

    def set_state(self, sid, r=None, q=None, v=None, w=None):
        arrs = [None if a is None else _f64(a) for a in (r, q, v, w)]
        self._check(self.lib.rbh_scene_set_state(self.h, int(sid), *[_ptr(a) for a in arrs]))

    def set_step_index(self, sid, k):
        self._check(self.lib.rbh_scene_set_step_index(self.h, int(sid), int(k)))

    def get_state(self, sid, n):
        r, q, v, w = np.empty((n, 3)), np.empty((n, 4)), np.empty((n, 3)), np.empty((n, 3))
        self._check(self.lib.rbh_scene_get_state(self.h, int(sid), _ptr(r), _ptr(q), _ptr(v), _ptr(w)))
        return r, q, v, w

    def step(self, scene_ids, n_steps=1, dt=0.0, flags=0):
        ids = _i32(scene_ids)
        self._check(self.lib.rbh_step(self.h, _ptr(ids), len(ids), int(n_steps), float(dt), int(flags)))

    def sync(self):
        self._check(self.lib.rbh_sync(self.h))

    def get_log(self, sid, max_iterations):
        lg = StepLog()
        mi = max(1, int(max_iterations))
        res = np.empty(2 * mi)
        self._check(self.lib.rbh_get_log(self.h, int(sid), C.byref(lg), _ptr(res), len(res)))
        d = {k: getattr(lg, k) for k, _ in StepLog._fields_}
        d["residuals"] = res[:int(max_iterations)]
        d["post_residuals"] = res[mi:mi + int(max_iterations)]
        return d

    def get_contacts(self, sid):
        cnt = C.c_int()
        self._check(self.lib.rbh_get_contacts(self.h, int(sid), 0, C.byref(cnt), None, None, None, None, None,
                                              None, None, None))
        K = cnt.value
        out = dict(ia=np.empty(K, np.int32), ib=np.empty(K, np.int32), tri=np.empty(K, np.int32),
                   p=np.empty((K, 3)), n=np.empty((K, 3)), gap=np.empty(K), impulse=np.empty((K, 4)),
                   color=np.empty(K, np.int32))
        if K:
            self._check(self.lib.rbh_get_contacts(self.h, int(sid), K, C.byref(cnt), *[_ptr(out[k]) for k in (
                "ia", "ib", "tri", "p", "n", "gap", "impulse", "color")]))
        return out

    def get_timings(self):
        ms = np.zeros(len(TIMING_KEYS))
        n = self._check(self.lib.rbh_get_timings(self.h, _ptr(ms), len(ms)))
        return dict(zip(TIMING_KEYS[:n], ms[:n] / 1e3))

    def get_counters(self):
        c = np.zeros(4, dtype=np.int64)
        self._check(self.lib.rbh_get_counters(self.h, _ptr(c), 4))
        return dict(zip(["body_pairs", "candidate_slots", "candidates", "contacts"], c.tolist()))

    def set_tuning(self, **kw):
        """Launch shapes of the step (rbh_tuning fields; omitted ones keep their value)."""
        t = Tuning()
        self._check(self.lib.rbh_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        self._check(self.lib.rbh_set_tuning(self.h, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.rbh_get_tuning(self.h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in Tuning._fields_}

    TOTAL_KEYS = ("steps", "contact_iters", "contacts", "candidates", "fw_cell_evals", "sweeps")

    def scene_totals(self, sid, clear=False):
        """Running work totals of a scene (kept on the device)."""
        t = np.zeros(len(self.TOTAL_KEYS), dtype=np.uint64)
        self._check(self.lib.rbh_scene_totals(self.h, int(sid), _ptr(t), len(t), 1 if clear else 0))
        return dict(zip(self.TOTAL_KEYS, (int(x) for x in t)))
