"""ctypes bindings of the C ABI declared in include/xdem.h.

The HIP library (exadem_amd/lib/libxdem_hip.so) is the only compute path: when it is missing
or cannot be loaded, every entry point raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

XDEM_OK = 0
XDEM_FORM_FP64 = 0
XDEM_FORM_FP32 = 1
ERR_NAMES = {-1: "XDEM_ERR_ARG", -2: "XDEM_ERR_HIP", -3: "XDEM_ERR_STATE", -4: "XDEM_ERR_CAPACITY",
             -5: "XDEM_ERR_UNSUPPORTED"}

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_fp = C.POINTER(C.c_float)


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("id", _u64p), ("type", _u32p), ("group", _u32p)] + [
        (name, _dp) for name in (
            "rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "radius", "homothety",
            "qw", "qx", "qy", "qz", "momx", "momy", "momz", "vrotx", "vroty", "vrotz",
            "arotx", "aroty", "arotz", "inx", "iny", "inz")]


class Domain(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("periodic", C.c_int32 * 3),
                ("cell_size", C.c_double)]


class ContactParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("kn", "kt", "kr", "mu", "damp_rate", "fc", "dncut", "gamma")]


DRIVER_SURFACE, DRIVER_CYLINDER, DRIVER_BALL = 0, 1, 2


class Surface(C.Structure):
    _fields_ = [("offset", C.c_double), ("normal", C.c_double * 3), ("center", C.c_double * 3),
                ("vel", C.c_double * 3), ("vrot", C.c_double * 3), ("kind", C.c_int32), ("pad", C.c_int32),
                ("radius", C.c_double), ("axis", C.c_double * 3)]


class DriverMotion(C.Structure):
    _fields_ = [("motion_type", C.c_int32), ("pad", C.c_int32), ("motion_vector", C.c_double * 3),
                ("const_vel", C.c_double), ("motion_start_threshold", C.c_double),
                ("motion_end_threshold", C.c_double), ("amplitude", C.c_double), ("omega", C.c_double),
                ("shaker_dir", C.c_double * 3)]


MOTION_TYPES = {"STATIONARY": 0, "LINEAR_MOTION": 1, "SHAKER": 7}


class Shape(C.Structure):
    _fields_ = [("nv", C.c_int32), ("ne", C.c_int32), ("nf", C.c_int32), ("pad", C.c_int32),
                ("minkowski", C.c_double), ("volume", C.c_double), ("inertia_on_mass", C.c_double * 3),
                ("obb_extent", C.c_double * 3), ("vertices", (C.c_double * 3) * 8), ("edges", (C.c_int32 * 2) * 12),
                ("face_size", C.c_int32 * 8), ("faces", (C.c_int32 * 4) * 8)]


class Scheme(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3), ("gravity_on", C.c_int32),
                ("form", C.c_int32), ("cp", ContactParams), ("cpd", ContactParams),
                ("rcut_inc", C.c_double), ("poll", C.c_int32)]


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("rebuilds", C.c_int64), ("polls", C.c_int64)]


class GnnConfig(C.Structure):
    _fields_ = [("r0", C.c_double), ("v0", C.c_double), ("a0", C.c_double), ("m0", C.c_double),
                ("ov_scale", C.c_double), ("substeps", C.c_int32), ("train", C.c_int32)]


class SurrogateStats(C.Structure):
    _fields_ = [("surrogate_steps", C.c_int64), ("dem_steps", C.c_int64), ("checks", C.c_int64),
                ("fallbacks", C.c_int64), ("rebuilds", C.c_int64), ("max_err", C.c_double),
                ("max_err_free", C.c_double), ("rms_err", C.c_double), ("rms_err_free", C.c_double)]


class SimState(C.Structure):
    _fields_ = [("kinetic_energy", C.c_double * 3), ("rotation_energy", C.c_double * 3), ("mass", C.c_double),
                ("volume", C.c_double), ("min_dn", C.c_double), ("n_particles", C.c_int64),
                ("n_interactions", C.c_int64), ("n_active", C.c_int64)]


GNN_NPARAMS = 205446


# field name -> (numpy dtype) for the xdem_particles members
FIELDS = ["id", "type", "group", "rx", "ry", "rz", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "radius",
          "homothety", "qw", "qx", "qy", "qz", "momx", "momy", "momz", "vrotx", "vroty", "vrotz",
          "arotx", "aroty", "arotz", "inx", "iny", "inz"]
DTYPES = {"id": np.uint64, "type": np.uint32, "group": np.uint32}

# exported symbols of libxdem_hip.so, one per declaration in include/xdem.h
HIP_SYMBOLS = [
    "xdem_ctx_create", "xdem_ctx_destroy", "xdem_last_error", "xdem_synchronize", "xdem_set_domain",
    "xdem_set_surfaces", "xdem_particles_upload", "xdem_particles_download", "xdem_num_particles",
    "xdem_combined_compute_prolog", "xdem_combined_compute_epilog", "xdem_reset_force_moment",
    "xdem_gravity_force", "xdem_force_to_accel", "xdem_particle_displ_over", "xdem_move_particles",
    "xdem_chunk_neighbors_contact", "xdem_nbh_sphere", "xdem_classify_interactions", "xdem_classifier_waves",
    "xdem_classifier_wave_dump", "xdem_contact_sphere",
    "xdem_run_steps", "xdem_interactions_count", "xdem_interactions_dump", "xdem_interactions_load", "xdem_poly_interactions_load", "xdem_interactions_moments",
    "xdem_poly_interactions_moments", "xdem_simulation_state", "xdem_simulation_state_noalloc",
    "xdem_set_driver_motion", "xdem_push_f_v_r_driver", "xdem_push_f_v_driver", "xdem_driver_displ_over",
    "xdem_drivers_download", "xdem_set_physical_time", "xdem_physical_time", "xdem_next_time_step",
    "xdem_set_contact_record", "xdem_contact_state_count", "xdem_contact_state_dump", "xdem_nbh_stats",
    "xdem_kernel_timing", "xdem_set_timing", "xdem_streams_dump",
    "xdem_set_shapes", "xdem_compute_vertices", "xdem_vertex_displ_over", "xdem_nbh_polyhedron",
    "xdem_contact_polyhedron", "xdem_poly_interactions_count", "xdem_poly_interactions_dump",
    "xdem_gnn_create", "xdem_gnn_destroy", "xdem_gnn_set_params", "xdem_gnn_get_params", "xdem_gnn_get_grads",
    "xdem_gnn_forward", "xdem_gnn_set_target", "xdem_gnn_backward", "xdem_gnn_adam", "xdem_gnn_train_window",
    "xdem_gnn_dump_graph", "xdem_surrogate_run", "xdem_set_mesh", "xdem_mesh_interactions",
    "xdem_mesh_interactions_load",
]


class XdemError(RuntimeError):
    pass


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def particles_struct(arrays, n):
    """Builds an xdem_particles view over a dict of contiguous numpy arrays (missing -> NULL)."""
    p = Particles()
    p.n = n
    keep = []
    for name in FIELDS:
        a = arrays.get(name)
        if a is None:
            continue
        dt = DTYPES.get(name, np.float64)
        assert a.dtype == dt and a.flags.c_contiguous and a.shape[0] >= n, name
        keep.append(a)
        ct = {np.uint64: C.c_uint64, np.uint32: C.c_uint32}.get(dt, C.c_double)
        setattr(p, name, _ptr(a, ct))
    return p, keep


_lib = None


def lib_dir():
    """Directory of the built libraries (XDEM_LIB_DIR overrides the in-tree exadem_amd/lib)."""
    return os.environ.get("XDEM_LIB_DIR", LIBDIR)


def lib_path():
    return os.path.join(lib_dir(), "libxdem_hip.so")


def load():
    """Loads libxdem_hip.so (raises XdemError when it is absent: there is no other compute path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise XdemError(f"{path} is missing: build it with exadem_amd/build.py (hipcc, gfx950)")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    vp = C.c_void_p
    sig = {
        "xdem_ctx_create": (C.c_int, [C.c_int32, C.POINTER(vp)]),
        "xdem_ctx_destroy": (None, [vp]),
        "xdem_last_error": (C.c_int, [vp, C.c_char_p, C.c_int64]),
        "xdem_synchronize": (C.c_int, [vp]),
        "xdem_set_domain": (C.c_int, [vp, C.POINTER(Domain)]),
        "xdem_set_surfaces": (C.c_int, [vp, C.POINTER(Surface), C.c_int32]),
        "xdem_particles_upload": (C.c_int, [vp, C.POINTER(Particles)]),
        "xdem_particles_download": (C.c_int, [vp, C.POINTER(Particles)]),
        "xdem_num_particles": (C.c_int64, [vp]),
        "xdem_combined_compute_prolog": (C.c_int, [vp, C.c_double]),
        "xdem_combined_compute_epilog": (C.c_int, [vp, C.c_double]),
        "xdem_reset_force_moment": (C.c_int, [vp]),
        "xdem_gravity_force": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "xdem_force_to_accel": (C.c_int, [vp]),
        "xdem_particle_displ_over": (C.c_int, [vp, C.c_double, C.POINTER(C.c_int32)]),
        "xdem_move_particles": (C.c_int, [vp]),
        "xdem_chunk_neighbors_contact": (C.c_int, [vp, C.c_double, C.c_double]),
        "xdem_nbh_sphere": (C.c_int, [vp, C.c_double, C.c_int32]),
        "xdem_classify_interactions": (C.c_int, [vp]),
        "xdem_classifier_waves": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "xdem_classifier_wave_dump": (C.c_int, [vp, C.c_int32, C.c_int64, _u64p, _u64p, C.POINTER(C.c_uint8), _u32p,
                                                _u32p, _u32p, C.POINTER(C.c_int64)]),
        "xdem_contact_sphere": (C.c_int, [vp, C.POINTER(ContactParams), C.POINTER(ContactParams), C.c_double,
                                          C.c_int32]),
        "xdem_run_steps": (C.c_int, [vp, C.POINTER(Scheme), C.c_int64, C.POINTER(RunStats)]),
        "xdem_interactions_count": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "xdem_interactions_dump": (C.c_int, [vp, C.c_int64, _u64p, _u64p, C.POINTER(C.c_uint16), _dp,
                                             C.POINTER(C.c_int64)]),
        "xdem_interactions_load": (C.c_int, [vp, C.c_int64, _u64p, _u64p, C.POINTER(C.c_uint16), _dp, _dp,
                                             C.POINTER(C.c_int64)]),
        "xdem_poly_interactions_load": (C.c_int, [vp, C.c_int64, _u64p, _u64p, C.POINTER(C.c_uint16),
                                                  C.POINTER(C.c_uint8), _u32p, _u32p, _dp, _dp, C.POINTER(C.c_int64)]),
        "xdem_interactions_moments": (C.c_int, [vp, C.c_int64, _dp, C.POINTER(C.c_int64)]),
        "xdem_poly_interactions_moments": (C.c_int, [vp, C.c_int64, _dp, C.POINTER(C.c_int64)]),
        "xdem_simulation_state": (C.c_int, [vp, C.POINTER(SimState)]),
        "xdem_simulation_state_noalloc": (C.c_int, [vp, C.POINTER(SimState)]),
        "xdem_set_driver_motion": (C.c_int, [vp, C.POINTER(DriverMotion), C.c_int32]),
        "xdem_push_f_v_r_driver": (C.c_int, [vp, C.c_double, C.c_double]),
        "xdem_push_f_v_driver": (C.c_int, [vp, C.c_double]),
        "xdem_driver_displ_over": (C.c_int, [vp, C.c_double, C.POINTER(C.c_int32)]),
        "xdem_drivers_download": (C.c_int, [vp, C.POINTER(Surface), C.c_int32]),
        "xdem_set_physical_time": (C.c_int, [vp, C.c_double]),
        "xdem_physical_time": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "xdem_next_time_step": (C.c_int, [vp, C.c_double]),
        "xdem_set_mesh": (C.c_int, [vp, C.c_int32, _dp, C.c_int64, _u32p, C.c_int64, _u32p, C.c_int64, C.c_double,
                                    _dp, _dp, C.c_int32]),
        "xdem_mesh_interactions": (C.c_int, [vp, C.c_int64, _u64p, _u32p, C.POINTER(C.c_uint16), _dp, _dp,
                                             C.POINTER(C.c_int64)]),
        "xdem_mesh_interactions_load": (C.c_int, [vp, C.c_int64, _u64p, _u32p, C.POINTER(C.c_uint16), _dp, _dp,
                                                  C.POINTER(C.c_int64)]),
        "xdem_set_contact_record": (C.c_int, [vp, C.c_int32]),
        "xdem_contact_state_count": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "xdem_contact_state_dump": (C.c_int, [vp, C.c_int64, _u64p, _u64p, _u32p, _u32p, C.POINTER(C.c_uint16),
                                              _dp, C.POINTER(C.c_int64)]),
        "xdem_streams_dump": (C.c_int, [vp, C.c_int64, _u64p, _u64p, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_uint8), C.POINTER(C.c_int64)]),
        "xdem_nbh_stats": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
        "xdem_kernel_timing": (C.c_int, [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "xdem_set_timing": (C.c_int, [vp, C.c_int32]),
        "xdem_set_shapes": (C.c_int, [vp, C.POINTER(Shape), C.c_int32]),
        "xdem_compute_vertices": (C.c_int, [vp]),
        "xdem_vertex_displ_over": (C.c_int, [vp, C.c_double, C.POINTER(C.c_int32)]),
        "xdem_nbh_polyhedron": (C.c_int, [vp, C.c_double]),
        "xdem_contact_polyhedron": (C.c_int, [vp, C.POINTER(ContactParams), C.POINTER(ContactParams), C.c_double]),
        "xdem_poly_interactions_count": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "xdem_poly_interactions_dump": (C.c_int, [vp, C.c_int64, _u64p, _u64p, C.POINTER(C.c_uint16),
                                                  C.POINTER(C.c_uint8), _u32p, _u32p, _dp, C.POINTER(C.c_int64)]),
        "xdem_gnn_create": (C.c_int, [vp, C.POINTER(GnnConfig), C.POINTER(vp)]),
        "xdem_gnn_destroy": (C.c_int, [vp]),
        "xdem_gnn_set_params": (C.c_int, [vp, _fp, C.c_int64]),
        "xdem_gnn_get_params": (C.c_int, [vp, _fp, C.c_int64]),
        "xdem_gnn_get_grads": (C.c_int, [vp, _fp, C.c_int64]),
        "xdem_gnn_forward": (C.c_int, [vp, _fp]),
        "xdem_gnn_set_target": (C.c_int, [vp, _fp]),
        "xdem_gnn_backward": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "xdem_gnn_adam": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
        "xdem_gnn_train_window": (C.c_int, [vp, C.POINTER(Scheme), C.c_double, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int32)]),
        "xdem_gnn_dump_graph": (C.c_int, [vp, C.c_int64, _u32p, _u32p, C.POINTER(C.c_uint16),
                                          C.POINTER(C.c_uint16), C.POINTER(C.c_int64)]),
        "xdem_surrogate_run": (C.c_int, [vp, C.POINTER(Scheme), C.c_int64, C.c_int32, C.c_double,
                                         C.POINTER(SurrogateStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def contact_params(kn=1e5, kt=None, kr=0.0, mu=0.3, damp_rate=0.9, fc=0.0, dncut=0.0, gamma=0.0):
    return ContactParams(kn, 0.8 * kn if kt is None else kt, kr, mu, damp_rate, fc, dncut, gamma)


def domain(lo, hi, periodic, cell_size):
    d = Domain()
    for a in range(3):
        d.lo[a] = lo[a]
        d.hi[a] = hi[a]
        d.periodic[a] = 1 if periodic[a] else 0
    d.cell_size = cell_size
    return d


def surfaces(walls):
    """walls: list of driver dicts -> ctypes array. A plane: {offset, normal[, center, vel, vrot]}; a
    cylinder: {"kind": "cylinder", radius, axis, center[, vel, vrot]}; a ball: {"kind": "ball", radius,
    center[, vel, vrot]}."""
    kinds = {"surface": DRIVER_SURFACE, "cylinder": DRIVER_CYLINDER, "ball": DRIVER_BALL}
    arr = (Surface * max(1, len(walls)))()
    for k, w in enumerate(walls):
        arr[k].kind = kinds[w.get("kind", "surface")]
        arr[k].offset = w.get("offset", 0.0)
        arr[k].radius = w.get("radius", 0.0)
        for a in range(3):
            arr[k].normal[a] = w.get("normal", (0, 0, 0))[a]
            arr[k].center[a] = w.get("center", (0, 0, 0))[a]
            arr[k].vel[a] = w.get("vel", (0, 0, 0))[a]
            arr[k].vrot[a] = w.get("vrot", (0, 0, 0))[a]
            arr[k].axis[a] = w.get("axis", (0, 0, 0))[a]
    return arr


def has_motion(walls):
    return any(w.get("motion", {}).get("motion_type", "STATIONARY") != "STATIONARY" for w in walls)


def motions(walls):
    """walls: list of driver dicts -> ctypes array of DriverMotion from each dict's "motion" entry (the
    params slot of register_*: motion_type name, motion_vector, const_vel, motion_start_threshold,
    motion_end_threshold, amplitude, omega, shaker_dir); no entry = STATIONARY."""
    arr = (DriverMotion * max(1, len(walls)))()
    for k, w in enumerate(walls):
        m = w.get("motion", {})
        arr[k].motion_type = MOTION_TYPES[m.get("motion_type", "STATIONARY")]
        arr[k].const_vel = m.get("const_vel", 0.0)
        arr[k].motion_start_threshold = m.get("motion_start_threshold", 0.0)
        arr[k].motion_end_threshold = m.get("motion_end_threshold", 1e300)
        arr[k].amplitude = m.get("amplitude", 0.0)
        arr[k].omega = m.get("omega", 0.0)
        for a in range(3):
            arr[k].motion_vector[a] = m.get("motion_vector", (0, 0, 0))[a]
            arr[k].shaker_dir[a] = m.get("shaker_dir", (0, 0, 1))[a]
    return arr


def driver_rows(arr, n):
    """xdem_surface rows -> list of dicts (offset, normal, center, vel, vrot, kind, radius, axis)."""
    return [{"offset": arr[k].offset, "normal": np.array(arr[k].normal[:]), "center": np.array(arr[k].center[:]),
             "vel": np.array(arr[k].vel[:]), "vrot": np.array(arr[k].vrot[:]), "kind": int(arr[k].kind),
             "radius": arr[k].radius, "axis": np.array(arr[k].axis[:])} for k in range(n)]


class Context:
    """One xdem_ctx on one GPU. Methods map 1:1 to the operator entry points of include/xdem.h."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.xdem_ctx_create(device, C.byref(h))
        if rc != XDEM_OK:
            raise XdemError(f"xdem_ctx_create failed: {ERR_NAMES.get(rc, rc)}")
        self.h = h

    def close(self):
        if self.h:
            self.lib.xdem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != XDEM_OK:
            buf = C.create_string_buffer(512)
            self.lib.xdem_last_error(self.h, buf, 512)
            raise XdemError(f"{what}: {ERR_NAMES.get(rc, rc)}: {buf.value.decode()}")

    def call(self, name, *args):
        rc = getattr(self.lib, name)(self.h, *args)
        self._check(rc, name)

    # ---- data
    def set_domain(self, dom):
        self.call("xdem_set_domain", C.byref(dom))

    def set_surfaces(self, walls):
        arr = surfaces(walls)
        self.call("xdem_set_surfaces", arr, len(walls))
        self.nwalls = len(walls)
        if has_motion(walls):
            self.call("xdem_set_driver_motion", motions(walls), len(walls))

    def set_driver_motion(self, walls):
        self.call("xdem_set_driver_motion", motions(walls), len(walls))

    def push_f_v_r_driver(self, time, dt):
        self.call("xdem_push_f_v_r_driver", time, dt)

    def push_f_v_driver(self, dt):
        self.call("xdem_push_f_v_driver", dt)

    def driver_displ_over(self, threshold):
        r = C.c_int32()
        self.call("xdem_driver_displ_over", threshold, C.byref(r))
        return bool(r.value)

    def drivers(self):
        n = getattr(self, "nwalls", 0)
        arr = (Surface * max(1, n))()
        self.call("xdem_drivers_download", arr, n)
        return driver_rows(arr, n)

    def physical_time(self):
        t = C.c_double()
        self.call("xdem_physical_time", C.byref(t))
        return t.value

    def set_physical_time(self, t):
        self.call("xdem_set_physical_time", t)

    def next_time_step(self, dt):
        self.call("xdem_next_time_step", dt)

    def upload(self, arrays):
        n = int(arrays["rx"].shape[0])
        p, keep = particles_struct(arrays, n)
        self.call("xdem_particles_upload", C.byref(p))

    def download(self):
        n = int(self.lib.xdem_num_particles(self.h))
        out = {name: np.zeros(n, dtype=DTYPES.get(name, np.float64)) for name in FIELDS}
        p, keep = particles_struct(out, n)
        self.call("xdem_particles_download", C.byref(p))
        return out

    def num_particles(self):
        return int(self.lib.xdem_num_particles(self.h))

    def synchronize(self):
        self.call("xdem_synchronize")

    # ---- operators
    def combined_compute_prolog(self, dt):
        self.call("xdem_combined_compute_prolog", dt)

    def combined_compute_epilog(self, dt):
        self.call("xdem_combined_compute_epilog", dt)

    def reset_force_moment(self):
        self.call("xdem_reset_force_moment")

    def gravity_force(self, g):
        self.call("xdem_gravity_force", (C.c_double * 3)(*g))

    def force_to_accel(self):
        self.call("xdem_force_to_accel")

    def particle_displ_over(self, threshold):
        r = C.c_int32()
        self.call("xdem_particle_displ_over", threshold, C.byref(r))
        return bool(r.value)

    def move_particles(self):
        self.call("xdem_move_particles")

    def chunk_neighbors_contact(self, rcut_inc, nbh_dist=-1.0):
        self.call("xdem_chunk_neighbors_contact", rcut_inc, nbh_dist)

    def nbh_sphere(self, rcut_inc, symetric=True):
        self.call("xdem_nbh_sphere", rcut_inc, 1 if symetric else 0)

    def classify_interactions(self):
        self.call("xdem_classify_interactions")

    def classifier_waves(self):
        """Wave size per interaction type 0-12 (xdem_classifier_waves), as an int64 array."""
        out = (C.c_int64 * 13)()
        self.call("xdem_classifier_waves", out)
        return np.array(list(out), np.int64)

    def classifier_wave(self, t):
        """Rows of polyhedron wave t (0-3) in wave order: id_i, id_j, swap, sub_i, sub_j, feature index."""
        m = int(self.classifier_waves()[t])
        a = [np.zeros(m, np.uint64), np.zeros(m, np.uint64), np.zeros(m, np.uint8), np.zeros(m, np.uint32),
             np.zeros(m, np.uint32), np.zeros(m, np.uint32)]
        got = C.c_int64()
        self.call("xdem_classifier_wave_dump", t, m, _ptr(a[0], C.c_uint64), _ptr(a[1], C.c_uint64),
                  _ptr(a[2], C.c_uint8), _ptr(a[3], C.c_uint32), _ptr(a[4], C.c_uint32), _ptr(a[5], C.c_uint32),
                  C.byref(got))
        return tuple(x[:got.value] for x in a)

    def contact_sphere(self, cp, cpd=None, dt=1e-5, form=XDEM_FORM_FP64):
        self.call("xdem_contact_sphere", C.byref(cp), C.byref(cpd) if cpd is not None else None, dt, form)

    def run_steps(self, scheme, nsteps):
        st = RunStats()
        self.call("xdem_run_steps", C.byref(scheme), nsteps, C.byref(st))
        return {"steps": st.steps, "rebuilds": st.rebuilds, "polls": st.polls}

    def set_timing(self, on, every=1):
        """HIP-event timing of the run's integrate and force launches, of every `every`-th step."""
        self.call("xdem_set_timing", max(1, int(every)) if on else 0)

    def kernel_timing(self, which, reset=True):
        ms = C.c_double()
        cnt = C.c_int64()
        self.call("xdem_kernel_timing", which, 1 if reset else 0, C.byref(ms), C.byref(cnt))
        return ms.value, cnt.value

    # ---- inspection
    def interactions(self):
        npp, npw = C.c_int64(), C.c_int64()
        self.call("xdem_interactions_count", C.byref(npp), C.byref(npw))
        cap = npp.value + npw.value
        ii = np.zeros(cap, np.uint64)
        jj = np.zeros(cap, np.uint64)
        tt = np.zeros(cap, np.uint16)
        fr = np.zeros(3 * cap, np.float64)
        wr = C.c_int64()
        self.call("xdem_interactions_dump", cap, _ptr(ii, C.c_uint64), _ptr(jj, C.c_uint64),
                  _ptr(tt, C.c_uint16), _ptr(fr, C.c_double), C.byref(wr))
        k = wr.value
        return ii[:k], jj[:k], tt[:k], fr[:3 * k].reshape(-1, 3)

    def interactions_load(self, ii, jj, tt, friction, moment=None):
        """Sets the history of the current interactions from saved rows (restart); returns the number
        of rows applied."""
        ii = np.ascontiguousarray(ii, np.uint64)
        jj = np.ascontiguousarray(jj, np.uint64)
        tt = np.ascontiguousarray(tt, np.uint16)
        fr = np.ascontiguousarray(friction, np.float64).reshape(-1)
        mo = None if moment is None else np.ascontiguousarray(moment, np.float64).reshape(-1)
        assert fr.size == 3 * ii.size and (mo is None or mo.size == fr.size)
        got = C.c_int64()
        self.call("xdem_interactions_load", ii.size, _ptr(ii, C.c_uint64), _ptr(jj, C.c_uint64), _ptr(tt, C.c_uint16),
                  _ptr(fr, C.c_double), None if mo is None else _ptr(mo, C.c_double), C.byref(got))
        return got.value

    def interaction_moments(self):
        """Rolling-moment history [rows][3] of the rows interactions() returns, in the same order."""
        npp, npw = C.c_int64(), C.c_int64()
        self.call("xdem_interactions_count", C.byref(npp), C.byref(npw))
        cap = max(1, npp.value + npw.value)
        mo = np.zeros(3 * cap)
        wr = C.c_int64()
        self.call("xdem_interactions_moments", cap, _ptr(mo, C.c_double), C.byref(wr))
        return mo[:3 * wr.value].reshape(-1, 3)

    def set_mesh(self, mesh_id, vertices, edges, faces, minkowski, center=(0.0, 0.0, 0.0),
                 quat=(1.0, 0.0, 0.0, 0.0), max_candidates=32):
        """Registers a stationary triangle-mesh driver acting on spheres or polyhedra (vertices [nv][3] in
        the mesh frame, edges [ne][2], triangles [nf][3]); takes effect at the next neighbour build.
        max_candidates: slots per particle (polyhedra meet more mesh features than spheres)."""
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(edges, np.uint32).reshape(-1, 2)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        c = np.ascontiguousarray(center, np.float64).reshape(3)
        q = np.ascontiguousarray(quat, np.float64).reshape(4)
        self.call("xdem_set_mesh", int(mesh_id), _ptr(v, C.c_double), len(v), _ptr(e, C.c_uint32), len(e),
                  _ptr(f, C.c_uint32), len(f), float(minkowski), _ptr(c, C.c_double), _ptr(q, C.c_double),
                  int(max_candidates))

    def mesh_interactions(self):
        """Mesh candidates with history: (owner id, feature, type 7/8/9, friction [k][3], moment [k][3]),
        sorted by (owner id, type, feature)."""
        wr = C.c_int64()
        self.call("xdem_mesh_interactions", 1 << 62, None, None, None, None, None, C.byref(wr))
        n = wr.value
        ii = np.zeros(max(n, 1), np.uint64)
        sj = np.zeros(max(n, 1), np.uint32)
        tt = np.zeros(max(n, 1), np.uint16)
        fr = np.zeros(3 * max(n, 1))
        mo = np.zeros(3 * max(n, 1))
        self.call("xdem_mesh_interactions", n, _ptr(ii, C.c_uint64), _ptr(sj, C.c_uint32), _ptr(tt, C.c_uint16),
                  _ptr(fr, C.c_double), _ptr(mo, C.c_double), C.byref(wr))
        return ii[:n], sj[:n], tt[:n], fr[:3 * n].reshape(-1, 3), mo[:3 * n].reshape(-1, 3)

    def mesh_interactions_load(self, ii, sj, tt, fr, mo=None):
        """Hands saved mesh histories (rows of mesh_interactions) to the current candidate lists by key;
        returns the number of rows matched."""
        ii = np.ascontiguousarray(ii, np.uint64)
        sj = np.ascontiguousarray(sj, np.uint32)
        tt = np.ascontiguousarray(tt, np.uint16)
        fr = np.ascontiguousarray(fr, np.float64).reshape(-1)
        mop = None if mo is None else _ptr(np.ascontiguousarray(mo, np.float64).reshape(-1), C.c_double)
        hit = C.c_int64()
        self.call("xdem_mesh_interactions_load", len(ii), _ptr(ii, C.c_uint64), _ptr(sj, C.c_uint32),
                  _ptr(tt, C.c_uint16), _ptr(fr, C.c_double), mop, C.byref(hit))
        return hit.value

    def poly_interaction_moments(self):
        """Rolling-moment history [rows][3] of the rows poly_interactions() returns, in the same order."""
        n = C.c_int64()
        self.call("xdem_poly_interactions_count", C.byref(n))
        cap = max(1, n.value)
        mo = np.zeros(3 * cap)
        wr = C.c_int64()
        self.call("xdem_poly_interactions_moments", cap, _ptr(mo, C.c_double), C.byref(wr))
        return mo[:3 * wr.value].reshape(-1, 3)

    def poly_interactions_load(self, ii, jj, tt, sw, si, sj, friction, moment=None):
        """Polyhedra: sets the history of the current feature and vertex-wall interactions from saved
        rows (restart); returns the number of rows applied."""
        ii = np.ascontiguousarray(ii, np.uint64)
        jj = np.ascontiguousarray(jj, np.uint64)
        tt = np.ascontiguousarray(tt, np.uint16)
        sw = np.ascontiguousarray(sw, np.uint8)
        si = np.ascontiguousarray(si, np.uint32)
        sj = np.ascontiguousarray(sj, np.uint32)
        fr = np.ascontiguousarray(friction, np.float64).reshape(-1)
        mo = None if moment is None else np.ascontiguousarray(moment, np.float64).reshape(-1)
        assert fr.size == 3 * ii.size and (mo is None or mo.size == fr.size)
        got = C.c_int64()
        self.call("xdem_poly_interactions_load", ii.size, _ptr(ii, C.c_uint64), _ptr(jj, C.c_uint64),
                  _ptr(tt, C.c_uint16), _ptr(sw, C.c_uint8), _ptr(si, C.c_uint32), _ptr(sj, C.c_uint32),
                  _ptr(fr, C.c_double), None if mo is None else _ptr(mo, C.c_double), C.byref(got))
        return got.value

    def simulation_state(self, noalloc=False):
        """Thermodynamic state (xdem_simulation_state, or xdem_simulation_state_noalloc, which allocates no
        device memory) as a dict."""
        st = SimState()
        self.call("xdem_simulation_state_noalloc" if noalloc else "xdem_simulation_state", C.byref(st))
        return {"kinetic_energy": list(st.kinetic_energy), "rotation_energy": list(st.rotation_energy),
                "mass": st.mass, "volume": st.volume, "min_dn": st.min_dn, "n_particles": st.n_particles,
                "n_interactions": st.n_interactions, "n_active": st.n_active}

    def set_contact_record(self, on=True):
        """Record the contact state of FP64 sphere force passes (outside the engine's integrating step)."""
        self.call("xdem_set_contact_record", 1 if on else 0)

    def contact_state(self, subs=False):
        """The recorded rows of the last force pass, as the reference's interaction dump selects them:
        first id, second id (wall index), type (0 VV / sphere pair, 1 VE, 2 VF, 3 EE, 5 wall) and per
        row 16 values: dn, contact point, f - ft, ft, first position, second position. subs=True
        also returns the two sub-feature columns (after the ids)."""
        cnt = C.c_int64()
        self.call("xdem_contact_state_count", C.byref(cnt))
        cap = max(1, cnt.value)
        ii = np.zeros(cap, np.uint64)
        jj = np.zeros(cap, np.uint64)
        si = np.zeros(cap, np.uint32)
        sj = np.zeros(cap, np.uint32)
        tt = np.zeros(cap, np.uint16)
        vals = np.zeros(16 * cap, np.float64)
        wr = C.c_int64()
        self.call("xdem_contact_state_dump", cap, _ptr(ii, C.c_uint64), _ptr(jj, C.c_uint64), _ptr(si, C.c_uint32),
                  _ptr(sj, C.c_uint32), _ptr(tt, C.c_uint16), _ptr(vals, C.c_double), C.byref(wr))
        k = wr.value
        if subs:
            return ii[:k], jj[:k], si[:k], sj[:k], tt[:k], vals[:16 * k].reshape(-1, 16)
        return ii[:k], jj[:k], tt[:k], vals[:16 * k].reshape(-1, 16)

    def streams(self):
        pairs, slots = C.c_int64(), C.c_int64()
        dims = (C.c_int32 * 3)()
        self.call("xdem_nbh_stats", C.byref(pairs), C.byref(slots), dims)
        cap = max(1, slots.value)
        a = np.zeros(cap, np.uint64)
        b = np.zeros(cap, np.uint64)
        code = np.zeros(cap, np.int32)
        own = np.zeros(cap, np.uint8)
        wr = C.c_int64()
        self.call("xdem_streams_dump", cap, _ptr(a, C.c_uint64), _ptr(b, C.c_uint64), _ptr(code, C.c_int32),
                  _ptr(own, C.c_uint8), C.byref(wr))
        k = wr.value
        return a[:k], b[:k], code[:k], own[:k]

    # ---- polyhedra
    def set_shapes(self, shapes):
        arr = (Shape * len(shapes))(*shapes)
        self.call("xdem_set_shapes", arr, len(shapes))

    def compute_vertices(self):
        self.call("xdem_compute_vertices")

    def vertex_displ_over(self, threshold):
        r = C.c_int32()
        self.call("xdem_vertex_displ_over", threshold, C.byref(r))
        return bool(r.value)

    def nbh_polyhedron(self, rcut_inc):
        self.call("xdem_nbh_polyhedron", rcut_inc)

    def contact_polyhedron(self, cp, cpd=None, dt=1e-5):
        self.call("xdem_contact_polyhedron", C.byref(cp), C.byref(cpd) if cpd is not None else None, dt)

    def poly_interactions_count(self):
        n = C.c_int64()
        self.call("xdem_poly_interactions_count", C.byref(n))
        return n.value

    def poly_interactions(self):
        n = C.c_int64()
        self.call("xdem_poly_interactions_count", C.byref(n))
        cap = max(1, n.value)
        ii, jj = np.zeros(cap, np.uint64), np.zeros(cap, np.uint64)
        tt, sw = np.zeros(cap, np.uint16), np.zeros(cap, np.uint8)
        si, sj = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        fr = np.zeros(3 * cap)
        wr = C.c_int64()
        self.call("xdem_poly_interactions_dump", cap, _ptr(ii, C.c_uint64), _ptr(jj, C.c_uint64),
                  _ptr(tt, C.c_uint16), _ptr(sw, C.c_uint8), _ptr(si, C.c_uint32), _ptr(sj, C.c_uint32),
                  _ptr(fr, C.c_double), C.byref(wr))
        k = wr.value
        return ii[:k], jj[:k], tt[:k], sw[:k], si[:k], sj[:k], fr[:3 * k].reshape(-1, 3)

    def nbh_stats(self):
        pairs, slots = C.c_int64(), C.c_int64()
        dims = (C.c_int32 * 3)()
        self.call("xdem_nbh_stats", C.byref(pairs), C.byref(slots), dims)
        return {"pairs": pairs.value, "slots": slots.value, "dims": tuple(dims)}


def bf16_to_f32(bits):
    """bfloat16 bit patterns (uint16) -> float32."""
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


class Gnn:
    """The learned surrogate's network bound to a Context (xdem_gnn_* of include/xdem.h)."""

    def __init__(self, ctx, r0, v0, a0, m0, ov_scale=50.0, substeps=10, train=False):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = GnnConfig(r0, v0, a0, m0, ov_scale, substeps, 1 if train else 0)
        self.cfg = cfg
        h = C.c_void_p()
        ctx._check(self.lib.xdem_gnn_create(ctx.h, C.byref(cfg), C.byref(h)), "xdem_gnn_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.xdem_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        self.ctx._check(getattr(self.lib, name)(self.h, *args), name)

    def set_params(self, flat):
        a = np.ascontiguousarray(flat, np.float32)
        assert a.shape == (GNN_NPARAMS,)
        self.call("xdem_gnn_set_params", _ptr(a, C.c_float), GNN_NPARAMS)

    def get_params(self):
        a = np.zeros(GNN_NPARAMS, np.float32)
        self.call("xdem_gnn_get_params", _ptr(a, C.c_float), GNN_NPARAMS)
        return a

    def get_grads(self):
        a = np.zeros(GNN_NPARAMS, np.float32)
        self.call("xdem_gnn_get_grads", _ptr(a, C.c_float), GNN_NPARAMS)
        return a

    def forward(self):
        n = self.ctx.num_particles()
        out = np.zeros((max(n, 1), 6), np.float32)
        self.call("xdem_gnn_forward", _ptr(out, C.c_float))
        return out[:n]

    def set_target(self, t):
        a = np.ascontiguousarray(t, np.float32)
        self.call("xdem_gnn_set_target", _ptr(a, C.c_float))

    def backward(self):
        loss = C.c_double()
        self.call("xdem_gnn_backward", C.byref(loss))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.call("xdem_gnn_adam", lr, beta1, beta2, eps)

    def train_window(self, scheme, lr):
        loss = C.c_double()
        used = C.c_int32()
        self.call("xdem_gnn_train_window", C.byref(scheme), lr, C.byref(loss), C.byref(used))
        return loss.value, bool(used.value)

    def dump_graph(self):
        """(src, dst, edge inputs [E][16] float32, node inputs [n][8] float32), device order."""
        n = self.ctx.num_particles()
        st = self.ctx.nbh_stats()
        cap = int(st["slots"]) + 1
        src = np.zeros(cap, np.uint32)
        dst = np.zeros(cap, np.uint32)
        e = np.zeros((cap, 16), np.uint16)
        x = np.zeros((max(n, 1), 8), np.uint16)
        k = C.c_int64()
        self.call("xdem_gnn_dump_graph", cap, _ptr(src, C.c_uint32), _ptr(dst, C.c_uint32), _ptr(e, C.c_uint16),
                  _ptr(x, C.c_uint16), C.byref(k))
        k = k.value
        return src[:k], dst[:k], bf16_to_f32(e[:k]), bf16_to_f32(x[:n])

    def surrogate_run(self, scheme, nsteps, check_every=20, tol=0.01):
        st = SurrogateStats()
        self.call("xdem_surrogate_run", C.byref(scheme), nsteps, check_every, tol, C.byref(st))
        return {f: getattr(st, f) for f, _ in SurrogateStats._fields_}


# ------------------------------------------------------------------------------ host library
class DeckStats(C.Structure):
    _fields_ = [("particles", C.c_int64), ("iterations", C.c_int64), ("rebuilds", C.c_int64),
                ("operator_calls", C.c_int64), ("seconds", C.c_double)]


HOST_SYMBOLS = ["xdem_gen_uniform", "xdem_gen_normal", "xdem_gen_pack_periodic", "xdem_deck_load",
                "xdem_deck_set_global", "xdem_deck_run", "xdem_deck_last_error", "xdem_deck_destroy"]
CONFIG_DIR = os.path.join(HERE, "config")
_host = None


def load_host():
    """libxdem_host.so (deck runner, generators); needs libxdem_hip.so loaded first."""
    global _host
    if _host is not None:
        return _host
    load()
    path = os.path.join(lib_dir(), "libxdem_host.so")
    if not os.path.exists(path):
        raise XdemError(f"{path} is missing: build it with exadem_amd/build.py")
    h = C.CDLL(path)
    vp = C.c_void_p
    h.xdem_deck_load.restype = C.c_int
    h.xdem_deck_load.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_int32, C.POINTER(vp)]
    h.xdem_deck_set_global.restype = C.c_int
    h.xdem_deck_set_global.argtypes = [vp, C.c_char_p, C.c_char_p]
    h.xdem_deck_run.restype = C.c_int
    h.xdem_deck_run.argtypes = [vp, vp, C.POINTER(DeckStats)]
    h.xdem_deck_last_error.restype = C.c_int
    h.xdem_deck_last_error.argtypes = [vp, C.c_char_p, C.c_int64]
    h.xdem_deck_destroy.restype = None
    h.xdem_deck_destroy.argtypes = [vp]
    _host = h
    return h


class Deck:
    """An exaDEM-style deck (.msp) loaded by the C++ runner of libxdem_host.so."""

    def __init__(self, path, search_dirs=()):
        self.lib = load_host()
        dirs = [CONFIG_DIR, *search_dirs]
        arr = (C.c_char_p * len(dirs))(*[d.encode() for d in dirs])
        h = C.c_void_p()
        rc = self.lib.xdem_deck_load(str(path).encode(), arr, len(dirs), C.byref(h))
        self.h = h
        if rc != XDEM_OK:
            msg = self.last_error()
            self.close()
            raise XdemError(f"xdem_deck_load: {ERR_NAMES.get(rc, rc)}: {msg}")

    def last_error(self):
        buf = C.create_string_buffer(1024)
        if self.h:
            self.lib.xdem_deck_last_error(self.h, buf, 1024)
        return buf.value.decode()

    def set_global(self, name, value):
        self.lib.xdem_deck_set_global(self.h, name.encode(), str(value).encode())

    def run(self, ctx):
        st = DeckStats()
        rc = self.lib.xdem_deck_run(self.h, ctx.h, C.byref(st))
        if rc != XDEM_OK:
            raise XdemError(f"xdem_deck_run: {ERR_NAMES.get(rc, rc)}: {self.last_error()}")
        return {f: getattr(st, f) for f, _ in DeckStats._fields_}

    def close(self):
        if self.h:
            self.lib.xdem_deck_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
