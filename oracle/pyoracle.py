"""ctypes wrapper of the CPU oracle (oracle/libxdem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the cpu_baseline
leg of bench.py, as the checker. Nothing in the exadem_amd package imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libxdem_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    from exadem_amd.native import Particles, Domain, Surface, ContactParams, Scheme  # POD layouts only
    lib = C.CDLL(LIB)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    u64p = C.POINTER(C.c_uint64)
    lib.xo_create.restype = vp
    lib.xo_create.argtypes = [C.POINTER(Particles), C.POINTER(Domain)]
    lib.xo_destroy.argtypes = [vp]
    lib.xo_set_mutation.argtypes = [C.c_int]
    lib.xo_set_surfaces.argtypes = [vp, C.POINTER(Surface), C.c_int]
    lib.xo_rebuild.argtypes = [vp, C.c_double]
    lib.xo_set_mesh.argtypes = [vp, C.c_uint64, dp, C.c_int64, C.POINTER(C.c_uint32), C.c_int64,
                                C.POINTER(C.c_uint32), C.c_int64, C.c_double, dp, dp]
    lib.xo_mesh_interactions.argtypes = [vp, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint16), dp, dp]
    lib.xo_mesh_interactions.restype = C.c_int64
    lib.xo_move_particles.argtypes = [vp]
    lib.xo_chunk_neighbors.argtypes = [vp, C.c_double, C.c_double]
    lib.xo_nbh_sphere.argtypes = [vp, C.c_double]
    lib.xo_reset_force_moment.argtypes = [vp]
    lib.xo_gravity.argtypes = [vp, dp]
    lib.xo_contact.argtypes = [vp, C.POINTER(ContactParams), C.POINTER(ContactParams), C.c_double]
    lib.xo_force_to_accel.argtypes = [vp]
    lib.xo_prolog.argtypes = [vp, C.c_double]
    lib.xo_epilog.argtypes = [vp, C.c_double]
    lib.xo_displ_over.argtypes = [vp, C.c_double]
    lib.xo_displ_over.restype = C.c_int
    lib.xo_first_iteration.argtypes = [vp, C.POINTER(Scheme)]
    lib.xo_run_steps.argtypes = [vp, C.POINTER(Scheme), C.c_int64]
    lib.xo_run_steps.restype = C.c_int64
    lib.xo_num_particles.argtypes = [vp]
    lib.xo_num_particles.restype = C.c_int64
    lib.xo_get_particles.argtypes = [vp, C.POINTER(Particles)]
    lib.xo_interactions_count.argtypes = [vp]
    lib.xo_interactions_count.restype = C.c_int64
    lib.xo_interactions_dump.argtypes = [vp, C.c_int64, u64p, u64p, C.POINTER(C.c_uint16), dp, dp]
    lib.xo_interactions_dump.restype = C.c_int64
    lib.xo_interactions_restore.argtypes = [vp, C.c_int64, u64p, u64p, C.POINTER(C.c_uint16), C.POINTER(C.c_uint8),
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), dp, dp]
    lib.xo_interactions_restore.restype = C.c_int64
    from exadem_amd.native import DriverMotion
    lib.xo_set_driver_motion.argtypes = [vp, C.POINTER(DriverMotion), C.c_int]
    lib.xo_set_driver_motion.restype = C.c_int
    lib.xo_get_drivers.argtypes = [vp, C.POINTER(Surface), C.c_int]
    lib.xo_set_time.argtypes = [vp, C.c_double]
    lib.xo_time.argtypes = [vp]
    lib.xo_time.restype = C.c_double
    lib.xo_streams_count.argtypes = [vp]
    lib.xo_streams_count.restype = C.c_int64
    lib.xo_streams_dump.argtypes = [vp, C.c_int64, u64p, u64p, C.POINTER(C.c_int32)]
    lib.xo_streams_dump.restype = C.c_int64
    lib.xo_pair_force.argtypes = [dp, dp, dp, C.c_double, C.c_double, dp, dp, dp, C.c_double, C.c_double,
                                  C.POINTER(ContactParams), C.c_double, dp, dp]
    lib.xo_angular_accel.argtypes = [dp, dp, dp, dp, dp]
    lib.xo_quat_rotate.argtypes = [dp, dp, dp]
    from exadem_amd.native import Shape
    sp = C.POINTER(Shape)
    lib.xo_set_shapes.argtypes = [vp, sp, C.c_int]
    lib.xo_compute_vertices.argtypes = [vp]
    lib.xo_rebuild_poly.argtypes = [vp, C.c_double]
    lib.xo_contact_poly.argtypes = [vp, C.POINTER(ContactParams), C.POINTER(ContactParams), C.c_double]
    lib.xo_vertex_displ_over.argtypes = [vp, C.c_double]
    lib.xo_vertex_displ_over.restype = C.c_int
    lib.xo_first_iteration_poly.argtypes = [vp, C.POINTER(Scheme)]
    lib.xo_run_steps_poly.argtypes = [vp, C.POINTER(Scheme), C.c_int64]
    lib.xo_run_steps_poly.restype = C.c_int64
    lib.xo_poly_interactions_dump.argtypes = [vp, C.c_int64, u64p, u64p, C.POINTER(C.c_uint16),
                                              C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), dp]
    lib.xo_poly_interactions_dump.restype = C.c_int64
    lib.xo_poly_detect.argtypes = [sp, dp, dp, sp, dp, dp, C.c_int, C.c_int, C.c_int, dp]
    lib.xo_poly_detect.restype = C.c_int
    lib.xo_obb_intersect.argtypes = [sp, dp, dp, sp, dp, dp, C.c_double]
    lib.xo_obb_intersect.restype = C.c_int
    _lib = lib
    return lib


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """The CPU restatement, holding one particle system (same entry-point order as the operators)."""

    def __init__(self, particles, dom, walls=()):
        from exadem_amd.native import particles_struct, surfaces
        self.lib = load()
        n = int(particles["rx"].shape[0])
        self._keep = {k: np.ascontiguousarray(v) for k, v in particles.items()}
        p, _ = particles_struct(self._keep, n)
        self.h = self.lib.xo_create(C.byref(p), C.byref(dom))
        self.nwalls = len(walls)
        if walls:
            self.lib.xo_set_surfaces(self.h, surfaces(list(walls)), len(walls))
            from exadem_amd.native import has_motion, motions
            if has_motion(list(walls)):
                if self.lib.xo_set_driver_motion(self.h, motions(list(walls)), len(walls)) != 0:
                    raise ValueError("driver motion rejected (zero motion vector, amplitude or omega)")

    def drivers(self):
        from exadem_amd.native import Surface, driver_rows
        arr = (Surface * max(1, self.nwalls))()
        self.lib.xo_get_drivers(self.h, arr, self.nwalls)
        return driver_rows(arr, self.nwalls)

    def time(self):
        return float(self.lib.xo_time(self.h))

    def set_time(self, t):
        self.lib.xo_set_time(self.h, t)

    def close(self):
        if self.h:
            self.lib.xo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rebuild(self, rcut_inc):
        self.lib.xo_rebuild(self.h, rcut_inc)

    def move_particles(self):
        self.lib.xo_move_particles(self.h)

    def chunk_neighbors(self, rcut_inc, nbh_dist):
        self.lib.xo_chunk_neighbors(self.h, rcut_inc, nbh_dist)

    def nbh_sphere(self, rcut_inc):
        self.lib.xo_nbh_sphere(self.h, rcut_inc)

    def reset_force_moment(self):
        self.lib.xo_reset_force_moment(self.h)

    def gravity(self, g):
        self.lib.xo_gravity(self.h, (C.c_double * 3)(*g))

    def contact(self, cp, cpd, dt):
        self.lib.xo_contact(self.h, C.byref(cp), C.byref(cpd), dt)

    def force_to_accel(self):
        self.lib.xo_force_to_accel(self.h)

    def prolog(self, dt):
        self.lib.xo_prolog(self.h, dt)

    def epilog(self, dt):
        self.lib.xo_epilog(self.h, dt)

    def displ_over(self, thr):
        return bool(self.lib.xo_displ_over(self.h, thr))

    def first_iteration(self, scheme):
        self.lib.xo_first_iteration(self.h, C.byref(scheme))

    def run_steps(self, scheme, nsteps):
        return int(self.lib.xo_run_steps(self.h, C.byref(scheme), nsteps))

    def particles(self):
        from exadem_amd.native import FIELDS, DTYPES, particles_struct
        n = int(self.lib.xo_num_particles(self.h))
        out = {name: np.zeros(n, dtype=DTYPES.get(name, np.float64)) for name in FIELDS}
        p, _ = particles_struct(out, n)
        self.lib.xo_get_particles(self.h, C.byref(p))
        return out

    def interactions(self):
        cap = int(self.lib.xo_interactions_count(self.h))
        ii = np.zeros(cap, np.uint64)
        jj = np.zeros(cap, np.uint64)
        tt = np.zeros(cap, np.uint16)
        fr = np.zeros(3 * cap)
        mo = np.zeros(3 * cap)
        k = self.lib.xo_interactions_dump(self.h, cap, ii.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          jj.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          tt.ctypes.data_as(C.POINTER(C.c_uint16)),
                                          fr.ctypes.data_as(C.POINTER(C.c_double)),
                                          mo.ctypes.data_as(C.POINTER(C.c_double)))
        return ii[:k], jj[:k], tt[:k], fr[:3 * k].reshape(-1, 3), mo[:3 * k].reshape(-1, 3)

    def restore_interactions(self, ii, jj, tt, fr, mo=None, sw=None, si=None, sj=None):
        """Restart: saved interaction rows (a checkpoint's dump) become their owners' interaction storage;
        the next rebuild hands the active ones to the new interactions by key. Returns the rows kept."""
        m = len(ii)
        ii = np.ascontiguousarray(ii, np.uint64)
        jj = np.ascontiguousarray(jj, np.uint64)
        tt = np.ascontiguousarray(tt, np.uint16)
        fr = np.ascontiguousarray(fr, np.float64).reshape(-1)
        arrs = {"mo": None if mo is None else np.ascontiguousarray(mo, np.float64).reshape(-1),
                "sw": None if sw is None else np.ascontiguousarray(sw, np.uint8),
                "si": None if si is None else np.ascontiguousarray(si, np.uint32),
                "sj": None if sj is None else np.ascontiguousarray(sj, np.uint32)}
        ptr = {"mo": C.c_double, "sw": C.c_uint8, "si": C.c_uint32, "sj": C.c_uint32}
        p = {k: (None if v is None else v.ctypes.data_as(C.POINTER(ptr[k]))) for k, v in arrs.items()}
        return int(self.lib.xo_interactions_restore(self.h, m, ii.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                    jj.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                    tt.ctypes.data_as(C.POINTER(C.c_uint16)), p["sw"], p["si"],
                                                    p["sj"], fr.ctypes.data_as(C.POINTER(C.c_double)), p["mo"]))

    # ---- triangle-mesh driver
    def set_mesh(self, mesh_id, vertices, edges, faces, minkowski, center=(0.0, 0.0, 0.0), quat=(1.0, 0.0, 0.0, 0.0)):
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(edges, np.uint32).reshape(-1, 2)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        c = np.ascontiguousarray(center, np.float64)
        q = np.ascontiguousarray(quat, np.float64)
        u32 = C.POINTER(C.c_uint32)
        self.lib.xo_set_mesh(self.h, mesh_id, v.ctypes.data_as(C.POINTER(C.c_double)), len(v), e.ctypes.data_as(u32),
                             len(e), f.ctypes.data_as(u32), len(f), minkowski, c.ctypes.data_as(C.POINTER(C.c_double)),
                             q.ctypes.data_as(C.POINTER(C.c_double)))

    def mesh_interactions(self):
        """(owner id, feature, type, friction [k][3], moment [k][3]) sorted by (owner id, type, feature)."""
        n = self.lib.xo_mesh_interactions(self.h, 0, None, None, None, None, None)
        ii = np.zeros(n, np.uint64)
        sj = np.zeros(n, np.uint32)
        tt = np.zeros(n, np.uint16)
        fr = np.zeros(3 * n)
        mo = np.zeros(3 * n)
        self.lib.xo_mesh_interactions(self.h, n, ii.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      sj.ctypes.data_as(C.POINTER(C.c_uint32)), tt.ctypes.data_as(C.POINTER(C.c_uint16)),
                                      fr.ctypes.data_as(C.POINTER(C.c_double)), mo.ctypes.data_as(C.POINTER(C.c_double)))
        return ii, sj, tt, fr.reshape(-1, 3), mo.reshape(-1, 3)

    # ---- polyhedra (config_polyhedra.msp order)
    def set_shapes(self, shapes):
        from exadem_amd.native import Shape
        self._shapes = (Shape * len(shapes))(*shapes)
        self.lib.xo_set_shapes(self.h, self._shapes, len(shapes))

    def compute_vertices(self):
        self.lib.xo_compute_vertices(self.h)

    def rebuild_poly(self, rcut_inc):
        self.lib.xo_rebuild_poly(self.h, rcut_inc)

    def contact_poly(self, cp, cpd, dt):
        self.lib.xo_contact_poly(self.h, C.byref(cp), C.byref(cpd), dt)

    def vertex_displ_over(self, thr):
        return bool(self.lib.xo_vertex_displ_over(self.h, thr))

    def first_iteration_poly(self, scheme):
        self.lib.xo_first_iteration_poly(self.h, C.byref(scheme))

    def run_steps_poly(self, scheme, nsteps):
        return int(self.lib.xo_run_steps_poly(self.h, C.byref(scheme), nsteps))

    def poly_interactions(self):
        """(id_i, id_j, type, swap, sub_i, sub_j, friction) per interaction in storage order."""
        cap = int(self.lib.xo_interactions_count(self.h))
        ii = np.zeros(cap, np.uint64)
        jj = np.zeros(cap, np.uint64)
        tt = np.zeros(cap, np.uint16)
        sw = np.zeros(cap, np.uint8)
        si = np.zeros(cap, np.uint32)
        sj = np.zeros(cap, np.uint32)
        fr = np.zeros(3 * max(cap, 1))
        k = self.lib.xo_poly_interactions_dump(self.h, cap, ii.ctypes.data_as(C.POINTER(C.c_uint64)),
                                               jj.ctypes.data_as(C.POINTER(C.c_uint64)),
                                               tt.ctypes.data_as(C.POINTER(C.c_uint16)),
                                               sw.ctypes.data_as(C.POINTER(C.c_uint8)),
                                               si.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               sj.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               fr.ctypes.data_as(C.POINTER(C.c_double)))
        return ii[:k], jj[:k], tt[:k], sw[:k], si[:k], sj[:k], fr[:3 * k].reshape(-1, 3)

    def streams(self):
        cap = int(self.lib.xo_streams_count(self.h))
        a = np.zeros(cap, np.uint64)
        b = np.zeros(cap, np.uint64)
        code = np.zeros(cap, np.int32)
        k = self.lib.xo_streams_dump(self.h, cap, a.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     b.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     code.ctypes.data_as(C.POINTER(C.c_int32)))
        return a[:k], b[:k], code[:k]


def set_mutation(m):
    load().xo_set_mutation(m)


def pair_force(pi, vi, wi, ri, mi, pj, vj, wj, rj, mj, cp, dt, friction=(0.0, 0.0, 0.0)):
    """One owner-side sphere-sphere contact evaluation. Returns dict(f, friction, dn, n, cp, contact)."""
    lib = load()
    fr = np.array(friction, np.float64)
    out = np.zeros(14)
    lib.xo_pair_force(_d(pi), _d(vi), _d(wi), ri, mi, _d(pj), _d(vj), _d(wj), rj, mj, C.byref(cp), dt,
                      fr.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
    return {"f": out[0:3], "friction": out[3:6], "dn": out[6], "n": out[7:10], "cp": out[10:13],
            "contact": bool(out[13])}


def angular_accel(q, mom, vrot, inertia):
    out = np.zeros(3)
    load().xo_angular_accel(_d(q), _d(mom), _d(vrot), _d(inertia), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def quat_rotate(q, v):
    out = np.zeros(3)
    load().xo_quat_rotate(_d(q), _d(v), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def poly_detect(sa, pa, qa, sb, pb, qb, ftype, sub_i, sub_j):
    """One feature detection (0 VV, 1 VE, 2 VF, 3 EE) between shape sa at (pa, qa) and sb at (pb, qb).
    Returns dict(contact, dn, n, cp)."""
    out = np.zeros(8)
    load().xo_poly_detect(C.byref(sa), _d(pa), _d(qa), C.byref(sb), _d(pb), _d(qb), ftype, sub_i, sub_j,
                          out.ctypes.data_as(C.POINTER(C.c_double)))
    return {"contact": bool(out[0]), "dn": out[1], "n": out[2:5], "cp": out[5:8]}


def obb_intersect(sa, pa, qa, sb, pb, qb, enlarge=0.0):
    return bool(load().xo_obb_intersect(C.byref(sa), _d(pa), _d(qa), C.byref(sb), _d(pb), _d(qb), enlarge))
