"""Operator-order driver of the sphere path on one GPU.

`first_iteration` and `step` call the operators one by one in the order of the reference decks
(data/config/main-config.msp:55-65 and data/config/config_numerical_schemes.msp:6-28, with
compute_force = [gravity_force, contact_sphere] and the rebuild branch of
config_move_particles.msp:141-195). `run` hands whole blocks of steps to the fused engine
(xdem_run_steps), which produces the same state with two kernels per step.
"""
import copy

import numpy as np

from . import native


def scheme(workload, form=native.XDEM_FORM_FP64, poll=16):
    s = native.Scheme()
    s.dt = workload.dt
    for a in range(3):
        s.gravity[a] = workload.gravity[a]
    s.gravity_on = 1 if any(g != 0 for g in workload.gravity) else 0
    s.form = form
    s.cp = workload.cp
    s.cpd = workload.cpd
    s.rcut_inc = workload.rcut_inc
    s.poll = poll
    return s


class SphereSim:
    def __init__(self, workload, device=0, form=native.XDEM_FORM_FP64, ctx=None, mesh=None):
        """mesh: keyword arguments of Context.set_mesh (a triangle-mesh driver), or None."""
        self.w = workload
        self.form = form
        self.ctx = ctx or native.Context(device)
        self.ctx.set_domain(workload.domain)
        self.ctx.set_surfaces(workload.walls)
        self.mesh = mesh
        if mesh is not None:
            self.ctx.set_mesh(**mesh)
        self.ctx.upload(workload.particles)
        self.rebuilds = 0

    def rebuild(self):
        """update_particles_full: move_particles, chunk_neighbors_contact, nbh_sphere, classify."""
        c = self.ctx
        c.move_particles()
        c.chunk_neighbors_contact(self.w.rcut_inc)
        c.nbh_sphere(self.w.rcut_inc)
        c.classify_interactions()

    def compute_force(self):
        c = self.ctx
        if any(g != 0 for g in self.w.gravity):
            c.gravity_force(self.w.gravity)
        c.contact_sphere(self.w.cp, self.w.cpd, self.w.dt, self.form)

    def first_iteration(self):
        """init_particles (rebuild), reset_force_moment, compute_force, next_time_step; no force_to_accel."""
        self.rebuild()
        self.ctx.reset_force_moment()
        self.compute_force()
        self.ctx.next_time_step(self.w.dt)

    def drivers_prolog(self):
        """push_f_v_r_driver at the step's physical time, then push_f_v_driver (dt_scale 0.5)."""
        c, w = self.ctx, self.w
        c.push_f_v_r_driver(c.physical_time(), w.dt)
        c.push_f_v_driver(0.5 * w.dt)

    def step(self):
        """One numerical_scheme_verlet_macro iteration, operator by operator, then next_time_step."""
        c, w = self.ctx, self.w
        c.combined_compute_prolog(w.dt)
        self.drivers_prolog()
        moved = c.particle_displ_over(0.5 * w.rcut_inc)
        if c.driver_displ_over(0.5 * w.rcut_inc) or moved:
            self.rebuild()
            self.rebuilds += 1
        c.reset_force_moment()
        self.compute_force()
        c.force_to_accel()
        c.push_f_v_driver(0.5 * w.dt)
        c.combined_compute_epilog(w.dt)
        c.next_time_step(w.dt)

    def run(self, nsteps, poll=16):
        st = self.ctx.run_steps(scheme(self.w, self.form, poll), nsteps)
        self.rebuilds += st["rebuilds"]
        return st

    def checkpoint(self, path):
        """Restart file (numpy .npz): every particle field, in device order, and the interactions with
        their friction history (the layout of xdem_interactions_dump) - what the reference's restart
        dump carries (sim_dump_writer_interaction.cpp:25-46), with the rolling-moment history."""
        p = self.ctx.download()
        ii, jj, tt, fr = self.ctx.interactions()
        mo = self.ctx.interaction_moments()
        mesh = {}
        if self.mesh is not None:  # mesh-driver histories
            mi, ms, mt, mf, mm = self.ctx.mesh_interactions()
            mesh = {"m_ii": mi, "m_sj": ms, "m_tt": mt, "m_fr": mf, "m_mo": mm}
        np.savez(path, ii=ii, jj=jj, tt=tt, fr=fr, mo=mo, **mesh, **self._driver_state(),
                 **{"p_" + k: v for k, v in p.items()})

    def _driver_state(self):
        """physical time and the drivers' state (offset, centre, velocity), as dump_driver keeps them."""
        d = self.ctx.drivers()
        return {"time": np.array(self.ctx.physical_time()),
                "d_offset": np.array([x["offset"] for x in d]), "d_center": np.array([x["center"] for x in d]),
                "d_vel": np.array([x["vel"] for x in d])}

    @staticmethod
    def _restore_drivers(w, d):
        """the workload's drivers moved to the saved state (offset, centre, velocity)."""
        if "d_offset" not in d.files or len(w.walls) == 0:
            return
        walls = []
        for k, x in enumerate(w.walls):
            x = dict(x)
            x["offset"] = float(d["d_offset"][k])
            x["center"] = tuple(float(v) for v in d["d_center"][k])
            x["vel"] = tuple(float(v) for v in d["d_vel"][k])
            walls.append(x)
        w.walls = walls

    @classmethod
    def restart(cls, workload, path, device=0, form=native.XDEM_FORM_FP64, history=True, mesh=None):
        """A simulation continued from a checkpoint: the saved particles (domain, walls and parameters
        from the workload), lists rebuilt at the saved positions, the saved history handed to the
        current interactions by key (xdem_interactions_load)."""
        d = np.load(path)
        w = copy.copy(workload)
        w.particles = {k[2:]: d[k] for k in d.files if k.startswith("p_")}
        cls._restore_drivers(w, d)
        sim = cls(w, device=device, form=form, mesh=mesh)
        if "time" in d.files:
            sim.ctx.set_physical_time(float(d["time"]))
        sim.rebuild()
        if history:
            sim.ctx.interactions_load(d["ii"], d["jj"], d["tt"], d["fr"], d["mo"])
            if mesh is not None and "m_ii" in d.files:
                sim.ctx.mesh_interactions_load(d["m_ii"], d["m_sj"], d["m_tt"], d["m_fr"], d["m_mo"])
        return sim


class PolySim(SphereSim):
    """The polyhedra scheme (data/config/config_polyhedra.msp): compute_vertices on the fast path,
    vertex_displ_over as the rebuild trigger, nbh_polyhedron and contact_polyhedron."""

    def __init__(self, workload, device=0, ctx=None, mesh=None):
        super().__init__(workload, device, native.XDEM_FORM_FP64, ctx, mesh=mesh)
        self.ctx.set_shapes(workload.shapes)

    def checkpoint(self, path):
        """Restart file: every particle field and the feature and vertex-wall interactions with their
        friction and rolling-moment history (the layout of xdem_poly_interactions_dump)."""
        p = self.ctx.download()
        ii, jj, tt, sw, si, sj, fr = self.ctx.poly_interactions()
        mo = self.ctx.poly_interaction_moments()
        np.savez(path, ii=ii, jj=jj, tt=tt, sw=sw, si=si, sj=sj, fr=fr, mo=mo, **self._driver_state(),
                 **{"p_" + k: v for k, v in p.items()})

    @classmethod
    def restart(cls, workload, path, device=0, history=True, mesh=None):
        """A polyhedra simulation continued from a checkpoint (see SphereSim.restart); the mesh driver's
        interactions are rows of the same table (types 7-12, partner = the mesh id)."""
        d = np.load(path)
        w = copy.copy(workload)
        w.particles = {k[2:]: d[k] for k in d.files if k.startswith("p_")}
        cls._restore_drivers(w, d)
        sim = cls(w, device=device, mesh=mesh)
        if "time" in d.files:
            sim.ctx.set_physical_time(float(d["time"]))
        sim.rebuild()
        if history:
            sim.ctx.poly_interactions_load(d["ii"], d["jj"], d["tt"], d["sw"], d["si"], d["sj"], d["fr"], d["mo"])
        return sim

    def rebuild(self):
        c = self.ctx
        c.move_particles()
        c.chunk_neighbors_contact(self.w.rcut_inc)
        c.nbh_polyhedron(self.w.rcut_inc)
        c.classify_interactions()

    def compute_force(self):
        c = self.ctx
        if any(g != 0 for g in self.w.gravity):
            c.gravity_force(self.w.gravity)
        c.contact_polyhedron(self.w.cp, self.w.cpd, self.w.dt)

    def step(self):
        c, w = self.ctx, self.w
        c.combined_compute_prolog(w.dt)
        self.drivers_prolog()
        moved = c.vertex_displ_over(0.5 * w.rcut_inc)
        if c.driver_displ_over(0.5 * w.rcut_inc) or moved:
            self.rebuild()
            self.rebuilds += 1
        else:
            c.compute_vertices()
        c.reset_force_moment()
        self.compute_force()
        c.force_to_accel()
        c.push_f_v_driver(0.5 * w.dt)
        c.combined_compute_epilog(w.dt)
        c.next_time_step(w.dt)
