"""Python mirror of the reference's per-step call surface (``FluidSim``,
src/soft_body/mod.rs:23-275) on top of the C ABI.

    sim = FluidSim(scene)          # FluidSim::new  (scene upload, static grids)
    sim.init()                     # FluidSim::init (explicit RK keeps no history)
    sim.update()                   # FluidSim::update: subticks x RK step
    p = sim.particles()            # FluidSim::particles, std430 Particle records

Per-stage drop-ins (``compute_forces``, ``lincomb``, ``velocity``, ``neighbors_update``)
take torch tensors holding std430 records on the context's device.
"""
import ctypes as C
import weakref

import numpy as np

from . import tableau as _tab
from . import wire
from ._lib import ArpError, check, lib

__all__ = ["FluidSim", "Batch", "ArpError"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class FluidSim:
    def __init__(self, scene, device=0, integrator=None, timestep=None, subticks=None):
        L = lib()
        self.scene = scene
        self.ctx = C.c_void_p()
        rc = L.arp_create(device, 0, C.byref(self.ctx))
        if rc != 0:
            raise ArpError("arp_create failed (%d)" % rc)
        s = scene
        self._keep = (np.ascontiguousarray(s.particles), np.ascontiguousarray(s.solids),
                      np.ascontiguousarray(s.boundary), np.ascontiguousarray(s.materials),
                      np.ascontiguousarray(s.interactions), np.ascontiguousarray(s.bounds))
        p, so, b, m, it, bd = self._keep
        check(self.ctx, L.arp_scene_upload(self.ctx, _p(p), len(p), _p(so), len(so), _p(b), len(b), _p(m),
                                           len(m), _p(it), _p(bd), s.h, s.g, s.alpha))
        self.set_integrator(integrator or s.integrator, timestep if timestep is not None else s.dt,
                            subticks if subticks is not None else s.subticks)

    # ---- integration ----------------------------------------------------------------
    def set_integrator(self, integrator, timestep, subticks=1):
        t = _tab.tableau(integrator) if isinstance(integrator, str) else integrator
        self.tableau = t
        self.timestep = float(timestep)
        self.subticks = int(subticks)
        check(self.ctx, lib().arp_set_integrator(self.ctx, _p(np.ascontiguousarray(t)), self.timestep, self.subticks))

    def init(self):
        """FluidSim::init: explicit RK methods need no state history."""
        return self

    def _touch(self):
        """Counts the calls that replace the state or the recording (see autograd.rollout)."""
        self.record_generation = getattr(self, "record_generation", 0) + 1

    def update(self, nframes=1):
        self._touch()
        check(self.ctx, lib().arp_update(self.ctx, nframes))

    def step(self, dt, nsteps=1):
        self._touch()
        check(self.ctx, lib().arp_step(self.ctx, float(dt), int(nsteps)))

    @property
    def time(self):
        return lib().arp_time(self.ctx)

    def synchronize(self):
        check(self.ctx, lib().arp_synchronize(self.ctx))

    def set_stream(self, stream_handle):
        check(self.ctx, lib().arp_set_stream(self.ctx, C.c_void_p(stream_handle)))

    # ---- state ----------------------------------------------------------------------
    def particles(self):
        out = wire.particles(lib().arp_num_particles(self.ctx))
        check(self.ctx, lib().arp_download(self.ctx, _p(out), len(out), None, 0))
        return out

    def solids(self):
        out = wire.solids(lib().arp_num_solids(self.ctx))
        check(self.ctx, lib().arp_download(self.ctx, None, 0, _p(out), len(out)))
        return out

    @property
    def kernel_radius(self):
        return lib().arp_kernel_radius(self.ctx)

    def num_materials(self):
        return lib().arp_num_materials(self.ctx)

    def add_particles(self, material, particles):
        """FluidSim::add_particles with generated particles (Particles::add_particles):
        returns True when a material was added."""
        m = np.ascontiguousarray(material, wire.MATERIAL)
        p = np.ascontiguousarray(particles, wire.PARTICLE)
        added = C.c_uint32(0)
        check(self.ctx, lib().arp_add_particles(self.ctx, _p(m), _p(p), len(p), C.byref(added)))
        return bool(added.value)

    def add_object(self, scene_file, index, offset=None):
        """FluidSim::add_particles of object ``index`` of an arp.config.SceneFile (the way
        the example runner spawns its on_click objects), optionally offset."""
        off = None if offset is None else np.zeros(4, np.float32)
        if off is not None:
            off[: len(offset)] = offset
        added = C.c_uint32(0)
        check(self.ctx, lib().arp_scene_spawn(self.ctx, scene_file._h, index, _p(off), C.byref(added)))
        return bool(added.value)

    def upload_state(self, particles, solids=None):
        particles = np.ascontiguousarray(particles)
        so = np.ascontiguousarray(solids) if solids is not None else None
        self._touch()
        check(self.ctx, lib().arp_upload_state(self.ctx, _p(particles), len(particles), _p(so),
                                               len(so) if so is not None else 0))

    # ---- per-stage drop-ins (device std430 buffers as torch tensors) -------------------
    @staticmethod
    def _dp(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    def compute_forces(self, d_state, d_solids=None, d_deriv=None, d_solid_deriv=None):
        check(self.ctx, lib().arp_compute_forces(self.ctx, self._dp(d_state), self._dp(d_solids),
                                                 self._dp(d_deriv), self._dp(d_solid_deriv)))

    def neighbors_update(self, d_state):
        check(self.ctx, lib().arp_neighbors_update(self.ctx, self._dp(d_state)))

    def lincomb(self, coefs, terms, out, solid_terms=None, solid_out=None, noalloc=False):
        """out = sum_t coefs[t] * terms[t] (device std430 buffers); noalloc: the variant that
        keeps its chunk partials in the library's fixed per-GPU scratch array."""
        n = len(terms)
        co = np.asarray(coefs, np.float32)
        tp = (C.c_void_p * n)(*[t.data_ptr() for t in terms])
        sp = (C.c_void_p * n)(*[t.data_ptr() for t in solid_terms]) if solid_terms else None
        fn = lib().arp_lincomb_noalloc if noalloc else lib().arp_lincomb
        check(self.ctx, fn(self.ctx, n, _p(co), C.cast(tp, C.c_void_p),
                           C.cast(sp, C.c_void_p) if sp is not None else None,
                           self._dp(out), self._dp(solid_out)))

    def velocity(self, r, d_in, d_out, d_solid_in=None, d_solid_out=None):
        check(self.ctx, lib().arp_velocity(self.ctx, float(r), self._dp(d_in), self._dp(d_solid_in),
                                           self._dp(d_out), self._dp(d_solid_out)))

    # ---- reverse mode (device std430 buffers as torch tensors) ---------------------------
    def adjoint_record(self, dt, nsteps, checkpoint_every=None):
        """Advance nsteps steps of length dt, recording for adjoint_backward (a checkpoint
        every checkpoint_every steps; default: the whole trajectory in one segment)."""
        k = nsteps if checkpoint_every is None else checkpoint_every
        # a new recording replaces the previous one: its generation tells holders of the old
        # one (a rollout's backward) that they must record again
        self._touch()
        check(self.ctx, lib().arp_adjoint_record(self.ctx, float(dt), int(nsteps), int(k)))

    def adjoint_backward(self, d_grad_out, d_grad_in, want_params=True):
        """d_grad_out / d_grad_in: device tensors of n Particle records (dL/d den, pos, vel
        after / before the recorded steps).  Returns an (nmat, 6) array of dL/d
        (bulk_viscosity, viscosity, normal_stiffness, shear_stiffness, normal_damp,
        shear_damp) per material, or None."""
        nm = lib().arp_num_materials(self.ctx)
        gp = np.zeros(6 * nm, np.float32) if want_params else None
        check(self.ctx, lib().arp_adjoint_backward(self.ctx, self._dp(d_grad_out), self._dp(d_grad_in), _p(gp)))
        return gp.reshape(nm, 6) if want_params else None

    def set_stiffness_field(self, scale):
        """Per-solid stiffness scale of (normal, shear) stiffness; None: 1 (the reference)."""
        if scale is None:
            check(self.ctx, lib().arp_set_stiffness_field(self.ctx, None, 0))
            return
        a = np.ascontiguousarray(scale, np.float32)
        self._kscale = a
        check(self.ctx, lib().arp_set_stiffness_field(self.ctx, _p(a), len(a)))

    def stiffness_grad(self):
        """dL/dk_s per solid after adjoint_backward."""
        out = np.zeros(lib().arp_num_solids(self.ctx), np.float32)
        check(self.ctx, lib().arp_adjoint_stiffness_grad(self.ctx, _p(out), len(out)))
        return out

    # ---- neighbour lists kept across stages --------------------------------------------
    def set_neighbor_skin(self, skin=0.2, cap=0, always_build=False, speculate=True):
        """Lists of the candidates within h*(1+skin) kept across RK stages while every
        particle stays within skin*h/2 of its position at the search (0 disables; cap 0:
        48 entries per particle).  always_build rebuilds at every stage (testing);
        speculate=False issues the rebuild chain at every stage (same results)."""
        flags = (1 if always_build else 0) | (0 if speculate else 2)
        self._touch()
        check(self.ctx, lib().arp_set_neighbor_skin(self.ctx, float(skin), int(cap), flags))

    def set_neighbor_max_age(self, stages):
        """Rebuild kept lists older than this many RK stages (0: only when invalid)."""
        self._touch()
        check(self.ctx, lib().arp_set_neighbor_max_age(self.ctx, int(stages)))

    def set_neighbor_short_skin(self, short_skin):
        """Two-level lists: the walks use a short list of the candidates within
        h*(1+short_skin), refreshed from the long list (no grid search) when a particle moves
        short_skin*h/2 from its position at the refresh (0 disables; below the skin)."""
        self._touch()
        check(self.ctx, lib().arp_set_neighbor_short_skin(self.ctx, float(short_skin)))

    def tune_lists(self, dt, trial_steps=100):
        """Autotune the neighbour lists: times trial_steps steps from the current state for
        each (skin, short skin) pair of a fixed table (one and two levels), restoring the
        state after every trial, and sets the fastest; returns (skin, short_skin)."""
        self._touch()
        a, b = C.c_float(), C.c_float()
        check(self.ctx, lib().arp_tune_lists(self.ctx, float(dt), int(trial_steps), C.byref(a), C.byref(b)))
        return float(a.value), float(b.value)

    def neighbor_refreshes(self):
        """Short-list refreshes since the list settings last changed."""
        r = C.c_uint32(0)
        check(self.ctx, lib().arp_neighbor_refreshes(self.ctx, C.byref(r)))
        return int(r.value)

    def neighbor_stats(self):
        """(stages that searched and built lists, stages that reused them)."""
        b, r = C.c_uint32(0), C.c_uint32(0)
        check(self.ctx, lib().arp_neighbor_stats(self.ctx, C.byref(b), C.byref(r)))
        return int(b.value), int(r.value)

    def speculation_stats(self):
        """(speculative steps completed, speculative steps run again with the rebuild chain)."""
        s, a = C.c_uint64(0), C.c_uint64(0)
        check(self.ctx, lib().arp_speculation_stats(self.ctx, C.byref(s), C.byref(a)))
        return int(s.value), int(a.value)

    # ---- diagnostics ------------------------------------------------------------------
    def grid_info(self):
        g = np.zeros((), wire.GRID_INFO)
        check(self.ctx, lib().arp_grid_info(self.ctx, _p(g)))
        return g

    def neighbor_dump(self):
        n = lib().arp_num_particles(self.ctx)
        cell = np.zeros(n, np.uint32)
        cc = np.zeros(n, np.uint32)
        check(self.ctx, lib().arp_neighbor_dump(self.ctx, _p(cell), _p(cc), None, 0))
        tot = int(cc.sum(dtype=np.uint64))
        cand = np.zeros(max(tot, 1), np.uint32)
        check(self.ctx, lib().arp_neighbor_dump(self.ctx, None, _p(cc), _p(cand), tot))
        return dict(cell=cell, cand_count=cc, cand=cand[:tot])

    def cell_counts(self, first=0, count=None):
        gi = self.grid_info()
        if count is None:
            count = int(gi["ncells"]) - first
        out = np.zeros(3 * count, np.uint32)
        check(self.ctx, lib().arp_cell_counts(self.ctx, first, count, _p(out)))
        return out

    # ---- measurement -------------------------------------------------------------------
    def profile(self, enable=True):
        check(self.ctx, lib().arp_profile(self.ctx, 1 if enable else 0))

    def profile_select(self, kernels=None):
        """Time only these kernels (list of names; None: all)."""
        names = ",".join(kernels) if kernels else ""
        check(self.ctx, lib().arp_profile_select(self.ctx, names.encode()))

    def profile_reset(self):
        check(self.ctx, lib().arp_profile_reset(self.ctx))

    def profile_query(self, kernel):
        ms = C.c_double()
        cnt = C.c_uint64()
        check(self.ctx, lib().arp_profile_query(self.ctx, kernel.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


    def profile_durations(self, kernel):
        """Event time (ms) of each launch of one kernel since the last reset, in launch order."""
        n = C.c_uint64(0)
        check(self.ctx, lib().arp_profile_durations(self.ctx, kernel.encode(), None, 0, C.byref(n)))
        buf = (C.c_float * max(int(n.value), 1))()
        check(self.ctx, lib().arp_profile_durations(self.ctx, kernel.encode(), buf, n.value, C.byref(n)))
        return np.array(buf[: int(n.value)], dtype=np.float64)
    def close(self):
        # a batch holding this context is destroyed first (garbage collection of a cycle may
        # finalize the sim before the batch that refers to it)
        for b in list(getattr(self, "_batches", ())):
            b.close()
        if self.ctx:
            lib().arp_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent scenes stepped together (arp_batch_*): one launch per kernel and RK stage
    for all of them, replayed from a hipGraph when graph=True."""

    def __init__(self, sims, graph=True):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.ctx.value for s in self.sims])
        self.b = C.c_void_p()
        rc = lib().arp_batch_create(C.cast(arr, C.c_void_p), len(self.sims), 1 if graph else 0, C.byref(self.b))
        if rc != 0:
            msg = lib().arp_last_error(self.sims[0].ctx)
            raise ArpError("arp_batch_create failed (%d): %s" % (rc, msg.decode() if msg else ""))
        for s in self.sims:
            if not hasattr(s, "_batches"):
                s._batches = weakref.WeakSet()
            s._batches.add(self)

    def _check(self, rc):
        if rc != 0:
            msg = lib().arp_batch_last_error(self.b)
            raise ArpError("arp batch error %d: %s" % (rc, msg.decode() if msg else ""))

    def step(self, dt, nsteps=1):
        self._check(lib().arp_batch_step(self.b, float(dt), int(nsteps)))

    def tune(self, dt, trial_steps=2):
        """Autotune the batch's knobs; returns the chosen particles per row-sort workgroup
        (tuning() gives every choice)."""
        d = C.c_uint32()
        self._check(lib().arp_batch_tune(self.b, float(dt), int(trial_steps), C.byref(d)))
        return d.value

    def tune_lists(self, dt, trial_steps=100):
        """Autotune the neighbour lists of every scene (one (skin, short skin) pair for the
        batch, FluidSim.tune_lists's table); returns (skin, short_skin)."""
        a, b = C.c_float(), C.c_float()
        self._check(lib().arp_batch_tune_lists(self.b, float(dt), int(trial_steps), C.byref(a), C.byref(b)))
        return float(a.value), float(b.value)

    def tuning(self):
        """(particles per row-sort workgroup, list walks buffered in LDS: 0 or 1)."""
        d, w = C.c_uint32(), C.c_uint32()
        self._check(lib().arp_batch_tuning(self.b, C.byref(d), C.byref(w)))
        return d.value, w.value

    def adjoint_record(self, dt, nsteps, checkpoint_every=None):
        """Record nsteps steps of every scene for adjoint_backward (a checkpoint every
        checkpoint_every steps; default: one segment)."""
        k = nsteps if checkpoint_every is None else checkpoint_every
        for s in self.sims:
            s._touch()
        self._check(lib().arp_batch_adjoint_record(self.b, float(dt), int(nsteps), int(k)))

    def adjoint_backward(self, d_grad_out, d_grad_in, want_params=True):
        """Per scene: device tensors of n_s Particle records (dL/d state after / before the
        recorded steps).  Returns one (nmat, 6) parameter-gradient array per scene, or None."""
        n = len(self.sims)
        go = (C.c_void_p * n)(*[t.data_ptr() for t in d_grad_out])
        gi = (C.c_void_p * n)(*[t.data_ptr() for t in d_grad_in])
        gps = [np.zeros(6 * lib().arp_num_materials(s.ctx), np.float32) for s in self.sims] if want_params else None
        gpp = (C.c_void_p * n)(*[g.ctypes.data for g in gps]) if want_params else None
        self._check(lib().arp_batch_adjoint_backward(self.b, C.cast(go, C.c_void_p), C.cast(gi, C.c_void_p),
                                                     C.cast(gpp, C.c_void_p) if gpp is not None else None))
        return [g.reshape(-1, 6) for g in gps] if want_params else None

    def synchronize(self):
        for s in self.sims:
            s.synchronize()

    def close(self):
        if self.b:
            lib().arp_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_device(arr, device="cuda"):
    """std430 record array -> uint8 torch tensor on the GPU (for the per-stage API)."""
    import torch

    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).to(device)


def from_device(t, dtype):
    return t.cpu().numpy().view(dtype)
