"""torch.autograd over the fused SPH step: a rollout of nsteps RK steps whose backward is the
hand-written adjoint (arp_adjoint_record / arp_adjoint_backward).

    y = rollout(sim, x0, dt, nsteps)          # x0, y: (n, 7) float32 cuda tensors
    loss = f(y); loss.backward()              # x0.grad = dL/dx0

Columns of the state tensor: density, position x/y/z, velocity x/y/z, rows in the scene's
particle order.  Optional ``params`` ((nmat, 2): bulk_viscosity, viscosity per material)
receives dL/dparams; its values must equal the scene's (the forward runs on the uploaded
materials).

A sim holds one recording.  Rollouts chained on one sim (y1 = rollout(sim, x0, ...),
y2 = rollout(sim, y1, ...)) are still differentiated correctly: a backward whose recording
was replaced records its rollout again from the saved input, which leaves the sim's live
state at that rollout's end.
"""
import numpy as np

from . import wire

__all__ = ["rollout", "state_tensor", "records_from_state"]

_COLS = (0, 4, 5, 6, 8, 9, 10)  # den, pos.xyz, vel.xyz as float32 columns of a Particle record


def records_from_state(x, mat=None):
    """(n, 7) float tensor -> (n, 12) float32 tensor laid out as n Particle records."""
    import torch

    r = torch.zeros((x.shape[0], 12), dtype=torch.float32, device=x.device)
    r[:, list(_COLS)] = x.to(torch.float32)
    if mat is not None:
        r[:, 1] = torch.as_tensor(np.ascontiguousarray(mat, np.uint32).view(np.float32), device=x.device)
    return r


def state_tensor(particles, device="cuda"):
    """Particle records (numpy) -> (n, 7) float32 tensor."""
    import torch

    a = np.ascontiguousarray(particles).view(np.float32).reshape(len(particles), 12)
    return torch.from_numpy(a[:, list(_COLS)].copy()).to(device)


def _record(sim, x0h, dt, nsteps, checkpoint_every):
    """Upload x0 (n, 7 host array) into the sim's state and record nsteps steps from it."""
    p = sim.scene.particles.copy()
    a = p.view(np.float32).reshape(len(p), 12)
    a[:, list(_COLS)] = x0h
    sim.upload_state(p)
    sim.adjoint_record(dt, nsteps, checkpoint_every)


def _rollout_fn():
    import torch

    class Rollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, params, sim, dt, nsteps, checkpoint_every):
            sc = sim.scene
            if params is not None:
                want = np.stack([sc.materials["bulk_visc"], sc.materials["visc"]], 1).astype(np.float32)
                if not np.array_equal(params.detach().cpu().numpy().astype(np.float32), want):
                    raise ValueError("params differ from the scene's materials (upload a new scene to change them)")
            x0h = x0.detach().cpu().numpy().copy()
            _record(sim, x0h, dt, nsteps, checkpoint_every)
            y = state_tensor(sim.particles(), x0.device)
            ctx.sim = sim
            ctx.has_params = params is not None
            # the sim holds one recording: a later rollout on it replaces this one, and the
            # backward then records this rollout again from its saved input
            ctx.replay = (x0h, dt, nsteps, checkpoint_every)
            ctx.generation = sim.record_generation
            return y

        @staticmethod
        def backward(ctx, gy):
            sim = ctx.sim
            if sim.record_generation != ctx.generation:
                _record(sim, *ctx.replay)
                ctx.generation = sim.record_generation
            n = gy.shape[0]
            g_out = records_from_state(gy.contiguous())
            g_in = torch.zeros((n, 12), dtype=torch.float32, device=gy.device)
            torch.cuda.synchronize(gy.device)
            gp = sim.adjoint_backward(g_out, g_in, want_params=ctx.has_params)
            sim.synchronize()
            gx = g_in[:, list(_COLS)].clone()
            gpar = torch.as_tensor(gp[:, :2].copy(), device=gy.device) if ctx.has_params else None
            return gx, gpar, None, None, None, None

    return Rollout


_FN = None


def rollout(sim, x0, dt, nsteps, params=None, checkpoint_every=None):
    """y = state after nsteps RK steps of length dt from x0 (differentiable in x0, params)."""
    global _FN
    if _FN is None:
        _FN = _rollout_fn()
    return _FN.apply(x0, params, sim, dt, nsteps, checkpoint_every)
