"""Reverse mode of the fused step (arp_adjoint_record / arp_adjoint_backward) against central
finite differences of the FP64 oracle (needs a GPU).

The adjoint is checked through directional derivatives <dL/dx0, u> with L = <g, y_N(x0)>:
the cotangent g and the direction u are chosen on disjoint fields (velocity vs position,
density vs velocity, ...) or, for the same field, on disjoint particles, so the identity part
of the step Jacobian drops out and only the force / density-rate Jacobian is compared.  Central differences in FP64 with a relative
step of 1e-6 carry an error far below the FP32 adjoint's; tolerance: relative error 2e-3 of
the directional derivative (FP32 trajectory and Jacobian vs FP64)."""
import dataclasses

import numpy as np
import pytest

from arp import scene as S
from arp import tableau as T
from arp import wire

pytestmark = pytest.mark.gpu

TOL = 2e-3
FIELDS = {"den": (0, 1), "pos": (1, 4), "vel": (4, 7)}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _scene(viscous):
    sc = S.small_fluid_3d(n_side=8, box=(0.06, 0.06, 0.06))
    rng = np.random.default_rng(7)
    p = sc.particles.copy()
    p["den"] = (1000.0 * (1.0 + 0.01 * rng.standard_normal(len(p)))).astype(np.float32)
    p["vel"][:, :3] = (0.05 * rng.standard_normal((len(p), 3))).astype(np.float32)
    mats = sc.materials.copy()
    if viscous:
        mats["visc"][1] = np.float32(2e-3)
    inter = S.resolve_interactions(mats, sc.h)
    return dataclasses.replace(sc, particles=p, materials=mats, interactions=inter)


def _x_of(parts):
    return np.concatenate([parts["den"][:, None], parts["pos"][:, :3], parts["vel"][:, :3]], 1).astype(np.float64)


def _oracle_L(O, sc, x, g, tab, dt, nsteps, mats=None):
    scn = dataclasses.replace(sc, materials=mats) if mats is not None else sc
    o = O.Oracle(scn)
    st = o.state("d")
    st["den"][:] = x[:, 0]
    st["pos"][:, :3] = x[:, 1:4]
    st["vel"][:, :3] = x[:, 4:7]
    y = o.step(st, tab, dt, nsteps)
    yv = np.concatenate([y["den"][:, None], y["pos"][:, :3], y["vel"][:, :3]], 1)
    # <g, y_N - x>: same derivative when g and the direction are disjoint, far less rounding
    return float(np.sum(g * (yv - x)))


def _same_state(a, b):
    """Bitwise equality of every field of two Particle record arrays (padding ignored)."""
    return all(np.array_equal(np.ascontiguousarray(a[f]).view(np.uint32), np.ascontiguousarray(b[f]).view(np.uint32))
               for f in ("den", "mat", "solid_id", "pos", "vel"))


def _records(torch, arr7):
    r = np.zeros((len(arr7), 12), np.float32)
    r[:, [0, 4, 5, 6, 8, 9, 10]] = arr7
    return torch.from_numpy(r).cuda()


def _gpu_adjoint(torch, sc, integ, dt, nsteps, K, g):
    from arp.sim import FluidSim

    sim = FluidSim(sc, integrator=integ, timestep=dt)
    sim.upload_state(sc.particles)
    sim.adjoint_record(dt, nsteps, K)
    d_out = _records(torch, g.astype(np.float32))
    d_in = torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda")
    gp = sim.adjoint_backward(d_out, d_in)
    sim.synchronize()
    gi = d_in.cpu().numpy()[:, [0, 4, 5, 6, 8, 9, 10]].astype(np.float64)
    return sim, gi, gp


def _masked(rng, n, field, scale, parity=None):
    a = np.zeros((n, 7))
    lo, hi = FIELDS[field]
    a[:, lo:hi] = rng.standard_normal((n, hi - lo)) * scale
    if parity is not None:  # keep every other particle (neighbours in the lattice order)
        a[np.arange(n) % 2 != parity] = 0.0
    return a


# per-field scales: density ~1000, positions ~h, velocities ~0.1
SCALE = {"den": 1.0, "pos": 1e-4, "vel": 1e-2}
PAIRS = [("vel", "pos"), ("den", "vel"), ("vel", "den"), ("den", "pos"), ("pos", "vel"), ("vel", "vel")]


@pytest.mark.parametrize("integ,nsteps,K,viscous", [("euler", 1, 1, True), ("rk2", 3, 2, False),
                                                      ("rk4", 2, 1, True), ("heun2", 2, 2, True)])
def test_adjoint_directional_vs_fd(torch_cuda, oracle_mod, integ, nsteps, K, viscous):
    sc = _scene(viscous)
    dt = 5e-5
    tab = T.tableau(integ)
    n = sc.n
    x0 = _x_of(sc.particles)
    rng = np.random.default_rng(11)
    for gf, uf in PAIRS:
        # same-field pairs: g and u on disjoint particles, so the identity block drops out
        same = gf == uf
        g = _masked(rng, n, gf, 1.0, 0 if same else None)
        print("case adjoint", integ, nsteps, K, viscous, "g", gf, "u", uf, flush=True)
        sim, gi, _ = _gpu_adjoint(torch_cuda, sc, integ, dt, nsteps, K, g)
        sim.close()
        u = _masked(rng, n, uf, SCALE[uf], 1 if same else None)
        eps = 1e-3  # u is already scaled to ~1e-3 of each field's natural size
        fd = (_oracle_L(oracle_mod, sc, x0 + eps * u, g, tab, dt, nsteps) -
              _oracle_L(oracle_mod, sc, x0 - eps * u, g, tab, dt, nsteps)) / (2 * eps)
        ad = float(np.sum(gi * u))
        rel = abs(ad - fd) / max(abs(fd), 1e-30)
        print("  ad %.9e fd %.9e rel %.2e" % (ad, fd, rel), flush=True)
        assert fd != 0.0
        assert rel < TOL


def test_adjoint_param_grads_vs_fd(torch_cuda, oracle_mod):
    sc = _scene(True)
    dt, nsteps, integ = 5e-5, 2, "rk2"
    tab = T.tableau(integ)
    rng = np.random.default_rng(3)
    # velocity cotangent only: the density's FP64 rounding (ulp ~1e-13 at 1000) would swamp
    # these small parameter derivatives
    g = np.concatenate([np.zeros((sc.n, 4)), rng.standard_normal((sc.n, 3))], 1)
    sim, _, gp = _gpu_adjoint(torch_cuda, sc, integ, dt, nsteps, 1, g)
    sim.close()
    x0 = _x_of(sc.particles)
    for m, field, col in [(1, "bulk_visc", 0), (1, "visc", 1)]:
        base = float(sc.materials[field][m])
        # L is affine in each viscosity up to terms far below the tolerance over two steps,
        # so a wide central difference (+-25%) is accurate and clear of rounding noise
        eps = 0.25 * base
        mp, mm = sc.materials.copy(), sc.materials.copy()
        mp[field][m] = base + eps
        mm[field][m] = base - eps
        ep = float(mp[field][m]) - float(mm[field][m])  # the FP32 values actually used
        fd = (_oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, mats=mp) -
              _oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, mats=mm)) / ep
        ad = float(gp[m, col])
        rel = abs(ad - fd) / max(abs(fd), 1e-30)
        print("param", field, "ad %.6e fd %.6e rel %.2e" % (ad, fd, rel), flush=True)
        assert abs(fd) > 0
        assert rel < 5e-3


def test_adjoint_leaves_state_and_replays_exactly(torch_cuda):
    from arp.sim import FluidSim

    sc = _scene(True)
    dt = 5e-5
    g = np.random.default_rng(5).standard_normal((sc.n, 7))
    sim = FluidSim(sc, integrator="rk2", timestep=dt)
    sim.set_neighbor_skin(0.0)  # the search at every stage, as a plain step (compared below)
    sim.upload_state(sc.particles)
    sim.adjoint_record(dt, 5, 2)
    end = sim.particles().copy()
    d_out = _records(torch_cuda, g.astype(np.float32))
    outs = []
    for _ in range(2):
        d_in = torch_cuda.zeros((sc.n, 12), dtype=torch_cuda.float32, device="cuda")
        gp = sim.adjoint_backward(d_out, d_in)
        sim.synchronize()
        outs.append((d_in.cpu().numpy().copy(), gp.copy()))
        assert _same_state(sim.particles(), end)
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    # one segment (no replay) gives the same gradient as checkpoints every 2 steps: bitwise
    # with the search at every stage; with kept neighbour lists (which restart with every
    # segment) up to the order of the per-particle sums
    for skin in (0.0, 0.2):
        s1 = FluidSim(sc, integrator="rk2", timestep=dt)
        s1.set_neighbor_skin(skin)
        s1.upload_state(sc.particles)
        s1.adjoint_record(dt, 5, 2)
        d_in1 = torch_cuda.zeros((sc.n, 12), dtype=torch_cuda.float32, device="cuda")
        s1.adjoint_backward(d_out, d_in1)
        d_in1b = torch_cuda.zeros((sc.n, 12), dtype=torch_cuda.float32, device="cuda")
        s1.adjoint_backward(d_out, d_in1b)  # replays repeat the recording exactly
        assert torch_cuda.equal(d_in1, d_in1b)
        sim2 = FluidSim(sc, integrator="rk2", timestep=dt)
        sim2.set_neighbor_skin(skin)
        sim2.upload_state(sc.particles)
        sim2.adjoint_record(dt, 5, 5)
        d_in2 = torch_cuda.zeros((sc.n, 12), dtype=torch_cuda.float32, device="cuda")
        sim2.adjoint_backward(d_out, d_in2)
        sim2.synchronize()
        g1, g2 = d_in1.cpu().numpy(), d_in2.cpu().numpy()
        if skin == 0.0:
            assert np.array_equal(g2, g1)
        else:
            assert np.max(np.abs(g2 - g1)) <= 1e-4 * np.max(np.abs(g1))
        s1.close()
        if skin == 0.0:
            sim2.close()
    # continuing after the backward equals continuing a plain run
    sim.step(dt, 3)
    sim3 = FluidSim(sc, integrator="rk2", timestep=dt)
    sim3.set_neighbor_skin(0.0)
    sim3.upload_state(sc.particles)
    sim3.step(dt, 8)
    assert _same_state(sim.particles(), sim3.particles())
    for s in (sim, sim2, sim3):
        s.close()


@pytest.mark.parametrize("integ,sigma,nsteps,K", [("rk2", 1.0, 20, 20), ("rk4", 1.0, 24, 12), ("heun3", 0.05, 18, 9)])
def test_speculative_recording_matches_the_chain(torch_cuda, integ, sigma, nsteps, K):
    """Recorded steps issued speculatively (each stage as its list walk alone, also in the
    replays of the reverse sweep) leave bitwise the state, list statistics and gradients of
    issuing the rebuild chain at every stage, also when the lists of a speculated step turn
    out invalid and the step is recorded again with the chain (fast particles: lists rebuilt
    every few steps)."""
    from arp.sim import FluidSim

    sc = _scene(False)
    rng = np.random.default_rng(11)
    sc.particles["vel"][:, :3] = (sigma * rng.standard_normal((sc.n, 3))).astype(np.float32)
    dt = 5e-5
    g = np.random.default_rng(5).standard_normal((sc.n, 7))
    d_out = _records(torch_cuda, g.astype(np.float32))
    print("case speculative recording", integ, sigma, nsteps, K, flush=True)
    res = []
    for spec in (True, False):
        sim = FluidSim(sc, integrator=integ, timestep=dt)
        sim.set_neighbor_skin(0.2, speculate=spec)
        sim.upload_state(sc.particles)
        sim.adjoint_record(dt, nsteps, K)
        end = sim.particles().copy()
        stats = sim.neighbor_stats()
        grads = []
        for _ in range(2):  # (a second sweep replays the segments again)
            d_in = torch_cuda.zeros((sc.n, 12), dtype=torch_cuda.float32, device="cuda")
            gp = sim.adjoint_backward(d_out, d_in)
            sim.synchronize()
            grads.append((d_in.cpu().numpy().copy(), gp.copy()))
        res.append((end, stats, grads, sim.time, sim.speculation_stats()))
        sim.close()
    (ea, sa, ga, ta, spa), (eb, sb, gb, tb, spb) = res
    print("  lists %s, speculative steps %d, run again %d" % (sa, spa[0], spa[1]), flush=True)
    assert spb == (0, 0)
    assert spa[0] > 0
    if sigma >= 1.0:
        assert spa[1] > 0
    assert _same_state(ea, eb)
    assert sa == sb and ta == tb
    for (da, pa), (db, pb) in zip(ga, gb):
        assert np.isfinite(da).all()
        assert np.array_equal(da, db) and np.array_equal(pa, pb)


def test_autograd_rollout(torch_cuda):
    from arp.autograd import rollout, state_tensor
    from arp.sim import FluidSim

    torch = torch_cuda
    sc = _scene(True)
    dt = 5e-5
    sim = FluidSim(sc, integrator="rk2", timestep=dt)
    x0 = state_tensor(sc.particles).requires_grad_(True)
    params = torch.tensor(np.stack([sc.materials["bulk_visc"], sc.materials["visc"]], 1), device="cuda",
                          requires_grad=True)
    y = rollout(sim, x0, dt, 3, params=params, checkpoint_every=2)
    loss = (y[:, 4:7] ** 2).sum() + y[:, 0].sum()
    loss.backward()
    g = torch.zeros_like(y)
    g[:, 4:7] = 2 * y[:, 4:7].detach()
    g[:, 0] = 1.0
    _, gi, gp = _gpu_adjoint(torch, sc, "rk2", dt, 3, 2, g.cpu().numpy().astype(np.float64))
    assert np.allclose(x0.grad.cpu().numpy(), gi, rtol=1e-6, atol=1e-6 * np.abs(gi).max())
    assert np.allclose(params.grad.cpu().numpy(), gp[:, :2], rtol=1e-6, atol=1e-9)
    sim.close()


def test_autograd_chained_rollouts_on_one_sim(torch_cuda):
    """Two rollouts chained on one sim: the first one's backward must not use the second
    one's recording (it records its own rollout again).  Checked against the same chain
    on two separate sims."""
    from arp.autograd import rollout, state_tensor
    from arp.sim import FluidSim

    torch = torch_cuda
    sc = _scene(True)
    dt = 5e-5

    def chain(sim_a, sim_b):
        x0 = state_tensor(sc.particles).requires_grad_(True)
        y1 = rollout(sim_a, x0, dt, 2)
        y2 = rollout(sim_b, y1, dt, 2)
        loss = (y1[:, 4:7] ** 2).sum() + (y2[:, 4:7] ** 3).sum() + y2[:, 1].sum()
        loss.backward()
        return x0.grad.cpu().numpy()

    shared = FluidSim(sc, integrator="rk2", timestep=dt)
    g_shared = chain(shared, shared)
    a = FluidSim(sc, integrator="rk2", timestep=dt)
    b = FluidSim(sc, integrator="rk2", timestep=dt)
    g_sep = chain(a, b)
    assert np.isfinite(g_sep).all() and np.abs(g_sep).max() > 0
    assert np.allclose(g_shared, g_sep, rtol=1e-6, atol=1e-6 * np.abs(g_sep).max())
    for s in (shared, a, b):
        s.close()


def test_backward_refuses_changed_neighbor_settings(torch_cuda):
    """A replay must walk the recorded steps with the list capacity they were recorded with:
    changing the list settings after a recording drops it."""
    from arp._lib import ArpError
    from arp.sim import FluidSim

    torch = torch_cuda
    sc = _scene(False)
    dt = 5e-5
    sim = FluidSim(sc, integrator="rk2", timestep=dt)
    sim.set_neighbor_skin(0.2)
    sim.adjoint_record(dt, 4, 2)
    sim.set_neighbor_skin(0.2, cap=1024)
    d_out = torch.ones((sc.n, 12), dtype=torch.float32, device="cuda")
    d_in = torch.zeros_like(d_out)
    with pytest.raises(ArpError):
        sim.adjoint_backward(d_out, d_in)
    # recorded again with the new capacity, the sweep runs (with replays)
    sim.adjoint_record(dt, 4, 2)
    sim.adjoint_backward(d_out, d_in)
    sim.synchronize()
    assert torch.isfinite(d_in).all()
    sim.close()
