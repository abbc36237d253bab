"""Reverse mode through the solid path (elastic force, hourglass, strain/stress, fluid-solid
contact) against central differences of the FP64 oracle (needs a GPU).

Same scheme as the fluid adjoint tests: directional derivatives <dL/dx0, u> of
L = <g, y_N - x0> with g and u on disjoint fields or particles; material gradients
(normal/shear stiffness and damping, bulk viscosity) and the per-solid stiffness-field
gradient by perturbing the parameter in the oracle.  Tolerance 2e-3 relative (FP32 adjoint
on an FP32 trajectory vs FP64 differences)."""
import dataclasses

import numpy as np
import pytest

from arp import scene as S
from arp import tableau as T

pytestmark = pytest.mark.gpu

TOL = 2e-3
FIELDS = {"den": (0, 1), "pos": (1, 4), "vel": (4, 7)}
SCALE = {"den": 1.0, "pos": 1e-4, "vel": 1e-2}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _cantilever():
    sc = S.cantilever_3d(nx=4, ny=4, nz=10)
    rng = np.random.default_rng(1)
    p = sc.particles.copy()
    p["pos"][:, :3] += (2e-4 * rng.standard_normal((len(p), 3))).astype(np.float32)
    p["vel"][:, :3] = (0.05 * rng.standard_normal((len(p), 3))).astype(np.float32)
    mats = sc.materials.copy()
    mats["normal_damp"][1] = np.float32(1e-3 * mats["normal_stiffness"][1])
    mats["shear_damp"][1] = np.float32(1e-3 * mats["shear_stiffness"][1])
    return dataclasses.replace(sc, particles=p, materials=mats)


def _contact():
    """The contact scene at small size: a 6^3 fluid block and a 68-particle elastic ball (every
    particle with a full reference support) one spacing above it, moving down; the fluid
    stirred and the ball's positions perturbed."""
    sc = S.contact_3d(n_fluid_side=6, n_sphere=68)
    p = sc.particles.copy()
    fl = p["mat"] == 2
    rng = np.random.default_rng(2)
    p["vel"][fl, :3] = (0.05 * rng.standard_normal((int(fl.sum()), 3))).astype(np.float32)
    p["pos"][~fl, :3] += (2e-4 * rng.standard_normal((int((~fl).sum()), 3))).astype(np.float32)
    return dataclasses.replace(sc, particles=p, name="contact_ball")


def _x_of(parts):
    return np.concatenate([parts["den"][:, None], parts["pos"][:, :3], parts["vel"][:, :3]], 1).astype(np.float64)


def _oracle_L(O, sc, x, g, tab, dt, nsteps, mats=None, kscale=None):
    scn = dataclasses.replace(sc, materials=mats) if mats is not None else sc
    o = O.Oracle(scn, kscale=kscale)
    st = o.state("d")
    st["den"][:] = x[:, 0]
    st["pos"][:, :3] = x[:, 1:4]
    st["vel"][:, :3] = x[:, 4:7]
    y = o.step(st, tab, dt, nsteps)
    yv = np.concatenate([y["den"][:, None], y["pos"][:, :3], y["vel"][:, :3]], 1)
    return float(np.sum(g * (yv - x)))


def _records(torch, arr7):
    r = np.zeros((len(arr7), 12), np.float32)
    r[:, [0, 4, 5, 6, 8, 9, 10]] = arr7
    return torch.from_numpy(r).cuda()


def _gpu_adjoint(torch, sc, integ, dt, nsteps, K, g, kscale=None):
    from arp.sim import FluidSim

    sim = FluidSim(sc, integrator=integ, timestep=dt)
    if kscale is not None:
        sim.set_stiffness_field(kscale)
    sim.upload_state(sc.particles)
    sim.adjoint_record(dt, nsteps, K)
    d_in = torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda")
    gp = sim.adjoint_backward(_records(torch, g.astype(np.float32)), d_in)
    sim.synchronize()
    gi = d_in.cpu().numpy()[:, [0, 4, 5, 6, 8, 9, 10]].astype(np.float64)
    ks = sim.stiffness_grad()
    sim.close()
    return gi, gp, ks


def _masked(rng, n, field, scale, parity=None):
    a = np.zeros((n, 7))
    lo, hi = FIELDS[field]
    a[:, lo:hi] = rng.standard_normal((n, hi - lo)) * scale
    if parity is not None:
        a[np.arange(n) % 2 != parity] = 0.0
    return a


PAIRS = [("vel", "pos"), ("vel", "vel"), ("pos", "vel"), ("den", "pos"), ("vel", "den")]


@pytest.mark.parametrize("scene,integ,nsteps,K", [("cantilever", "rk2", 2, 1), ("cantilever", "rk4", 1, 1),
                                                   ("contact", "rk2", 2, 2)])
def test_solid_adjoint_directional_vs_fd(torch_cuda, oracle_mod, scene, integ, nsteps, K):
    sc = _cantilever() if scene == "cantilever" else _contact()
    dt = 1e-5
    tab = T.tableau(integ)
    x0 = _x_of(sc.particles)
    rng = np.random.default_rng(11)
    for gf, uf in PAIRS:
        same = gf == uf
        g = _masked(rng, sc.n, gf, 1.0, 0 if same else None)
        u = _masked(rng, sc.n, uf, SCALE[uf], 1 if same else None)
        print("case solid adjoint", scene, integ, nsteps, K, "g", gf, "u", uf, flush=True)
        gi, _, _ = _gpu_adjoint(torch_cuda, sc, integ, dt, nsteps, K, g)
        eps = 1e-3
        fd = (_oracle_L(oracle_mod, sc, x0 + eps * u, g, tab, dt, nsteps) -
              _oracle_L(oracle_mod, sc, x0 - eps * u, g, tab, dt, nsteps)) / (2 * eps)
        ad = float(np.sum(gi * u))
        rel = abs(ad - fd) / max(abs(fd), 1e-30)
        print("  ad %.9e fd %.9e rel %.2e" % (ad, fd, rel), flush=True)
        assert fd != 0.0
        assert rel < TOL


def test_solid_adjoint_material_and_field_grads(torch_cuda, oracle_mod):
    sc = _cantilever()
    dt, nsteps, integ = 1e-5, 2, "rk2"
    tab = T.tableau(integ)
    rng = np.random.default_rng(3)
    g = _masked(rng, sc.n, "vel", 1.0)
    # a seeded lognormal stiffness field (sigma 0.1)
    ks = np.exp(0.1 * rng.standard_normal(sc.ns)).astype(np.float32)
    _, gp, kg = _gpu_adjoint(torch_cuda, sc, integ, dt, nsteps, 1, g, kscale=ks)
    x0 = _x_of(sc.particles)
    for field, col in [("normal_stiffness", 2), ("shear_stiffness", 3), ("normal_damp", 4), ("shear_damp", 5)]:
        base = float(sc.materials[field][1])
        mp, mm = sc.materials.copy(), sc.materials.copy()
        mp[field][1] = base * 1.01
        mm[field][1] = base * 0.99
        ep = float(mp[field][1]) - float(mm[field][1])
        fd = (_oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, mats=mp, kscale=ks) -
              _oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, mats=mm, kscale=ks)) / ep
        ad = float(gp[1, col])
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print("param", field, "ad %.6e fd %.6e rel %.2e" % (ad, fd, rel), flush=True)
        assert fd != 0.0
        assert rel < 5e-3
    for s in (0, 7, sc.ns - 1):
        kp, km = ks.copy(), ks.copy()
        kp[s] += np.float32(0.01)
        km[s] -= np.float32(0.01)
        fd = (_oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, kscale=kp) -
              _oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, kscale=km)) / (float(kp[s]) - float(km[s]))
        rel = abs(kg[s] - fd) / max(abs(fd), 1e-300)
        print("field solid", s, "ad %.6e fd %.6e rel %.2e" % (kg[s], fd, rel), flush=True)
        assert fd != 0.0
        assert rel < 5e-3


def test_contact_initial_velocity_and_viscosity_grads(torch_cuda, oracle_mod):
    sc = _contact()
    dt, nsteps, integ = 1e-5, 2, "rk2"
    tab = T.tableau(integ)
    x0 = _x_of(sc.particles)
    solid = sc.particles["mat"] == 1
    # loss: the cube's centroid height at the end; gradient w.r.t. its initial velocity
    # (a uniform shift of the solid particles' velocities)
    g = np.zeros((sc.n, 7))
    g[solid, 2] = 1.0 / solid.sum()
    gi, _, _ = _gpu_adjoint(torch_cuda, sc, integ, dt, nsteps, 2, g)
    for axis in range(3):
        u = np.zeros((sc.n, 7))
        u[solid, 4 + axis] = 1e-2
        fd = (_oracle_L(oracle_mod, sc, x0 + 1e-3 * u, g, tab, dt, nsteps) -
              _oracle_L(oracle_mod, sc, x0 - 1e-3 * u, g, tab, dt, nsteps)) / 2e-3
        ad = float(np.sum(gi * u))
        err = abs(ad - fd)
        print("centroid vs initial velocity axis", axis, "ad %.6e fd %.6e" % (ad, fd), flush=True)
        # the y component is the leading term (2 dt / N per unit shift); x, z only see the
        # contact coupling
        scale = abs(fd) if axis == 1 else 2 * dt * 1e-2
        assert err < TOL * scale
    # the fluid's artificial viscosity (bulk_visc of material 2), velocity loss
    rng = np.random.default_rng(5)
    g = _masked(rng, sc.n, "vel", 1.0)
    _, gp, _ = _gpu_adjoint(torch_cuda, sc, integ, dt, nsteps, 2, g)
    base = float(sc.materials["bulk_visc"][2])
    mp, mm = sc.materials.copy(), sc.materials.copy()
    mp["bulk_visc"][2] = base * 1.25
    mm["bulk_visc"][2] = base * 0.75
    ep = float(mp["bulk_visc"][2]) - float(mm["bulk_visc"][2])
    fd = (_oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, mats=mp) -
          _oracle_L(oracle_mod, sc, x0, g, tab, dt, nsteps, mats=mm)) / ep
    ad = float(gp[2, 0])
    rel = abs(ad - fd) / max(abs(fd), 1e-300)
    print("fluid bulk viscosity ad %.6e fd %.6e rel %.2e" % (ad, fd, rel), flush=True)
    assert fd != 0.0
    assert rel < 5e-3
