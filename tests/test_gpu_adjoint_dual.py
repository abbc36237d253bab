"""Reverse mode against the exact FP64 derivative of the oracle (needs a GPU).

The oracle's forward-mode build (oracle/oracle_dual.cpp, REAL = an FP64 dual number) gives
J u exactly for any direction u, and dy/dp for any material parameter p.  The GPU's gradient
G = dL/dx0 of L = <g, y_N - x0> is checked through <G, u> = <g, J u - u> for many random
directions at once (no finite-difference step), and its parameter gradients against
<g, dy/dp>.  g and u lie on disjoint fields or particles, so the identity block drops out.
Tolerance: 1e-3 relative (an FP32 adjoint on an FP32 trajectory against exact FP64)."""
import numpy as np
import pytest

from arp import tableau as T
from test_gpu_adjoint import _records, _scene
from test_gpu_adjoint_solid import _cantilever, _contact
from helpers import deformed_solid_scene, self_contact_scene

pytestmark = pytest.mark.gpu

TOL = 1e-3
FIELDS = {"den": (0, 1), "pos": (1, 4), "vel": (4, 7)}
SCALE = {"den": 1.0, "pos": 1e-4, "vel": 1e-2}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _state_from(o, parts):
    return o.state("d", particles=parts)


def _tangent(o, st, u7):
    t = {k: np.zeros_like(v) for k, v in st.items()}
    t["den"][:] = u7[:, 0]
    t["pos"][:, :3] = u7[:, 1:4]
    t["vel"][:, :3] = u7[:, 4:7]
    return t


def _flat(y):
    return np.concatenate([y["den"][:, None], y["pos"][:, :3], y["vel"][:, :3]], 1)


def _gpu_grad(torch, sc, integ, dt, nsteps, K, g, short_skin=0.0):
    from arp.sim import FluidSim

    sim = FluidSim(sc, integrator=integ, timestep=dt)
    if short_skin:
        sim.set_neighbor_skin(0.3)
        sim.set_neighbor_short_skin(short_skin)
    sim.upload_state(sc.particles)
    sim.adjoint_record(dt, nsteps, K)
    d_in = torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda")
    gp = sim.adjoint_backward(_records(torch, g.astype(np.float32)), d_in)
    sim.synchronize()
    gi = d_in.cpu().numpy()[:, [0, 4, 5, 6, 8, 9, 10]].astype(np.float64)
    sim.close()
    return gi, gp


def _masked(rng, n, field, scale, parity=None):
    a = np.zeros((n, 7))
    lo, hi = FIELDS[field]
    a[:, lo:hi] = rng.standard_normal((n, hi - lo)) * scale
    if parity is not None:
        a[np.arange(n) % 2 != parity] = 0.0
    return a


def _cantilever_order(order):
    """The small cantilever with another Seth-Hill order (solid.rs:105-168): 0 the log
    series, odd orders the Denman-Beavers root, negative orders the inverse.  Stiffness and
    damping 1e5 times the cantilever's (stepped with dt 1e-7), so that the stress carries
    about 60 % of <g, J u> and the orders differ in it by up to 22 % (measured on the
    oracle): a wrong strain derivative cannot hide behind the hourglass term."""
    import dataclasses

    sc = _cantilever()
    mats = sc.materials.copy()
    mats["strain_order"][1] = order
    for f in ("normal_stiffness", "shear_stiffness", "normal_damp", "shear_damp"):
        mats[f][1] = np.float32(mats[f][1] * 1e5)
    return dataclasses.replace(sc, materials=mats)


def _fluid_large():
    """A viscous 24^3 block (13,824 particles, jittered, 1 % density noise, random velocities
    of 0.5 m/s per axis), so that the short lists are refreshed within the recorded steps."""
    import dataclasses

    from arp import scene as S

    sc = S.small_fluid_3d(n_side=24, box=(0.2, 0.2, 0.2))
    rng = np.random.default_rng(9)
    p = sc.particles.copy()
    p["den"] = (1000.0 * (1.0 + 0.01 * rng.standard_normal(len(p)))).astype(np.float32)
    p["vel"][:, :3] = (0.5 * rng.standard_normal((len(p), 3))).astype(np.float32)
    mats = sc.materials.copy()
    mats["visc"][1] = np.float32(2e-3)
    return dataclasses.replace(sc, particles=p, materials=mats, interactions=S.resolve_interactions(mats, sc.h))


def _plastic(seed):
    """The deformed plastic block (plastic flow active at every solid, |S| > A) with its
    stiffness and damping 1e5 times larger and relaxation time 4e3 s, stepped with dt 1e-7: the
    plastic strain then moves the stress within the two steps, and <g, J u> differs by 11-14 %
    from the same scene without plastic flow (measured on the oracle), so a wrong plastic-rate
    derivative cannot hide behind the hourglass term."""
    import dataclasses

    sc = deformed_solid_scene(dim=3, plastic=True, seed=seed)
    m = sc.materials.copy()
    for f in ("normal_stiffness", "shear_stiffness", "normal_damp", "shear_damp"):
        m[f][2] = np.float32(m[f][2] * 1e5)
    m["relaxation_time"][2] = 4e3
    return dataclasses.replace(sc, materials=m)


CASES = {
    "fluid_rk2_replay": (lambda: _scene(True), "rk2", 5e-5, 3, 2, {1: ("bulk_visc", "visc")}),
    "fluid_rk4": (lambda: _scene(False), "rk4", 5e-5, 2, 2, {1: ("bulk_visc",)}),
    # 13,824 particles, 12 steps with a checkpoint every 4 (two segments replayed), on
    # two-level kept lists (skins 0.3 / 0.03 h) recorded speculatively
    "fluid_large_rk2_lists": (_fluid_large, "rk2", 5e-5, 12, 4, {1: ("bulk_visc", "visc")}, 0.03),
    "cantilever_rk2": (_cantilever, "rk2", 1e-5, 2, 1,
                       {1: ("normal_stiffness", "shear_stiffness", "normal_damp", "shear_damp")}),
    "contact_rk2": (_contact, "rk2", 1e-5, 2, 2, {1: ("normal_stiffness", "shear_stiffness"), 2: ("bulk_visc",)}),
    "self_contact_rk2": (self_contact_scene, "rk2", 1e-5, 2, 1, {2: ("normal_stiffness", "shear_damp")}),
    # checkpoints every step in plastic_rk2 (a replay restores the plastic strain).  Its
    # dL/d normal_damp is 430 times smaller than dL/d shear_damp (a sum that nearly cancels;
    # the oracle's value agrees with its central differences to 3e-5, the FP32 sweep's to
    # 6e-3), so the normal damping is checked on the unscaled block (plastic_soft_rk2)
    "plastic_rk2": (lambda: _plastic(0), "rk2", 1e-7, 2, 1,
                    {2: ("normal_stiffness", "shear_stiffness", "shear_damp"), 1: ("bulk_visc",)}),
    "plastic_soft_rk2": (lambda: deformed_solid_scene(dim=3, plastic=True), "rk2", 1e-6, 2, 1,
                         {2: ("normal_damp", "normal_stiffness")}),
    "plastic_rk4": (lambda: _plastic(3), "rk4", 1e-7, 2, 2, {2: ("shear_stiffness", "shear_damp")}),
    "cantilever_order2": (lambda: _cantilever_order(2), "rk2", 1e-7, 2, 1, {1: ("normal_stiffness",)}),
    "cantilever_order0": (lambda: _cantilever_order(0), "rk2", 1e-7, 2, 1, {1: ("normal_stiffness", "shear_damp")}),
    "cantilever_order1": (lambda: _cantilever_order(1), "rk2", 1e-7, 2, 1, {1: ("shear_stiffness",)}),
    "cantilever_order3": (lambda: _cantilever_order(3), "rk2", 1e-7, 2, 1, {1: ("normal_stiffness", "normal_damp")}),
    "cantilever_order-2": (lambda: _cantilever_order(-2), "rk2", 1e-7, 2, 1, {1: ("shear_stiffness", "shear_damp")}),
}
PAIRS = [("vel", "pos"), ("den", "vel"), ("pos", "vel"), ("vel", "den"), ("vel", "vel"), ("pos", "pos")]


@pytest.mark.parametrize("case", sorted(CASES))
def test_adjoint_vs_exact_derivative(torch_cuda, oracle_mod, case):
    make, integ, dt, nsteps, K, params = CASES[case][:6]
    short_skin = CASES[case][6] if len(CASES[case]) > 6 else 0.0
    sc = make()
    tab = T.tableau(integ)
    o = oracle_mod.Oracle(sc)
    st = _state_from(o, sc.particles)
    rng = np.random.default_rng(17)
    worst = 0.0
    for gf, uf in PAIRS:
        same = gf == uf
        g = _masked(rng, sc.n, gf, 1.0, 0 if same else None)
        gi, gp = _gpu_grad(torch_cuda, sc, integ, dt, nsteps, K, g, short_skin)
        for rep in range(3):  # several directions for one gradient
            u = _masked(rng, sc.n, uf, SCALE[uf], 1 if same else None)
            y, yd = o.jvp(st, _tangent(o, st, u), tab, dt, nsteps)
            exact = float(np.sum(g * (_flat(yd) - u)))
            ad = float(np.sum(gi * u))
            rel = abs(ad - exact) / max(abs(exact), 1e-300)
            worst = max(worst, rel)
            print("case %s g %s u %s #%d ad %.9e exact %.9e rel %.2e" % (case, gf, uf, rep, ad, exact, rel),
                  flush=True)
            assert exact != 0.0
            assert rel < TOL
        if gf == "vel" and uf == "pos":  # parameter gradients of this loss
            zero = _tangent(o, st, np.zeros((sc.n, 7)))
            cols = {"bulk_visc": 0, "visc": 1, "normal_stiffness": 2, "shear_stiffness": 3, "normal_damp": 4,
                    "shear_damp": 5}
            for m, names in params.items():
                for name in names:
                    _, yd = o.jvp(st, zero, tab, dt, nsteps, param=(m, name))
                    exact = float(np.sum(g * _flat(yd)))
                    ad = float(gp[m, cols[name]])
                    rel = abs(ad - exact) / max(abs(exact), 1e-300)
                    print("case %s param %d %s ad %.9e exact %.9e rel %.2e" % (case, m, name, ad, exact, rel),
                          flush=True)
                    assert exact != 0.0
                    assert rel < TOL
    print("case %s worst directional rel %.2e" % (case, worst), flush=True)
