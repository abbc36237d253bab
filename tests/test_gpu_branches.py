"""Reference branches beyond the Tait fluid and Green-Lagrange solid, on the GPU against the
FP32 oracle through the C ABI (needs a GPU):

- every strain measure the reference scenes use: log strain (strain_order 0, solid.rs:105-116,
  examples/solid_test_3D.toml), Denman-Beavers orders 1 and 3 (solid.rs:119-133), the
  inverse order -2, and order 4;
- the ideal-gas and constant equations of state (fluid.rs:45-52) and the constant, Tait,
  Lennard-Jones and Hertz contact potentials (fluid.rs:54-67) between two fluids;
- reduced scene files in the style of examples/raindrop.toml (ideal gas around a drop, strong
  contact) and examples/solid_test_3D.toml (log-strain ball on a floor, velocity Verlet);
- the 2-D dam break's 1000-step trajectory: every 100 steps one step of the oracle from the
  GPU's state.

Tolerance: normwise relative error max|gpu - oracle| / max|oracle| <= 1e-5 per field."""
import numpy as np
import pytest

from arp import config as K
from arp import scene as S
from arp import tableau as T
from arp import wire
from helpers import box_scene, deformed_solid_scene, uniform_block

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _check(tag, oracle_mod, got, ref, fields=("den", "pos", "vel"), tol=TOL):
    errs = {k: oracle_mod.rel_err(got[k], ref[k]) for k in fields}
    print(tag, {k: "%.2e" % e for k, e in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e <= tol, (tag, k, e)


@pytest.mark.parametrize("order", [0, 1, 3, -2, 4])
def test_strain_orders(torch_cuda, oracle_mod, order):
    from arp.sim import FluidSim, from_device, to_device

    sc = deformed_solid_scene(dim=3, order=order, seed=order + 5)
    print("case strain order", order, sc.n, sc.ns, flush=True)
    o = oracle_mod.Oracle(sc)
    sim = FluidSim(sc, integrator="rk2", timestep=1e-5)
    d32 = o.compute_forces(o.state("f"))
    din, sin = to_device(sc.particles), to_device(sc.solids)
    dout, sout = to_device(wire.particles(sc.n)), to_device(wire.solids(sc.ns))
    sim.compute_forces(din, sin, dout, sout)
    torch_cuda.cuda.synchronize()
    got = from_device(dout, wire.PARTICLE)
    _check("forces order %d" % order, oracle_mod, {k: got[k] for k in ("den", "pos", "vel")}, d32)
    y32 = o.step(o.state("f"), T.tableau("rk2"), 1e-5, 2)
    sim.update(2)
    _check("2 steps order %d" % order, oracle_mod, sim.particles(), y32)
    sim.close()


def _two_fluids(pot, eos2):
    """Two 3-D fluid blocks side by side: material 1 Tait (or ideal gas), material 2 with EOS
    eos2; their interaction has potential `pot`.  fluid.rs:177 passes the interaction's radius
    as the potential's strength and its strength as the range r0, so (strength, radius) below
    are (r0, scale): a range about the spacing and a scale that keeps the forces moderate."""
    sp = 0.01
    h = 1.3 * sp
    A = uniform_block(7, sp, 3, jitter=0.1, seed=1, origin=(0.3, 0.2, 0.3))
    B = uniform_block(7, sp, 3, jitter=0.1, seed=2, origin=(0.37, 0.2, 0.3))
    P = np.concatenate([A, B])
    f1 = S.fluid_material(h, 0.05, density=1000.0, speed_of_sound=10.0, viscosity=0.01)
    if eos2 == "ideal":
        f2 = S.fluid_material(h, 0.05, density=1000.0, speed_of_sound=10.0, start_density=1000.0,
                              target_density=500.0, viscosity=0.02)
    else:
        f2 = S.fluid_material(h, 0.05, density=1000.0, speed_of_sound=10.0, state=wire.EQ_CONSTANT)
    for m in (f1, f2):
        m["mass"] = 1000.0 * sp**3
    strength, radius = {wire.POT_CONSTANT: (2 * h, 300.0), wire.POT_TAIT: (0.008, 1000.0),
                        wire.POT_LENNARD_JONES: (0.008, 100.0), wire.POT_HERTZ: (0.01, 1000.0)}[pot]
    it = {(1, 2): wire.interaction(pot, float(np.float32(strength)), radius, 0.02, 0.1)}
    Bd = uniform_block(12, sp, 3, origin=(0.25, 0.12, 0.25))
    Bd = Bd[Bd[:, 1] < 0.2]
    sc = box_scene(P, bpos=Bd, bounds=((0.0, 0.0, 0.0), (0.8, 0.8, 0.8)), h=h, fluid=f1,
                   mats_extra=[f2], interactions=it)
    sc.particles["mat"][len(A):] = 2
    rng = np.random.default_rng(pot)
    sc.particles["vel"][:, :3] = rng.normal(0, 0.05, (sc.n, 3))
    sc.particles["den"] *= (1 + rng.uniform(-5e-3, 5e-3, sc.n)).astype(np.float32)
    return sc


@pytest.mark.parametrize("pot", ["constant", "tait", "lennard_jones", "hertz"])
@pytest.mark.parametrize("eos2", ["ideal", "constant"])
def test_eos_and_contact_potentials(torch_cuda, oracle_mod, pot, eos2):
    from arp.sim import FluidSim

    p = {"constant": wire.POT_CONSTANT, "tait": wire.POT_TAIT, "lennard_jones": wire.POT_LENNARD_JONES,
         "hertz": wire.POT_HERTZ}[pot]
    sc = _two_fluids(p, eos2)
    assert int(sc.materials["state_eq"][2]) == (wire.EQ_IDEAL_GAS if eos2 == "ideal" else wire.EQ_CONSTANT)
    assert int(sc.interactions[1 * sc.materials.shape[0] + 2]["potential"]) == p
    print("case eos/contact", pot, eos2, sc.n, flush=True)
    o = oracle_mod.Oracle(sc)
    d32 = o.compute_forces(o.state("f"))
    assert np.isfinite(d32["vel"]).all()
    sim = FluidSim(sc, integrator="rk4", timestep=5e-5)
    sim.set_neighbor_skin(0.0)  # the plain per-stage search
    y32 = o.step(o.state("f"), T.tableau("rk4"), 5e-5, 2)
    sim.update(2)
    _check("eos/contact %s %s" % (pot, eos2), oracle_mod, sim.particles(), y32)
    sim.close()


RAINDROP_LIKE = """\
name = "Drop in gas"
kernel_radius = 0.005
gravity = 3.1
time_step = 0.0001
subticks = 4
integrator = "RK4"
min = [-0.06, -0.3]
dim = [0.12, 0.16]

[objects.drop]
packing = 0.5
speed_of_sound = 1000.0
bulk_viscocity = 100.0
density = 20.0
region = {center = [0.0, -0.2], radii = [0.02, 0.02]}
gas = {contact_strength = 20000.0}
wall = {contact_strength = 20000.0}

[objects.gas]
packing = 0.5
state_eq = "Ideal_Gas"
speed_of_sound = 1000.0
start_density = 0.05
target_density = 0.0
bulk_viscocity = 100.0
difference = [
  {min = [-0.05, -0.29], dim = [0.1, 0.14]},
  {center = [0.0, -0.2], radii = [0.025, 0.025]}
]
wall = {contact_strength = 20000.0}

[boundaries.wall]
friction = 0.0
region = {min = [-0.06, -0.3], dim = [0.12, 0.16], border = 0.01}
"""

LOG_BALL = """\
name = "Log-strain ball"
kernel_radius = 0.02
gravity = 5.0
time_step = 0.00016
subticks = 1
artificial_viscocity = 0.0
integrator = "Verlet"
min = [-0.3, -0.6, -0.5]
dim = [1.0, 1.0, 1.0]

[objects.ball]
packing = 0.7
density = 1.0
strain_order = 0
normal_stiffness = -80000000.0
shear_stiffness = -20000000.0
shear_dampening = -70000.0
region = {center = [0.0, -0.5, -0.3], dim = [0.07, 0.07, 0.07]}
velocity = [0.0, -0.5, 0.0]
floor = {potential = "linear", contact_strength = 10000.0, friction = 0.5}

[boundaries.floor]
packing = 0.5
region = {min = [-0.1, -0.6, -0.4], dim = [0.2, 0.06, 0.2]}
"""


@pytest.mark.parametrize("name,text,frames", [("ideal_gas_drop_2d", RAINDROP_LIKE, 1), ("log_strain_ball_3d", LOG_BALL, 3)])
def test_reference_style_scene_files(torch_cuda, oracle_mod, name, text, frames):
    from arp.sim import FluidSim

    sc = K.loads(text)
    print("case scene", name, sc.integrator, sc.n, sc.nb, sc.ns, flush=True)
    assert sc.n > 50 and sc.nb > 50
    eqs = set(int(e) for e in sc.materials["state_eq"])
    if name.startswith("ideal"):
        assert wire.EQ_IDEAL_GAS in eqs and wire.EQ_TAIT in eqs
    else:
        assert 0 in set(int(o) for o in sc.materials["strain_order"][sc.materials["normal_stiffness"] != 0])
        assert sc.ns > 50
    sim = FluidSim(sc)
    dt = sc.dt / sc.subticks
    sim.update(frames)
    o = oracle_mod.Oracle(sc)
    ref = o.step(o.state("f"), T.tableau(sc.integrator), dt, frames * sc.subticks)
    _check("scene " + name, oracle_mod, sim.particles(), ref)
    sim.close()


def test_dam_break_trajectory(torch_cuda, oracle_mod):
    """The 2-D dam break (4,096 particles, RK4, dt 1e-4) over its 1000 steps: at every 100th
    step the oracle takes one step from the GPU's state, and the GPU's own next step matches
    it; the run stays finite."""
    from arp.sim import FluidSim

    sc = S.dam_break_2d()
    sim = FluidSim(sc, integrator="rk4", timestep=1e-4)
    o = oracle_mod.Oracle(sc)
    tab = T.tableau("rk4")
    done = 0
    for chk in range(10):
        sim.update(99)
        p = sim.particles()
        ref = o.step(o.state("f", particles=p), tab, 1e-4, 1)
        sim.update(1)
        done += 100
        got = sim.particles()
        assert np.isfinite(got["pos"]).all()
        _check("dam break step %d" % done, oracle_mod, got, ref)
    sim.close()
