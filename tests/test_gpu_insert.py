"""Particle insertion (FluidSim::add_particles) on the GPU (needs a GPU).

- The state before the insertion carries over bitwise; the new particles are the spawned
  object's lattice, offset; the materials and interaction matrix grow as the oracle's
  restatement of Particles::add_particles says (an equal fluid material is reused, a stiff
  one is always new).
- Steps after the insertion match the oracle stepped from the merged scene (normwise
  relative error <= 1e-5 per field).
- Solids inserted into a scene without solids replace its placeholder solid and run the
  solid path.
"""
import numpy as np
import pytest

from arp import config as K
from arp import tableau as T
from arp import wire
from arp.scene import Scene
from oracle import scene_ref as R

pytestmark = pytest.mark.gpu
tomli = pytest.importorskip("tomli")

SCENE = """\
name = "Spawn"
kernel_radius = 0.02
time_step = 0.0001
gravity = 9.81
artificial_viscocity = 0.05
integrator = "RK2"
min = [-0.1, -0.1]
dim = [1.2, 1.0]

[boundaries.tank]
spacing = 0.01
region = {min = [-0.1, -0.1], dim = [1.2, 1.0], border = 0.03}

[objects.water]
density = 1000.0
speed_of_sound = 20.0
spacing = 0.01
region = {min = [0.0, 0.0], dim = [0.3, 0.2]}

[objects.drop]
on_click = "Relative"
density = 1000.0
speed_of_sound = 20.0
spacing = 0.01
region = {min = [0.0, 0.0], dim = [0.3, 0.2]}

[objects.gas]
on_click = "Absolute"
density = 1.2
speed_of_sound = 20.0
spacing = 0.01
velocity = [0.0, -0.5]
region = {center = [0.6, 0.6], radii = [0.05, 0.05]}

[objects.rubber]
on_click = "Absolute"
density = 1000.0
normal_stiffness = -30000.0
shear_stiffness = -30000.0
spacing = 0.01
region = {center = [0.8, 0.3], radii = [0.05, 0.05]}
"""


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _rel(got, ref):
    return float(np.max(np.abs(got.astype(np.float64) - ref))) / max(float(np.max(np.abs(ref))), 1e-300)


def _spawned(cfg, name, h, offset):
    o = [x for x in cfg["on_click"] if x["name"] == name][0]
    parts, mass, _ = R.gen_object(o, h, 0)
    new = wire.particles(len(parts["pos"]))
    for k in ("den", "pos", "vel"):
        new[k] = parts[k]
    if offset is not None:
        off = np.zeros(4, np.float32)
        off[: len(offset)] = offset
        new["pos"] = (new["pos"] + off).astype(np.float32)
    m = np.zeros((), wire.MATERIAL)
    for k, v in o["mat"].items():
        m[k] = v
    m["mass"] = mass
    return m, new


def test_insert_fluid_and_solid(torch_cuda, oracle_mod):
    from arp.sim import FluidSim

    f = K.SceneFile(SCENE)
    sc = f.generate()
    cfg = R.parse(tomli.loads(SCENE))
    names = [bytes(o["name"]).rstrip(b"\0").decode() for o in sc.objects]
    idx = {nm: i for i, nm in enumerate(names)}
    print("case insert", sc.n, sc.nb, names, flush=True)
    sim = FluidSim(sc)
    dt = sc.dt
    sim.step(dt, 3)
    before = sim.particles()
    P, S_, M, I = before, sim.solids(), sc.materials.copy(), sc.interactions.copy()
    expect_added = {"drop": False, "gas": True, "rubber": True}
    for name, off in (("drop", (0.4, 0.5)), ("gas", None), ("rubber", None)):
        mat, new = _spawned(cfg, name, np.float32(sc.h), off)
        P, S_, M, I, mid, added = R.add_particles(P, S_, M, I, mat, new, np.float32(sc.h))
        got_added = sim.add_object(f, idx[name], off)
        print("  spawned %s: %d particles, material %d, added %s" % (name, len(new), mid, got_added), flush=True)
        assert got_added == added == expect_added[name]
        assert sim.num_materials() == len(M)
        cur = sim.particles()
        assert len(cur) == len(P)
        for k in ("den", "mat", "solid_id", "pos", "vel"):
            assert np.array_equal(cur[k], P[k]), (name, k)
    so = sim.solids()
    assert len(so) == len(S_) and np.array_equal(so["part_id"], S_["part_id"])
    assert np.array_equal(so["part_id"], np.nonzero(P["solid_id"] != wire.INVALID)[0].astype(np.uint32))
    assert abs(sim.time - 3 * dt) < 1e-9
    # the merged scene stepped by the oracle
    merged = Scene(P.copy(), sc.boundary.copy(), M, I, sc.bounds.copy(), sc.h, sc.g, sc.alpha, solids=S_.copy())
    o = oracle_mod.Oracle(merged)
    st = o.state("f")
    ref = o.step(st, T.tableau("rk2"), dt, 2)
    sim.step(dt, 2)
    got = sim.particles()
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], ref[k])
        print("  after insertion %s %.2e" % (k, e), flush=True)
        assert e <= 1e-5, (k, e)
    sim.close()
    f.close()
