"""Neighbour lists kept across RK stages (arp_set_neighbor_skin) through the C ABI (needs a
GPU).

- A stage that builds lists searches with radius h + skin and evaluates only the pairs
  within h, in the order of the plain search: with a rebuild at every stage the state is
  bitwise the plain step's.
- Stages that reuse the lists must see the neighbours within h the reference sees: the
  state stays within the parity tolerance of the oracle, which re-sorts at every stage.
- A bucket that would exceed its 128 slots during reuse is caught on the device and the
  stage sorts again (the oracle refuses the same particles).
- Batched ensembles keep lists per scene and match the scenes stepped alone."""
import numpy as np
import pytest

from arp import scene as S
from arp import tableau as T
from helpers import box_scene, random_grid_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(a[f]).view(np.uint32), np.ascontiguousarray(b[f]).view(np.uint32))
               for f in ("den", "mat", "solid_id", "pos", "vel"))


def _rel(got, ref):
    return float(np.max(np.abs(got.astype(np.float64) - ref))) / max(float(np.max(np.abs(ref))), 1e-300)


def _grid_edge_3d():
    sc = random_grid_scene(1, 3)
    ok = np.all(np.isfinite(sc.particles["pos"]), 1)
    return box_scene(sc.particles["pos"][ok, :3], bpos=sc.boundary["pos"][:, :3], bounds=((0, 0, 0), (1.05, 1.05, 1.05)),
                     h=sc.h)


def _moving(sc, sigma, seed):
    rng = np.random.default_rng(seed)
    sc.particles["vel"][:, : sc.dim] = rng.normal(0, sigma, (sc.n, sc.dim)).astype(np.float32)
    return sc


ALWAYS = [
    ("dam_break_2d", "rk4", 1e-4, 3, lambda: _moving(S.dam_break_2d(), 0.1, 1)),
    ("column_3d", "rk2", 5e-5, 6, lambda: _moving(S.small_fluid_3d(n_side=14, box=(0.14, 0.1, 0.14)), 0.05, 2)),
    ("column_3d_euler", "euler", 5e-5, 5, lambda: _moving(S.small_fluid_3d(n_side=10, box=(0.1, 0.1, 0.1)), 0.05, 3)),
    ("grid_edge_3d", "rk2", 1e-4, 4, lambda: _moving(_grid_edge_3d(), 0.2, 4)),
    ("column_3d_verlet", "verlet", 5e-5, 6, lambda: _moving(S.small_fluid_3d(n_side=10, box=(0.1, 0.1, 0.1)), 0.05, 5)),
]


@pytest.mark.parametrize("name,integ,dt,nsteps,make", ALWAYS, ids=[a[0] for a in ALWAYS])
def test_build_every_stage_is_the_plain_step(torch_cuda, name, integ, dt, nsteps, make):
    from arp.sim import FluidSim

    sc = make()
    print("case nl always-build", name, integ, sc.n, sc.nb, flush=True)
    plain = FluidSim(sc, integrator=integ, timestep=dt)
    plain.set_neighbor_skin(0.0)
    nl = FluidSim(sc, integrator=integ, timestep=dt)
    nl.set_neighbor_skin(0.3, always_build=True)
    plain.step(dt, nsteps)
    nl.step(dt, nsteps)
    assert _same(plain.particles(), nl.particles())
    builds, reused = nl.neighbor_stats()
    assert (builds, reused) == (nsteps * max(int(T.tableau(integ)["stages"]), 1), 0)
    assert plain.neighbor_stats() == (0, 0)
    plain.close()
    nl.close()


@pytest.mark.parametrize("skin", [0.1, 0.3])
def test_reuse_stays_with_the_oracle(torch_cuda, oracle_mod, skin):
    from arp.sim import FluidSim

    sc = S.small_fluid_3d(n_side=12, box=(0.12, 0.1, 0.12))
    print("case nl reuse", skin, sc.n, sc.nb, flush=True)
    sim = FluidSim(sc, integrator="rk2", timestep=5e-5)
    sim.set_neighbor_skin(skin)
    o = oracle_mod.Oracle(sc)
    y32 = o.step(o.state("f"), T.tableau("rk2"), 5e-5, 20)
    sim.update(20)
    got = sim.particles()
    builds, reused = sim.neighbor_stats()
    print("  builds %d reused %d" % (builds, reused), flush=True)
    assert builds + reused == 40 and reused > 0
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], y32[k])
        print("  %s %.2e" % (k, e), flush=True)
        assert e <= 1e-4, (k, e)


def test_verlet_reuse_stays_with_the_oracle(torch_cuda, oracle_mod):
    """Velocity Verlet's force point moves ahead of the state: the displacement check
    follows the point the next step evaluates at."""
    from arp.sim import FluidSim

    sc = _moving(S.small_fluid_3d(n_side=12, box=(0.12, 0.1, 0.12)), 0.3, 6)
    print("case nl verlet", sc.n, sc.nb, flush=True)
    sim = FluidSim(sc, integrator="verlet", timestep=5e-5)
    sim.set_neighbor_skin(0.1)
    o = oracle_mod.Oracle(sc)
    y32 = o.step(o.state("f"), T.tableau("verlet"), 5e-5, 30)
    sim.update(30)
    got = sim.particles()
    builds, reused = sim.neighbor_stats()
    print("  builds %d reused %d" % (builds, reused), flush=True)
    assert builds + reused == 30 and reused > 0 and builds > 1
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], y32[k])
        print("  %s %.2e" % (k, e), flush=True)
        assert e <= 1e-4, (k, e)


def test_reuse_with_fast_particles(torch_cuda, oracle_mod):
    """Particles fast enough to leave their skin within a few stages: rebuilds are forced
    by the displacement check, and the result stays with the oracle."""
    from arp.sim import FluidSim

    sc = _moving(S.dam_break_2d(), 1.5, 7)
    print("case nl fast", sc.n, flush=True)
    sim = FluidSim(sc, integrator="rk4", timestep=1e-4)
    sim.set_neighbor_skin(0.1)
    o = oracle_mod.Oracle(sc)
    y32 = o.step(o.state("f"), T.tableau("rk4"), 1e-4, 6)
    sim.update(6)
    got = sim.particles()
    builds, reused = sim.neighbor_stats()
    print("  builds %d reused %d" % (builds, reused), flush=True)
    assert builds > 1 and builds + reused == 24
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], y32[k])
        print("  %s %.2e" % (k, e), flush=True)
        assert e <= 1e-4, (k, e)


def test_bucket_cap_during_reuse(torch_cuda, oracle_mod):
    """A stream of non-interacting particles (zero EOS, no viscosity, no gravity) runs into a
    bucket that already holds 124: once it would hold more than 128 the reference refuses the
    surplus, which changes the density rates of everyone nearby.  The lists must give way to
    a sort at that stage."""
    from arp.sim import FluidSim

    rng = np.random.default_rng(9)
    h = 0.1
    c = 1.05 / 10  # bucket size (10 buckets per axis)
    lo = 5 * c
    cluster = lo + rng.uniform(0.05, 0.95, (124, 3)) * c
    stream = np.stack([lo - 0.004 - 0.003 * np.arange(16), np.full(16, lo + 0.5 * c), np.full(16, lo + 0.5 * c)], 1)
    stream[:, 1:] += rng.uniform(-0.2, 0.2, (16, 2)) * c
    pos = np.concatenate([stream, cluster]).astype(np.float32)
    vel = np.zeros_like(pos)
    vel[:16, 0] = 2.0
    fluid = S.fluid_material(h, 0.0, density=1000.0)
    sc = box_scene(pos, vel=vel, bounds=((0, 0, 0), (1.05, 1.05, 1.05)), h=h, g=0.0, fluid=fluid)
    dt, nsteps = 1e-3, 30
    sim = FluidSim(sc, integrator="rk2", timestep=dt)
    sim.set_neighbor_skin(0.2, cap=256)  # every cluster particle lists the whole cluster
    o = oracle_mod.Oracle(sc)
    y32 = o.step(o.state("f"), T.tableau("rk2"), dt, nsteps)
    sim.update(nsteps)
    got = sim.particles()
    builds, reused = sim.neighbor_stats()
    print("case nl cap: builds %d reused %d" % (builds, reused), flush=True)
    assert reused > 0 and builds > 1
    # the cluster bucket ends over its cap: some particles are refused at the end
    sim2 = FluidSim(sc, integrator="rk2", timestep=dt)
    from arp.sim import to_device

    sim2.neighbors_update(to_device(got))
    info = sim2.grid_info()
    assert int(info["n_overflow"][2]) > 0
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], y32[k])
        print("  %s %.2e" % (k, e), flush=True)
        assert e <= 1e-5, (k, e)
    for s in (sim, sim2):
        s.close()


def test_batch_keeps_lists_per_scene(torch_cuda):
    from arp.sim import Batch, FluidSim

    dt = 5e-5
    scenes = [S.fluid_block_3d(n, alpha=a, seed=11 + k, name="nl%d" % k) for k, (n, a) in
              enumerate([(900, 0.02), (2000, 0.05), (128, 0.08)])]
    solo = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    sims = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    for s in solo:
        s.step(dt, 12)
    B = Batch(sims, graph=True)
    B.step(dt, 12)
    B.synchronize()
    for k, (a, b) in enumerate(zip(sims, solo)):
        assert _same(a.particles(), b.particles()), "scene %d differs" % k
        assert a.neighbor_stats() == b.neighbor_stats()
        assert a.neighbor_stats()[1] > 0
    B.close()
    for s in sims + solo:
        s.close()


def test_reading_the_state_between_steps_changes_nothing(torch_cuda):
    """A list-reusing step of several stages leaves y only in the sorted buffers; reading
    the state (which brings the landing buffers up to date), recording a checkpoint of it or
    switching to a one-stage method in between must not change the trajectory."""
    from arp.sim import FluidSim

    sc = _moving(S.dam_break_2d(), 1.0, 12)  # fast enough to rebuild between the reads
    a = FluidSim(sc, integrator="rk4", timestep=1e-4)
    b = FluidSim(sc, integrator="rk4", timestep=1e-4)
    for s in (a, b):
        s.set_neighbor_skin(0.1)
    a.update(12)
    for _ in range(4):
        b.update(3)
        b.particles()
    pa, pb = a.particles(), b.particles()
    assert _same(pa, pb)
    assert a.neighbor_stats() == b.neighbor_stats() and a.neighbor_stats()[0] > 1
    # a one-stage step after a lazy step reads the landing buffers
    a.set_integrator("euler", 1e-4)
    a.update(2)
    c = FluidSim(sc, integrator="rk4", timestep=1e-4)
    c.set_neighbor_skin(0.1)
    c.update(12)
    c.particles()
    c.set_integrator("euler", 1e-4)
    c.update(2)
    assert _same(a.particles(), c.particles())
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("integ,sigma", [("rk4", 1.0), ("rk2", 2.0), ("heun3", 0.1)])
def test_speculative_steps_match_the_rebuild_chain(torch_cuda, integ, sigma):
    """Steps issued speculatively (each stage as its walk alone) give bitwise the state and
    statistics of issuing the rebuild chain at every stage, also when lists turn out invalid
    in the middle of a step and the step is run again."""
    from arp.sim import FluidSim

    sc = _moving(S.dam_break_2d(), sigma, 21)
    print("case nl speculation", integ, sigma, flush=True)
    a = FluidSim(sc, integrator=integ, timestep=1e-4)
    b = FluidSim(sc, integrator=integ, timestep=1e-4)
    a.set_neighbor_skin(0.1)
    b.set_neighbor_skin(0.1, speculate=False)
    a.update(9)
    a.update(5)
    b.update(14)
    assert _same(a.particles(), b.particles())
    assert a.neighbor_stats() == b.neighbor_stats()
    builds, reused = a.neighbor_stats()
    print("  builds %d reused %d" % (builds, reused), flush=True)
    assert reused > 0
    assert a.time == pytest.approx(b.time)
    for s in (a, b):
        s.close()
