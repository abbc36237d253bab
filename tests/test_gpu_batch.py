"""Batched ensembles (arp_batch_*) against the same scenes stepped one by one (needs a GPU).

A batch launch runs the same kernel bodies on every scene's own buffers, so the results
must be bitwise identical to stepping each scene alone, with and without graph replay,
for an odd and an even number of steps, and after the autotuner's trials."""
import numpy as np
import pytest

from arp import scene as S

pytestmark = pytest.mark.gpu

SIZES = [(700, 0.02), (1500, 0.05), (300, 0.08), (2200, 0.03), (64, 0.01)]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _scenes():
    return [S.fluid_block_3d(n, alpha=a, seed=3 + k, name="b%d" % k) for k, (n, a) in enumerate(SIZES)]


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(a[f]).view(np.uint32), np.ascontiguousarray(b[f]).view(np.uint32))
               for f in ("den", "mat", "solid_id", "pos", "vel"))


def _sims(scenes, integ, dt):
    from arp.sim import FluidSim

    return [FluidSim(sc, integrator=integ, timestep=dt) for sc in scenes]


@pytest.mark.parametrize("integ,graph,nsteps", [("rk2", True, 7), ("rk2", False, 6), ("rk4", True, 4), ("euler", True, 5)])
def test_batch_matches_individual(torch_cuda, integ, graph, nsteps):
    from arp.sim import Batch

    dt = 5e-5
    scenes = _scenes()
    print("case batch", integ, graph, nsteps, [sc.n for sc in scenes], flush=True)
    solo = _sims(scenes, integ, dt)
    for s in solo:
        s.step(dt, nsteps)
    sims = _sims(scenes, integ, dt)
    B = Batch(sims, graph=graph)
    B.step(dt, nsteps)
    B.synchronize()
    for k, (a, b) in enumerate(zip(sims, solo)):
        assert _same(a.particles(), b.particles()), "scene %d differs" % k
        assert abs(a.time - b.time) < 1e-12
    # the contexts keep working on their own after a batch call, and batches resume
    for a, b in zip(sims, solo):
        a.step(dt, 1)
        b.step(dt, 1)
    B.step(dt, 3)
    for s in solo:
        s.step(dt, 3)
    for k, (a, b) in enumerate(zip(sims, solo)):
        assert _same(a.particles(), b.particles()), "scene %d differs after resuming" % k
    B.close()
    for s in sims + solo:
        s.close()


def _two_fluids(sc):
    """The scene with every other fluid particle of a second, lighter fluid material."""
    import dataclasses

    mats = np.concatenate([sc.materials, sc.materials[1:2]])
    mats["target_den"][2] = mats["start_den"][2] = np.float32(800.0)
    mats["mass"][2] = np.float32(mats["mass"][1] * 0.8)
    p = sc.particles.copy()
    p["mat"][1::2] = 2
    return dataclasses.replace(sc, particles=p, materials=mats, interactions=S.resolve_interactions(mats, sc.h))


def _swapped(sc):
    """The scene with its two materials in the other order (fluid 0, boundary 1)."""
    import dataclasses

    mats = sc.materials[::-1].copy()
    p, b = sc.particles.copy(), sc.boundary.copy()
    p["mat"] = 1 - p["mat"]
    b["mat"] = 1 - b["mat"]
    return dataclasses.replace(sc, particles=p, boundary=b, materials=mats,
                               interactions=S.resolve_interactions(mats, sc.h))


def test_batch_of_mixed_material_forms(torch_cuda):
    """A batch takes the fast-material walk only when every scene has one mobile and one
    boundary material (each scene's own material indices); a scene with two fluids puts the
    whole batch on the general walk.  Every batch equals the scenes stepped alone (which take
    their own forms)."""
    from arp.sim import Batch

    dt = 5e-5
    base = _scenes()
    for scenes in (base, base[:1] + [_swapped(base[1])] + base[2:], base[:2] + [_two_fluids(base[2])] + base[3:]):
        print("case batch material forms", [len(sc.materials) for sc in scenes], flush=True)
        solo = _sims(scenes, "rk2", dt)
        for s in solo:
            s.step(dt, 6)
        sims = _sims(scenes, "rk2", dt)
        B = Batch(sims, graph=True)
        B.step(dt, 6)
        B.synchronize()
        for k, (a, b) in enumerate(zip(sims, solo)):
            assert _same(a.particles(), b.particles()), "scene %d differs" % k
        B.close()
        for s in sims + solo:
            s.close()


def test_batch_tune_leaves_states(torch_cuda):
    from arp.sim import Batch

    dt = 5e-5
    scenes = _scenes()
    sims = _sims(scenes, "rk2", dt)
    solo = _sims(scenes, "rk2", dt)
    for s in sims + solo:
        s.step(dt, 2)
    B = Batch(sims, graph=True)
    before = [s.particles().copy() for s in sims]
    t0 = [s.time for s in sims]
    div = B.tune(dt, 2)
    print("tuned row-sort workgroup size", div, flush=True)
    assert div in (16, 32, 64, 128, 512)
    for k, s in enumerate(sims):
        assert _same(s.particles(), before[k])
        assert s.time == t0[k]
    # the trials' restores drop the kept neighbour lists (the next stage searches again):
    # drop the solo scenes' lists too, so both continue from the same structures
    for s in solo:
        s.set_neighbor_skin(0.2)
    B.step(dt, 5)
    for s in solo:
        s.step(dt, 5)
    for k, (a, b) in enumerate(zip(sims, solo)):
        assert _same(a.particles(), b.particles()), "scene %d differs after tuning" % k
    B.close()
    for s in sims + solo:
        s.close()


def test_batch_rejects_mixed_integrators(torch_cuda):
    from arp._lib import ArpError
    from arp.sim import Batch

    scenes = _scenes()[:2]
    sims = [_sims([scenes[0]], "rk2", 5e-5)[0], _sims([scenes[1]], "rk4", 5e-5)[0]]
    with pytest.raises(ArpError):
        Batch(sims)
    for s in sims:
        s.close()
