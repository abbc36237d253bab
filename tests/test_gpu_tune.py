"""The list autotuner (arp_tune_lists, arp_batch_tune_lists) through the C ABI (needs a GPU).

The tuner times a few steps for each (skin, short skin) pair of its table, restores the state
after every trial and keeps the fastest pair, so:

- the state and the time are those from before the call, bitwise;
- the pair is one of the table's, and the context then steps exactly as a context given that
  pair by hand (bitwise);
- a batch keeps the same pair on every scene, and its scenes step as the scenes alone with
  that pair;
- scenes with solids are refused."""
import numpy as np
import pytest

from arp import scene as S

pytestmark = pytest.mark.gpu

PAIRS = {(0.1, 0.0), (0.2, 0.0), (0.3, 0.0), (0.2, 0.04), (0.3, 0.03), (0.3, 0.05), (0.4, 0.04), (0.5, 0.05)}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(a[f]).view(np.uint32), np.ascontiguousarray(b[f]).view(np.uint32))
               for f in ("den", "mat", "solid_id", "pos", "vel"))


def _pair(skin, sskin):
    return (round(skin, 4), round(sskin, 4))


def _moving(sc, sigma, seed):
    rng = np.random.default_rng(seed)
    sc.particles["vel"][:, : sc.dim] = rng.normal(0, sigma, (sc.n, sc.dim)).astype(np.float32)
    return sc


def test_tune_lists_restores_and_sets_its_pair(torch_cuda):
    from arp.sim import FluidSim

    dt = 1e-4
    sc = _moving(S.dam_break_2d(), 0.5, 3)
    print("case tune_lists dam_break_2d", sc.n, flush=True)
    a = FluidSim(sc, integrator="rk2", timestep=dt)
    b = FluidSim(sc, integrator="rk2", timestep=dt)
    before = a.particles().copy()
    skin, sskin = a.tune_lists(dt, 12)
    print("  tuned skin %.2f short %.2f" % (skin, sskin), flush=True)
    assert _pair(skin, sskin) in PAIRS
    assert _same(a.particles(), before) and a.time == 0.0
    assert a.neighbor_stats() == (0, 0) and a.neighbor_refreshes() == 0
    b.set_neighbor_skin(skin)
    b.set_neighbor_short_skin(sskin)
    a.update(20)
    b.update(20)
    assert _same(a.particles(), b.particles())
    assert a.neighbor_stats() == b.neighbor_stats()
    assert a.neighbor_refreshes() == b.neighbor_refreshes()
    for s in (a, b):
        s.close()


def test_batch_tune_lists_equals_the_scenes_alone(torch_cuda):
    from arp.sim import Batch, FluidSim

    dt = 5e-5
    scenes = [_moving(S.fluid_block_3d(n, alpha=0.03, seed=5 + k, name="tune_%d" % k), 0.5, 60 + k)
              for k, n in enumerate((700, 1800, 256))]
    print("case batch tune_lists", [sc.n for sc in scenes], flush=True)
    sims = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    solo = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    for s in sims + solo:
        s.step(dt, 3)
    before = [s.particles().copy() for s in sims]
    t0 = [s.time for s in sims]
    B = Batch(sims, graph=True)
    skin, sskin = B.tune_lists(dt, 6)
    print("  tuned skin %.2f short %.2f" % (skin, sskin), flush=True)
    assert _pair(skin, sskin) in PAIRS
    for k, s in enumerate(sims):
        assert _same(s.particles(), before[k]) and s.time == t0[k]
    for s in solo:
        s.set_neighbor_skin(skin)
        s.set_neighbor_short_skin(sskin)
        s.step(dt, 9)
    B.step(dt, 9)
    B.synchronize()
    for k, (a, b) in enumerate(zip(sims, solo)):
        assert _same(a.particles(), b.particles()), "scene %d differs" % k
        assert a.neighbor_stats() == b.neighbor_stats()
    B.close()
    for s in sims + solo:
        s.close()


def test_tune_lists_refuses_solids(torch_cuda):
    from arp._lib import ArpError
    from arp.sim import FluidSim

    sim = FluidSim(S.contact_3d(n_fluid_side=12, n_sphere=512), integrator="rk2", timestep=2e-5)
    with pytest.raises(ArpError):
        sim.tune_lists(2e-5, 4)
    sim.close()
