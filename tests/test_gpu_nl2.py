"""Two-level neighbour lists (arp_set_neighbor_short_skin) through the C ABI (needs a GPU).

The list walks use a short list of the candidates within h (1 + short skin), refreshed from
the long list (no grid search) whenever a particle has moved half the short skin since the
last refresh; the long list is rebuilt by the grid search as with one level.  The
neighbours within h stay the reference's, so:

- the state stays within the parity tolerance of the oracle, which sorts at every stage,
  with refreshes and long rebuilds both happening during the run;
- speculative steps give bitwise the state and statistics of the rebuild chain;
- batches give bitwise the scenes stepped alone;
- recordings replay exactly, and speculative recordings match the chain."""
import numpy as np
import pytest

from arp import scene as S
from arp import tableau as T

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


def _moving(sc, sigma, seed):
    rng = np.random.default_rng(seed)
    sc.particles["vel"][:, : sc.dim] = rng.normal(0, sigma, (sc.n, sc.dim)).astype(np.float32)
    return sc


@pytest.mark.parametrize("integ,sigma,nsteps", [("rk2", 0.3, 30), ("rk4", 1.0, 16), ("verlet", 0.3, 30)])
def test_two_level_stays_with_the_oracle(torch_cuda, oracle_mod, integ, sigma, nsteps):
    from arp.sim import FluidSim

    sc = _moving(S.small_fluid_3d(n_side=12, box=(0.12, 0.1, 0.12)), sigma, 31)
    print("case two-level", integ, sigma, nsteps, sc.n, sc.nb, flush=True)
    sim = FluidSim(sc, integrator=integ, timestep=5e-5)
    sim.set_neighbor_skin(0.2)
    sim.set_neighbor_short_skin(0.04)
    o = oracle_mod.Oracle(sc)
    ref = o.step(o.state("f"), T.tableau(integ), 5e-5, nsteps)
    sim.update(nsteps)
    got = sim.particles()
    builds, reused = sim.neighbor_stats()
    refreshes = sim.neighbor_refreshes()
    print("  builds %d reused %d refreshes %d" % (builds, reused, refreshes), flush=True)
    assert refreshes > 0 and reused > refreshes
    if sigma >= 1.0:
        assert builds > 1
    for k in ("den", "pos", "vel"):
        e = _rel(got[k], ref[k])
        print("  %s %.2e" % (k, e), flush=True)
        assert e <= 1e-4, (k, e)
    sim.close()


@pytest.mark.parametrize("integ,sigma", [("rk4", 1.0), ("rk2", 2.0)])
def test_two_level_speculation_matches_the_chain(torch_cuda, integ, sigma):
    from arp.sim import FluidSim

    sc = _moving(S.dam_break_2d(), sigma, 21)
    print("case two-level speculation", integ, sigma, flush=True)
    a = FluidSim(sc, integrator=integ, timestep=1e-4)
    b = FluidSim(sc, integrator=integ, timestep=1e-4)
    a.set_neighbor_skin(0.2)
    b.set_neighbor_skin(0.2, speculate=False)
    for s in (a, b):
        s.set_neighbor_short_skin(0.05)
    a.update(11)
    a.update(9)
    b.update(20)
    assert _same(a.particles(), b.particles())
    assert a.neighbor_stats() == b.neighbor_stats()
    assert a.neighbor_refreshes() == b.neighbor_refreshes() > 0
    print("  lists %s refreshes %d speculation %s" % (a.neighbor_stats(), a.neighbor_refreshes(),
                                                     a.speculation_stats()), flush=True)
    for s in (a, b):
        s.close()


def test_two_level_batch_equals_the_scenes_alone(torch_cuda):
    from arp.sim import Batch, FluidSim

    dt = 5e-5
    scenes = [_moving(S.fluid_block_3d(n, alpha=a, seed=11 + k, name="nl2_%d" % k), 0.5, 40 + k)
              for k, (n, a) in enumerate([(900, 0.02), (2000, 0.05), (128, 0.08)])]
    solo = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    sims = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    for s in solo + sims:
        s.set_neighbor_short_skin(0.04)
    for s in solo:
        s.step(dt, 14)
    B = Batch(sims, graph=True)
    B.step(dt, 14)
    B.synchronize()
    for k, (a, b) in enumerate(zip(sims, solo)):
        assert _same(a.particles(), b.particles()), "scene %d differs" % k
        assert a.neighbor_stats() == b.neighbor_stats()
        assert a.neighbor_refreshes() == b.neighbor_refreshes()
    assert sum(s.neighbor_refreshes() for s in sims) > 0
    B.close()
    for s in sims + solo:
        s.close()


def test_two_level_recording_replays_and_matches_the_chain(torch_cuda):
    """A recording with two-level lists: the reverse sweep's replays repeat it exactly (two
    sweeps give the same gradients), and speculative recording gives the chain's state,
    statistics and gradients bitwise."""
    from arp.sim import FluidSim

    sc = _moving(S.small_fluid_3d(n_side=8, box=(0.06, 0.06, 0.06)), 1.0, 7)
    dt = 5e-5
    g = np.random.default_rng(5).standard_normal((sc.n, 7)).astype(np.float32)
    r = np.zeros((sc.n, 12), np.float32)
    r[:, [0, 4, 5, 6, 8, 9, 10]] = g
    d_out = torch_cuda.from_numpy(r).cuda()
    res = []
    for spec in (True, False):
        sim = FluidSim(sc, integrator="rk2", timestep=dt)
        sim.set_neighbor_skin(0.2, speculate=spec)
        sim.set_neighbor_short_skin(0.04)
        sim.upload_state(sc.particles)
        sim.adjoint_record(dt, 16, 8)
        end = sim.particles().copy()
        stats = (sim.neighbor_stats(), sim.neighbor_refreshes())
        grads = []
        for _ in range(2):
            d_in = torch_cuda.zeros((sc.n, 12), dtype=torch_cuda.float32, device="cuda")
            sim.adjoint_backward(d_out, d_in)
            sim.synchronize()
            grads.append(d_in.cpu().numpy().copy())
        res.append((end, stats, grads))
        sim.close()
    (ea, sa, ga), (eb, sb, gb) = res
    print("  lists %s refreshes %d" % sa, flush=True)
    assert sa[1] > 0
    assert _same(ea, eb) and sa == sb
    assert np.isfinite(ga[0]).all()
    assert np.array_equal(ga[0], ga[1]) and np.array_equal(gb[0], gb[1]) and np.array_equal(ga[0], gb[0])


def test_short_skin_arguments(torch_cuda):
    from arp._lib import ArpError
    from arp.sim import FluidSim

    sim = FluidSim(S.dam_break_2d(), integrator="rk2", timestep=1e-4)
    sim.set_neighbor_skin(0.2)
    for bad in (-0.1, 0.005, 0.2, 0.5):
        with pytest.raises(ArpError):
            sim.set_neighbor_short_skin(bad)
    sim.set_neighbor_short_skin(0.05)
    sim.set_neighbor_skin(0.04)  # the short skin no longer lies below: one level again
    sim.update(10)
    assert sim.neighbor_refreshes() == 0 and sim.neighbor_stats()[1] > 0
    sim.close()
