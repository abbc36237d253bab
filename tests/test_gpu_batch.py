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
    assert div in (16, 32, 64, 128, 512, 2048)
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
    _, walk_buffered = B.tuning()
    print("tuned list-walk form: %s" % ("LDS-buffered" if walk_buffered else "registers"), flush=True)
    for k, (a, b) in enumerate(zip(sims, solo)):
        pa, pb = a.particles(), b.particles()
        if not walk_buffered:  # the solo runs walk in registers too: bitwise
            assert _same(pa, pb), "scene %d differs after tuning" % k
        else:  # the other walk form sums the same pairs in the same order, compiled apart
            for f in ("den", "pos", "vel"):
                d = np.abs(pa[f].astype(np.float64) - pb[f]).max() / max(np.abs(pb[f]).max(), 1e-300)
                assert d <= 1e-6, (k, f, d)
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


def test_mixed_fluid_solid_batch_vs_oracle(torch_cuda, oracle_mod):
    """A batch of fluid scenes and scenes with elastic (one plastic) solids and fluid-solid
    contact: every kernel of a stage runs once for all of them (the fluid scenes take no
    part in the solid path's launches).  Each scene is bitwise the same scene stepped alone
    and within 1e-5 of the FP32 oracle per field, plastic strain included."""
    from arp import tableau as T
    from arp.sim import Batch
    from helpers import deformed_solid_scene

    dt, nsteps = 1e-5, 3
    scenes = [S.fluid_block_3d(700, alpha=0.02, seed=11, name="f0"),
              deformed_solid_scene(dim=3, seed=4),
              S.fluid_block_3d(300, alpha=0.06, seed=12, name="f1"),
              deformed_solid_scene(dim=3, seed=5, plastic=True),
              S.contact_3d(n_fluid_side=8, n_sphere=100)]
    print("case mixed batch", [(sc.n, sc.ns) for sc in scenes], flush=True)
    solo = _sims(scenes, "rk2", dt)
    for s in solo:
        # a batch with solid scenes searches at every stage (its fluid scenes keep no lists):
        # so do the solo runs it is compared with bitwise
        s.set_neighbor_skin(0.0)
        s.step(dt, nsteps)
    sims = _sims(scenes, "rk2", dt)
    B = Batch(sims, graph=True)
    B.step(dt, nsteps)
    B.synchronize()
    tab = T.tableau("rk2")
    for k, (a, b, sc) in enumerate(zip(sims, solo, scenes)):
        pa = a.particles()
        assert _same(pa, b.particles()), "scene %d differs from its solo run" % k
        o = oracle_mod.Oracle(sc)
        ref = o.step(o.state("f"), tab, dt, nsteps)
        for f in ("den", "pos", "vel"):
            e = oracle_mod.rel_err(pa[f], ref[f])
            print("  scene %d %s %.2e" % (k, f, e), flush=True)
            assert e <= 1e-5, (k, f, e)
        if sc.ns > 1:
            so = a.solids()
            assert np.array_equal(np.ascontiguousarray(so).view(np.uint8), np.ascontiguousarray(b.solids()).view(np.uint8))
            e = oracle_mod.rel_err(so["stress"].reshape(-1, 16), ref["stress"])
            assert e <= 1e-5 or not np.abs(ref["stress"]).max() > 0, e
    B.close()
    for s in sims + solo:
        s.close()


def test_fluid_batch_vs_oracle(torch_cuda, oracle_mod):
    """Every scene of a fluid batch (lists kept per scene, graph replay) within 1e-4 of the
    oracle after 12 RK2 steps (the kept lists change the order of the per-particle sums, as
    in tests/test_gpu_nl.py)."""
    from arp import tableau as T
    from arp.sim import Batch

    dt, nsteps = 5e-5, 12
    scenes = _scenes()
    sims = _sims(scenes, "rk2", dt)
    B = Batch(sims, graph=True)
    B.step(dt, nsteps)
    B.synchronize()
    tab = T.tableau("rk2")
    for k, (a, sc) in enumerate(zip(sims, scenes)):
        o = oracle_mod.Oracle(sc)
        ref = o.step(o.state("f"), tab, dt, nsteps)
        pa = a.particles()
        for f in ("den", "pos", "vel"):
            e = oracle_mod.rel_err(pa[f], ref[f])
            assert e <= 1e-4, (k, f, e)
    B.close()
    for s in sims:
        s.close()


def test_batched_adjoint_kinetic_energy(torch_cuda, oracle_mod):
    """Batched reverse mode (arp_batch_adjoint_*) of a per-scene kinetic-energy loss
    L_s = sum_i m_i |v_i|^2 / 2 at the end, for fluid scenes and a solid scene: each scene's
    gradient is bitwise its solo sweep's, and matches the oracle's exact FP64 derivative along
    random position and density directions (within 1e-3)."""
    import torch

    from arp import tableau as T
    from arp.sim import Batch, FluidSim, to_device
    from helpers import deformed_solid_scene

    dt, nsteps, K = 1e-5, 4, 2
    scenes = [S.fluid_block_3d(500, alpha=0.05, seed=21, name="f0"), deformed_solid_scene(dim=3, seed=6),
              S.fluid_block_3d(900, alpha=0.02, seed=22, name="f1")]
    for sc in scenes:
        rng = np.random.default_rng(sc.n)
        sc.particles["vel"][:, :3] += (0.05 * rng.standard_normal((sc.n, 3))).astype(np.float32)
    sims = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    B = Batch(sims, graph=False)
    B.adjoint_record(dt, nsteps, K)

    def ke_cotangent(sim, sc):
        p = sim.particles()
        m = sc.materials["mass"][p["mat"]].astype(np.float64)
        g = np.zeros((sc.n, 7))
        g[:, 4:7] = m[:, None] * p["vel"][:, :3]
        return g

    gs = [ke_cotangent(s, sc) for s, sc in zip(sims, scenes)]
    from test_gpu_adjoint import _records

    gout = [_records(torch, g.astype(np.float32)) for g in gs]
    gin = [torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda") for sc in scenes]
    gps = B.adjoint_backward(gout, gin)
    B.synchronize()
    tab = T.tableau("rk2")
    rng = np.random.default_rng(5)
    for k, (sc, g, gi_t, gp) in enumerate(zip(scenes, gs, gin, gps)):
        gi = gi_t.cpu().numpy()[:, [0, 4, 5, 6, 8, 9, 10]].astype(np.float64)
        # the same scene alone
        solo = FluidSim(sc, integrator="rk2", timestep=dt)
        solo.adjoint_record(dt, nsteps, K)
        gin1 = torch.zeros_like(gi_t)
        gp1 = solo.adjoint_backward(_records(torch, g.astype(np.float32)), gin1)
        solo.synchronize()
        assert torch.equal(gin1, gi_t) and np.array_equal(gp1, gp), k
        solo.close()
        o = oracle_mod.Oracle(sc)
        st = o.state("d")
        for rep in range(4):
            u = np.zeros((sc.n, 7))
            if rep % 2:
                u[:, 0] = rng.standard_normal(sc.n)
            else:
                u[:, 1:4] = 1e-4 * rng.standard_normal((sc.n, 3))
            t = {kk: np.zeros_like(v) for kk, v in st.items()}
            t["den"][:] = u[:, 0]
            t["pos"][:, :3] = u[:, 1:4]
            _, yd = o.jvp(st, t, tab, dt, nsteps)
            exact = float(np.sum(g[:, 4:7] * yd["vel"][:, :3]))  # <g, J u> (g on velocities, u not)
            ad = float(np.sum(gi * u))
            rel = abs(ad - exact) / max(abs(exact), 1e-300)
            print("batch adjoint scene %d dir %d ad %.9e exact %.9e rel %.2e" % (k, rep, ad, exact, rel), flush=True)
            assert exact != 0.0 and rel < 1e-3
    B.close()
    for s in sims:
        s.close()
