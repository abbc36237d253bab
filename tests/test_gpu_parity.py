"""HIP path vs the oracle, through the C ABI (needs a GPU).

Tolerances: neighbour structure bit-exact; per-field normwise relative error
max|gpu - oracle| / max|oracle| <= 1e-5 for one stage derivative and one step (FP32 vs
the FP32 oracle); the FP64 oracle error is reported alongside."""
import numpy as np
import pytest

from arp import scene as S
from arp import tableau as T
from arp import wire
from helpers import box_scene, deformed_solid_scene as _solid_scene, random_grid_scene, self_contact_scene, split_cand, \
    uniform_block

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _sim(sc, **kw):
    from arp.sim import FluidSim

    return FluidSim(sc, **kw)


def _state_of(parts):
    return dict(den=parts["den"], pos=parts["pos"], vel=parts["vel"])


def _report(tag, gpu, o32, o64):
    errs = {}
    for k in ("den", "pos", "vel"):
        e32 = float(np.max(np.abs(gpu[k].astype(np.float64) - o32[k]))) / max(float(np.max(np.abs(o32[k]))), 1e-300)
        e64 = float(np.max(np.abs(gpu[k].astype(np.float64) - o64[k]))) / max(float(np.max(np.abs(o64[k]))), 1e-300)
        errs[k] = (e32, e64)
    print(tag, {k: "%.2e/%.2e" % v for k, v in errs.items()}, flush=True)
    return errs


SCENES = {
    "dam_break_2d": lambda: S.dam_break_2d(),
    "column_3d": lambda: S.small_fluid_3d(n_side=14, box=(0.14, 0.1, 0.14)),
    "grid_edge_2d": lambda: random_grid_scene(0, 2),
    "grid_edge_3d": lambda: random_grid_scene(1, 3),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_neighbour_structure_bit_exact(torch_cuda, oracle_mod, name):
    from arp.sim import to_device

    sc = SCENES[name]()
    print("case", name, sc.n, sc.nb, flush=True)
    sim = _sim(sc)
    d = to_device(sc.particles)
    sim.neighbors_update(d)
    g = sim.neighbor_dump()
    o = oracle_mod.Oracle(sc).grid_dump(with_counts=True)
    assert np.array_equal(g["cell"], o["cell"])
    assert np.array_equal(g["cand_count"], o["cand_count"])
    assert np.array_equal(g["cand"], o["cand"])
    cc = sim.cell_counts()
    assert np.array_equal(cc, o["counts"])
    info = sim.grid_info()
    valid = int((o["cell"] != wire.INVALID).sum())
    assert int(info["n_valid"][2]) == valid
    sim.close()


@pytest.mark.parametrize("name", ["dam_break_2d", "column_3d", "grid_edge_3d"])
def test_compute_forces_parity(torch_cuda, oracle_mod, name):
    from arp.sim import from_device, to_device

    sc = SCENES[name]()
    if name.startswith("grid_edge"):
        # finite positions only: the reference reads buckets[INVALID] for dropped particles
        ok = np.all(np.isfinite(sc.particles["pos"]), 1)
        sc = box_scene(sc.particles["pos"][ok, :3], bpos=sc.boundary["pos"][:, :3], bounds=((0, 0, 0), (1, 1, 1)), h=sc.h)
    rng = np.random.default_rng(5)
    sc.particles["vel"][:, : sc.dim] = rng.normal(0, 0.05, (sc.n, sc.dim))
    sc.particles["den"] *= (1 + rng.uniform(-2e-3, 2e-3, sc.n)).astype(np.float32)
    print("case", name, sc.n, flush=True)
    sim = _sim(sc)
    o = oracle_mod.Oracle(sc)
    d32 = o.compute_forces(o.state("f"))
    d64 = o.compute_forces(o.state("d"))
    din = to_device(sc.particles)
    dout = to_device(wire.particles(sc.n))
    sim.compute_forces(din, None, dout, None)
    torch_cuda.cuda.synchronize()
    got = from_device(dout, wire.PARTICLE)
    errs = _report("forces " + name, _state_of(got), d32, d64)
    for k, (e32, _) in errs.items():
        assert e32 <= TOL, (k, e32)
    assert np.array_equal(got["mat"], sc.particles["mat"])
    assert np.array_equal(got["solid_id"], sc.particles["solid_id"])
    sim.close()


@pytest.mark.parametrize("integ", ["rk4", "rk2", "heun2", "rk_3_8", "euler", "verlet"])
def test_fused_step_parity(torch_cuda, oracle_mod, integ):
    sc = S.dam_break_2d()
    rng = np.random.default_rng(1)
    sc.particles["vel"][:, :2] = rng.normal(0, 0.1, (sc.n, 2))
    print("case step", integ, flush=True)
    sim = _sim(sc, integrator=integ, timestep=1e-4)
    o = oracle_mod.Oracle(sc)
    tab = T.tableau(integ)
    y32 = o.step(o.state("f"), tab, 1e-4, 1)
    y64 = o.step(o.state("d"), tab, 1e-4, 1)
    sim.update(1)
    got = sim.particles()
    errs = _report("step " + integ, _state_of(got), y32, y64)
    for k, (e32, _) in errs.items():
        assert e32 <= TOL, (k, e32)
    # the derivative part of the step, relative to the step's own change
    dv = got["vel"].astype(np.float64) - sc.particles["vel"]
    dv_ref = y64["vel"] - sc.particles["vel"]
    assert np.max(np.abs(dv - dv_ref)) <= 1e-4 * np.max(np.abs(dv_ref))
    assert sim.time == pytest.approx(1e-4)
    sim.close()


def test_multi_step_3d_close_to_oracle(torch_cuda, oracle_mod):
    sc = S.small_fluid_3d(n_side=12, box=(0.12, 0.1, 0.12))
    print("case multistep", sc.n, sc.nb, flush=True)
    sim = _sim(sc, integrator="rk2", timestep=5e-5)
    o = oracle_mod.Oracle(sc)
    y32 = o.step(o.state("f"), T.tableau("rk2"), 5e-5, 20)
    y64 = o.step(o.state("d"), T.tableau("rk2"), 5e-5, 20)
    sim.update(20)
    got = sim.particles()
    errs = _report("20 steps rk2", _state_of(got), y32, y64)
    for k, (e32, _) in errs.items():
        assert e32 <= 1e-4, (k, e32)


def test_verlet_multi_step_and_resume(torch_cuda, oracle_mod):
    """Velocity Verlet keeps the last force evaluation between steps: 20 steps in one call,
    and 7 + 13 in two, match the oracle's scheme; a state set from the host evaluates the
    forces afresh."""
    sc = S.small_fluid_3d(n_side=12, box=(0.12, 0.1, 0.12))
    print("case verlet multistep", sc.n, sc.nb, flush=True)
    o = oracle_mod.Oracle(sc)
    y32 = o.step(o.state("f"), T.tableau("verlet"), 5e-5, 20)
    y64 = o.step(o.state("d"), T.tableau("verlet"), 5e-5, 20)
    a = _sim(sc, integrator="verlet", timestep=5e-5)
    a.update(20)
    b = _sim(sc, integrator="verlet", timestep=5e-5)
    b.update(7)
    b.update(13)
    ga, gb = a.particles(), b.particles()
    for k in ("den", "pos", "vel"):
        assert np.array_equal(ga[k], gb[k]), k
    errs = _report("20 steps verlet", _state_of(ga), y32, y64)
    for k, (e32, _) in errs.items():
        assert e32 <= 1e-4, (k, e32)
    # resetting the state drops the kept evaluation: the oracle restarted at step 7
    c = _sim(sc, integrator="verlet", timestep=5e-5)
    c.update(7)
    mid = c.particles()
    c.upload_state(mid)
    c.update(13)
    o2 = oracle_mod.Oracle(sc)
    st = o2.state("f")
    st7 = o2.step(st, T.tableau("verlet"), 5e-5, 7)
    ref = o2.step(st7, T.tableau("verlet"), 5e-5, 13)
    gc = c.particles()
    for k in ("den", "pos", "vel"):
        e = oracle_mod.rel_err(gc[k], ref[k])
        print("  restart %s %.2e" % (k, e), flush=True)
        assert e <= 1e-4, (k, e)
    for s in (a, b, c):
        s.close()


def test_lincomb_and_velocity(torch_cuda, oracle_mod):
    from arp.sim import from_device, to_device

    sc = S.dam_break_2d()
    rng = np.random.default_rng(3)
    sim = _sim(sc)
    o = oracle_mod.Oracle(sc)
    terms_h, terms_d = [], []
    for _ in range(7):
        p = sc.particles.copy()
        p["den"] = rng.normal(size=sc.n)
        p["pos"][:, :2] = rng.normal(size=(sc.n, 2))
        p["vel"][:, :2] = rng.normal(size=(sc.n, 2))
        terms_h.append(p)
        terms_d.append(to_device(p))
    co = rng.normal(size=7).astype(np.float32)
    out = to_device(wire.particles(sc.n))
    sim.lincomb(co, terms_d, out)
    torch_cuda.cuda.synchronize()
    got = from_device(out, wire.PARTICLE)
    ref = o.lincomb(co, [o.state("f", particles=p) for p in terms_h])
    for k in ("den", "pos", "vel"):
        # the GPU fuses multiply-adds: a few ulp from the unfused oracle
        assert oracle_mod.rel_err(got[k], ref[k]) <= 1e-6, k
    vout = to_device(wire.particles(sc.n))
    sim.velocity(0.5, terms_d[0], vout)
    torch_cuda.cuda.synchronize()
    gv = from_device(vout, wire.PARTICLE)
    rv = o.velocity(0.5, o.state("f", particles=terms_h[0]))
    for k in ("den", "pos", "vel"):
        assert np.array_equal(gv[k], rv[k]), k
    assert np.array_equal(gv["mat"], terms_h[0]["mat"])


def test_deterministic_runs(torch_cuda):
    sc = S.small_fluid_3d(n_side=16, box=(0.16, 0.12, 0.16))
    a = _sim(sc, integrator="rk4", timestep=5e-5)
    a.update(10)
    pa = a.particles()
    b = _sim(sc, integrator="rk4", timestep=5e-5)
    b.update(10)
    pb = b.particles()
    assert pa.tobytes() == pb.tobytes()


def test_full_size_column_properties(torch_cuda):
    """1M-particle column at its real shape: conservation and grid bookkeeping."""
    sc = S.fluid_column_3d()
    print("case full", sc.n, sc.nb, flush=True)
    sim = _sim(sc, integrator="rk2", timestep=5e-5)
    p0 = sim.particles()
    assert np.array_equal(p0["pos"], sc.particles["pos"])
    sim.update(5)
    p = sim.particles()
    assert np.all(np.isfinite(p["pos"])) and np.all(np.isfinite(p["vel"]))
    assert np.array_equal(p["mat"], sc.particles["mat"])
    # particles fall under gravity: mean vertical velocity ~ -g t (pressure is weak)
    t = 5 * 5e-5
    vy = float(p["vel"][:, 1].mean())
    assert -9.81 * t * 1.5 < vy < 0
    from arp.sim import to_device

    sim.neighbors_update(to_device(p))
    info = sim.grid_info()
    assert int(info["n_valid"][2]) + int(info["n_overflow"][2]) + int(info["n_nonfinite"][2]) == sc.n


# ---------------- solid path (strain.rs, solid.rs) ------------------------------------
@pytest.mark.parametrize("case", ["3d", "2d", "3d_plastic", "3d_solid_only"])
def test_solid_compute_forces_parity(torch_cuda, oracle_mod, case):
    from arp.sim import from_device, to_device

    dim = 2 if case == "2d" else 3
    sc = _solid_scene(dim=dim, plastic="plastic" in case, with_fluid="only" not in case)
    assert sc.ns > 1
    print("case solid forces", case, sc.n, sc.ns, sc.nb, flush=True)
    sim = _sim(sc)
    o = oracle_mod.Oracle(sc)
    d32 = o.compute_forces(o.state("f"))
    d64 = o.compute_forces(o.state("d"))
    din, sin = to_device(sc.particles), to_device(sc.solids)
    dout, sout = to_device(wire.particles(sc.n)), to_device(wire.solids(sc.ns))
    sim.compute_forces(din, sin, dout, sout)
    torch_cuda.cuda.synchronize()
    got = from_device(dout, wire.PARTICLE)
    gs = from_device(sout, wire.SOLID)
    errs = _report("solid forces " + case, _state_of(got), d32, d64)
    for k, (e32, _) in errs.items():
        assert e32 <= TOL, (k, e32)
    e = oracle_mod.rel_err(gs["stress"].reshape(-1, 16), d32["stress"])
    print("plastic rate rel err %.2e (max %.3e)" % (e, float(np.abs(d32["stress"]).max())), flush=True)
    assert e <= TOL
    assert np.array_equal(gs["part_id"], sc.solids["part_id"])
    assert np.all(gs["ref_pos"] == 0)


@pytest.mark.parametrize("integ", ["rk4", "rk2", "verlet"])
def test_solid_step_parity(torch_cuda, oracle_mod, integ):
    sc = _solid_scene(dim=3, plastic=True)
    print("case solid step", integ, flush=True)
    sim = _sim(sc, integrator=integ, timestep=1e-5)
    o = oracle_mod.Oracle(sc)
    tab = T.tableau(integ)
    y32 = o.step(o.state("f"), tab, 1e-5, 2)
    y64 = o.step(o.state("d"), tab, 1e-5, 2)
    sim.update(2)
    got = sim.particles()
    errs = _report("solid step " + integ, _state_of(got), y32, y64)
    for k, (e32, _) in errs.items():
        assert e32 <= TOL, (k, e32)
    so = sim.solids()
    e = oracle_mod.rel_err(so["stress"].reshape(-1, 16), y32["stress"])
    print("plastic strain rel err %.2e" % e, flush=True)
    assert e <= TOL
    assert np.array_equal(so["ref_pos"], sc.solids["ref_pos"])


@pytest.mark.parametrize("case", ["inside", "outside"])
def test_self_contact_parity(torch_cuda, oracle_mod, case):
    """Self contact (solid.rs:426-450) of two blocks of one material, far apart in the
    reference state and touching now: one force evaluation and two RK2 steps against the
    oracle.  "outside": the grid ends inside one block, whose outer particles sit in clamped
    cells of the stencil."""
    from arp.sim import from_device, to_device

    sc = self_contact_scene(outside=case == "outside")
    print("case self contact", case, sc.n, sc.ns, flush=True)
    o = oracle_mod.Oracle(sc)
    d32 = o.compute_forces(o.state("f"))
    d64 = o.compute_forces(o.state("d"))
    # the contact is part of the forces checked: without it block A's forces change
    st = o.state("d")
    st["pos"][125:, 0] += 0.3
    far = o.compute_forces(st)
    effect = float(np.abs(far["vel"][:125] - d64["vel"][:125]).max())
    assert effect > 0.5, effect  # m/s^2 (gravity's term is 2g here)
    sim = _sim(sc)
    din, sin = to_device(sc.particles), to_device(sc.solids)
    dout, sout = to_device(wire.particles(sc.n)), to_device(wire.solids(sc.ns))
    sim.compute_forces(din, sin, dout, sout)
    torch_cuda.cuda.synchronize()
    errs = _report("self contact forces " + case, _state_of(from_device(dout, wire.PARTICLE)), d32, d64)
    for k, (e32, _) in errs.items():
        assert e32 <= TOL, (k, e32)
    sim.close()
    sim = _sim(sc, integrator="rk2", timestep=1e-5)
    tab = T.tableau("rk2")
    y32 = o.step(o.state("f"), tab, 1e-5, 2)
    y64 = o.step(o.state("d"), tab, 1e-5, 2)
    sim.update(2)
    errs = _report("self contact step " + case, _state_of(sim.particles()), y32, y64)
    for k, (e32, _) in errs.items():
        assert e32 <= TOL, (k, e32)
    sim.close()


@pytest.mark.parametrize("nterms", [3, 9])
def test_lincomb_and_velocity_with_solids(torch_cuda, oracle_mod, nterms):
    """s_lin_comb_N (particles.rs:309-313) and vel_solid (particles.rs:206-250) on a scene with
    solids: plastic strain and reference positions combine like den/pos/vel; 9 terms take the
    chunk-of-5 path.  Solid outputs are left alone when no solid terms are given."""
    from arp.sim import from_device, to_device

    sc = _solid_scene(dim=3, plastic=True)
    rng = np.random.default_rng(11)
    sim = _sim(sc)
    o = oracle_mod.Oracle(sc)
    terms_h, sterms_h, terms_d, sterms_d = [], [], [], []
    for _ in range(nterms):
        p = sc.particles.copy()
        p["den"] = rng.normal(size=sc.n)
        p["pos"][:, :3] = rng.normal(size=(sc.n, 3))
        p["vel"][:, :3] = rng.normal(size=(sc.n, 3))
        s = sc.solids.copy()
        s["ref_pos"][:, :3] = rng.normal(size=(sc.ns, 3))
        s["stress"][:] = rng.normal(size=s["stress"].shape)
        terms_h.append(p)
        sterms_h.append(s)
        terms_d.append(to_device(p))
        sterms_d.append(to_device(s))
    co = rng.normal(size=nterms).astype(np.float32)
    out, sout = to_device(wire.particles(sc.n)), to_device(wire.solids(sc.ns))
    sim.lincomb(co, terms_d, out, solid_terms=sterms_d, solid_out=sout)
    torch_cuda.cuda.synchronize()
    got, gs = from_device(out, wire.PARTICLE), from_device(sout, wire.SOLID)
    ref = o.lincomb(co, [o.state("f", particles=p, solids=s) for p, s in zip(terms_h, sterms_h)])
    for k in ("den", "pos", "vel"):
        assert oracle_mod.rel_err(got[k], ref[k]) <= 1e-6, k
    assert oracle_mod.rel_err(gs["ref_pos"], ref["ref"]) <= 1e-6
    assert oracle_mod.rel_err(gs["stress"].reshape(-1, 16), ref["stress"]) <= 1e-6
    assert np.array_equal(gs["part_id"], sterms_h[0]["part_id"])
    # no solid terms: the solid output is not touched (and nothing reads a null term)
    sout2 = to_device(wire.solids(sc.ns))
    sim.lincomb(co, terms_d, out, solid_terms=None, solid_out=sout2)
    torch_cuda.cuda.synchronize()
    assert not sout2.any().item()
    # vel_solid
    vout, svout = to_device(wire.particles(sc.n)), to_device(wire.solids(sc.ns))
    sim.velocity(0.25, terms_d[0], vout, d_solid_in=sterms_d[0], d_solid_out=svout)
    torch_cuda.cuda.synchronize()
    gv, gsv = from_device(vout, wire.PARTICLE), from_device(svout, wire.SOLID)
    rv = o.velocity(0.25, o.state("f", particles=terms_h[0], solids=sterms_h[0]))
    for k in ("den", "pos", "vel"):
        assert np.array_equal(gv[k], rv[k]), k
    assert np.array_equal(gsv["ref_pos"], rv["ref"]) and np.array_equal(gsv["stress"].reshape(-1, 16), rv["stress"])
    assert np.array_equal(gsv["part_id"], sterms_h[0]["part_id"])
    sim.close()
