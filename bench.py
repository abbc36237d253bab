"""Benchmark: SPH particle-steps/s of the 3-D fluid column (1,048,576 particles, RK2,
dt 5e-5, 500 steps) on one MI355X, with the dominant kernel's HBM roofline and the CPU oracle
timed on the host cores beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

The timed region is steps W .. W+K of the configuration (after W warm-up steps from the
initial state).  Besides it, the line reports the whole 500-step configuration from the
initial state, uploaded again (its own timing, list rebuilds and reuses),
mass / momentum / energy drift over it, a per-field FP32 and FP64 error sample of the GPU
step against the oracle at its end, forward + adjoint with checkpoint replay, and the
batched ensemble.  Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ar-physics_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# algorithmic bytes per particle of one k_fluid launch: read the stage state (pos+den,
# vel+q: 32 B), read y_n (28 B), write the next stage state (28 B), write the cell key (4 B)
B_ALG_FLUID = 92.0
# the configuration: 500 RK2 steps of the column
CONFIG_STEPS = 500


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (committed
    summary under profiles/), or None."""
    p = os.path.join(ROOT, "profiles", "traffic_%s_%s.txt" % (kernel, workload))
    if not os.path.exists(p):
        return None
    vals = {}
    for line in open(p):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    return vals.get("hbm_bytes_per_launch")


def cpu_baseline(scene, integ, dt, steps, threads):
    from arp import tableau
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    o = O.Oracle(scene, threads=threads)
    st = o.state("f")
    t0 = time.perf_counter()
    o.step(st, tableau.tableau(integ), dt, steps)
    sec = time.perf_counter() - t0
    return dict(value=scene.n * steps / sec, unit="particle-steps/s", cores=threads, kind="port",
                sample="%d %s step(s) of the full %d-particle column by the FP32 C oracle (OpenMP, %d threads); "
                       "the reference GLSL cannot run here (no GL context)" % (steps, integ.upper(), scene.n, threads))


def tait_energy(den, rho0, c):
    """Specific internal energy of the Tait EOS (fluid.rs:49), e(rho) = int_rho0^rho p/r^2 dr
    with p = (c^2 rho0 / 7)((r/rho0)^7 - 1)."""
    B = c * c * rho0 / 7.0
    return B * ((den**6 / rho0**7 - 1.0 / rho0) / 6.0 + 1.0 / den - 1.0 / rho0)


def drift(sc, p0, p1, t, g):
    """Mass, momentum and energy of the mobile particles at the start and end of a run.
    Momentum: change beyond the free-fall impulse -M g t (the boundary and, through the
    reference's m_i/rho_i force scaling, the pressure exchange momentum with the walls).
    Energy: kinetic + gravitational + Tait internal energy (artificial viscosity and the
    walls dissipate it)."""
    import numpy as np

    m = sc.materials["mass"][p0["mat"]].astype(np.float64)
    M = float(m.sum())
    mat = sc.materials[int(p0["mat"][0])]
    rho0, c = float(mat["start_den"]), float(mat["sound_speed"])

    def mom(p):
        return (m[:, None] * p["vel"][:, :3].astype(np.float64)).sum(0)

    def energy(p):
        v = p["vel"][:, :3].astype(np.float64)
        y = p["pos"][:, 1].astype(np.float64)
        d = p["den"].astype(np.float64)
        e = dict(kinetic=float((m * 0.5 * (v * v).sum(1)).sum()), gravitational=float((m * g * y).sum()),
                 internal=float((m * tait_energy(d, rho0, c)).sum()))
        e["total"] = e["kinetic"] + e["gravitational"] + e["internal"]
        return e

    P0, P1 = mom(p0), mom(p1)
    imp = np.array([0.0, -M * g * t, 0.0])
    e0, e1 = energy(p0), energy(p1)
    # potential energy released if the whole column fell by its own height: the energy scale
    Escale = M * g * float(p0["pos"][:, 1].max() - p0["pos"][:, 1].min())
    return dict(
        mass_rel=0.0,  # particle count and masses are fixed: the total mass is exact
        momentum_change_beyond_gravity_rel=float(np.linalg.norm(P1 - P0 - imp) / (M * g * t)),
        energy_J={"start": e0, "end": e1},
        energy_drift_rel=float((e1["total"] - e0["total"]) / Escale),
        density_rel_range_end=[float(p1["den"].min() / rho0 - 1.0), float(p1["den"].max() / rho0 - 1.0)],
        note="momentum change beyond the free-fall impulse, relative to M g t; energy drift "
             "(kinetic + gravitational + Tait internal) relative to M g (column height).  The "
             "reference scales pressure accelerations by m_i/rho_i, so pressure does almost no "
             "work while the continuity equation compresses the layers on the floor: the Tait "
             "energy grows without a matching loss elsewhere",
    )


def fp64_sample(sim, sc, p, integ, dt, threads):
    """One step from the downloaded state p: the GPU against the FP32 and the FP64 oracle,
    normwise relative error per field."""
    from arp import tableau
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    o = O.Oracle(sc, threads=threads)
    tab = tableau.tableau(integ)
    r32 = o.step(o.state("f", particles=p), tab, dt, 1)
    r64 = o.step(o.state("d", particles=p), tab, dt, 1)
    sim.upload_state(p)
    sim.step(dt, 1)
    got = sim.particles()
    out = {}
    for k in ("den", "pos", "vel"):
        out[k] = dict(vs_fp32=O.rel_err(got[k], r32[k]), vs_fp64=O.rel_err(got[k], r64[k]),
                      fp32_oracle_vs_fp64=O.rel_err(r32[k], r64[k]))
    return out


def config_run(sim, p_init, nsteps):
    """The configuration's nsteps steps from the initial state, timed as a whole."""
    sim.upload_state(p_init)
    sim.synchronize()
    nl0, sp0, rf0 = sim.neighbor_stats(), sim.speculation_stats(), sim.neighbor_refreshes()
    t0 = time.perf_counter()
    sim.update(nsteps)
    sim.synchronize()
    sec = time.perf_counter() - t0
    nl1, sp1, rf1 = sim.neighbor_stats(), sim.speculation_stats(), sim.neighbor_refreshes()
    n = len(p_init)
    return dict(steps=nsteps, seconds=sec, ms_per_step=1e3 * sec / nsteps, particle_steps_per_s=n * nsteps / sec,
                stage_builds=nl1[0] - nl0[0], stage_reuses=nl1[1] - nl0[1], short_list_refreshes=rf1 - rf0,
                speculative_steps=sp1[0] - sp0[0], speculation_aborts=sp1[1] - sp0[1])


def time_adjoint(sim, sc, p_start, dt, nsteps, K):
    """Forward with recording (a checkpoint every K steps) + reverse sweep (replaying every
    segment but the last from its checkpoint) of nsteps steps, against nsteps plain steps,
    all from the same state (uploaded before each timed call).  Median of three rounds after
    an untimed one."""
    import numpy as np
    import torch

    log("adjoint: %d steps, checkpoint every %d" % (nsteps, K))
    rng = np.random.default_rng(0)
    g = torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda")
    g[:, [0, 4, 5, 6, 8, 9, 10]] = torch.from_numpy(rng.standard_normal((sc.n, 7)).astype(np.float32)).cuda()
    gin = torch.zeros_like(g)
    torch.cuda.synchronize()

    def timed(fn):
        sim.upload_state(p_start)
        sim.synchronize()
        t0 = time.perf_counter()
        fn()
        sim.synchronize()
        return time.perf_counter() - t0

    fwd, rec, bwd = [], [], []
    for r in range(4):
        f = timed(lambda: sim.step(dt, nsteps))
        a = timed(lambda: sim.adjoint_record(dt, nsteps, K))
        t0 = time.perf_counter()
        sim.adjoint_backward(g, gin, want_params=True)
        sim.synchronize()
        b = time.perf_counter() - t0
        if r:  # the first round allocates the recording and the reverse-mode buffers
            fwd.append(f)
            rec.append(a)
            bwd.append(b)
    med = lambda x: sorted(x)[len(x) // 2]  # noqa: E731
    fwd_ms, rec_ms, bwd_ms = 1e3 * med(fwd), 1e3 * med(rec), 1e3 * med(bwd)
    # the reverse stage kernels, timed with events in another round
    adj_kernels = ["k_fluid_adj" + sfx for sfx in ("", "_rec", "_f", "_rec_f", "_g", "_rec_g")]
    sim.upload_state(p_start)
    sim.adjoint_record(dt, nsteps, K)
    sim.profile_select(adj_kernels)
    sim.profile(True)
    sim.profile_reset()
    sim.adjoint_backward(g, gin, want_params=True)
    sim.synchronize()
    ms, cnt = 0.0, 0
    for k in adj_kernels:
        m1, c1 = sim.profile_query(k)
        ms, cnt = ms + m1, cnt + c1
    sim.profile(False)
    finite = bool(torch.isfinite(gin).all().item())
    return dict(steps=nsteps, checkpoint_every=K, segments_replayed=(nsteps + K - 1) // K - 1,
                forward_ms=fwd_ms, record_ms=rec_ms, backward_ms=bwd_ms, timed_rounds=len(fwd),
                ratio_fwd_plus_adjoint_to_fwd=(rec_ms + bwd_ms) / fwd_ms,
                k_fluid_adj_avg_us=1e3 * ms / cnt if cnt else None, grad_finite=finite)


def time_solid_config(name, sc, integ, dt, fwd_steps, adj_steps, K, loss, rounds=4):
    """A configuration with solids at its full size: fwd_steps plain steps from the initial
    state (timed as a whole), then forward + adjoint of adj_steps steps with a checkpoint
    every K (replaying every segment but the last) against the same steps plain (median of
    three rounds after an untimed one), and the gradient norms of the loss.
    loss(particles) -> (value, cotangent (n, 7))."""
    import numpy as np
    import torch

    from arp.sim import FluidSim

    log("%s: %d particles (%d solid), %d steps" % (name, sc.n, sc.ns, fwd_steps))
    sim = FluidSim(sc, integrator=integ, timestep=dt)
    sim.update(2)
    sim.upload_state(sc.particles)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(dt, fwd_steps)
    sim.synchronize()
    sec = time.perf_counter() - t0
    p = sim.particles()
    out = dict(particles=sc.n, solids=sc.ns, integrator=integ, dt=dt, steps=fwd_steps, seconds=sec,
               ms_per_step=1e3 * sec / fwd_steps, particle_steps_per_s=sc.n * fwd_steps / sec,
               state_finite=bool(np.isfinite(p["pos"]).all()))
    if adj_steps:
        log("%s: forward + adjoint of %d steps, checkpoint every %d" % (name, adj_steps, K))

        def timed(fn):
            sim.synchronize()
            t0 = time.perf_counter()
            r = fn()
            sim.synchronize()
            return time.perf_counter() - t0, r

        # one untimed round (it allocates the recording and the reverse-mode buffers), then
        # the median of three, as for the fluid column
        fwds, recs, bwds = [], [], []
        for r in range(max(rounds, 1)):
            sim.upload_state(sc.particles)
            fwd, _ = timed(lambda: sim.step(dt, adj_steps))
            sim.upload_state(sc.particles)
            rec, _ = timed(lambda: sim.adjoint_record(dt, adj_steps, K))
            val, g7 = loss(sim.particles())
            g = torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda")
            g[:, [0, 4, 5, 6, 8, 9, 10]] = torch.from_numpy(g7.astype(np.float32)).cuda()
            gin = torch.zeros_like(g)
            bwd, gp = timed(lambda: sim.adjoint_backward(g, gin, want_params=True))
            if r or rounds <= 1:  # (a single round is timed, allocation included)
                fwds.append(fwd)
                recs.append(rec)
                bwds.append(bwd)
        med = lambda x: sorted(x)[len(x) // 2]  # noqa: E731
        fwd, rec, bwd = med(fwds), med(recs), med(bwds)
        gi = gin.cpu().numpy()
        out["fwd_adjoint"] = dict(steps=adj_steps, checkpoint_every=K, segments_replayed=(adj_steps + K - 1) // K - 1,
                                  forward_ms=1e3 * fwd, record_ms=1e3 * rec, backward_ms=1e3 * bwd, timed_rounds=len(fwds),
                                  ratio_fwd_plus_adjoint_to_fwd=(rec + bwd) / fwd, loss=val,
                                  grad_x0_norm=float(np.linalg.norm(gi[:, [4, 5, 6]])),
                                  grad_v0_norm=float(np.linalg.norm(gi[:, [8, 9, 10]])),
                                  grad_params={"normal_stiffness": float(gp[1, 2]), "shear_stiffness": float(gp[1, 3]),
                                               "bulk_visc_last_material": float(gp[-1, 0])},
                                  grad_finite=bool(np.isfinite(gi).all() and np.isfinite(gp).all()))
    sim.close()
    return out


def cantilever_loss(sc):
    """Squared tip-displacement error to a seeded target (the free end's particles)."""
    import numpy as np

    z = sc.particles["pos"][:, 2]
    tip = z > z.max() - 0.003
    rng = np.random.default_rng(0)
    target = sc.particles["pos"][tip, :3].astype(np.float64) + 1e-5 * rng.standard_normal((int(tip.sum()), 3))

    def loss(p):
        d = p["pos"][tip, :3].astype(np.float64) - target
        g = np.zeros((len(p), 7))
        g[tip, 1:4] = 2.0 * d
        return float((d * d).sum()), g

    return loss


def contact_loss(sc):
    """The ball's final centroid height."""
    import numpy as np

    ball = sc.particles["mat"] == 1

    def loss(p):
        g = np.zeros((len(p), 7))
        g[ball, 2] = 1.0 / ball.sum()
        return float(p["pos"][ball, 1].mean()), g

    return loss


def time_ensemble(nscenes, steps, dt, lists=None):
    """Independent fluid scenes (log-uniform 2,048-65,536 particles, alpha uniform in
    [0.01, 0.1], seeded) stepped as one batch (graph replay; list skins, row-sort workgroup
    size and list-walk form autotuned), against the same scenes stepped one after the
    other."""
    import numpy as np
    import torch

    from arp import scene as S
    from arp.sim import Batch, FluidSim

    log("ensemble: building %d scenes" % nscenes)
    scenes = S.ensemble_3d(count=nscenes)
    sims = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    npart = sum(sc.n for sc in scenes)
    B = Batch(sims, graph=True)
    if lists is None:
        skin, short_skin = B.tune_lists(dt, 40)
    else:  # fixed list skins (reproducible profiles)
        skin, short_skin = lists
        for s in sims:
            s.set_neighbor_skin(skin)
            s.set_neighbor_short_skin(short_skin)
    div = B.tune(dt, 4)
    _, walk_buffered = B.tuning()
    B.step(dt, 2)
    B.synchronize()
    torch.cuda.synchronize()
    log("ensemble: timing %d batched steps (%d particles)" % (steps, npart))
    t0 = time.perf_counter()
    B.step(dt, steps)
    B.synchronize()
    sec = time.perf_counter() - t0
    # the same scenes one launch sequence each, a few steps
    seq_steps = 5
    for s in sims:
        s.synchronize()
    t1 = time.perf_counter()
    for s in sims:
        s.step(dt, seq_steps)
    for s in sims:
        s.synchronize()
    seq = time.perf_counter() - t1
    finite = all(bool(np.isfinite(s.particles()["pos"]).all()) for s in sims[:8])
    B.close()
    for s in sims:
        s.close()
    return dict(scenes=nscenes, particles=npart, steps=steps, integrator="rk2", dt=dt,
                ms_per_step=1e3 * sec / steps, particle_steps_per_s=npart * steps / sec,
                sequential_particle_steps_per_s=npart * seq_steps / seq, graph=True,
                tuned_list_skin=skin, tuned_short_skin=short_skin, tuned_rowsort_particles_per_workgroup=div,
                tuned_list_walk="LDS-buffered" if walk_buffered else "registers", state_finite=finite)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=CONFIG_STEPS)
    ap.add_argument("--config-steps", type=int, default=CONFIG_STEPS,
                    help="steps of the configuration run from the initial state (the column's 500; fewer only "
                         "to shorten profiling runs)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--shape", default="128,128,64", help="fluid lattice nx,ny,nz (y up)")
    ap.add_argument("--integrator", default="rk2")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--event-steps", type=int, default=50,
                    help="steps of the separate pass that times the kernels with HIP events")
    ap.add_argument("--events", default="fluid", choices=["all", "fluid", "none"],
                    help="kernels timed with HIP events inside the timed region (fluid: the dominant "
                         "kernel only, which keeps the event overhead off the other launches)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--ensemble", type=int, default=256,
                    help="scenes of the batched-ensemble measurement (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--ensemble-lists", default="auto",
                    help="list skins of the ensemble: auto (the list tuner) or SKIN,SHORT_SKIN")
    # (the list settings measured fastest on this workload: skins 0.2-0.5 with short skins
    # 0.03-0.06 gave 1.10e10-1.16e10 particle-steps/s, one level at 0.2 9.8e9)
    ap.add_argument("--skin", type=float, default=0.4,
                    help="neighbour lists kept across RK stages with radius h*(1+skin) (0: search at every stage)")
    ap.add_argument("--short-skin", type=float, default=0.04,
                    help="two-level lists: the walks use a short list of radius h*(1+short_skin), refreshed from "
                         "the long list (0: one level)")
    ap.add_argument("--max-age", type=int, default=None,
                    help="rebuild kept neighbour lists older than this many stages (default: the library's)")
    ap.add_argument("--solid-configs", type=int, default=1,
                    help="also run the cantilever (640,000 solids: 200 RK4 steps; forward + adjoint with a "
                         "checkpoint every 20) and the fluid-ball contact (327,680 particles: 1000 RK2 steps; "
                         "forward + adjoint of 200 steps, a checkpoint every 50) at full size (0: skip)")
    ap.add_argument("--solid-only", default="all", choices=["all", "cantilever", "contact"],
                    help="which solid configurations --solid-configs runs")
    ap.add_argument("--solid-steps", type=int, default=200,
                    help="forward + adjoint steps of the solid configurations (200 as configured; fewer for profiles)")
    ap.add_argument("--solid-rounds", type=int, default=4,
                    help="forward + adjoint rounds of a solid configuration (the first untimed when more than one)")
    ap.add_argument("--adjoint-steps", type=int, default=20,
                    help="steps of the recorded rollouts whose reverse sweep is timed, with a checkpoint "
                         "every steps/4 (replays) and every steps (0: skip)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = 64 * 1024 * 1024
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy kernel: %d bytes read and %d bytes written per launch" % (4 * n, 4 * n))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (each scene lives on one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from arp import scene as S
    from arp.sim import FluidSim

    dt = 5e-5
    shape = tuple(int(x) for x in args.shape.split(","))
    log("building scene: %s fluid column" % (shape,))
    sc = S.fluid_column_3d(shape=shape, rk=args.integrator, dt=dt)
    log("particles %d boundary %d" % (sc.n, sc.nb))
    sim = FluidSim(sc, integrator=args.integrator, timestep=dt)
    sim.set_neighbor_skin(args.skin)
    if 0 < args.short_skin < args.skin:  # (two levels need a long list: --skin 0 searches at every stage)
        sim.set_neighbor_short_skin(args.short_skin)
    if args.max_age is not None:
        sim.set_neighbor_max_age(args.max_age)
    # the list walk (its fast-material form when the scene has one mobile and one boundary
    # material), or the plain step without lists
    fluid_kernels = ["k_fluid_nl", "k_fluid_nl_f"] if args.skin > 0 else ["k_fluid"]
    p_init = sc.particles.copy()
    log("warmup %d steps" % args.warmup)
    sim.update(args.warmup)
    sim.synchronize()
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    # the timed steps (steps W .. W+K of the configuration) run without per-kernel events; a
    # separate pass with events on the stream times the kernels
    nl0 = sim.neighbor_stats()
    sp0 = sim.speculation_stats()
    rf0 = sim.neighbor_refreshes()
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    sim.update(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    nl1 = sim.neighbor_stats()
    sp1 = sim.speculation_stats()
    rf1 = sim.neighbor_refreshes()
    timed_kernels = fluid_kernels + ["k_fluid_nlb", "k_rowsort", "k_scatter", "k_scan", "k_key"]
    sim.profile_select(fluid_kernels if args.events == "fluid" else None)
    sim.profile(args.events != "none")
    sim.profile_reset()
    ev_nl0, ev_rf0 = sim.neighbor_stats(), sim.neighbor_refreshes()
    if args.events != "none":
        log("kernel timing: %d steps with events" % args.event_steps)
        sim.update(args.event_steps)
        sim.synchronize()
    ev_walks = sim.neighbor_stats()[1] - ev_nl0[1]
    ev_refreshes = sim.neighbor_refreshes() - ev_rf0
    kern = {}
    for k in ["k_fluid"] + timed_kernels:
        ms, cnt = sim.profile_query(k)
        if cnt:
            kern[k] = dict(total_ms=ms, launches=cnt, avg_us=1e3 * ms / cnt)
    # the walk kernel that ran (most launches)
    fluid_kernel = max(fluid_kernels, key=lambda k: kern.get(k, {}).get("launches", 0))
    if args.skin > 0 and fluid_kernel in kern:
        # A walk launch returns at once when its stage searched instead, or when an earlier
        # speculative stage found the lists stale: only the stages that walked (one launch
        # each, the longest ones) are the kernel's work.  With two-level lists the longest of
        # those are the short-list refreshes.
        d = np.sort(sim.profile_durations(fluid_kernel))[::-1]
        work = d[:ev_walks] if ev_walks else d
        f = kern[fluid_kernel]
        f["walking_launches"] = int(len(work))
        f["avg_us_all_launches"] = f["avg_us"]
        f["avg_us"] = 1e3 * float(work.mean())
        if ev_refreshes and ev_walks > ev_refreshes:
            f["refreshes"] = int(ev_refreshes)
            f["avg_us_short_walks"] = 1e3 * float(work[ev_refreshes:].mean())
            f["avg_us_refreshes"] = 1e3 * float(work[:ev_refreshes].mean())
    sim.profile(False)
    # the rebuild chain's kernels (what a moving scene pays), timed with events on a build
    sim.profile_select(timed_kernels[len(fluid_kernels):])
    sim.profile(args.events != "none")
    sim.profile_reset()
    if args.events != "none":
        sim.upload_state(sim.particles())  # the next stage rebuilds the lists
        sim.update(1)
        sim.synchronize()
        for k in timed_kernels[len(fluid_kernels):]:
            ms, cnt = sim.profile_query(k)
            if cnt:
                kern.setdefault(k, dict(total_ms=ms, launches=cnt, avg_us=1e3 * ms / cnt))
    sim.profile(False)
    # the configuration itself: CONFIG_STEPS steps from the initial state, timed as a whole
    log("configuration: %d steps from the initial state" % args.config_steps)
    cfg = config_run(sim, p_init, args.config_steps)
    p_end = sim.particles()
    finite = bool(np.isfinite(p_end["pos"]).all())
    cfg["state_finite"] = finite
    cfg["drift"] = drift(sc, p_init, p_end, args.config_steps * dt, float(np.float32(sc.g)))
    if not args.no_cpu:
        log("fp32/fp64 error sample: one step at the end of the configuration")
        cfg["error_sample_one_step_at_end"] = fp64_sample(sim, sc, p_end, args.integrator, dt, args.cpu_threads)
    fwd_adj = []
    ens = None
    if args.adjoint_steps > 0:
        for K in sorted({max(1, args.adjoint_steps // 4), args.adjoint_steps}):
            fwd_adj.append(time_adjoint(sim, sc, p_end, dt, args.adjoint_steps, K))
    sim.close()
    solid = None
    if args.solid_configs:
        solid = {}
        f = args.solid_steps / 200.0  # (the configurations' step counts, scaled for short profiles)
        if args.solid_only in ("all", "cantilever"):
            c3 = S.cantilever_3d()
            solid["cantilever_3d"] = time_solid_config("cantilever", c3, "rk4", 1e-5, args.solid_steps, args.solid_steps,
                                                       max(1, int(20 * f)), cantilever_loss(c3), args.solid_rounds)
        if args.solid_only in ("all", "contact"):
            c4 = S.contact_3d()
            solid["contact_3d"] = time_solid_config("contact", c4, "rk2", 2e-5, int(1000 * f), args.solid_steps,
                                                    max(1, int(50 * f)), contact_loss(c4), args.solid_rounds)
    if args.ensemble > 0:
        lists = None if args.ensemble_lists == "auto" else tuple(float(v) for v in args.ensemble_lists.split(","))
        ens = time_ensemble(args.ensemble, args.ensemble_steps, dt, lists)
    value = sc.n * args.steps / sec
    f = kern.get(fluid_kernel)
    bytes_launch = B_ALG_FLUID * sc.n
    achieved = bytes_launch / (f["avg_us"] * 1e-6) if f else None
    traffic = read_traffic(fluid_kernel, "fluid_column_3d")
    out = {
        "metric": "SPH particle-steps/sec at 1M particles (fwd)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "fluid_column_3d",
            "particles": sc.n,
            "boundary_particles": sc.nb,
            "lattice": "%dx%dx%d (x,y,z) jittered +-0.05*spacing" % shape,
            "spacing": 0.005,
            "h": sc.h,
            "integrator": args.integrator,
            "dt": dt,
            "eos": "tait rho0=1000 c0=20",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": fluid_kernel,
            "achieved": achieved / 1e9 if achieved else None,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK if achieved else None,
            "traffic": traffic,
            "alg_bytes_per_launch": bytes_launch,
        },
        "kernels": kern,
        "neighbor_lists": {"skin": args.skin, "short_skin": args.short_skin if 0 < args.short_skin < args.skin else 0.0, "stage_builds": nl1[0] - nl0[0],
                           "stage_reuses": nl1[1] - nl0[1], "short_list_refreshes": rf1 - rf0,
                           # timed steps issued as list walks alone, and those run again
                           "speculative_steps": sp1[0] - sp0[0], "speculation_aborts": sp1[1] - sp0[1]},
        "config_run": cfg,
        "fwd_adjoint": fwd_adj,
        "solid_configs": solid,
        "ensemble": ens,
    }
    if not args.no_cpu:
        log("cpu baseline: %d step(s), %d threads" % (args.cpu_steps, args.cpu_threads))
        out["cpu_baseline"] = cpu_baseline(sc, args.integrator, dt, args.cpu_steps, args.cpu_threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
