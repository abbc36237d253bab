"""Benchmark: SPH particle-steps/s of the 3-D fluid column (1,048,576 particles, RK2,
dt 5e-5) on one MI355X, with the dominant kernel's HBM roofline and the CPU oracle timed on
the host cores beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line on stdout; progress goes to stderr.
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
# algorithmic bytes per particle per RK2 step of the whole pipeline (key pass, cell table,
# reorder, gather, fused RK; cell-key counting sort design)
B_ALG_STEP = {"rk2": 352.0, "rk4": 872.0}
# algorithmic bytes per particle of one k_fluid launch: read the stage state (pos+den,
# vel+q: 32 B), read y_n (28 B), write the next stage state (28 B), write the cell key (4 B)
B_ALG_FLUID = 92.0


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


def time_adjoint(sim, sc, dt, nsteps, fwd_ms_per_step):
    """Forward with recording + reverse sweep of nsteps steps (one segment, no replay) on the
    same workload, against nsteps plain steps."""
    import numpy as np
    import torch

    log("adjoint: record %d steps + reverse sweep" % nsteps)
    rng = np.random.default_rng(0)
    g = torch.zeros((sc.n, 12), dtype=torch.float32, device="cuda")
    g[:, [0, 4, 5, 6, 8, 9, 10]] = torch.from_numpy(rng.standard_normal((sc.n, 7)).astype(np.float32)).cuda()
    gin = torch.zeros_like(g)
    torch.cuda.synchronize()
    # one untimed round of the timed calls (allocates the recording and the reverse-mode
    # buffers), then the median of three timed rounds
    sim.adjoint_record(dt, nsteps, nsteps)
    sim.adjoint_backward(g, gin, want_params=True)
    sim.synchronize()
    rounds = []
    for _ in range(3):
        t0 = time.perf_counter()
        sim.adjoint_record(dt, nsteps, nsteps)
        sim.synchronize()
        t1 = time.perf_counter()
        sim.adjoint_backward(g, gin, want_params=True)
        sim.synchronize()
        t2 = time.perf_counter()
        rounds.append((t1 - t0, t2 - t1))
    rec_s = sorted(r[0] for r in rounds)[1]
    bwd_s = sorted(r[1] for r in rounds)[1]
    # the reverse stage (on the recorded neighbours, or searching after a record overflow;
    # fluid-only / one-material / general table forms), timed with events in another round
    adj_kernels = ["k_fluid_adj" + sfx for sfx in ("", "_rec", "_f", "_rec_f", "_g", "_rec_g")]
    sim.adjoint_record(dt, nsteps, nsteps)
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
    rec_ms, bwd_ms = 1e3 * rec_s, 1e3 * bwd_s
    return dict(steps=nsteps, checkpoint_every=nsteps, record_ms=rec_ms, backward_ms=bwd_ms, timed_rounds=3,
                forward_ms=fwd_ms_per_step * nsteps,
                ratio_fwd_plus_adjoint_to_fwd=(rec_ms + bwd_ms) / (fwd_ms_per_step * nsteps),
                k_fluid_adj_avg_us=1e3 * ms / cnt if cnt else None, grad_finite=finite)


def time_ensemble(nscenes, steps, dt):
    """Independent fluid scenes (log-uniform 2,048-65,536 particles, alpha uniform in
    [0.01, 0.1], seeded) stepped as one batch (graph replay, tuned), against the same scenes
    stepped one after the other."""
    import numpy as np
    import torch

    from arp import scene as S
    from arp.sim import Batch, FluidSim

    log("ensemble: building %d scenes" % nscenes)
    scenes = S.ensemble_3d(count=nscenes)
    sims = [FluidSim(sc, integrator="rk2", timestep=dt) for sc in scenes]
    npart = sum(sc.n for sc in scenes)
    B = Batch(sims, graph=True)
    div = B.tune(dt, 4)
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
                tuned_rowsort_particles_per_workgroup=div, state_finite=finite)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
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
    ap.add_argument("--skin", type=float, default=0.2,
                    help="neighbour lists kept across RK stages with radius h*(1+skin) (0: search at every stage)")
    ap.add_argument("--max-age", type=int, default=None,
                    help="rebuild kept neighbour lists older than this many stages (default: the library's)")
    ap.add_argument("--adjoint-steps", type=int, default=10,
                    help="steps of the recorded rollout whose reverse sweep is timed (0: skip)")
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
    if args.max_age is not None:
        sim.set_neighbor_max_age(args.max_age)
    # the list walk (its fast-material form when the scene has one mobile and one boundary
    # material), or the plain step without lists
    fluid_kernels = ["k_fluid_nl", "k_fluid_nl_f"] if args.skin > 0 else ["k_fluid"]
    log("warmup %d steps" % args.warmup)
    sim.update(args.warmup)
    sim.synchronize()
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    # the timed steps run without per-kernel events; a separate pass with events on the
    # stream times the kernels
    nl0 = sim.neighbor_stats()
    sp0 = sim.speculation_stats()
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    sim.update(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    nl1 = sim.neighbor_stats()
    sp1 = sim.speculation_stats()
    sim.profile_select(fluid_kernels if args.events == "fluid" else None)
    sim.profile(args.events != "none")
    sim.profile_reset()
    if args.events != "none":
        log("kernel timing: %d steps with events" % args.event_steps)
        sim.update(args.event_steps)
        sim.synchronize()
    kern = {}
    for k in ("k_fluid", "k_fluid_nl", "k_fluid_nl_f", "k_fluid_nlb", "k_rowsort", "k_scatter", "k_scan", "k_key"):
        ms, cnt = sim.profile_query(k)
        if cnt:
            kern[k] = dict(total_ms=ms, launches=cnt, avg_us=1e3 * ms / cnt)
    # the walk kernel that ran (most launches)
    fluid_kernel = max(fluid_kernels, key=lambda k: kern.get(k, {}).get("launches", 0))
    if fluid_kernel in kern:
        # the timed region's bound on one launch (its steps hold little but these walks, back
        # to back): the step time over the stages per step
        from arp import tableau

        stages = int(tableau.tableau(args.integrator)["stages"])
        if stages > 0:
            kern[fluid_kernel]["timed_region_bound_us"] = 1e6 * sec / args.steps / stages
    sim.profile(False)
    fwd_adj = None
    ens = None
    if args.adjoint_steps > 0:
        fwd_adj = time_adjoint(sim, sc, dt, args.adjoint_steps, 1e3 * sec / args.steps)
    p = sim.particles()
    finite = bool(np.isfinite(p["pos"]).all())
    sim.close()
    if args.ensemble > 0:
        ens = time_ensemble(args.ensemble, args.ensemble_steps, dt)
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
        "step_roofline_frac": B_ALG_STEP.get(args.integrator, 0.0) * value / HBM_PEAK,
        "kernels": kern,
        "neighbor_lists": {"skin": args.skin, "stage_builds": nl1[0] - nl0[0],
                           "stage_reuses": nl1[1] - nl0[1],
                           # timed steps issued as list walks alone, and those run again
                           "speculative_steps": sp1[0] - sp0[0], "speculation_aborts": sp1[1] - sp0[1]},
        "state_finite": finite,
        "fwd_adjoint": fwd_adj,
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
