"""Benchmark of the 3-D electrostatic PIC step on one MI355X.

Workload: 256x256x512 mesh, 32M particles of a seeded KV-like beam, the
pipe and +10 kV annulus conductors (warp_amd/configs.py "c2").  One step is
the reference's step3d order on the GPU: gather E + two half kicks + drift
(padvnc3d fullv), particle boundaries, CIC deposit, multigrid solve to
mgtol = 1e-6 V warm-started from the previous phi.  Inputs are resident in
HBM before the timed region.

Prints one JSON line (steps/s, V-cycles and ms per solve, the roofline of
the dominant kernel, and the FP64 oracle timed on the host as the CPU
baseline).  ``--gpus N`` with N > 1 exits: the path runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--config", default="c2")
    p.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--guess", default="previous", choices=["previous", "linear"],
                   help="initial guess of each solve: the previous phi (the reference's) or "
                        "its linear extrapolation in time")
    p.add_argument("--mixed-precision", default="bf16", choices=["fp32", "bf16", "none"],
                   help="precision of the second run reported in the 'mixed' block "
                        "(none: no second run)")
    p.add_argument("--mixed-guess", default="linear", choices=["previous", "linear"])
    p.add_argument("--sort-interval", type=int, default=25,
                   help="steps between particle sorts by cell (0: only at set-up)")
    p.add_argument("--sorform", type=int, default=0, choices=[0, 1, 2],
                   help="SOR pass form: 0 two half sweeps per pass, 1 both colours in one "
                        "launch, 2 that on the finest level only (the same bits)")
    p.add_argument("--cpu-particles", type=int, default=4_000_000,
                   help="particles in the CPU baseline's sample")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(args, cfg, precision, guess, dev, keep=False):
    """One complete run of the workload: set-up, initial solve, warm-up
    steps, then the timed step(K) call.  Returns a dict of the figures and,
    with keep, the solver and particles."""
    import torch
    from warp_amd import configs, registersolver, unregistersolver, top
    from warp_amd import stepper
    from warp_amd.stepper import step, generate

    top.it, top.time, top.lspecial, top.nspecial = 0, 0., False, 0
    t0 = time.time()
    solver = configs.make_solver(cfg, precision=precision, device=dev, mgverbose=-1,
                                 guess=guess, sorform=args.sorform)
    pg = configs.load_beam(cfg, device=dev)
    top.pgroup = pg
    top.dt = configs.DT
    configs.set_particle_boundaries(cfg)
    registersolver(solver)
    npart0 = pg.nplive
    # particles ordered by cell once (particlesortxyzwithcopy) and again
    # every particlesortinterval steps; the beam drifts as a whole, so the
    # order stays nearly sorted in between
    solver.sortparticles(pg)
    top.particlesortinterval = args.sort_interval
    log("[%s/%s] set-up %.1f s: %d particles, mesh %s"
        % (precision, guess, time.time() - t0, npart0, cfg["n"]))
    t0 = time.time()
    generate()
    torch.cuda.synchronize()
    cold_iters = solver.mgiters
    log("[%s/%s] first solve (cold start) %d V-cycles, %.2f s incl. conductor set-up"
        % (precision, guess, cold_iters, time.time() - t0))
    if args.warmup > 0:
        step(args.warmup)
    torch.cuda.synchronize()

    # --- timed region: one step(K) call, as a run script makes it; the
    # reference's cadence of special steps (synchv half kick + moments every
    # top.nhist = 5 steps and on the last step of the call) is included
    solver.timers = dict(deposit=0., solve=0., fetche=0.)
    nspecial0 = top.nspecial
    nsolve0 = len(solver.mgiters_history)
    counts = []
    rec = lambda: counts.append(pg.nplive)
    stepper.installafterstep(rec)
    it0 = top.it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stepper.uninstallafterstep(rec)
    iters = solver.mgiters_history[nsolve0:]
    nspecial = top.nspecial - nspecial0
    ms_per_step = 1e3 * elapsed / args.steps
    ms_solve = 1e3 * solver.timers["solve"] / args.steps
    log("[%s/%s] timed %d steps (it %d..%d, %d special): %.3f ms/step, %.3f ms/solve, "
        "V-cycles %s, particles %d -> %d" % (precision, guess, args.steps, it0 + 1, top.it,
                                             nspecial, ms_per_step, ms_solve, iters,
                                             counts[0], counts[-1]))
    out = dict(ms_per_step=ms_per_step, ms_per_solve=ms_solve, iters=iters,
               cold_iters=cold_iters, nspecial=nspecial, npart0=npart0,
               particles_mean=float(sum(counts)) / len(counts), particles_at_end=counts[-1],
               steps_timed=(it0 + 1, top.it))
    if keep:
        out["solver"], out["pg"] = solver, pg
    else:
        unregistersolver(solver)
        top.pgroup = None
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (the solver has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from warp_amd import configs, unregistersolver, top, _native as N
    from warp_amd.stepper import step

    dev = torch.device("cuda", 0)
    cfg = configs.CONFIGS[args.config]
    r = run(args, cfg, args.precision, args.guess, dev, keep=True)
    solver, pg = r.pop("solver"), r.pop("pg")
    ms_per_step, iters = r["ms_per_step"], r["iters"]

    # --- dominant kernel: finest-level SOR half sweep, HIP events on the
    # solver's stream around each launch (one more step, kernels launched
    # one V-cycle at a time while timing).
    N.call("wm_mg3d_set_timing", solver._h, 1)
    phi_before = solver.phi.clone()
    step(1)
    torch.cuda.synchronize()
    ms_tot, nl, fused = C.c_double(0.), C.c_int64(0), C.c_int64(0)
    N.call("wm_mg3d_get_timing", solver._h, C.byref(ms_tot), C.byref(nl))
    N.call("wm_mg3d_timing_fused", solver._h, C.byref(fused))
    N.call("wm_mg3d_set_timing", solver._h, 0)
    nx, ny, nz = cfg["n"]
    nodes = (nx + 1) * (ny + 1) * (nz + 1)
    s_bytes = {"fp64": 8, "fp32": 4, "bf16": 2}[args.precision]  # finest-level storage
    if fused.value:
        # full pass, both colours: read phi (s) + rho (s), write phi (s) per node
        kname, alg_bytes = "k_sor_fused (finest-level SOR pass, both colours)", 3 * s_bytes * nodes
        tf = os.path.join(ROOT, "profiles", "sor_fused_256x256x512_traffic.json")
    else:
        # half sweep: read phi (s) + rho/2, write phi/2 per node
        kname, alg_bytes = "k_sor_zmarch (finest-level SOR half sweep)", 2 * s_bytes * nodes
        tf = os.path.join(ROOT, "profiles", "sor_zmarch_256x256x512_traffic.json")
    avg_ms = ms_tot.value / max(nl.value, 1)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    log("%s: %d launches, %.4f ms avg, %.0f GB/s" % (kname, nl.value, avg_ms, achieved))
    traffic = None
    # the counters were read for the FP64 kernel of this configuration only
    if args.precision == "fp64" and args.config == "c2" and os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    # Passes of known size for scaling a profiler's FETCH_SIZE / WRITE_SIZE
    # counters: 2 GiB read only, then 2 GiB written only (wm_stream_probe).
    a = torch.zeros(1 << 28, dtype=torch.float64, device=dev)
    o = torch.zeros(1, dtype=torch.float64, device=dev)
    N.call("wm_stream_probe", a.data_ptr(), o.data_ptr(), a.numel(), 0, N.stream_of(dev))
    N.call("wm_stream_probe", a.data_ptr(), o.data_ptr(), a.numel(), 1, N.stream_of(dev))
    torch.cuda.synchronize()
    del a, o

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, cfg, solver, pg, phi_before, iters,
                           r["nspecial"] / args.steps)

    mgtol = solver.mgtol
    mixed = None
    if args.mixed_precision != "none":
        # the same workload with the mixed-precision solve and the better
        # initial guess; phi and the grid E of its last step against the
        # FP64 run's
        phi64 = solver.phi.clone()
        e64 = solver.getselfe().clone()
        nstep64 = top.it
        unregistersolver(solver)
        top.pgroup = None
        del solver, pg
        torch.cuda.empty_cache()
        m = run(args, cfg, args.mixed_precision, args.mixed_guess, dev, keep=True)
        step(nstep64 - top.it)  # the FP64 run's extra timing step
        torch.cuda.synchronize()
        ph, e = m["solver"].phi, m["solver"].getselfe()
        rel = lambda u, v: float(torch.linalg.vector_norm(u - v) / torch.linalg.vector_norm(v))
        mixed = {"precision": args.mixed_precision, "initial_guess": args.mixed_guess,
                 "ms_per_step": m["ms_per_step"], "steps_per_s": 1e3 / m["ms_per_step"],
                 "ms_per_solve": m["ms_per_solve"],
                 "vcycles_per_solve": float(np.mean(m["iters"])),
                 "speedup_vs_fp64": ms_per_step / m["ms_per_step"],
                 "phi_rel_l2_vs_fp64": rel(ph, phi64), "e_rel_l2_vs_fp64": rel(e, e64),
                 "particles_mean": m["particles_mean"]}
        log("mixed: %s" % json.dumps(mixed))
        unregistersolver(m["solver"])
        top.pgroup = None

    out = {
        "metric": "PIC steps/s at 256^2x512, 32M ptcls; V-cycles & ms per solve to mgtol=1e-6",
        "value": 1e3 / ms_per_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 V-cycles in f64 defect correction",
                  "bf16": "bf16 finest level, f32 coarse levels, f64 defect correction"
                  }[args.precision],
        "data": "synthetic (seeded KV beam, pipe + 10 kV annulus)",
        "config": {"workload": "ES PIC step %dx%dx%d, %d particles, subgrid conductors"
                               % (nx, ny, nz, r["npart0"]),
                   "mesh": [nx, ny, nz], "particles": r["npart0"],
                   "particles_mean_timed": r["particles_mean"],
                   "particles_at_end": r["particles_at_end"],
                   "timed_steps_it": list(r["steps_timed"]),
                   "mgtol": mgtol, "particle_sort_interval": args.sort_interval,
                   "precision": args.precision, "initial_guess": args.guess,
                   "sor_form": ["two half sweeps", "fused pass",
                                "fused pass on the finest level"][args.sorform],
                   "parallelism": "1 GPU"},
        "vcycles_per_solve": float(np.mean(iters)),
        "vcycles_cold_solve": r["cold_iters"],
        "ms_per_solve": r["ms_per_solve"],
        "roofline": {"kernel": kname, "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
        "mixed": mixed,
    }
    print(json.dumps(out), flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the CPU baseline: the process's CPU set, capped by
    OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(args, cfg, solver, pg, phi_before, iters, fspecial):
    """The FP64 oracle (the C restatement of the reference; its independent
    loops run under OpenMP with unchanged results, the deposit stays
    serial) on the host: one step of the same workload with a particle
    sample of --cpu-particles (deposit and push times scaled to the full
    count), the full-mesh solve warm-started from the same phi as the GPU's
    last solve, and the synchv half kick + beam moments of special steps
    weighted by their share fspecial of the timed steps.  Timed on all the
    process's threads (the value) and on one thread."""
    import ctypes as C
    import numpy as np
    from oracle import oracle as O
    from warp_amd import configs
    nx, ny, nz = cfg["n"]
    x0, x1, y0, y1, z0, z1 = configs.DOMAIN
    dx, dy, dz = (x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz
    live = pg.species_slice(0)
    npart = live.stop - live.start
    stride = max(1, -(-npart // args.cpu_particles))
    sl = slice(live.start, live.stop, stride)  # every stride-th particle: the whole beam
    cols0 = [getattr(pg, c)[sl].cpu().numpy().copy() for c in ("xp", "yp", "zp", "uxp", "uyp",
                                                                "uzp")]
    ns = len(cols0[0])
    q, m, w = pg.sq[0], pg.sm[0], pg.sw[0]
    co = O.install_conductors(configs.conductors(cfg["conductors"], O), nx, ny, nz, x0, x1,
                              y0, y1, z0, z1)
    phi0 = phi_before.cpu().numpy().copy()
    rho0 = solver.rho.cpu().numpy().copy()
    lib = O.lib()
    lib.oc_set_threads.argtypes = [C.c_int64]

    def one_step(threads):
        lib.oc_set_threads(threads)
        x, y, z, ux, uy, uz = [a.copy() for a in cols0]
        gi = np.ones(ns)
        phi = phi0.copy()
        rho = np.zeros((nz + 1, ny + 1, nx + 1))
        t = time.perf_counter()
        ex, ey, ez = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        O.sete3d(phi, x, y, z, ex, ey, ez, dx, dy, dz, x0, y0, z0)
        O.epush3d(ux, uy, uz, ex, ey, ez, q, m, 0.5 * configs.DT)
        O.epush3d(ux, uy, uz, ex, ey, ez, q, m, 0.5 * configs.DT)
        O.xpush3d(x, y, z, ux, uy, uz, gi, configs.DT)
        # particleboundaries3d + clearpart (fill method 1)
        O.particleboundaries3d(x, y, z, gi, dx, dy, dz, nx, ny, nz, x0, y0, z0)
        nk = O.clearpart([x, y, z, ux, uy, uz, gi], 1)
        t_push = time.perf_counter() - t
        t = time.perf_counter()
        O.setrho3d(rho, x[:nk], y[:nk], z[:nk], q, w, dx, dy, dz, x0, y0, z0)
        t_dep = time.perf_counter() - t
        xs, ys, zs, uxs, uys, uzs = x[:nk], y[:nk], z[:nk], ux[:nk], uy[:nk], uz[:nk]
        t = time.perf_counter()
        es = [np.zeros(len(xs)) for _ in range(3)]
        O.sete3d(phi, xs, ys, zs, *es, dx, dy, dz, x0, y0, z0)
        O.epush3d(uxs, uys, uzs, *es, q, m, 0.5 * configs.DT)
        O.getzmmnt(xs, ys, zs, uxs, uys, uzs, np.ones(len(xs)))
        t_synch = time.perf_counter() - t
        rho_full = rho0.copy()
        t = time.perf_counter()
        it, err = O.multigrid3dsolve(phi, rho_full, dx, dy, dz, conductors=co)
        t_solve = time.perf_counter() - t
        step_s = (t_push + t_dep + fspecial * t_synch) * (npart / ns) + t_solve
        log("CPU oracle, %d thread(s): push %.2f s + deposit %.2f s + synchv/moments %.2f s for "
            "%d particles, solve %.2f s (%d V-cycles)" % (threads, t_push, t_dep, t_synch, ns,
                                                         t_solve, it))
        return step_s, t_solve, it

    nthr = cpu_threads()
    step_n, solve_n, it = one_step(nthr)
    step_1, solve_1, _ = one_step(1)
    scale = npart / ns
    return {"value": 1. / step_n, "unit": "steps/s", "cores": nthr, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "one step: gather+push and deposit of every %d-th particle (%d of %d, times scaled "
                      "by %.2f; the deposit serial), synchv kick + moments weighted %.2f, full "
                      "%dx%dx%d solve warm-started from the GPU's phi (%d V-cycles, %.2f s on %d "
                      "threads)" % (stride, ns, npart, scale, fspecial, nx, ny, nz, it, solve_n,
                                    nthr),
            "ms_per_solve": 1e3 * solve_n, "vcycles": it,
            "value_1thread": 1. / step_1, "ms_per_solve_1thread": 1e3 * solve_1}


if __name__ == "__main__":
    main()
