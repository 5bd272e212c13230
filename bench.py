"""Benchmark: argon CCP PIC-MCC timesteps on one MI355X.

Workload (default): 256x256 cells, dx = 0.2 mm, 200 V at 13.56 MHz between
the top and bottom node rows, Neumann sides, Ar 50 mTorr, 200 macro-particles
per cell per species (26.2 M), dt = T_RF/800, seeded Maxwellian loading.
A step is one pass of the full loop (field solve, E field, push, walls, MCC,
rebucketing, deposit, accumulators) over all particles, inputs resident in
HBM.  The timed path is FP32 particles/fields with FP64 refinement of the
field solve; the same workload is also timed on the FP64 plain path, and the
CPU restatement (the oracle) is timed on a bounded sample on the host.

Prints one JSON line (rank 0; this build runs on one GPU only).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hybrid-final_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s, FP32-refined vs FP64 plain; p50/p99 ensemble case latency"


def run_gpu(deck, precision, steps, warmup, timing_steps=0, field_solver=0):
    """Warm up, time `steps` steps with no instrumentation, then (optionally)
    `timing_steps` more steps with HIP events around every stage kernel on the
    case's stream (the per-kernel times; the events add launch gaps, so they
    are kept out of the throughput run).  field_solver: 0 the default of the
    precision, 1 Jacobi PCG, 2 the separable FP64 solve in either precision."""
    from picmcc import Case
    c = Case(deck, precision, field_solver=field_solver)
    c.step(warmup)
    if os.environ.get("PIC_BENCH_DEBUG"):
        print("debug: last field solve", c.last_field_solve(), file=sys.stderr, flush=True)
    c.sync()
    inf0 = c.info()
    t0 = time.perf_counter()
    c.step(steps)
    c.sync()
    dt = time.perf_counter() - t0
    inf1 = c.info()
    stages, nst = {}, 0
    if timing_steps > 0:
        c.stage_timing(True)
        c.step(timing_steps)
        c.sync()
        stages, nst = c.stage_times()
        c.stage_timing(False)
    it, outer, rel = c.last_field_solve()
    out = {
        "seconds": dt,
        "np0": (inf0.np[0], inf0.np[1]),
        "np1": (inf1.np[0], inf1.np[1]),
        "stages_ms": {k: v / max(nst, 1) for k, v in stages.items()},
        "field_iters": it, "field_cycles": outer, "field_rel_res": rel,
        "solver_kind": c.field_solver_kind(),
        "solver_shape": c.field_solver_shape(),
    }
    c.close()
    return out


SOLVER_LABEL = {0: "FP64 Jacobi PCG (the reference algorithm in FP64), tol 1e-10 relative",
                1: "FP64 separable fast-Poisson inverse (closed-form DCT/DST eigenbases, FP64 matrix-core GEMMs) "
                   "in FP64 refinement, tol 1e-10 relative"}
REFINED_PCG_LABEL = "FP32 Jacobi PCG in FP64 refinement, tol 1e-10 relative"


def solver_label(r, precision):
    from picmcc import FP32_REFINED
    if r["solver_kind"] == 1:
        return SOLVER_LABEL[1]
    return REFINED_PCG_LABEL if precision == FP32_REFINED else SOLVER_LABEL[0]


def profile_check(deck, cycles, sample_every=40, emax=40.0):
    """Time-averaged profiles over `cycles` RF cycles on the FP32-refined path
    and on the FP64 path with the same (separable, FP64) field solve, from the
    same loaded particles and Philox stream: the x-averaged electron and ion
    density and potential profiles and the EEDF (sampled every sample_every
    steps), as the largest difference relative to the FP64 profile's peak."""
    import numpy as np
    from picmcc import Case, FP32_REFINED, FP64_PLAIN
    steps = cycles * int(deck["SimulationMethod"]["TimeStep_PIC"])
    out = {}
    for name, prec in (("fp32", FP32_REFINED), ("fp64", FP64_PLAIN)):
        c = Case(deck, prec, field_solver=2)
        for _ in range(steps // sample_every):
            c.step(sample_every)
            c.eedf_sample(emax)
        c.sync()
        inf = c.info()
        ngx, ngy, nsp = inf.ngx, inf.ngy, inf.nsp
        den = c.array("sum_den").reshape(nsp, ngx, ngy) / steps
        phi = c.array("sum_phi").reshape(ngx, ngy) / steps
        eedf = c.array("eedf").astype(np.float64)
        out[name] = (den, phi, eedf / eedf.sum())
        c.close()
    (d32, p32, f32), (d64, p64, f64) = out["fp32"], out["fp64"]

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return {"cycles": cycles, "steps": steps,
            "electron_density": rel(d32[0].mean(axis=0), d64[0].mean(axis=0)),
            "ion_density": rel(d32[1].mean(axis=0), d64[1].mean(axis=0)),
            "potential": rel(p32.mean(axis=0), p64.mean(axis=0)),
            "eedf": rel(f32, f64),
            "measure": "max |FP32 - FP64| / peak of the FP64 profile; density and potential averaged over x "
                       "(the Neumann direction), EEDF sampled every %d steps over 0-%g eV" % (sample_every, emax)}


def particles_avg(r):
    return 0.5 * (sum(r["np0"]) + sum(r["np1"]))


def progress(msg):
    """One line on stderr per finished leg (the JSON line goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    """Threads the CPU baseline uses: this process's CPU set, at most 16
    (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(deck, steps, threads, repeats=3):
    """The FP64 oracle's OpenMP step (cell tiles; bit-identical to the
    sequential oracle) on `steps` steps of the workload, `repeats` times from
    the same loaded state; returns the median rate and the run times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    rates, times = [], []
    for _ in range(repeats):
        o = Oracle(deck, "f64_omp")
        o.orc_deposit()
        n0 = o.scalar("np0") + o.scalar("np1")
        t0 = time.perf_counter()
        rc = o.orc_run_par(steps, threads)
        dt = time.perf_counter() - t0
        n1 = o.scalar("np0") + o.scalar("np1")
        o.close()
        if rc != 0:
            raise RuntimeError(f"oracle run failed with {rc}")
        rates.append(0.5 * (n0 + n1) * steps / dt)
        times.append(dt)
    return float(statistics.median(rates)), times


def run_pbatch(n, reps=3):
    """n independent 5-point problems (sides log-uniform 64-512, Gaussian
    charge blobs), solved together to 1e-10 relative residual in FP64; the
    right-hand sides are resident on the device before the timed solves."""
    from picmcc import PoissonBatch, poisson_cases
    sizes, bcs, rhs = poisson_cases.make(n)
    pb = PoissonBatch(sizes, bcs)
    pb.set_rhs(rhs)
    pb.solve(1e-10, 3)
    pb.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        pb.solve(1e-10, 3)
    pb.sync()
    dt = (time.perf_counter() - t0) / reps
    _, rel, bad = pb.result()
    out = {"problems": n, "unknowns": pb.N, "ms_per_batch": dt * 1e3, "unknowns_per_s": pb.N / dt,
           "max_rel_residual": float(rel.max()), "unconverged": bad,
           "method": "FP64 separable fast-Poisson inverse (4 batched GEMMs per cycle) in FP64 refinement, tol 1e-10"}
    pb.close()
    return out


def run_o2(steps, warmup=2):
    """O2 chemistry (gas type 1), staged steps: 256x256 cells, dx=0.2 mm,
    100 V, 50 mTorr, 50 particles/cell for e and O2+ and 10 for O+ and O-,
    ions colliding too; FP32 particles with the FP64-refined field solve.
    Particles are resident in HBM before the timed steps."""
    from picmcc import Case, FP32_REFINED, decks
    deck = decks.oxygen_deck(gas=1, ncx=256, ncy=256, dx=2e-4, volts=100.0, pres_torr=0.05, ppc=50,
                             steps_per_cycle=800, ion_collisions=True)
    c = Case(deck, FP32_REFINED)
    c.step(warmup)
    c.sync()
    n = int(sum(c.info().np[s] for s in range(4)))
    t0 = time.perf_counter()
    c.step(steps)
    c.sync()
    dt = (time.perf_counter() - t0) / steps
    c.close()
    return {"workload": "O2 CCP 256x256 cells dx=0.2mm, 100 V, 50 mTorr, 50/50/10/10 particles per cell "
                        "(e, O2+, O+, O-), ions colliding, dt=T_RF/800",
            "particles": n, "steps": steps, "ms_per_step": dt * 1e3, "particle_steps_per_s": n / dt,
            "path": "staged stages (the fused two-pass step is built for argon)"}


def run_aro2(steps, warmup=2):
    """Ar/O2 chemistry (gas type 2), staged steps on the 512x256 workload:
    O2 20 %, 30 mTorr, 150 V, dx=0.2 mm, 100 particles/cell for e, Ar+ and
    O2+ and 20 for O+ and O- (44.6 M particles), dt=T_RF/800; FP32 particles
    with the FP64-refined field solve.  Particles are resident in HBM before
    the timed steps."""
    from picmcc import Case, FP32_REFINED, decks
    c = Case(decks.config3(), FP32_REFINED)
    c.step(warmup)
    c.sync()
    inf0 = c.info()
    n0 = int(sum(inf0.np[s] for s in range(inf0.nsp)))
    t0 = time.perf_counter()
    c.step(steps)
    c.sync()
    dt = (time.perf_counter() - t0) / steps
    inf1 = c.info()
    n1 = int(sum(inf1.np[s] for s in range(inf1.nsp)))
    c.close()
    n = 0.5 * (n0 + n1)
    return {"workload": "Ar/O2 CCP 512x256 cells dx=0.2mm, O2 20 %, 30 mTorr, 150 V, 100/100/100/20/20 particles "
                        "per cell (e, Ar+, O2+, O+, O-), dt=T_RF/800",
            "particles": n1, "steps": steps, "ms_per_step": dt * 1e3, "particle_steps_per_s": n / dt,
            "path": "staged stages with the oxygen collision step (the fused two-pass step is built for argon)"}


def run_ensemble(n, rate_hz):
    """A served ensemble of n argon CCP cases (grids 32^2-96^2, 20-60
    particles/cell, 10-500 mTorr, 50-300 V, 1-3 RF cycles of 400 steps,
    Poisson arrivals): the concurrent server (continuous admission, each
    case on its own stream), the same requests served first come first
    served with one case resident, and the same cases run one by one."""
    from picmcc import serve
    reqs = serve.ar_ensemble(n, rate_hz=rate_hz)
    # one by one first: it also takes the process's one-time costs (first
    # graph captures and kernel loads) out of the served runs
    alone = [serve.run_alone(r) for r in reqs]
    progress(f"ensemble: {n} cases one by one")
    out = serve.summarize(serve.EnsembleServer(max_active=16, setup_threads=2).run(reqs), alone)
    progress("ensemble: concurrent server")
    out["mode"] = ("concurrent: continuous admission by device-memory budget (at most 16 resident), one stream "
                   "per case, graph replays, case setup on 2 worker threads")
    srv = serve.EnsembleServer(max_active=16, setup_threads=2, batch=True)
    bat = serve.summarize(srv.run(reqs))
    out["batched"] = {k: bat[k] for k in ("p50_latency_s", "p99_latency_s", "makespan_s",
                                          "aggregate_particle_steps_per_s")}
    progress("ensemble: batched server")
    out["batch_stepping"] = run_batch_stepping()
    progress("ensemble: batch stepping")
    out["batched"]["mode"] = ("concurrent admission as above; the resident cases stepped together through one "
                              "pic_batch (each kernel of the step launched once for all of them)")
    out["batched"]["batch_rebuilds"] = srv.rebuilds
    seq = serve.summarize(serve.SequentialServer().run(reqs))
    progress("ensemble: sequential server")
    out["sequential"] = {k: seq[k] for k in ("p50_latency_s", "p99_latency_s", "makespan_s", "busy_fraction",
                                             "aggregate_particle_steps_per_s")}
    out["workload"] = (f"{n} Ar CCP cases, grids 32^2-96^2, 20-60 particles/cell, 10-500 mTorr log-uniform, "
                       f"50-300 V, 1-3 RF cycles of 400 steps, Poisson arrivals {rate_hz:g}/s, latency = finish - "
                       "arrival incl. setup and queueing")
    return out


def run_batch_stepping(steps=100):
    """Stepping only (no setup, no admission): 16 argon cases of one size,
    each on its own stream, against the same cases in one pic_batch."""
    from picmcc import Batch, Case, FP32_REFINED, decks
    out = []
    for n, ncx, ppc in [(16, 32, 20), (16, 48, 30), (16, 64, 50), (8, 96, 50)]:
        ds = [decks.ccp_deck(ncx=ncx, ncy=ncx, ppc=ppc, volts=150.0 + 5 * i, seed=100 + i, philox_seed=200 + i)
              for i in range(n)]
        cs = [Case(d, FP32_REFINED) for d in ds]
        for c in cs:
            c.step(2)
        for c in cs:
            c.sync()
        npart = sum(sum(c.info().np[:2]) for c in cs)
        t0 = time.perf_counter()
        for c in cs:
            c.step(steps)
        for c in cs:
            c.sync()
        ts = time.perf_counter() - t0
        b = Batch(cs)
        b.step(2)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        tb = time.perf_counter() - t0
        b.close()
        for c in cs:
            c.close()
        out.append({"cases": n, "grid": f"{ncx}x{ncx}", "particles_per_cell": ppc, "particles": npart,
                    "streams_ms_per_step": ts / steps * 1e3, "batch_ms_per_step": tb / steps * 1e3,
                    "speedup": ts / tb})
    return out


def run_serve5(k, rate_hz=8.0):
    """The first k requests of the serving workload (512 Poisson arrivals at
    rate_hz of argon and Ar/O2 cases, grids 64-256 cells a side, 20-200
    particles per cell, 10-500 mTorr, 50-300 V, 3-20 RF cycles of 400 steps;
    picmcc.serve.config5_requests), served concurrently, against the same
    cases run one by one."""
    from picmcc import serve
    reqs = serve.config5_requests(512, rate_hz=rate_hz)[:k]
    alone = []
    for r in reqs:
        alone.append(serve.run_alone(r))
        progress(f"serving workload: case {r.rid} alone, {alone[-1]['seconds']:.1f} s")
    res = serve.EnsembleServer(max_active=16, setup_threads=2).run(reqs)
    progress("serving workload: concurrent server")
    out = serve.summarize(res, alone)
    srv = serve.EnsembleServer(max_active=16, setup_threads=2, batch=True)
    bat = serve.summarize(srv.run(reqs))
    out["batched"] = {k: bat[k] for k in ("p50_latency_s", "p99_latency_s", "makespan_s",
                                          "aggregate_particle_steps_per_s", "failed")}
    out["batched"]["batch_rebuilds"] = srv.rebuilds
    progress("serving workload: batched server")
    out["gas_mix"] = {"argon": sum(1 for r in reqs if serve.case_gas(r.deck) == 0),
                      "ar_o2": sum(1 for r in reqs if serve.case_gas(r.deck) == 2)}
    out["one_by_one_failed"] = sum(1 for r in alone if r["error"])
    out["workload"] = (f"first {k} of 512 Poisson arrivals at {rate_hz:g}/s: Ar or Ar/O2 (70-95 % Ar) CCP cases, "
                       "grids 64-256 cells a side (dx=0.2mm), 20-200 particles/cell, 10-500 mTorr log-uniform, "
                       "50-300 V at 13.56 MHz, 3-20 RF cycles of 400 steps; latency = finish - arrival incl. "
                       "setup and queueing")
    out["mode"] = "concurrent server, at most 16 resident, case setup on 2 worker threads"
    return out


def copy_calibration(mb):
    """Device-to-device copy of `mb` MiB through the HIP runtime's copy kernel.

    Profiled under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, it gives the
    counter reading of a kernel whose HBM traffic is known (mb MiB read and
    mb MiB written), which scales the counters of the particle kernels.
    """
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}), file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=20, help="steps timed on the FP64 plain path")
    ap.add_argument("--same-solver-steps", type=int, default=200,
                    help="steps timed on the FP64 path with the separable FP64 field solve (0 = skip)")
    ap.add_argument("--refined-pcg-steps", type=int, default=20,
                    help="steps timed on the FP32 path with the FP32 PCG in FP64 refinement (0 = skip)")
    ap.add_argument("--profile-cycles", type=int, default=10,
                    help="RF cycles of the FP32 vs FP64 time-averaged profile comparison (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=10, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-config1-steps", type=int, default=0,
                    help="also time the CPU baseline on the 64x64 argon deck for this many steps (0 = skip)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--pbatch", type=int, default=256, metavar="N",
                    help="independent Poisson problems (64^2-512^2) in the batched-solve leg (0 = skip)")
    ap.add_argument("--o2-steps", type=int, default=20, metavar="K",
                    help="staged O2-chemistry steps timed on a 256x256 O2 case (0 = skip)")
    ap.add_argument("--aro2-steps", type=int, default=10, metavar="K",
                    help="staged Ar/O2 steps timed on the 512x256 Ar/O2 workload (0 = skip)")
    ap.add_argument("--ensemble", type=int, default=16, metavar="N",
                    help="cases in the served argon ensemble (Poisson arrivals; concurrent server, first-come "
                         "first-served server and one by one; 0 = skip)")
    ap.add_argument("--ensemble-rate", type=float, default=32.0, metavar="HZ",
                    help="mean arrival rate of the ensemble's cases")
    ap.add_argument("--serve5", type=int, default=12, metavar="K",
                    help="serve the first K requests of the 512-case argon and Ar/O2 serving workload, and run "
                         "them one by one (0 = skip)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MB",
                    help="first run a device copy of MB MiB (a kernel of known traffic, for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs one discharge case on one GPU; --gpus must be 1 "
                                   "(ensembles shard one case per GPU with no communication)"}))
        sys.exit(2)
    if args.copy_calib > 0:
        copy_calibration(args.copy_calib)

    from picmcc import FP32_REFINED, FP64_PLAIN, decks
    deck = decks.config2() if args.config == "c2" else decks.config1()

    # the served ensemble runs first, as on a freshly started server: after
    # the multi-GB allocations of the legs below, creating small cases is
    # measurably slower
    ensemble = run_ensemble(args.ensemble, args.ensemble_rate) if args.ensemble > 0 else None
    serve5 = run_serve5(args.serve5) if args.serve5 > 0 else None

    ref = run_gpu(deck, FP32_REFINED, args.steps, args.warmup, timing_steps=min(args.steps, 100))
    progress("FP32 refined steps")
    value = particles_avg(ref) * args.steps / ref["seconds"]
    plain, plain_value = None, 0.0
    if args.plain_steps > 0:
        plain = run_gpu(deck, FP64_PLAIN, args.plain_steps, 2)
        progress("FP64 plain steps")
        plain_value = particles_avg(plain) * args.plain_steps / plain["seconds"]
    # the same precision comparison with the same field solver on both sides
    same, same_value = None, 0.0
    if args.same_solver_steps > 0:
        same = run_gpu(deck, FP64_PLAIN, args.same_solver_steps, 5, timing_steps=min(args.same_solver_steps, 50),
                       field_solver=2)
        same_value = particles_avg(same) * args.same_solver_steps / same["seconds"]
        progress("FP64 steps on the separable solve")
    # the FP32 Jacobi PCG inside FP64 refinement (the solver for operators
    # that are not separable) on the same workload
    rpcg, rpcg_value = None, 0.0
    if args.refined_pcg_steps > 0:
        rpcg = run_gpu(deck, FP32_REFINED, args.refined_pcg_steps, 2, field_solver=1)
        progress("FP32 refined PCG steps")
        rpcg_value = particles_avg(rpcg) * args.refined_pcg_steps / rpcg["seconds"]

    # roofline of the dominant particle kernel (HIP events on the case stream)
    st = ref["stages_ms"]
    ne, ni = [0.5 * (a + b) for a, b in zip(ref["np0"], ref["np1"])]
    kernels = {
        # stage: (kernel, algorithmic bytes per launch, ms per launch); one
        # launch per stage; FP32 record = x, y, vx, vy, vz, id = 24 B
        "scatter_e": ("k_fused_scatter<float, electron>", 48.0 * ne, st.get("scatter_e", 0.0)),
        "scatter_i": ("k_fused_scatter<float, ion>", 48.0 * ni, st.get("scatter_i", 0.0)),
        "count_e": ("k_fused_count<float, electron>", 24.0 * ne, st.get("count_e", 0.0)),
        "count_i": ("k_fused_count<float, ion>", 24.0 * ni, st.get("count_i", 0.0)),
    }
    dom = max(kernels, key=lambda k: kernels[k][2])
    kname, bytes_, ms = kernels[dom]
    achieved = bytes_ / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    # measured HBM bytes per launch of that kernel: rocprofv3 FETCH_SIZE and
    # WRITE_SIZE of the same workload, scaled by a copy kernel of known traffic
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_ar_ccp_256x256.json")
    if args.config == "c2" and os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)["per_launch_bytes"].get(dom)
        traffic = t["total"] if t else None

    # the same kernel's shader counters (rocprofv3 SQ pass of the same workload)
    sq = None
    qpath = os.path.join(ROOT, "profiles", "sq_ar_ccp_256x256.json")
    if args.config == "c2" and os.path.exists(qpath):
        with open(qpath) as f:
            sq = json.load(f).get(dom)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ref["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Maxwellian loading of the deck; no datasets)",
        "config": {
            "workload": ("Ar CCP 256x256 cells dx=0.2mm, 200 V 13.56 MHz, 50 mTorr, 200 particles/cell/species, "
                         "dt=T_RF/800" if args.config == "c2" else
                         "Ar CCP 64x64 cells dx=0.5mm, 100 V 13.56 MHz, 100 mTorr, 50 particles/cell/species, "
                         "dt=T_RF/400"),
            "particles": int(sum(ref["np1"])),
            "field_solve": solver_label(ref, FP32_REFINED),
        },
        "fp64_plain": {"value": plain_value, "unit": "particle-steps/s",
                       "ms_per_step": plain["seconds"] / args.plain_steps * 1e3, "steps": args.plain_steps,
                       "field_solve": solver_label(plain, FP64_PLAIN),
                       "field_iters_last": plain["field_iters"],
                       "pcg_launch_tuned": {"blocks": plain["solver_shape"][0],
                                            "threads_per_block": plain["solver_shape"][1]}} if plain else None,
        # headline ratio: the FP32-refined path against the reference
        # algorithm in FP64 (its Jacobi PCG dominates the FP64 step)
        "refined_over_plain": value / plain_value if plain_value > 0 else None,
        "fp64_same_solver": {"value": same_value, "unit": "particle-steps/s",
                             "ms_per_step": same["seconds"] / args.same_solver_steps * 1e3,
                             "steps": args.same_solver_steps, "field_solve": solver_label(same, FP64_PLAIN),
                             "stage_ms_per_step": {k: round(v, 5) for k, v in same["stages_ms"].items()}}
        if same else None,
        # precision alone: FP32 against FP64 particles and fields, both with
        # the separable FP64 field solve
        "fp32_over_fp64_same_solver": value / same_value if same_value > 0 else None,
        "fp32_refined_pcg": {"value": rpcg_value, "unit": "particle-steps/s",
                             "ms_per_step": rpcg["seconds"] / args.refined_pcg_steps * 1e3,
                             "steps": args.refined_pcg_steps, "field_solve": solver_label(rpcg, FP32_REFINED),
                             "field_iters_last": rpcg["field_iters"], "field_cycles_last": rpcg["field_cycles"],
                             "field_rel_res_last": rpcg["field_rel_res"],
                             "pcg_launch_tuned": {"blocks": rpcg["solver_shape"][0],
                                                  "threads_per_block": rpcg["solver_shape"][1]}} if rpcg else None,
        "field_solve_last": {"cycles": ref["field_cycles"], "rel_residual": ref["field_rel_res"]},
        "stage_ms_per_step": {k: round(v, 5) for k, v in st.items()},
        # algorithmic HBM bytes of each particle stage over its time (particle
        # records only: 24 B read by a count pass, 24 B read + 24 B written by
        # a scatter, x and y of both species by the deposit)
        "stage_hbm_gbs": {k: round(b / (st[k] * 1e-3) / 1e9, 1) for k, b in
                          (("count_e", 24.0 * ne), ("count_i", 24.0 * ni), ("scatter_e", 48.0 * ne),
                           ("scatter_i", 48.0 * ni), ("deposit", 8.0 * (ne + ni))) if st.get(k, 0) > 0},
        "roofline": {"kernel": kname, "stage": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_, "ms_per_launch": ms},
    }
    if sq:
        # the measured limiter: an HBM-bound kernel below the HBM peak whose
        # waves mostly wait is limited by memory latency; one whose VALU is
        # mostly busy is limited by instruction issue
        lim = ("memory latency" if sq["wait_per_wave_cycle"] > sq["valu_per_wave_cycle"] else "VALU issue")
        line["roofline"]["counters"] = {"valu_busy": sq["valu_busy"], "valu_per_wave_cycle": sq["valu_per_wave_cycle"],
                                        "wait_per_wave_cycle": sq["wait_per_wave_cycle"],
                                        "source": "profiles/sq_ar_ccp_256x256.md"}
        line["roofline"]["limiter"] = lim
    if args.profile_cycles > 0:
        line["profile_error_vs_fp64"] = profile_check(deck, args.profile_cycles)
        progress("profile check")
    if args.pbatch > 0:
        line["batched_poisson"] = run_pbatch(args.pbatch)
        progress("batched Poisson")
    if args.o2_steps > 0:
        line["o2_chemistry"] = run_o2(args.o2_steps)
        progress("O2 steps")
    if args.aro2_steps > 0:
        line["aro2_chemistry"] = run_aro2(args.aro2_steps)
        progress("Ar/O2 steps")
    line["ensemble"] = ensemble if ensemble is not None else {"status": "not run (--ensemble 0)"}
    line["serving_workload"] = serve5 if serve5 is not None else {"status": "not run (--serve5 0)"}
    if args.cpu_steps > 0:
        th = cpu_threads()
        cv, ct = cpu_baseline(deck, args.cpu_steps, th)
        line["cpu_baseline"] = {"value": cv, "unit": "particle-steps/s", "cores": th, "kind": "port",
                                "sample": f"median of 3 runs of {args.cpu_steps} steps of the same workload on the "
                                          f"FP64 oracle's OpenMP step ({th} threads over cell tiles; run times "
                                          + ", ".join(f"{t:.1f}" for t in ct) + " s, setup excluded)",
                                "cpu": cpu_model(), "nproc": os.cpu_count()}
    if args.cpu_config1_steps > 0:
        th = cpu_threads()
        cv, ct = cpu_baseline(decks.config1(), args.cpu_config1_steps, th)
        line["cpu_baseline_64x64"] = {"value": cv, "unit": "particle-steps/s", "cores": th, "kind": "port",
                                      "sample": f"median of 3 runs of {args.cpu_config1_steps} steps of the 64x64 "
                                                "argon deck (409,600 particles)",
                                      "run_s": ct, "cpu": cpu_model(), "nproc": os.cpu_count()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
