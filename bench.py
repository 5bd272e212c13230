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
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hybrid-final_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s, FP32-refined vs FP64 plain; p50/p99 ensemble case latency"


def run_gpu(deck, precision, steps, warmup, timing_steps=0):
    """Warm up, time `steps` steps with no instrumentation, then (optionally)
    `timing_steps` more steps with HIP events around every stage kernel on the
    case's stream (the per-kernel times; the events add launch gaps, so they
    are kept out of the throughput run)."""
    from picmcc import Case
    c = Case(deck, precision)
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
    }
    c.close()
    return out


def particles_avg(r):
    return 0.5 * (sum(r["np0"]) + sum(r["np1"]))


def cpu_baseline(deck, steps):
    """Oracle (FP64 CPU restatement, one thread) on `steps` steps of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle(deck, "f64")
    o.orc_deposit()
    n0 = o.scalar("np0") + o.scalar("np1")
    t0 = time.perf_counter()
    rc = o.orc_run(steps)
    dt = time.perf_counter() - t0
    n1 = o.scalar("np0") + o.scalar("np1")
    if rc != 0:
        raise RuntimeError(f"oracle run failed with {rc}")
    return 0.5 * (n0 + n1) * steps / dt, dt


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
    out = serve.summarize(serve.EnsembleServer(max_active=16, setup_threads=2).run(reqs), alone)
    out["mode"] = ("concurrent: continuous admission by device-memory budget (at most 16 resident), one stream "
                   "per case, graph replays, case setup on 2 worker threads")
    seq = serve.summarize(serve.SequentialServer().run(reqs))
    out["sequential"] = {k: seq[k] for k in ("p50_latency_s", "p99_latency_s", "makespan_s", "busy_fraction",
                                             "aggregate_particle_steps_per_s")}
    out["workload"] = (f"{n} Ar CCP cases, grids 32^2-96^2, 20-60 particles/cell, 10-500 mTorr log-uniform, "
                       f"50-300 V, 1-3 RF cycles of 400 steps, Poisson arrivals {rate_hz:g}/s, latency = finish - "
                       "arrival incl. setup and queueing (Ar/O2 cases need the O2 collision step, not built)")
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
    ap.add_argument("--cpu-steps", type=int, default=10, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--pbatch", type=int, default=256, metavar="N",
                    help="independent Poisson problems (64^2-512^2) in the batched-solve leg (0 = skip)")
    ap.add_argument("--o2-steps", type=int, default=20, metavar="K",
                    help="staged O2-chemistry steps timed on a 256x256 O2 case (0 = skip)")
    ap.add_argument("--ensemble", type=int, default=16, metavar="N",
                    help="cases in the served argon ensemble (Poisson arrivals; concurrent server, first-come "
                         "first-served server and one by one; 0 = skip)")
    ap.add_argument("--ensemble-rate", type=float, default=32.0, metavar="HZ",
                    help="mean arrival rate of the ensemble's cases")
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

    ref = run_gpu(deck, FP32_REFINED, args.steps, args.warmup, timing_steps=min(args.steps, 100))
    value = particles_avg(ref) * args.steps / ref["seconds"]
    plain, plain_value = None, 0.0
    if args.plain_steps > 0:
        plain = run_gpu(deck, FP64_PLAIN, args.plain_steps, 2)
        plain_value = particles_avg(plain) * args.plain_steps / plain["seconds"]

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
            "field_solve": "FP32 separable inverse in FP64 refinement, tol 1e-10 relative"
            if ref["solver_kind"] == 1 else "FP32 Jacobi PCG in FP64 refinement, tol 1e-10 relative",
        },
        "fp64_plain": {"value": plain_value, "unit": "particle-steps/s",
                       "ms_per_step": plain["seconds"] / args.plain_steps * 1e3, "steps": args.plain_steps,
                       "field_solve": "FP64 Jacobi PCG (reference algorithm), tol 1e-10 relative",
                       "field_iters_last": plain["field_iters"]} if plain else None,
        "refined_over_plain": value / plain_value if plain_value > 0 else None,
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
    if args.pbatch > 0:
        line["batched_poisson"] = run_pbatch(args.pbatch)
    if args.o2_steps > 0:
        line["o2_chemistry"] = run_o2(args.o2_steps)
    line["ensemble"] = ensemble if ensemble is not None else {"status": "not run (--ensemble 0)"}
    if args.cpu_steps > 0:
        cv, cdt = cpu_baseline(deck, args.cpu_steps)
        line["cpu_baseline"] = {"value": cv, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                                "sample": f"{args.cpu_steps} steps of the same workload on the FP64 oracle "
                                          f"({cdt:.1f} s of CPU time, setup excluded)"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
