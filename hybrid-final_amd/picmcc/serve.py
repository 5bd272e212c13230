"""Continuous-batching ensemble server for discharge cases on one GPU.

Requests (a deck and a step count, with an arrival time) are admitted while
the device memory budget allows it, charged with a footprint estimate from
the deck.  An admitted case is created (host setup, uploads) on a worker
thread while the loop keeps feeding the running cases; the library's calls
release the interpreter lock, and step graphs are captured in thread-local
mode, so creations on other threads do not disturb a capture.  Every case
keeps its particles, fields and random-stream state resident in HBM on its
own HIP stream.

The loop feeds each active case a slice of fused steps whenever its
previous slice has finished.  The FP32 separable-solver path replays each
step as one captured graph, so the host cost of a slice is small and the
slices of all active cases overlap on the device.  A case leaves when its
last slice completes; its slot and memory go to the next arrival.

Each case's result is independent of what else runs: the step is
deterministic per case, so a case run alone and the same case run in the
mix give the same bits.

SequentialServer serves the same requests first come, first served, with
one case resident at a time.
"""
import collections
import ctypes as C
from concurrent.futures import ThreadPoolExecutor
import math
import random
import time

import numpy as np

from . import FP32_REFINED, Batch, Case, PicError, _check, decks, lib


def device_free_bytes(device=0):
    """Free and total device memory, through the library's own HIP runtime."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(lib().pic_device_memory(device, C.byref(f), C.byref(t)))
    return f.value, t.value


def footprint_estimate(deck, precision=FP32_REFINED):
    """Device bytes a case needs, from its deck (an upper estimate)."""
    geo = deck["Geometry"]["SystemSpec"]
    ngx, ngy = int(geo["NumGridx"]), int(geo["NumGridy"])
    G = ngx * ngy
    R = 4 if precision == FP32_REFINED else 8
    caps = [int(v.get("Max_np", 0)) for k, v in deck["GasSpecies"]["ChargeSpecies"].items() if isinstance(v, dict)]
    capmax = max(caps) if caps else 0
    per_slot = 2 * (5 * R + 4) + 1 + 4 + 4          # two particle buffers, move tag, sort key, cell_start share
    shared = capmax * (6 * R + 4)                    # product staging / stash, product ranks
    nodes = G * (30 * R + 40 * 8 + 9 * 4 * 12)      # node fields, FP64 sums, route and counting tables
    solve = 2 * (ngx * ngx + ngy * ngy + 3 * G) * 8  # separable-solver matrices and work arrays
    return sum(caps) * per_slot + shared + nodes + solve + (16 << 20)


Request = collections.namedtuple("Request", "rid arrival deck steps")


def ar_ensemble(n, seed=7, rate_hz=40.0, grid=(32, 96), ppc=(20, 60), torr=(0.01, 0.5), volts=(50.0, 300.0),
                cycles=(1, 3), steps_per_cycle=400):
    """Poisson arrivals of argon CCP cases with log-uniform pressure and
    uniform grid size, particles per cell, RF amplitude and cycle count."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.expovariate(rate_hz)
        ncx = rng.randint(grid[0] // 8, grid[1] // 8) * 8
        ncy = rng.randint(grid[0] // 8, grid[1] // 8) * 8
        p = math.exp(rng.uniform(math.log(torr[0]), math.log(torr[1])))
        d = decks.ccp_deck(ncx=ncx, ncy=ncy, dx=5e-4, volts=rng.uniform(*volts), pres_torr=p,
                           ppc=rng.randint(*ppc), steps_per_cycle=steps_per_cycle, cycles=rng.randint(*cycles),
                           seed=1000 + i, philox_seed=5000 + i)
        steps = d["SimulationMethod"]["RFCycles"] * steps_per_cycle
        out.append(Request(i, t, d, steps))
    return out


def config5_requests(n=512, seed=2026, rate_hz=8.0, ar_fraction=0.5, grid=(64, 256), ppc=(20, 200),
                     torr=(0.01, 0.5), volts=(50.0, 300.0), cycles=(3, 20), o2_share=(0.05, 0.30),
                     steps_per_cycle=400, dx=2e-4, capacity_factor=4.0):
    """The serving workload: Poisson arrivals (rate_hz, seeded) of discharge
    cases, each argon or Ar/O2 (70-95 % argon), grid sides uniform over 64-256
    cells (dx = 0.2 mm), 20-200 particles per cell, pressure log-uniform over
    10-500 mTorr, 50-300 V at 13.56 MHz, 3-20 RF cycles of steps_per_cycle
    steps, seeded Maxwellian loading.  Returns the Request list (arrival times
    in seconds from the start)."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.expovariate(rate_hz)
        ncx = rng.randint(grid[0], grid[1])
        ncy = rng.randint(grid[0], grid[1])
        p = math.exp(rng.uniform(math.log(torr[0]), math.log(torr[1])))
        kw = dict(ncx=ncx, ncy=ncy, dx=dx, volts=rng.uniform(*volts), pres_torr=p, ppc=rng.randint(*ppc),
                  steps_per_cycle=steps_per_cycle, cycles=rng.randint(*cycles), seed=1000 + i,
                  philox_seed=5000 + i)
        if rng.random() < ar_fraction:
            d = decks.ccp_deck(capacity_factor=capacity_factor, **kw)
        else:
            d = decks.oxygen_deck(gas=2, o2_ratio=rng.uniform(*o2_share), capacity_factor=capacity_factor, **kw)
        out.append(Request(i, t, d, d["SimulationMethod"]["RFCycles"] * steps_per_cycle))
    return out


def case_gas(deck):
    return int(deck["GasSpecies"]["Type(0:Ar,1:O2,2:Ar/O2)"])


def _sync_status(case):
    """Waits for the case; 0, or the library's status when the case failed
    (for example -30 when it outgrew its particle capacity)."""
    try:
        case.sync()
        return 0
    except PicError as e:
        return e.code


class _Active:
    __slots__ = ("req", "case", "remaining", "inflight", "admit", "footprint", "np0", "batchable")

    def __init__(self, req, case, admit, footprint):
        self.req, self.case, self.admit, self.footprint = req, case, admit, footprint
        self.remaining = req.steps
        self.inflight = False
        self.batchable = None   # decided when a batching server first sees it
        inf = case.info()
        self.np0 = sum(inf.np[:inf.nsp])


# Concurrent serving was switched off after two GPU memory faults in runs
# that stepped several cases at once.  Their cause: the host setup read past
# the node array for the structure map's top-right corner on grids taller
# than wide (the served decks had such grids), so a route of the corner cell
# could carry an arbitrary conductor, dielectric or cell code into the
# kernels.  That read is fixed; CONCURRENT is on again.
CONCURRENT = True
DISABLED_REASON = "concurrent ensemble serving is switched off (serve.CONCURRENT)"


class EnsembleServer:
    """mem_fraction: share of the free device memory the server may fill;
    slice_steps: steps enqueued per case per feed; max_active: slot count;
    setup_threads: host threads that create admitted cases (deck parsing,
    geometry, operator, Laplace solves, particle loading, uploads) while the
    loop keeps feeding the running cases; 0 creates them in the loop.

    Admission charges the budget with the deck's footprint estimate (device
    memory is allocated by several creations at once, so a free-memory
    difference would not belong to one case).

    batch: step the small active argon cases on the separable field solve
    (at most batch_max_particles particles at admission) together through
    one pic_batch (each kernel of the step launched once for all of them; see
    picmcc.Batch), rebuilt whenever one joins or leaves; the batch advances
    by the smallest remaining step count of its members, at most
    slice_steps.  Other cases get their own slices as before.  Small cases
    are launch-bound and gain from the shared launches; larger ones gain
    more from overlapping on their own streams, where one case's kernels
    run beside another's instead of after a grid-wide stage boundary.

    run() raises if CONCURRENT is off (DISABLED_REASON)."""

    def __init__(self, precision=FP32_REFINED, mem_fraction=0.8, slice_steps=50, max_active=16, setup_threads=2,
                 batch=False, batch_max_particles=400_000):
        self.precision = precision
        self.slice = slice_steps
        self.max_active = max_active
        self.setup_threads = setup_threads
        self.batch = batch
        self.batch_max_particles = batch_max_particles
        free, _ = device_free_bytes()
        self.budget = int(free * mem_fraction)
        self._b = None          # the current Batch
        self._members = []      # its _Active entries
        self.rebuilds = 0

    def run(self, requests, keep_state=None, on_retire=None):
        """Serve `requests` in real time; returns one record per case.
        keep_state(rid) -> bool: return that case's final particles and density.
        on_retire(record): called as each case finishes."""
        if not CONCURRENT:
            raise RuntimeError(DISABLED_REASON)
        self._on_retire = on_retire
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        active, creating, results = [], [], []
        used = 0
        pool = ThreadPoolExecutor(self.setup_threads) if self.setup_threads > 0 else None
        t0 = time.perf_counter()
        try:
            while pending or active or creating:
                now = time.perf_counter() - t0
                # admission: arrivals that fit in the slot count and the memory budget
                while pending and pending[0].arrival <= now and len(active) + len(creating) < self.max_active:
                    est = footprint_estimate(pending[0].deck, self.precision)
                    if (active or creating) and used + est > self.budget:
                        break
                    req = pending.popleft()
                    used += est
                    admit = time.perf_counter() - t0
                    if pool is None:
                        active.append(_Active(req, Case(req.deck, self.precision), admit, est))
                    else:
                        creating.append((req, pool.submit(Case, req.deck, self.precision), admit, est))
                # cases whose host setup has finished join the running set
                for item in list(creating):
                    req, fut, admit, est = item
                    if fut.done():
                        creating.remove(item)
                        active.append(_Active(req, fut.result(), admit, est))
                if self.batch:
                    used -= self._feed_batch(active, results, keep_state, t0)
                used -= self._feed(active, results, keep_state, t0)
                busy = any(a.inflight for a in active) or bool(creating)
                if not busy and pending:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival - (time.perf_counter() - t0))))
                elif busy:
                    time.sleep(0.0002)
        finally:
            if pool is not None:
                pool.shutdown(wait=True)
            if self._b is not None:
                self._b.close()
                self._b, self._members = None, []
        return results

    def _retire(self, a, active, results, keep_state, t0):
        err = _sync_status(a.case)
        inf = a.case.info()
        np1 = sum(inf.np[:inf.nsp])
        rec = {"rid": a.req.rid, "arrival": a.req.arrival, "admit": a.admit,
               "finish": time.perf_counter() - t0, "steps": a.req.steps,
               "particle_steps": 0.5 * (a.np0 + np1) * a.req.steps, "footprint": a.footprint,
               "cells": (inf.ngx - 1) * (inf.ngy - 1), "gas": case_gas(a.req.deck), "error": err}
        rec["latency"] = rec["finish"] - rec["arrival"]
        if keep_state and keep_state(a.req.rid):
            rec["state"] = {"particles": [a.case.particles(s) for s in range(inf.nsp)],
                            "den": a.case.array("den")}
        a.case.close()
        active.remove(a)
        results.append(rec)
        if getattr(self, "_on_retire", None):
            self._on_retire(rec)
        return a.footprint

    def _feed_batch(self, active, results, keep_state, t0):
        """Steps the batchable cases together; returns the footprint freed by
        the ones that finished."""
        if self._members and any(not m.case.done() for m in self._members):
            return 0   # the last batch step is still running
        for m in self._members:
            m.inflight = False
        for a in active:
            if a.batchable is None:
                a.batchable = (case_gas(a.req.deck) == 0 and a.case.field_solver_kind() == 1
                               and a.np0 <= self.batch_max_particles)
        want = [a for a in active if a.batchable]
        freed = 0
        done = [a for a in want if a.remaining == 0]
        keep = [a for a in want if a.remaining > 0]
        if done or [id(a) for a in keep] != [id(m) for m in self._members]:
            if self._b is not None:
                self._b.close()
                self._b = None
            for a in done:
                freed += self._retire(a, active, results, keep_state, t0)
            self._members = keep
            if keep:
                try:
                    self._b = Batch([a.case for a in keep])
                    self.rebuilds += 1
                except PicError:
                    # a case the batch refuses (for example one whose field
                    # solve runs another cycle count) is served on its own
                    for a in keep:
                        a.batchable = False
                    self._members = []
        if self._b is not None:
            k = min(self.slice, min(a.remaining for a in self._members))
            self._b.step(k)
            for a in self._members:
                a.remaining -= k
                a.inflight = True
        return freed

    def _feed(self, active, results, keep_state, t0):
        """Gives every idle case its next slice of steps and retires the cases
        that have finished; returns the footprint freed by the retired ones."""
        freed = 0
        for a in list(active):
            if a.batchable:
                continue   # stepped by the batch
            if a.inflight:
                if not a.case.done():
                    continue
                a.inflight = False
            if a.remaining == 0:
                freed += self._retire(a, active, results, keep_state, t0)
                continue
            k = min(self.slice, a.remaining)
            a.case.step(k)
            a.remaining -= k
            a.inflight = True
        return freed


class SequentialServer:
    """First-come first-served, one case resident at a time.

    Each request waits until the device is free, then its case is created,
    stepped to the end with fused steps (captured graphs on the FP32
    separable-solver path), synchronised, read and destroyed.  Latency is
    finish minus arrival, so it includes the queueing behind earlier cases."""

    def __init__(self, precision=FP32_REFINED):
        self.precision = precision

    def run(self, requests, keep_state=None):
        results = []
        t0 = time.perf_counter()
        for req in sorted(requests, key=lambda r: r.arrival):
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            admit = time.perf_counter() - t0
            f0, _ = device_free_bytes()
            case = Case(req.deck, self.precision)
            f1, _ = device_free_bytes()
            inf = case.info()
            np0 = sum(inf.np[:inf.nsp])
            case.step(req.steps)
            err = _sync_status(case)
            inf = case.info()
            np1 = sum(inf.np[:inf.nsp])
            rec = {"rid": req.rid, "arrival": req.arrival, "admit": admit, "finish": time.perf_counter() - t0,
                   "steps": req.steps, "particle_steps": 0.5 * (np0 + np1) * req.steps,
                   "footprint": max(f0 - f1, 0), "cells": (inf.ngx - 1) * (inf.ngy - 1),
                   "gas": case_gas(req.deck), "error": err}
            rec["latency"] = rec["finish"] - rec["arrival"]
            rec["service"] = rec["finish"] - rec["admit"]
            if keep_state and keep_state(req.rid):
                rec["state"] = {"particles": [case.particles(s) for s in range(inf.nsp)], "den": case.array("den")}
            case.close()
            results.append(rec)
        return results


def run_alone(req, precision=FP32_REFINED, keep_state=False):
    """One case by itself: creation, all steps, synchronisation."""
    t0 = time.perf_counter()
    case = Case(req.deck, precision)
    inf = case.info()
    np0 = sum(inf.np[:inf.nsp])
    case.step(req.steps)
    err = _sync_status(case)
    inf = case.info()
    np1 = sum(inf.np[:inf.nsp])
    rec = {"rid": req.rid, "seconds": time.perf_counter() - t0, "particle_steps": 0.5 * (np0 + np1) * req.steps,
           "error": err}
    if keep_state:
        rec["state"] = {"particles": [case.particles(s) for s in range(inf.nsp)], "den": case.array("den")}
    case.close()
    return rec


def summarize(results, one_by_one=None):
    lat = np.array([r["latency"] for r in results])
    ps = sum(r["particle_steps"] for r in results)
    span = max(r["finish"] for r in results) - min(r["arrival"] for r in results)
    out = {"cases": len(results), "failed": sum(1 for r in results if r.get("error")),
           "p50_latency_s": float(np.percentile(lat, 50)),
           "p99_latency_s": float(np.percentile(lat, 99)), "makespan_s": span,
           "busy_fraction": (sum(r["service"] for r in results) / span) if "service" in results[0] else None,
           "aggregate_particle_steps_per_s": ps / span,
           "peak_footprint_bytes": max(r["footprint"] for r in results)}
    if one_by_one:
        tot = sum(r["seconds"] for r in one_by_one)
        out["one_by_one_seconds"] = tot
        out["one_by_one_particle_steps_per_s"] = sum(r["particle_steps"] for r in one_by_one) / tot
    return out
