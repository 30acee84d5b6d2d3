"""Request loop with continuous batching of resident integrations.

Requests (one integration each: grid, stratification, shear and initial
packet drawn per member, a step count and a precision) arrive over time.  The
server admits them, in arrival order, as soon as the HBM budget allows,
builds each one on the device (grids, background, calc_parameter, E from the
separable initializer, dt at CFL 0.5) on a set-up thread while the device
steps the others, and advances all admitted integrations of one precision
together in a persistent-kernel ensemble step (rt6d_ensemble_step, two steps
per HIP graph replay).  Every tick advances each ensemble by
min(chunk, steps left of its members); members whose steps are done leave
their ensemble, their result (E, and the energy budget when asked) is read,
their memory freed, and queued requests take the room.  State stays on the
device from admission to completion, unless preemption is on: then a request
that has waited long enough at the head of the queue may preempt a resident
one with many steps left. The resident one is checkpointed (its E, in memory
or as a restart file in the reference's layout, raytracer_3d_amd.restart),
its device memory freed, and it queues again right behind the newcomer.
When it is re-admitted it is rebuilt, resumes from the checkpoint and
finishes its remaining steps with the same result.

    python -m raytracer_3d_amd.serve --rate 40 --duration 120 --budget-gb 200

prints one JSON line: p50/p99 completion latency (arrival to result), the
aggregate cell-updates/s of the served run, and the same requests run one at
a time with the plain binary64 step (``--baseline-seconds`` bounds that
sample; its latency is the FIFO single-server queue of the measured service
times under the same arrivals).

Results do not depend on the company a request keeps: the ensemble runs the
same column-kernel arithmetic as a single integration, so each result equals
the one computed alone (checked bit for bit by the GPU tests).
"""
from __future__ import annotations

import argparse
import collections
import dataclasses
import json
import math
import queue
import sys
import threading
import time

import numpy as np

from . import cases, rt6d

TUNE_STEPS = 24 * (2 + 2)   # rt6d_ensemble_tune with trial_steps = 2: 3 tile choices x 8 schedules x 4 steps


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    member: int             # member index of the generator
    n: tuple
    steps: int
    precision: int          # rt6d.FP64 or rt6d.FP32
    seed: int = 4242
    # filled by the server
    admitted: float = math.nan
    started: float = math.nan
    done: float = math.nan
    result: dict | None = None
    preemptions: int = 0
    resume: dict | None = None   # checkpoint of a preempted request: steps done / left, E or restart path
    resident: int = 0            # device bytes charged while the member is resident

    @property
    def cells(self) -> int:
        return int(np.prod(self.n))


def request_stream(rate: float, duration: float, seed: int = 4242, steps_range=(50, 1000), fp32_share=0.75,
                   n=None, first_member: int = 0):
    """Poisson arrivals at `rate` per second for `duration` seconds; each
    request is a member of the generator (grids from {12..24}^3 x {8..16}^3),
    with steps ~ U[steps_range] and FP32 with probability `fp32_share`."""
    out = []
    t = 0.0
    i = 0
    while True:
        u = cases.uniform(seed, first_member + i, np.arange(4, dtype=np.uint64), stream=11)
        t += -math.log1p(-u[0]) / rate
        if t >= duration:
            break
        spec_n = n if n is not None else cases.member_spec(seed, first_member + i).n
        steps = int(steps_range[0] + (steps_range[1] - steps_range[0] + 1) * u[1])
        steps = min(steps, steps_range[1])
        prec = rt6d.FP32 if u[2] < fp32_share else rt6d.FP64
        out.append(Request(i, t, first_member + i, tuple(spec_n), steps, prec, seed))
        i += 1
    return out


def _field_bytes(n, es):
    """Bytes of the 6D coefficient fields (six face-velocity fields, lambda)
    and of the homogeneous path's (z,k,l,m) tables and plane records."""
    cells = int(np.prod(n))
    faces = sum(cells // n[d] * (n[d] + 1) for d in range(6))
    n4 = cells // (n[0] * n[1])
    tables = 3 * n4 * 2 + n4 // n[2] * (n[2] + 1) + n4 // n[5] * (n[5] + 1) + n4
    return es * (faces + cells), es * (tables + 24 * n4)


def resident_bytes(n, precision, homogeneous=True) -> int:
    """Device bytes of an integration after calc_parameter: E twice, grids,
    coefficients, background, and either the homogeneous path's tables and
    plane records (its 6D fields are freed) or the six face-velocity fields
    and lambda."""
    cells = int(np.prod(n))
    es = 8 if precision == rt6d.FP64 else 4
    n3h = (n[0] + 4) * (n[1] + 4) * (n[2] + 4)
    fields, tables = _field_bytes(n, es)
    fixed = es * 2 * cells + 8 * (3 * n3h) + 8 * 7 * sum(v + 4 for v in n) + (1 << 16)
    return fixed + (tables if homogeneous else fields)


def peak_bytes(n, precision) -> int:
    """Device bytes while calc_parameter runs: the 6D fields (binary64 first,
    then rounded for FP32) and the tables exist together."""
    cells = int(np.prod(n))
    faces = sum(cells // n[d] * (n[d] + 1) for d in range(6))
    es = 8 if precision == rt6d.FP64 else 4
    fields, tables = _field_bytes(n, es)
    extra = 0 if precision == rt6d.FP64 else 8 * (faces + cells)
    return resident_bytes(n, precision, False) + tables + extra


def build_member(req: Request, cfl: float = 0.5):
    """Device-resident integration of a request (calc_parameter done, E set;
    a preempted request resumes from its checkpoint)."""
    spec = cases.member_spec(req.seed, req.member, n=req.n)
    m = rt6d.Model.from_case(spec, precision=req.precision)
    m.calc_parameter()
    m.dt = m.cfl_dt(cfl)
    if req.resume is not None:
        if "path" in req.resume:
            from . import restart
            itt = restart.load(m, req.resume["path"])
            if itt != req.resume["done"]:
                raise RuntimeError(f"request {req.rid}: restart file at step {itt}, expected {req.resume['done']}")
        else:
            m.upload_E(req.resume["E"])
    return m


class Server:
    """Continuous batching over one ensemble per precision."""

    def __init__(self, budget_bytes: float, chunk: int = 8, max_members: int = 1024, keep_E: bool = False,
                 diagnose: bool = False, tune: bool = True, tune_min_members: int = 8,
                 preempt_after: float | None = None, checkpoint_dir: str | None = None):
        """preempt_after: seconds a request may wait at the head of the queue
        for HBM room before it preempts a resident request (None: never).
        checkpoint_dir: write preempted requests' state as restart files there
        (else keep E in host memory)."""
        self.budget = float(budget_bytes)
        self.preempt_after = preempt_after
        self.checkpoint_dir = checkpoint_dir
        self.preemptions = 0
        self.chunk = int(chunk)
        self.max_members = int(max_members)
        self.keep_E = keep_E
        self.diagnose = diagnose
        self.tune = tune
        self.tune_min_members = int(tune_min_members)
        self.ens = {rt6d.FP64: rt6d.Ensemble(rt6d.FP64), rt6d.FP32: rt6d.Ensemble(rt6d.FP32)}
        self.tuned = {rt6d.FP64: False, rt6d.FP32: False}
        self.used = 0.0
        self.peak_used = 0.0
        self.ticks = 0
        self.device_ms = 0.0

    def close(self):
        for e in self.ens.values():
            e.close()

    def run(self, requests, log=None):
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = collections.deque(reqs)
        waiting = collections.deque()
        ready = queue.Queue()
        active = {rt6d.FP64: {}, rt6d.FP32: {}}   # slot -> (req, model, steps left)
        building = 0
        lock = threading.Lock()
        jobs = queue.Queue()

        def builder():
            while True:
                r = jobs.get()
                if r is None:
                    return
                try:
                    m = build_member(r)
                    ready.put((r, m, None))
                except Exception as e:  # noqa: BLE001 -- reported to the loop
                    ready.put((r, None, e))

        th = threading.Thread(target=builder, daemon=True)
        th.start()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        done = 0
        total = len(reqs)
        try:
            while done < total:
                t = now()
                while pending and pending[0].arrival <= t:
                    waiting.append(pending.popleft())
                # admission in arrival order against the HBM budget
                while waiting and building + sum(len(a) for a in active.values()) < self.max_members:
                    r = waiting[0]
                    need = peak_bytes(r.n, r.precision)
                    with lock:
                        fits = not (self.used + need > self.budget and (building or any(active.values())))
                    if not fits:
                        victim = self._victim(r, active, now())
                        if victim is None:
                            break
                        self._preempt(victim, active, lock)
                        waiting.insert(1, victim[1])
                        continue
                    with lock:
                        self.used += need
                        self.peak_used = max(self.peak_used, self.used)
                    waiting.popleft()
                    r.admitted = now()
                    building += 1
                    jobs.put(r)
                # members built since the last tick join their ensemble
                while True:
                    try:
                        r, m, err = ready.get_nowait()
                    except queue.Empty:
                        break
                    building -= 1
                    if err is not None:
                        raise err
                    r.resident = resident_bytes(r.n, r.precision, bool(m.query(rt6d.Q_HOMOGENEOUS)))
                    with lock:
                        self.used += r.resident - peak_bytes(r.n, r.precision)
                    slot = self.ens[r.precision].add(m)
                    if not np.isfinite(r.started):
                        r.started = now()
                    active[r.precision][slot] = [r, m, r.resume["left"] if r.resume else r.steps]
                if not any(active.values()):
                    if building:
                        r, m, err = ready.get()   # wait for the set-up thread
                        ready.put((r, m, err))
                        continue
                    if pending:
                        time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
                    continue
                launched = []
                for prec, act in active.items():
                    if not act:
                        continue
                    e = self.ens[prec]
                    if self.tune and not self.tuned[prec] and len(act) >= self.tune_min_members:
                        left = min(v[2] for v in act.values())
                        if left >= TUNE_STEPS + self.chunk:
                            e.tune(trial_steps=2)   # counted in the members' steps
                            for v in act.values():
                                v[2] -= TUNE_STEPS
                            self.tuned[prec] = True
                    s = min(self.chunk, min(v[2] for v in act.values()))
                    e.step_async(s)    # both precisions' ensembles run on the device together
                    launched.append((prec, s))
                for prec, s in launched:
                    e = self.ens[prec]
                    act = active[prec]
                    self.device_ms += e.wait()
                    self.ticks += 1
                    finished = []
                    for slot, v in act.items():
                        v[2] -= s
                        if v[2] == 0:
                            finished.append(slot)
                    for slot in finished:
                        r, m, _ = act.pop(slot)
                        e.remove(slot)
                        res = {"steps": m.query(rt6d.Q_STEPS) + (r.resume["done"] if r.resume else 0), "dt": m.dt}
                        if self.diagnose:
                            d = m.diagnose()
                            res.update(eint=d["eint"], residual=d["residual"])
                        if self.keep_E:
                            res["E"] = m.download("E")
                        r.result = res
                        m.close()
                        r.done = now()
                        with lock:
                            self.used -= r.resident
                        done += 1
                        if log and done % 50 == 0:
                            log(f"served {done}/{total} at t={r.done:.1f}s, resident "
                                f"{sum(len(a) for a in active.values())}, used {self.used / 2**30:.1f} GiB")
        finally:
            jobs.put(None)
            th.join(timeout=30)
        return now()

    def _victim(self, r: Request, active, t: float):
        """The resident request that the queue head r preempts, or None: r has
        waited preempt_after seconds, neither was preempted before, and the
        victim has more than two chunks of steps left (the one with the most)."""
        if self.preempt_after is None or r.preemptions or t - r.arrival < self.preempt_after:
            return None
        best = None
        for prec, act in active.items():
            for slot, v in act.items():
                q, _, left = v
                if q.preemptions or left <= 2 * self.chunk:
                    continue
                if best is None or left > best[2]:
                    best = (prec, q, left, slot)
        return None if best is None else (best[0], best[1], best[3])

    def _preempt(self, victim, active, lock):
        prec, q, slot = victim
        _, m, left = active[prec][slot]
        done = q.steps - left
        # checkpoint first: if it fails, the victim is still resident and
        # in its ensemble, and the server's accounting is unchanged
        if self.checkpoint_dir:
            import os
            from . import restart
            path = os.path.join(self.checkpoint_dir, f"restart_request{q.rid:06d}.dta")
            restart.save(m, path, done, ritt=False)
            resume = {"done": done, "left": left, "path": path}
        else:
            resume = {"done": done, "left": left, "E": m.download("E")}
        active[prec].pop(slot)
        self.ens[prec].remove(slot)
        q.resume = resume
        m.close()
        with lock:
            self.used -= q.resident
        q.preemptions += 1
        self.preemptions += 1


def run_alone(req: Request, precision=None):
    """The request run on its own (plain single-integration steps); returns
    (seconds, E)."""
    r = Request(req.rid, req.arrival, req.member, req.n, req.steps,
                req.precision if precision is None else precision, req.seed)   # none of the server's fields
    m = build_member(r)
    m.synchronize()
    t0 = time.perf_counter()
    m.integrate(r.steps)
    t = time.perf_counter() - t0
    E = m.download("E")
    m.close()
    return t, E


def fifo_latencies(arrivals, service):
    """Completion latencies of a FIFO single server."""
    free = 0.0
    out = []
    for a, s in zip(arrivals, service):
        start = max(a, free)
        free = start + s
        out.append(free - a)
    return out


def serve_benchmark(rate=40.0, duration=120.0, budget_gb=200.0, chunk=8, baseline_seconds=20.0, seed=4242,
                    log=None, preempt_after=None, checkpoint_dir=None):
    reqs = request_stream(rate, duration, seed=seed)
    work = sum(r.cells * r.steps for r in reqs)
    srv = Server(budget_gb * 2**30, chunk=chunk, preempt_after=preempt_after, checkpoint_dir=checkpoint_dir)
    try:
        wall = srv.run(reqs, log=log)
    finally:
        srv.close()
    lat = np.array([r.done - r.arrival for r in reqs])
    served = {
        "requests": len(reqs), "cell_updates": work, "wall_s": wall, "value": work / wall,
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()), "device_ms": srv.device_ms, "ticks": srv.ticks,
        "peak_hbm_gib": srv.peak_used / 2**30,
        "fp32_share": float(np.mean([r.precision == rt6d.FP32 for r in reqs])),
        "preemptions": srv.preemptions,
    }
    # the same requests one at a time with the plain binary64 step (bounded
    # sample): the step loop alone, and the whole request (set-up included, as
    # the served wall time includes it)
    svc, svc_full, wk = [], [], 0
    t_start = time.perf_counter()
    for r in reqs:
        if time.perf_counter() - t_start > baseline_seconds:
            break
        t0 = time.perf_counter()
        t, _ = run_alone(r, precision=rt6d.FP64)
        svc_full.append(time.perf_counter() - t0)
        svc.append(t)
        wk += r.cells * r.steps
    base_rate = wk / sum(svc) if svc else None
    base_rate_full = wk / sum(svc_full) if svc_full else None
    # the whole stream served one at a time at the sampled per-request rate
    est = [r.cells * r.steps / base_rate for r in reqs] if base_rate else []
    blat = np.array(fifo_latencies([r.arrival for r in reqs], est)) if est else np.array([math.nan])
    baseline = {
        "kind": "one at a time, plain binary64 step (rt6d_step); value: step loop only, "
                "value_with_setup: build + steps + read-back of each request",
        "sampled_requests": len(svc), "value": base_rate, "value_with_setup": base_rate_full,
        "p50_latency_s": float(np.percentile(blat, 50)), "p99_latency_s": float(np.percentile(blat, 99)),
    }
    served["speedup_vs_one_at_a_time_fp64"] = (served["value"] / base_rate) if base_rate else None
    served["speedup_vs_one_at_a_time_fp64_with_setup"] = (served["value"] / base_rate_full) if base_rate_full else None
    return {"served": served, "one_at_a_time_fp64": baseline,
            "scenario": {"rate_per_s": rate, "duration_s": duration, "budget_gib": budget_gb, "chunk": chunk,
                         "steps": "U[50,1000]", "fp32:fp64": "3:1", "members": "{12,16,20,24}^3 x {8,12,16}^3"}}


def ensemble_benchmark(members=96, steps=300, precision=rt6d.FP64, seed=4242, alone_members=None, log=None):
    """Batched ensemble: `members` generator members in one ensemble,
    autotuned, then `steps` timed steps; against the same members stepped
    one at a time (a sample of `alone_members`, all if None)."""
    reqs = [Request(i, 0.0, i, cases.member_spec(seed, i).n, steps, precision, seed) for i in range(members)]
    t0 = time.perf_counter()
    ms = [build_member(r) for r in reqs]
    setup = time.perf_counter() - t0
    cells = sum(r.cells for r in reqs)
    with rt6d.Ensemble(precision) as ens:
        for m in ms:
            ens.add(m)
        bpc, order, tiles = ens.tune(trial_steps=2)
        launches = ens.query(rt6d.EQ_LAUNCHES)
        ens.step(2)
        t1 = time.perf_counter()
        dev_ms = ens.step(steps)
        wall = time.perf_counter() - t1
    for m in ms:
        m.close()
    # the same members one at a time, on the same clock as the batched run
    # (host wall time from the first launch to the end of the last step)
    sample = reqs if alone_members is None else reqs[:alone_members]
    t_alone, d_alone, w_alone = 0.0, 0.0, 0
    for r in sample:
        m = build_member(r)
        m.integrate(2)
        m.synchronize()
        t0 = time.perf_counter()
        d_alone += m.integrate_timed(steps)
        m.synchronize()
        t_alone += time.perf_counter() - t0
        w_alone += r.cells * steps
        m.close()
    out = {"members": members, "steps": steps, "cells": cells, "setup_s": setup, "schedule": [bpc, order], "tiles": tiles,
           "launches_per_step": launches,
           "wall_s": wall, "device_ms": dev_ms, "value": cells * steps / wall,
           "one_at_a_time": w_alone / t_alone if t_alone else None, "one_at_a_time_members": len(sample),
           "one_at_a_time_wall_s": t_alone, "one_at_a_time_device_ms": d_alone,
           "clock": "host wall time around the timed steps, batched and one at a time alike"}
    out["speedup_vs_one_at_a_time"] = out["value"] / out["one_at_a_time"] if out["one_at_a_time"] else None
    if log:
        log(f"ensemble {members} members x {steps} steps ({'fp32' if precision else 'fp64'}): "
            f"{out['value']:.3e} cell-updates/s, one at a time {out['one_at_a_time']:.3e}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=40.0)
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--chunk", type=int, default=8)
    ap.add_argument("--baseline-seconds", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=4242)
    ap.add_argument("--preempt-after", type=float, default=None,
                    help="seconds a request may wait for HBM room before it preempts a resident one")
    ap.add_argument("--checkpoint-dir", default=None, help="write preempted state as restart files there")
    a = ap.parse_args(argv)
    out = serve_benchmark(a.rate, a.duration, a.budget_gb, a.chunk, a.baseline_seconds, a.seed,
                          log=lambda s: print("[serve] " + s, file=sys.stderr, flush=True),
                          preempt_after=a.preempt_after, checkpoint_dir=a.checkpoint_dir)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
