"""A served stream of registration requests with Poisson arrivals through `Server`: the mixed
workload of a registration service (full normalisations at 2, 1.5 or 1 mm with one or two
channels, and rigid+affine-only fits), submitted at their arrival times and measured end to end.

    python -m fastwarp_amd.serve_stream --requests 64 --load 0.8

prints one JSON line: registrations/hour over the stream, latency percentiles from arrival to
result (queueing included), and the count of each kind. The arrival rate is `load` times the rate
the server reaches on the same mix with every request queued at once (measured first, on a
prefix of the mix). Volumes come from the seeded phantom generator (a few distinct pairs per grid,
reused across requests); fields are downloaded to host memory for every full request.
"""
from __future__ import annotations

import argparse
import json
import sys
import threading
import time

import numpy as np

from .phantom import make_pair
from .register import cohort_plan
from .serve import Server

# 1 mm grid of the cohort, and its 1.5 and 2 mm versions (same field of view)
GRIDS = {1.0: (182, 218, 182), 1.5: (121, 145, 121), 2.0: (91, 109, 91)}


def make_mix(n: int, seed: int = 6, full_share: float = 0.7):
    """Request kinds in arrival order: (spacing, channels, full) with 70 % full normalisations at
    2, 1.5 or 1 mm and one or two channels, 30 % rigid+affine only."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        full = rng.random() < full_share
        out.append((float(rng.choice([2.0, 1.5, 1.0])), int(rng.choice([1, 2])) if full else 1, bool(full)))
    return out


def pairs_for(mix):
    """One phantom pair per (spacing, channels) the mix uses."""
    cache = {}
    for spacing, channels, _ in mix:
        key = (spacing, channels)
        if key not in cache:
            p = make_pair(GRIDS[spacing], spacing, seed=int(10 * spacing) + channels, channels=channels)
            cache[key] = (p.grid, [np.asarray(v, np.float32) for v in p.fixed], [np.asarray(v, np.float32) for v in p.moving])
    return cache


def plan_for(channels: int, full: bool):
    p = cohort_plan(channels)
    p.run_syn = 1 if full else 0
    return p


def run_stream(server: Server, mix, pairs, rate_per_s: float | None, seed: int = 5):
    """Submits the mix at Poisson arrival times (all at once when `rate_per_s` is None) and waits
    for every result, failed ones included; returns (seconds from the first submission to the last
    result, per-request latencies, failure messages). A request that could not be submitted counts
    as failed, and the stream still waits for every request it did submit."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, len(mix)) if rate_per_s else np.zeros(len(mix))
    tickets = [None] * len(mix)
    submitted = [0.0] * len(mix)
    latency = [0.0] * len(mix)
    failures: list[str] = []
    ready = threading.Semaphore(0)

    def collect():
        # the server times each request from its submission to its end
        for i in range(len(mix)):
            ready.acquire()
            if tickets[i] is None:
                continue
            try:
                latency[i] = server.wait(tickets[i]).timing.total_s
            except Exception as e:  # a failed request: counted, and the stream goes on
                failures.append(f"request {i}: {e}")

    waiter = threading.Thread(target=collect)
    waiter.start()
    t0 = time.perf_counter()
    due = t0
    released = 0
    try:
        for i, (spacing, channels, full) in enumerate(mix):
            due += gaps[i]
            pause = due - time.perf_counter()
            if pause > 0:
                time.sleep(pause)
            grid, f, m = pairs[(spacing, channels)]
            submitted[i] = time.perf_counter()
            try:
                tickets[i] = server.submit(plan_for(channels, full), grid, f, grid, m, fields=full)
            except Exception as e:
                failures.append(f"request {i} not submitted: {e}")
            ready.release()
            released += 1
    finally:
        # the collector waits for every submitted ticket, also when submission stopped early
        for _ in range(len(mix) - released):
            ready.release()
        waiter.join()
    wall = max(sub - t0 + lat for sub, lat in zip(submitted, latency))
    return wall, latency, failures


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--load", type=float, default=0.8, help="arrival rate as a share of the saturated rate")
    ap.add_argument("--slots", type=int, default=4)
    args = ap.parse_args(argv)
    mix = make_mix(args.requests)
    pairs = pairs_for(mix)
    with Server(slots=args.slots) as s:
        # warm every slot on every kind of request, then the saturated rate on a prefix of the mix
        _, _, warm_failures = run_stream(s, sorted(set(mix)), pairs, None)
        prefix = mix[: max(8, len(mix) // 4)]
        sat_wall, _, sat_failures = run_stream(s, prefix, pairs, None)
        saturated = len(prefix) / sat_wall
        wall, lat, failures = run_stream(s, mix, pairs, args.load * saturated)
    failures = warm_failures + sat_failures + failures
    lat = np.array(lat)
    kinds = {}
    for spacing, channels, full in mix:
        key = f"{'full' if full else 'linear'} {spacing} mm x{channels}"
        kinds[key] = kinds.get(key, 0) + 1
    print(json.dumps({"requests": len(mix), "slots": args.slots, "load": args.load,
                      "saturated_per_hour": round(3600.0 * saturated, 1),
                      "arrival_per_hour": round(3600.0 * args.load * saturated, 1),
                      "served_per_hour": round(3600.0 * len(mix) / wall, 1),
                      "latency_s": {q: round(float(np.percentile(lat, int(q[1:]))), 3) for q in ("p50", "p95", "p99")},
                      "kinds": dict(sorted(kinds.items())), "failed": len(failures),
                      "failures": failures[:8]}), flush=True)
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
