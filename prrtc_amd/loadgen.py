"""Poisson request streams through the continuous-batching server.

    python -m prrtc_amd.loadgen --rates 200,500,1000,2000 --seconds 10 --out serving_sweep.json

Requests are drawn in turn from a pool of seeded problems
(``workloads.serving_request``: Panda : Fetch : Baxter = 2 : 1 : 1, a fresh
4-128 primitive scene each, collision-free start and goal); the k-th pass
over the pool plans with seed k + 1.  Arrivals are exponential with mean
1 / rate; each request is submitted at its arrival time and its latency is
arrival -> result (admission wait + time resident).  A request refused with
EBUSY (queue full) is counted as rejected.  ``plans_per_s`` is solved
requests over the whole run (first arrival to last result), so an
overloaded stream reports the server's sustained capacity.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

# planner settings of the serving stream, per robot (pRRTC_settings fields)
SERVE_SETTINGS = dict(num_new_configs=64, granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
                      max_samples=50000, max_iters=3000)


def pct(a, q):
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else float("nan")


def request_pool(n: int, seed: int = 0) -> list:
    from . import workloads
    return [workloads.serving_request(i, seed=seed) for i in range(n)]


def run_stream(pool: list, rate: float, seconds: float, *, arrival_seed: int = 11, budget_bytes: int = 16 << 30,
               slice_us: int = 1000, young_rounds: int = 0, settings: dict | None = None,
               deadline_ms: int = 0, progress=None) -> dict:
    """One Poisson stream of rate * seconds requests; returns its summary
    plus the per-request (request, result, path, timing) tuples under
    "_outs".  progress(done, total) is called about every 20 s while the
    results are collected."""
    import prrtc_amd as P
    from .lib import PRRTC_EBUSY, PrrtcError
    st = dict(SERVE_SETTINGS, **(settings or {}))
    n_req = max(1, int(rate * seconds))
    arrivals = np.cumsum(np.random.default_rng(arrival_seed).exponential(1.0 / rate, n_req))
    submitted, rejected = [], 0
    with P.Server(budget_bytes=budget_bytes, slice_us=slice_us, young_rounds=young_rounds,
                  deadline_ms=deadline_ms) as srv:
        t_start = time.perf_counter()
        for i, a in enumerate(arrivals):
            delay = t_start + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            rq = pool[i % len(pool)]
            try:
                submitted.append((rq, srv.submit(rq, dict(st, seed=1 + i // len(pool)))))
            except PrrtcError as e:
                if e.code != PRRTC_EBUSY:
                    raise
                rejected += 1
        t_arrived = time.perf_counter()
        outs, t_note = [], t_arrived
        for rq, t in submitted:
            outs.append((rq, *srv.poll(t)))
            if progress and time.perf_counter() - t_note > 20.0:
                t_note = time.perf_counter()
                progress(len(outs), len(submitted))
        t_end = time.perf_counter()
        sstats = srv.stats()
    tm = [o[3] for o in outs]
    lat = [(t["queued_ns"] + t["resident_ns"]) / 1e6 for t in tm]
    queued = [t["queued_ns"] / 1e6 for t in tm]
    solved = [o[1]["solved"] for o in outs]
    lat_solved = [l for l, s in zip(lat, solved) if s]
    wall = t_end - t_start
    by_robot = {}
    for (rq, r, _, _), l in zip(outs, lat):
        d = by_robot.setdefault(rq["robot"], {"requests": 0, "solved": 0, "lat": []})
        d["requests"] += 1
        d["solved"] += int(r["solved"])
        d["lat"].append(l)
    return {
        "rate_per_s": rate, "seconds": seconds, "requests": n_req, "rejected": rejected, "solved": int(sum(solved)),
        "plans_per_s": float(sum(solved)) / wall, "attempted_per_s": len(outs) / wall,
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "p50_ms_solved": pct(lat_solved, 50), "p99_ms_solved": pct(lat_solved, 99),
        "admission_wait_p50_ms": pct(queued, 50), "admission_wait_p99_ms": pct(queued, 99),
        "drain_s": t_end - t_arrived, "wall_s": wall, "slices": sstats["slices"],
        "deadline_ms": deadline_ms, "timed_out": sstats["timed_out"],
        "peak_resident_mib": sstats["peak_resident_bytes"] / 2**20,
        "per_robot": {k: {"requests": v["requests"], "solved": v["solved"], "p50_ms": pct(v["lat"], 50),
                          "p99_ms": pct(v["lat"], 99)} for k, v in by_robot.items()},
        "settings": "g {granularity}, nnc {num_new_configs}, range {range}, balance {balance}, max_iters {max_iters}; "
                    "{slice} us slices, {budget} GiB budget, young_rounds {young}, deadline {dl}".format(
                        **st, slice=slice_us, budget=budget_bytes >> 30, young=young_rounds or 64,
                        dl=f"{deadline_ms} ms" if deadline_ms else "none"),
        "_outs": outs,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--rates", default="200,500,1000,2000", help="requests/s, comma-separated")
    ap.add_argument("--seconds", type=float, default=10.0, help="arrival window per rate")
    ap.add_argument("--pool", type=int, default=2000, help="distinct seeded problems")
    ap.add_argument("--young-rounds", type=int, default=0)
    ap.add_argument("--deadline-ms", type=int, default=0, help="per-request deadline (0: none)")
    ap.add_argument("--out", default=None, help="write the summaries as JSON")
    ap.add_argument("--tuned", default=None, help="tuned launch shapes to apply (default: prrtc_amd/tuned.json)")
    args = ap.parse_args(argv)
    import prrtc_amd as P
    from prrtc_amd import autotune
    P.load()
    autotune.apply(args.tuned or autotune.DEFAULT_OUT)
    print(f"[loadgen] building {args.pool} requests", file=sys.stderr, flush=True)
    pool = request_pool(args.pool)
    rows = []
    for rate in (float(r) for r in args.rates.split(",")):
        print(f"[loadgen] {rate:.0f} req/s for {args.seconds:.0f} s", file=sys.stderr, flush=True)
        r = run_stream(pool, rate, args.seconds, young_rounds=args.young_rounds, deadline_ms=args.deadline_ms,
                       progress=lambda d, n: print(f"[loadgen] {d}/{n} results", file=sys.stderr, flush=True))
        r.pop("_outs")
        print(f"[loadgen] {rate:.0f}/s: {r['solved']}/{r['requests']} solved ({r['rejected']} rejected), "
              f"{r['plans_per_s']:.1f} plans/s, p50 {r['p50_ms']:.1f} ms, p99 {r['p99_ms']:.1f} ms, "
              f"{r['timed_out']} timed out, "
              f"drain {r['drain_s']:.1f} s", file=sys.stderr, flush=True)
        rows.append(r)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(rows, fh, indent=1)
    print(json.dumps(rows))


if __name__ == "__main__":
    main()
