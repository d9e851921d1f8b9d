"""Serving benchmark: Poisson request arrivals into the native continuous-batching server.

Each request is one pair list of one engine kind drawn like the mixed batch: kind 50% byte Levenshtein (ACGT),
25% UTF-8 Levenshtein, 25% affine NW; pair count log-uniform in [1, max_pairs]; both lengths log-uniform in
[16, max_length]. Requests are sampled (with replacement) from a fixed pool of pre-generated templates so host
memory stays bounded. Latency = completion time - arrival time, measured on the host.

The baseline serves the same arrivals with one plain launch per request, in arrival order, without batching
across requests (``engine.pairs`` per request).

Templates are uploaded to HBM once (``to_device``) unless ``--host-tapes`` is given, so both loops score
device-resident tapes and only the results travel back.

Run on the GPU: ``python -m stringzillas_amd.serving --qps 200 1000 5000 --seconds 20``.
"""

from __future__ import annotations

import argparse
import json
import math
import queue
import sys
import threading
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

import stringzillas_amd as szs
from stringzillas_amd import workloads as W


@dataclass
class Template:
    kind: str
    firsts: "szs.Strs"
    seconds: "szs.Strs"
    cells: int


def _log_uniform_int(rng, low: int, high: int) -> int:
    return int(min(high, max(low, math.exp(rng.uniform(math.log(low), math.log(high + 1))))))


def make_templates(count: int = 48, max_pairs: int = 4096, max_length: int = 16384, seed: int = 0x535A0004,
                   device: bool = True):
    """Request payloads following the mixed-batch length and kind mix (seeded, numpy)."""
    import torch

    rng = np.random.default_rng(seed)
    templates = []
    for i in range(count):
        pairs = _log_uniform_int(rng, 1, max_pairs)
        kind = rng.choice(["levenshtein", "levenshtein_utf8", "needleman_wunsch"], p=[0.5, 0.25, 0.25])
        generator = torch.Generator()
        generator.manual_seed(seed + i)
        lengths = [torch.from_numpy(np.array([_log_uniform_int(rng, 16, max_length) for _ in range(pairs)]))
                   for _ in range(2)]
        if kind == "levenshtein_utf8":
            tapes = [W._torch_utf8_tape(generator, l) for l in lengths]
        else:
            tapes = [W._torch_random_tape(generator, W.DNA if kind == "levenshtein" else W.AMINO, l) for l in lengths]
        firsts = szs.Strs.from_tape(tapes[0][0].numpy(), tapes[0][1].numpy().astype(np.uint64))
        seconds = szs.Strs.from_tape(tapes[1][0].numpy(), tapes[1][1].numpy().astype(np.uint64))
        segment = W.Segment(kind, firsts.data, firsts.offsets, seconds.data, seconds.offsets)
        if device:
            firsts, seconds = szs.to_device(firsts), szs.to_device(seconds)
        templates.append(Template(kind, firsts, seconds, segment.cells()))
    return templates


def engines_by_kind():
    b2c, costs = W.protein_matrix(np.random.default_rng(W.SEED_BASE + 1))
    return {
        "levenshtein": szs.LevenshteinDistances(),
        "levenshtein_utf8": szs.LevenshteinDistancesUTF8(),
        "needleman_wunsch": szs.NeedlemanWunschScores(b2c, costs, -11, -1),
    }


def _arrivals(rng, qps: float, duration: float) -> List[float]:
    times, now = [], 0.0
    while True:
        now += rng.exponential(1.0 / qps)
        if now >= duration:
            return times
        times.append(now)


def _summary(latencies: List[float], cells: int, wall: float, offered: float, completed: int):
    ms = np.array(latencies) * 1e3
    return {
        "offered_qps": offered,
        "completed": completed,
        "achieved_qps": round(completed / wall, 2),
        "gcups": round(cells / wall / 1e9, 2),
        "p50_ms": round(float(np.percentile(ms, 50)), 2) if len(ms) else None,
        "p99_ms": round(float(np.percentile(ms, 99)), 2) if len(ms) else None,
        "wall_s": round(wall, 2),
    }


def run_server(templates, engines, qps: float, duration: float, hbm_budget: int = 64 << 30,
               inflight: int = 2, seed: int = 1, drain_timeout: float = 120.0):
    """Continuous batching: a submitter thread follows the arrival schedule; the main thread joins tickets in
    arrival order, blocking in the library (the interpreter lock is released), and takes each request's latency
    from the completion time the server stamped on it. Returns the summary and, for every completed request, its
    template index and results."""
    rng = np.random.default_rng(seed)
    arrivals = _arrivals(rng, qps, duration)
    picks = rng.integers(0, len(templates), size=len(arrivals))
    server = szs.Server(hbm_budget_bytes=hbm_budget, max_inflight_rounds=inflight)
    submitted = queue.Queue()
    start = time.perf_counter()

    def submit_all():
        for at, pick in zip(arrivals, picks):
            delay = start + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            template = templates[pick]
            ticket = server.submit(engines[template.kind], template.firsts, template.seconds)
            submitted.put((ticket, start + at, int(pick)))
        submitted.put(None)

    submitter = threading.Thread(target=submit_all)
    submitter.start()
    latencies, cells, done, late = [], 0, 0, 0
    deadline = start + duration + drain_timeout
    results = []
    last_completion = start
    while True:
        item = submitted.get()
        if item is None:
            break
        ticket, arrived, pick = item
        out, completed = server.wait_timed(ticket)
        last_completion = max(last_completion, completed)
        if completed > deadline:
            late += 1
            continue
        latencies.append(completed - arrived)
        cells += templates[pick].cells
        done += 1
        results.append((pick, out))
    submitter.join()
    stats = server.stats()
    server.close()
    wall = min(last_completion, deadline) - start
    summary = _summary(latencies, cells, wall, qps, done)
    summary["dropped_after_drain_timeout"] = late
    summary["rounds"] = stats["rounds"]
    summary["largest_round"] = stats["largest_round"]
    summary["prepare_s"] = round(stats["prepare_seconds"], 3)
    summary["device_s"] = round(stats["device_seconds"], 3)
    return summary, results


def verify(templates, engines, results) -> dict:
    """Scores every template used by `results` alone (one ``engine.pairs`` call each) and counts the served
    requests whose results equal it: each pair is scored independently, so a request's results must not depend on
    the round it joined or what else that round held."""
    alone = {}
    matched = 0
    for pick, out in results:
        if pick not in alone:
            template = templates[pick]
            alone[pick] = engines[template.kind].pairs(template.firsts, template.seconds)
        matched += int(np.array_equal(out, alone[pick]))
    return {"verified": len(results), "equal_to_alone": matched, "distinct_requests": len(alone)}


def run_baseline(templates, engines, qps: float, duration: float, seed: int = 1, drain_timeout: float = 120.0):
    """One plain launch per request, in arrival order, no cross-request batching."""
    rng = np.random.default_rng(seed)
    arrivals = _arrivals(rng, qps, duration)
    picks = rng.integers(0, len(templates), size=len(arrivals))
    start = time.perf_counter()
    latencies, cells, done = [], 0, 0
    for at, pick in zip(arrivals, picks):
        if time.perf_counter() > start + duration + drain_timeout:
            break
        delay = start + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        template = templates[pick]
        engines[template.kind].pairs(template.firsts, template.seconds)
        latencies.append(time.perf_counter() - (start + at))
        cells += template.cells
        done += 1
    wall = time.perf_counter() - start
    summary = _summary(latencies, cells, wall, qps, done)
    summary["dropped_after_drain_timeout"] = len(arrivals) - done
    return summary


def main(argv=None) -> None:
    parser = argparse.ArgumentParser(description="Poisson serving benchmark of the continuous-batching server")
    parser.add_argument("--qps", type=float, nargs="+", default=[200.0, 1000.0, 5000.0])
    parser.add_argument("--seconds", type=float, default=20.0, help="arrival window per offered rate")
    parser.add_argument("--drain", type=float, default=30.0, help="extra time allowed to finish queued requests")
    parser.add_argument("--templates", type=int, default=96)
    parser.add_argument("--max-pairs", type=int, default=4096)
    parser.add_argument("--max-length", type=int, default=16384)
    parser.add_argument("--inflight", type=int, default=2)
    parser.add_argument("--budget-gib", type=float, default=64.0)
    parser.add_argument("--host-tapes", action="store_true")
    parser.add_argument("--skip-baseline", action="store_true")
    args = parser.parse_args(argv)

    def log(message):
        print(f"[serving] {message}", file=sys.stderr, flush=True)

    start = time.time()
    templates = make_templates(args.templates, args.max_pairs, args.max_length, device=not args.host_tapes)
    engines = engines_by_kind()
    mean_cells = float(np.mean([t.cells for t in templates]))
    log(f"{len(templates)} templates, mean {mean_cells / 1e9:.2f} G cells per request ({time.time() - start:.1f}s)")
    # Warm both paths (kernel loading, first allocations).
    run_server(templates[:4], engines, qps=50.0, duration=0.2, hbm_budget=int(args.budget_gib * (1 << 30)))
    for template in templates[:4]:
        engines[template.kind].pairs(template.firsts, template.seconds)
    report = {"templates": len(templates), "mean_cells_per_request": mean_cells,
              "tapes": "host" if args.host_tapes else "device", "runs": []}
    for qps in args.qps:
        summary, results = run_server(templates, engines, qps, args.seconds,
                                      hbm_budget=int(args.budget_gib * (1 << 30)), inflight=args.inflight,
                                      drain_timeout=args.drain)
        summary.update(verify(templates, engines, results))
        log(f"server at {qps:g} req/s: {summary}")
        entry = {"offered_qps": qps, "server": summary}
        if not args.skip_baseline:
            baseline = run_baseline(templates, engines, qps, args.seconds, drain_timeout=args.drain)
            log(f"per-request launches at {qps:g} req/s: {baseline}")
            entry["per_request"] = baseline
        report["runs"].append(entry)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
