"""Benchmark: GCUPS of the batched persistent similarity engine on one MI355X, against the one-launch-per-pair
form and the reference's CPU path, on the uneven mixed batch (1,000,000 pairs with lengths log-uniform in
[16, 16384] plus 32 pairs of 65,536: 50% byte Levenshtein, 25% UTF-8 Levenshtein, 25% affine Needleman-Wunsch).

One step = one replay of the HIP graph that transcodes, materializes, bins and scores the whole batch, with the
tapes already resident in HBM. Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS on 1 MI355X (batched vs per-call) + p50/p99 latency in serving loop"
INT32_VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # CUs x lanes x clock: 39.3 T lane-ops/s
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the scoring kernel on the full mixed batch: FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000
# (factors from a plain 1 GiB device copy), recorded in profiles/persistent_score_kernel_mixed_1M_pairs.md.
MEASURED_TRAFFIC_BYTES = {"mixed_1000000_pairs_log_uniform_16_16384_plus_32x65536": 62.27e9}
# VALU issue rate of the same launch from SQ_INSTS_VALU (same file): wave64 instructions per second over the
# gfx950 peak of one per cycle per CU.
MEASURED_VALU_ISSUE = {"mixed_1000000_pairs_log_uniform_16_16384_plus_32x65536": 0.953}


def parse():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--pairs", type=int, default=1_000_000, help="pairs of the mixed batch (1,000,000)")
    parser.add_argument("--long-pairs", type=int, default=32)
    parser.add_argument("--per-call-sample", type=int, default=2000, help="pairs timed with one launch each")
    parser.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline duration")
    parser.add_argument("--check-pairs", type=int, default=48, help="pairs per segment checked against the oracle")
    parser.add_argument("--mode", choices=["graph", "persistent"], default="graph")
    parser.add_argument("--skip-cpu", action="store_true")
    parser.add_argument("--skip-per-call", action="store_true")
    parser.add_argument("--breakdown", action="store_true", help="also time each kind alone")
    parser.add_argument("--serve-qps", type=float, nargs="*", default=[200.0, 1000.0, 5000.0],
                        help="offered rates of the short serving run (none to skip)")
    parser.add_argument("--serve-seconds", type=float, default=8.0, help="arrival window per serving rate")
    parser.add_argument("--skip-fingerprints", action="store_true", help="skip the fingerprint engine's measurement")
    parser.add_argument("--fingerprint-mb", type=int, default=200, help="corpus size per fingerprint shape, MB")
    parser.add_argument("--only-fingerprints", action="store_true",
                        help="measure only the fingerprint engine and print its report (for profiling)")
    return parser.parse_args()


def log(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


class TensorTape:
    """A tape held by torch device tensors, handed to the C ABI in place."""

    def __init__(self, data, offsets):
        self.data, self.offsets = data.contiguous(), offsets.contiguous()

    def __len__(self):
        return self.offsets.numel() - 1

    def tape(self):
        from stringzillas_amd import _native

        return _native.U64Tape(self.data.data_ptr(), self.offsets.data_ptr(), len(self))


def engines_for(segments, szs):
    engines = []
    for segment in segments:
        if segment.kind == "levenshtein":
            engines.append(szs.LevenshteinDistances(segment.match, segment.mismatch, segment.open, segment.extend))
        elif segment.kind == "levenshtein_utf8":
            engines.append(szs.LevenshteinDistancesUTF8(segment.match, segment.mismatch, segment.open, segment.extend))
        elif segment.kind == "needleman_wunsch":
            engines.append(szs.NeedlemanWunschScores(segment.byte_to_class, segment.class_costs, segment.open,
                                                     segment.extend))
        else:
            engines.append(szs.SmithWatermanScores(segment.byte_to_class, segment.class_costs, segment.open,
                                                   segment.extend))
    return engines


def oracle_params(segment):
    from oracle import oracle as O

    kwargs = dict(match=segment.match, mismatch=segment.mismatch, open=segment.open, extend=segment.extend)
    if segment.byte_to_class is not None:
        kwargs.update(byte_to_class=segment.byte_to_class, class_costs=segment.class_costs)
    return O.Params(segment.kind, **kwargs)


def sample_indices(rng, count: int, size: int) -> np.ndarray:
    return np.sort(rng.choice(count, size=min(size, count), replace=False))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}): the pairs are independent, so a "
              "multi-GPU run is a sharded copy of this one (pairs dealt to GPUs by estimated cost, each GPU writing its "
              "own slice of the results, no collective beyond a final gather).", file=sys.stderr)
        sys.exit(2)
    import torch

    import stringzillas_amd as szs
    from stringzillas_amd import workloads as W

    if args.only_fingerprints:
        print(json.dumps({"fingerprints": fingerprints_run(args.fingerprint_mb * 1_000_000, args.steps,
                                                           not args.skip_cpu)}), flush=True)
        return
    device = torch.device("cuda", 0)
    t0 = time.time()
    workload = W.mixed_c3(args.pairs, args.long_pairs, device="cuda")
    torch.cuda.synchronize()
    log(f"generated {workload.name}: {workload.pairs} pairs in {time.time() - t0:.1f}s")
    cells = workload.cells()
    ops_per_cell = workload.weighted_ops_per_cell()
    log(f"{cells / 1e12:.3f} T cells, {ops_per_cell:.2f} VALU ops per cell (cell-weighted)")

    engines = engines_for(workload.segments, szs)
    tapes, outputs = [], []
    for segment in workload.segments:
        a = TensorTape(segment.firsts_data, segment.firsts_offsets)
        b = TensorTape(segment.seconds_data, segment.seconds_offsets)
        tapes.append((a, b))
        outputs.append(torch.zeros(segment.count, dtype=torch.int64, device=device))
    mode = szs.BATCH_GRAPH if args.mode == "graph" else szs.BATCH_PERSISTENT
    t0 = time.time()
    batch = szs.Batch([(e, a, b) for e, (a, b) in zip(engines, tapes)], mode=mode, outputs=outputs)
    log(f"prepared the batch in {time.time() - t0:.1f}s ({batch.cells / 1e12:.3f} T cells)")
    assert batch.cells == cells, (batch.cells, cells)

    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    kernel_ms = []
    start = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
        batch.wait()
        kernel_ms.append(batch.last_kernel_ms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    ms_per_step = elapsed / args.steps * 1e3
    gcups = cells / (elapsed / args.steps) / 1e9
    replay_ms_mean = float(np.mean(kernel_ms))
    # The scoring kernel alone, from HIP events on its stream, in the non-graph form of the same batch.
    scoring = szs.Batch([(e, a, b) for e, (a, b) in zip(engines, tapes)], mode=szs.BATCH_PERSISTENT, outputs=outputs)
    scoring.run()
    scoring_ms = []
    for _ in range(max(1, min(args.steps, 3))):
        scoring.run()
        scoring_ms.append(scoring.last_kernel_ms)
    kernel_ms_mean = float(np.mean(scoring_ms))
    del scoring
    log(f"batched: {ms_per_step:.1f} ms/step ({replay_ms_mean:.1f} ms replay), {gcups:.1f} GCUPS; "
        f"scoring kernel {kernel_ms_mean:.1f} ms")

    # Each kind alone (persistent form): where the scoring time goes.
    breakdown = {}
    if args.breakdown:
        for engine, (a, b), segment, output in zip(engines, tapes, workload.segments, outputs):
            alone = szs.Batch([(engine, a, b)], mode=szs.BATCH_PERSISTENT, outputs=[output])
            alone.run()
            alone.run()
            seg_cells = alone.cells
            breakdown[segment.kind] = {
                "pairs": segment.count,
                "cells": seg_cells,
                "kernel_ms": round(alone.last_kernel_ms, 2),
                "gcups": round(seg_cells / (alone.last_kernel_ms / 1e3) / 1e9, 1),
                "ops_per_cell": segment.ops_per_cell(),
            }
            del alone
        log(f"per kind: {breakdown}")

    # Parity spot check against the oracle, including one of the longest pairs of each kind.
    from oracle import oracle as O

    rng = np.random.default_rng(2024)
    checked = mismatches = 0
    host_segments = []
    for segment, output in zip(workload.segments, outputs):
        host = segment.host()
        host_segments.append(host)
        got = output.cpu().numpy()
        picks = sample_indices(rng, host.count, args.check_pairs)
        longest = int(np.argmax(host.lengths("firsts")))
        picks = np.unique(np.concatenate([picks, [longest]]))
        sub = host.subset(picks)
        expected = O.pairs(oracle_params(host), (sub.firsts_data, sub.firsts_offsets),
                           (sub.seconds_data, sub.seconds_offsets), threads=min(16, os.cpu_count() or 1))
        mismatches += int((got[picks] != expected).sum())
        checked += len(picks)
    log(f"parity: {checked} pairs checked against the oracle, {mismatches} mismatches")
    if mismatches:
        raise SystemExit("batched scores differ from the oracle")

    # One launch per pair on a uniform sample, and the persistent form on the same sample.
    per_call = None
    if not args.skip_per_call:
        rng = np.random.default_rng(7)
        total = workload.pairs
        sample_segments, sample_engines = [], []
        for segment, host, engine in zip(workload.segments, host_segments, engines):
            share = max(1, int(round(args.per_call_sample * segment.count / total)))
            sub = host.subset(sample_indices(rng, host.count, share))
            sample_segments.append(sub)
            sample_engines.append(engine)
        sample_cells = sum(s.cells() for s in sample_segments)
        results = {}
        for name, sample_mode in (("per_call", szs.BATCH_PER_PAIR), ("batched", szs.BATCH_PERSISTENT)):
            strs = [(e, szs.Strs.from_tape(s.firsts_data, s.firsts_offsets),
                     szs.Strs.from_tape(s.seconds_data, s.seconds_offsets))
                    for e, s in zip(sample_engines, sample_segments)]
            sample_batch = szs.Batch(strs, mode=sample_mode)
            sample_batch.run()
            t = time.perf_counter()
            sample_batch.run()
            results[name] = sample_cells / (time.perf_counter() - t) / 1e9
            results[name + "_outputs"] = [o.copy() for o in sample_batch.outputs]
        for a, b in zip(results["per_call_outputs"], results["batched_outputs"]):
            if not np.array_equal(a, b):
                raise SystemExit("per-call and batched scores differ")
        per_call = {
            "gcups": round(results["per_call"], 2),
            "batched_same_sample_gcups": round(results["batched"], 2),
            "pairs": int(sum(s.count for s in sample_segments)),
            "cells": int(sample_cells),
            "launch": "one 64-lane workgroup per pair, one launch per pair, same scorer",
        }
        log(f"per-call: {per_call['gcups']} GCUPS; batched on the same sample: {per_call['batched_same_sample_gcups']}")

    serving = None
    if args.serve_qps:
        serving = serving_run(args.serve_qps, args.serve_seconds)

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(host_segments, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    fingerprints = None
    if not args.skip_fingerprints:
        fingerprints = fingerprints_run(args.fingerprint_mb * 1_000_000, args.steps, not args.skip_cpu)

    achieved = cells * ops_per_cell / (kernel_ms_mean / 1e3) / 1e12  # T lane-ops/s over the scoring kernel
    line = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded, generated on the device)",
        "config": {
            "workload": workload.name,
            "pairs": workload.pairs,
            "cells": cells,
            "mix": "50% byte Levenshtein ACGT, 25% UTF-8 Levenshtein, 25% NW affine -11/-1 (21-class matrix)",
            "launch": args.mode,
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(INT32_VALU_PEAK_TOPS, 2),
            "unit": "TOP/s (INT32 VALU lane-ops)",
            "frac": round(achieved / INT32_VALU_PEAK_TOPS, 4),
            "traffic": MEASURED_TRAFFIC_BYTES.get(workload.name),
            "ops_per_cell": round(ops_per_cell, 3),
            "kernel_ms": round(kernel_ms_mean, 3),
            "note": "fixed DP ops per cell (linear 5, affine 9); unit-cost Levenshtein over bytes and runes runs "
                    "the bit-parallel tier (64 cells per word op) and affine NW runs packed 16-bit cells (about 4.3 "
                    "VALU per cell), which is why the fraction exceeds 1; valu_issue_frac is the measured share of "
                    "the VALU issue rate the kernel uses",
            "valu_issue_frac": MEASURED_VALU_ISSUE.get(workload.name),
        },
        "cpu_baseline": cpu,
        "per_call": per_call,
        "batched_over_per_call": round(gcups / per_call["gcups"], 1) if per_call else None,
        "parity": {"checked_pairs": checked, "mismatches": mismatches},
        "serving": serving,
        "fingerprints": fingerprints,
    }
    if breakdown:
        line["per_kind"] = breakdown
    print(json.dumps(line), flush=True)


# The reference's published fingerprint throughput on an H100 (include/stringzillas/fingerprints/README.md:15-42):
# equal-length ACGT documents, 64 dimensions, default window widths; MB there is 2^20 bytes.
FINGERPRINT_H100_MIB_S = {100: 438.7, 1000: 3400.0 * 1.024, 10000: 11200.0 * 1.024, 100000: 5600.0 * 1.024}
# Per (ndim, document bytes): VALU wave instructions per hash update (SQ_INSTS_VALU x 64 / updates) and HBM bytes
# per launch of the hashing kernel (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the factors of a plain device copy),
# recorded in profiles/fingerprint_kernels_acgt_documents.md.
FINGERPRINT_VALU_PER_UPDATE = {(64, 100): 13.6, (64, 1000): 11.3, (64, 10000): 11.6, (64, 100000): 11.4,
                               (1024, 10000): 10.4}
FINGERPRINT_TRAFFIC_BYTES = {(64, 10000): 673.7e6, (1024, 10000): 571.1e6}
VALU_ISSUE_PEAK = 256 * 2.4e9  # wave64 VALU instructions per second: one per cycle per CU


def fingerprints_run(corpus_bytes: int, steps: int, with_cpu: bool):
    """The fingerprint engine on equal-length ACGT documents resident in HBM (the reference README's shapes): 64
    dimensions over the default 8 window widths (the lane kernel), and 1,024 dimensions (128 per width, the wave
    kernel) on the 10,000-byte documents. Throughput counts text bytes; a hash update is one window of one
    dimension. Spot-checked against the oracle."""
    import torch

    import stringzillas_amd as szs
    from oracle import oracle as O

    letters = torch.tensor(list(b"ACGT"), dtype=torch.uint8, device="cuda")
    generator = torch.Generator(device="cuda")
    generator.manual_seed(64)
    report = {"corpus_bytes": corpus_bytes, "documents": "equal-length ACGT, generated on the device", "shapes": []}
    shapes = [(64, 100), (64, 1000), (64, 10000), (64, 100000), (1024, 10000)]
    for ndim, length in shapes:
        count = corpus_bytes // length
        data = letters[torch.randint(0, 4, (count * length + 16,), device="cuda", generator=generator)]
        offsets = torch.arange(count + 1, dtype=torch.int64, device="cuda") * length
        texts = TensorTape(data, offsets)
        torch.cuda.synchronize()
        engine = szs.Fingerprints(ndim=ndim)
        hashes = torch.empty((count, ndim), dtype=torch.int32, device="cuda")
        counts = torch.empty((count, ndim), dtype=torch.int32, device="cuda")
        engine(texts, out=(hashes, counts))
        torch.cuda.synchronize()
        kernel_ms, start = [], time.perf_counter()
        for _ in range(max(1, steps)):
            engine(texts, out=(hashes, counts))
            kernel_ms.append(engine.last_kernel_ms)
        torch.cuda.synchronize()
        wall_ms = (time.perf_counter() - start) / max(1, steps) * 1e3
        kernel = float(np.mean(kernel_ms))
        # Oracle check on 4 documents spread over the corpus.
        picks = np.linspace(0, count - 1, 4).astype(np.int64)
        host = data[: count * length].view(count, length)[torch.from_numpy(picks).cuda()].cpu().numpy()
        expected_hashes, expected_counts = O.fingerprints(O.tape([row.tobytes() for row in host]), ndim, threads=8)
        got_hashes = hashes[torch.from_numpy(picks).cuda()].cpu().numpy().view(np.uint32)
        got_counts = counts[torch.from_numpy(picks).cuda()].cpu().numpy().view(np.uint32)
        if not (np.array_equal(got_hashes, expected_hashes) and np.array_equal(got_counts, expected_counts)):
            raise SystemExit(f"fingerprints differ from the oracle (ndim {ndim}, {length}-byte documents)")
        total = count * length
        entry = {
            "ndim": ndim, "document_bytes": length, "documents": count,
            "kernel": "lane" if ndim == 64 else "wave",
            "kernel_ms": round(kernel, 3), "call_ms": round(wall_ms, 3),
            "gb_per_s": round(total / kernel / 1e6, 2),
            "mhash_per_s": round(count * ndim / kernel / 1e3, 1),
            "t_updates_per_s": round(total * ndim / kernel / 1e9, 3),
            "checked_documents": len(picks),
        }
        per_update = FINGERPRINT_VALU_PER_UPDATE.get((ndim, length))
        if per_update:
            issue = total * ndim * per_update / 64 / (kernel / 1e3)
            entry["roofline"] = {
                "bound": "valu", "achieved": round(issue / 1e9, 1), "peak": round(VALU_ISSUE_PEAK / 1e9, 1),
                "unit": "G wave64 VALU instructions/s", "frac": round(issue / VALU_ISSUE_PEAK, 3),
                "valu_per_update": per_update, "traffic": FINGERPRINT_TRAFFIC_BYTES.get((ndim, length)),
            }
        if ndim == 64 and length in FINGERPRINT_H100_MIB_S:
            h100 = FINGERPRINT_H100_MIB_S[length] * 1.048576  # MB/s
            entry["reference_h100_mb_per_s"] = round(h100, 1)
            entry["over_reference_h100"] = round(total / kernel / 1e3 / h100, 2)
        log(f"fingerprints {entry}")
        report["shapes"].append(entry)
        del data, offsets, texts, hashes, counts, engine
    if with_cpu:
        report["cpu_baseline"] = fingerprint_cpu_baseline()
        log(f"fingerprint cpu baseline: {report['cpu_baseline']}")
    return report


def fingerprint_cpu_baseline(target_seconds: float = 6.0):
    """The reference's own CPU fingerprint engine on 10,000-byte ACGT documents at 64 dimensions, sharded over
    threads; a bounded sample. At 64 dimensions over 8 widths `szs_fingerprints_init` builds the per-dimension
    `basic_rolling_hashers` engine (serial code), whichever SIMD build the shim comes from."""
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    variant = next((v for v in ("skylake", "serial") if O.reference_fingerprints_available(v)), None)
    rng = np.random.default_rng(5)
    documents = 2 * threads
    while True:
        texts = [bytes(rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), 10000)) for _ in range(documents)]
        tape = O.tape(texts)
        start = time.perf_counter()
        if variant:
            O.reference_fingerprints(tape, 64, threads=threads, variant=variant)
        else:
            O.fingerprints(tape, 64, threads=threads)
        seconds = time.perf_counter() - start
        if seconds >= target_seconds / 3 or documents >= 1 << 16:
            break
        documents = int(min(documents * max(2.0, 0.9 * target_seconds / max(seconds, 1e-3)), 1 << 16))
    return {
        "value": round(documents * 10000 / seconds / 1e6, 2), "unit": "MB/s", "cores": threads,
        "kind": "reference" if variant else "port",
        "sample": f"{documents} ACGT documents of 10,000 bytes at 64 dimensions ({seconds:.1f}s), engine="
                  f"{'reference basic_rolling_hashers (' + variant + ' build)' if variant else 'oracle'}",
    }


def serving_run(rates, seconds: float):
    """Poisson arrivals into the continuous-batching server, and one plain launch per request at the first rate.

    Requests follow the mixed batch's kind and length mix with 1..4096 pairs each, from HBM-resident templates;
    latency runs from a request's arrival to the completion of its round.
    """
    from stringzillas_amd import serving as S

    templates = S.make_templates(96)
    engines = S.engines_by_kind()
    S.run_server(templates[:4], engines, qps=50.0, duration=0.2, hbm_budget=16 << 30)  # warm-up
    report = {"requests": "log-uniform 1..4096 pairs of the mixed kinds, lengths log-uniform 16..16384",
              "mean_cells_per_request": round(float(np.mean([t.cells for t in templates]))),
              "seconds_per_rate": seconds, "server": [], "per_request": None}
    for qps in rates:
        summary, _ = S.run_server(templates, engines, qps, seconds, hbm_budget=16 << 30, drain_timeout=10.0)
        log(f"serving at {qps:g} req/s: {summary}")
        report["server"].append({k: summary[k] for k in ("offered_qps", "achieved_qps", "gcups", "p50_ms", "p99_ms",
                                                         "completed", "rounds")})
    baseline = S.run_baseline(templates, engines, rates[0], seconds, drain_timeout=4.0)
    log(f"one launch per request at {rates[0]:g} req/s: {baseline}")
    report["per_request"] = {k: baseline[k] for k in ("offered_qps", "achieved_qps", "gcups", "p50_ms", "p99_ms",
                                                      "completed", "dropped_after_drain_timeout")}
    return report


def cpu_baseline(host_segments, target_seconds: float):
    """The reference's own CPU scorers (Ice Lake AVX-512 when the host has it) on a bounded uniform sample."""
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    variant = "icelake" if O.reference_available("icelake") else "serial" if O.reference_available("serial") else None
    rng = np.random.default_rng(99)
    total = sum(s.count for s in host_segments)

    def run(sample_pairs):
        cells = 0
        start = time.perf_counter()
        for segment in host_segments:
            share = max(1, int(round(sample_pairs * segment.count / total)))
            sub = segment.subset(sample_indices(rng, segment.count, share))
            tapes = ((sub.firsts_data, sub.firsts_offsets), (sub.seconds_data, sub.seconds_offsets))
            if variant:
                O.reference_pairs(oracle_params(sub), *tapes, threads=threads, variant=variant)
            else:
                O.pairs(oracle_params(sub), *tapes, threads=threads)
            cells += sub.cells()
        return cells, time.perf_counter() - start

    sample_pairs = 64
    cells, seconds = run(sample_pairs)
    while seconds < target_seconds / 3 and sample_pairs < 200_000:
        sample_pairs = int(min(sample_pairs * max(2.0, 0.9 * target_seconds / max(seconds, 1e-3)), 200_000))
        cells, seconds = run(sample_pairs)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((line.split(":", 1)[1].strip() for line in f if line.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": round(cells / seconds / 1e9, 3),
        "unit": "GCUPS",
        "cores": threads,
        "kind": "reference" if variant else "port",
        "sample": f"{sample_pairs} pairs drawn uniformly from the same batch ({cells / 1e9:.1f} G cells, "
                  f"{seconds:.1f}s), scorer={'reference ' + variant if variant else 'oracle'}, cpu={model}",
    }


if __name__ == "__main__":
    main()
