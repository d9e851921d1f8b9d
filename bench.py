"""Benchmark: GCUPS of the batched persistent similarity engine on one MI355X, against the one-launch-per-pair
form and the reference's CPU path.

Headline (``value``): the uneven mixed batch, 1,000,000 pairs with lengths log-uniform in [16, 16384] plus 32 pairs
of 65,536 (50% byte Levenshtein, 25% UTF-8 Levenshtein, 25% affine Needleman-Wunsch). One step = one replay of the
HIP graph that transcodes, materializes, bins and scores the whole batch, tapes already resident in HBM.

Beside it, in the same JSON line: the roofline of the scoring kernel (core recurrence ops over its HIP-event
duration), each kind alone, the per-call form on a shared sample, the reference's CPU scorers on all cores of the
share and on one core, the Levenshtein (1,024 pairs), Needleman-Wunsch (16,384) and Smith-Waterman (100,000)
configurations, the reference's published all-pairs shapes (equal-length ACGT, 100 and 1,000 bytes), the serving
loop (20 s of Poisson arrivals per rate, every result verified) and the fingerprint engine.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS on 1 MI355X (batched vs per-call) + p50/p99 latency in serving loop"
INT32_VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # CUs x lanes x clock: 39.3 T lane-ops/s
VALU_ISSUE_PEAK = 256 * 2.4e9  # nominal wave64 VALU issue: one per cycle per CU (VOP3/VOP3P forms issue a little
# below it on gfx950, simple VOP2 forms such as v_xor_b32 above it)
MIXED = "mixed_1000000_pairs_log_uniform_16_16384_plus_32x65536"
# From profiles/persistent_score_kernel_mixed_1M_pairs.md (rocprofv3 --pmc on this workload): HBM bytes per launch
# (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the factors a plain 1 GiB device copy reads) and SQ_INSTS_VALU per
# launch (wave64 instructions), whose rate over the kernel's time is the measured VALU issue fraction.
MEASURED_TRAFFIC_BYTES = {MIXED: 41.98e9}
MEASURED_VALU_INSTRUCTIONS = {MIXED: 134.96e9}
# The reference's published all-pairs GCUPS on equal-length ACGT (include/stringzillas/similarities/README.md:8-26):
# Ice Lake AVX-512 on a 16-core Xeon4, and the H100 kernels (no UTF-8 kernel there).
PUBLISHED = {
    100: {"levenshtein": (114.6, 1499.8), "levenshtein_utf8": (6.7, None), "needleman_wunsch": (16.0, 187.3),
          "smith_waterman": (15.8, 164.8)},
    1000: {"levenshtein": (339.4, 5910.8), "levenshtein_utf8": (36.6, None), "needleman_wunsch": (97.1, 711.3),
           "smith_waterman": (85.0, 615.8)},
}


def parse():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--pairs", type=int, default=1_000_000, help="pairs of the mixed batch (1,000,000)")
    parser.add_argument("--long-pairs", type=int, default=32)
    parser.add_argument("--per-call-sample", type=int, default=2000, help="pairs timed with one launch each")
    parser.add_argument("--cpu-seconds", type=float, default=10.0, help="target duration of an all-cores CPU sample")
    parser.add_argument("--check-pairs", type=int, default=48, help="pairs per segment checked against the oracle")
    parser.add_argument("--mode", choices=["graph", "persistent"], default="graph")
    parser.add_argument("--skip-cpu", action="store_true")
    parser.add_argument("--skip-per-call", action="store_true")
    parser.add_argument("--skip-breakdown", action="store_true", help="skip timing each kind alone")
    parser.add_argument("--skip-configs", action="store_true", help="skip the 1,024 / 16,384 / 100,000-pair lines")
    parser.add_argument("--skip-cross", action="store_true", help="skip the published all-pairs shapes")
    parser.add_argument("--serve-qps", type=float, nargs="*", default=[200.0, 1000.0, 5000.0],
                        help="offered rates of the serving run (none to skip)")
    parser.add_argument("--serve-seconds", type=float, default=20.0, help="arrival window per serving rate")
    parser.add_argument("--baseline-seconds", type=float, default=20.0,
                        help="arrival window of the one-launch-per-request baseline")
    parser.add_argument("--skip-fingerprints", action="store_true", help="skip the fingerprint engine's measurement")
    parser.add_argument("--fingerprint-mb", type=int, default=200, help="corpus size per fingerprint shape, MB")
    parser.add_argument("--only-fingerprints", action="store_true",
                        help="measure only the fingerprint engine and print its report (for profiling)")
    parser.add_argument("--only-headline", action="store_true",
                        help="measure the mixed batch alone, no side measurements (for profiling)")
    parser.add_argument("--only-allpairs", type=int, nargs=2, metavar=("BYTES", "TILE"),
                        help="time one all-pairs call on TILE x TILE random ACGT strings of BYTES bytes (the "
                             "reference-shaped call, tapes and results in HBM) and print it (for profiling)")
    parser.add_argument("--allpairs-kind", choices=["levenshtein", "needleman_wunsch", "smith_waterman"],
                        default="levenshtein", help="engine of --only-allpairs (unit costs: 0/-1/-1 for NW and SW)")
    parser.add_argument("--only-kind", choices=["levenshtein", "levenshtein_utf8", "needleman_wunsch"],
                        help="time one kind of the mixed batch alone in the persistent form (for profiling)")
    return parser.parse_args()


def log(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def cpu_share() -> int:
    """Threads for CPU-side work: the cores this process may use, at most 16 (one GPU's share of the host)."""
    try:
        available = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        available = os.cpu_count() or 1
    return max(1, min(16, available))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((line.split(":", 1)[1].strip() for line in f if line.startswith("model name")), "")
    except OSError:
        return ""


class TensorTape:
    """A tape held by torch device tensors, handed to the C ABI in place."""

    def __init__(self, data, offsets):
        self.data, self.offsets = data.contiguous(), offsets.contiguous()

    def __len__(self):
        return self.offsets.numel() - 1

    @property
    def nbytes(self):
        return int(self.offsets[-1].item()) if len(self) else 0

    def tape(self):
        from stringzillas_amd import _native

        return _native.U64Tape(self.data.data_ptr(), self.offsets.data_ptr(), len(self))


def to_device_tape(data: np.ndarray, offsets: np.ndarray) -> TensorTape:
    import torch

    return TensorTape(torch.from_numpy(np.array(data, dtype=np.uint8)).cuda(),
                      torch.from_numpy(np.asarray(offsets).astype(np.int64)).cuda())


def engine_for(segment, szs):
    if segment.kind == "levenshtein":
        return szs.LevenshteinDistances(segment.match, segment.mismatch, segment.open, segment.extend)
    if segment.kind == "levenshtein_utf8":
        return szs.LevenshteinDistancesUTF8(segment.match, segment.mismatch, segment.open, segment.extend)
    if segment.kind == "needleman_wunsch":
        return szs.NeedlemanWunschScores(segment.byte_to_class, segment.class_costs, segment.open, segment.extend)
    return szs.SmithWatermanScores(segment.byte_to_class, segment.class_costs, segment.open, segment.extend)


def oracle_params(segment):
    from oracle import oracle as O

    kwargs = dict(match=segment.match, mismatch=segment.mismatch, open=segment.open, extend=segment.extend)
    if segment.byte_to_class is not None:
        kwargs.update(byte_to_class=segment.byte_to_class, class_costs=segment.class_costs)
    return O.Params(segment.kind, **kwargs)


def sample_indices(rng, count: int, size: int) -> np.ndarray:
    return np.sort(rng.choice(count, size=min(size, count), replace=False))


def nominal_ops(segment) -> int:
    """The fixed per-cell VALU count of the textbook recurrence: linear 5, affine global 9, affine local 10."""
    if segment.open == segment.extend:
        return 5
    return 10 if segment.kind == "smith_waterman" else 9


# --------------------------------------------------------------------------------------------------------------
# Timing helpers


def time_batch(batch, steps: int, warmup: int):
    """Runs a prepared batch `warmup` + `steps` times; returns (wall ms per step, HIP-event ms per step)."""
    import torch

    for _ in range(warmup):
        batch.run()
    torch.cuda.synchronize()
    event_ms = []
    start = time.perf_counter()
    for _ in range(steps):
        batch.launch()
        batch.wait()
        event_ms.append(batch.last_kernel_ms)
    torch.cuda.synchronize()
    return (time.perf_counter() - start) / steps * 1e3, float(np.mean(event_ms))


def roofline_of(core_ops: int, cells: int, kernel_ms: float, nominal_ops_per_cell: float):
    """Algorithmic VALU work rate of the scoring kernel: cells weighted by the core op count of the tier that
    scores each (a lower bound on what the kernel issues), over its HIP-event duration."""
    achieved = core_ops / (kernel_ms / 1e3) / 1e12
    return {
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": round(INT32_VALU_PEAK_TOPS, 2),
        "unit": "T INT32 VALU lane-ops/s (core recurrence ops per cell of each pair's tier)",
        "frac": round(achieved / INT32_VALU_PEAK_TOPS, 4),
        "core_ops_per_cell": round(core_ops / max(cells, 1), 4),
        "kernel_ms": round(kernel_ms, 3),
        # The old fixed-op figure (linear 5, affine 9 per cell), which the bit-parallel and packed tiers undercut.
        "nominal_dp_ratio": round(cells * nominal_ops_per_cell / (kernel_ms / 1e3) / 1e12 / INT32_VALU_PEAK_TOPS, 3),
    }


# --------------------------------------------------------------------------------------------------------------
# CPU baseline: the reference's own scorers (Ice Lake AVX-512 where the host has it), driven over threads


def cpu_baseline(host_segments, target_seconds: float, threads: int, seed: int = 99):
    from oracle import oracle as O

    variant = "icelake" if O.reference_available("icelake") else "serial" if O.reference_available("serial") else None
    rng = np.random.default_rng(seed)
    total = sum(s.count for s in host_segments)

    def run(sample_pairs, min_seconds):
        """Scores a sample drawn once; a sample that is every pair (small workloads) is scored again until it has
        run `min_seconds`, so the rate is not a timer's rounding."""
        subs = []
        for segment in host_segments:
            share = max(1, int(round(sample_pairs * segment.count / total)))
            subs.append(segment.subset(sample_indices(rng, segment.count, share)))
        cells = passes = 0
        start = time.perf_counter()
        while True:
            for sub in subs:
                tapes = ((sub.firsts_data, sub.firsts_offsets), (sub.seconds_data, sub.seconds_offsets))
                if variant:
                    O.reference_pairs(oracle_params(sub), *tapes, threads=threads, variant=variant)
                else:
                    O.pairs(oracle_params(sub), *tapes, threads=threads)
                cells += sub.cells()
            passes += 1
            seconds = time.perf_counter() - start
            if sample_pairs < total or seconds >= min_seconds or passes >= 10_000:
                return cells, seconds, passes

    sample_pairs = 16 * threads
    cells, seconds, passes = run(sample_pairs, target_seconds / 3)
    while seconds < target_seconds / 3 and sample_pairs < min(total, 400_000):
        sample_pairs = int(min(sample_pairs * max(2.0, 0.9 * target_seconds / max(seconds, 1e-3)), 400_000))
        cells, seconds, passes = run(sample_pairs, target_seconds / 3)
    sample_pairs = min(sample_pairs, total)
    return {
        "value": round(cells / seconds / 1e9, 3),
        "unit": "GCUPS",
        "cores": threads,
        "kind": "reference" if variant else "port",
        "sample": f"{sample_pairs} pairs drawn uniformly from the same pairs"
                  f"{f', scored {passes} times' if passes > 1 else ''} ({cells / 1e9:.1f} G cells, "
                  f"{seconds:.1f}s), scorer={'reference ' + variant if variant else 'oracle'}",
    }


def cpu_baselines(host_segments, target_seconds: float):
    """All cores of the share, then one core on a smaller sample."""
    threads = cpu_share()
    report = cpu_baseline(host_segments, target_seconds, threads)
    single = cpu_baseline(host_segments, max(2.0, target_seconds / 3), 1, seed=98)
    report["single_core_value"] = single["value"]
    report["single_core_sample"] = single["sample"]
    report["nproc"] = os.cpu_count()
    report["cpu"] = cpu_model()
    return report


# --------------------------------------------------------------------------------------------------------------
# Workload lines


def parity_check(host_segments, outputs, check_pairs: int, seed: int = 2024):
    """Oracle equality on a sample of every segment that includes one of its longest pairs."""
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    checked = mismatches = 0
    for host, output in zip(host_segments, outputs):
        got = output.cpu().numpy() if hasattr(output, "cpu") else np.asarray(output)
        picks = sample_indices(rng, host.count, check_pairs)
        longest = int(np.argmax(host.lengths("firsts") * host.lengths("seconds")))
        picks = np.unique(np.concatenate([picks, [longest]]))
        sub = host.subset(picks)
        expected = O.pairs(oracle_params(host), (sub.firsts_data, sub.firsts_offsets),
                           (sub.seconds_data, sub.seconds_offsets), threads=cpu_share())
        mismatches += int((got[picks].astype(np.int64) != expected).sum())
        checked += len(picks)
    return {"checked_pairs": checked, "mismatches": mismatches}


def per_call_line(szs, engines, host_segments, sample_pairs: int, seed: int = 7):
    """One launch per pair (the plain per-call form) and the persistent launch, on the same sample."""
    rng = np.random.default_rng(seed)
    total = sum(s.count for s in host_segments)
    subs = []
    for host in host_segments:
        share = max(1, int(round(sample_pairs * host.count / total)))
        subs.append(host.subset(sample_indices(rng, host.count, share)))
    cells = sum(s.cells() for s in subs)
    rates, outputs = {}, {}
    for name, mode in (("per_call", szs.BATCH_PER_PAIR), ("batched", szs.BATCH_PERSISTENT)):
        segments = [(e, to_device_tape(s.firsts_data, s.firsts_offsets), to_device_tape(s.seconds_data,
                                                                                        s.seconds_offsets))
                    for e, s in zip(engines, subs)]
        batch = szs.Batch(segments, mode=mode)
        batch.run()
        start = time.perf_counter()
        batch.run()
        rates[name] = cells / (time.perf_counter() - start) / 1e9
        outputs[name] = [o.copy() for o in batch.outputs]
        del batch
    for a, b in zip(outputs["per_call"], outputs["batched"]):
        if not np.array_equal(a, b):
            raise SystemExit("per-call and batched scores differ")
    return {
        "gcups": round(rates["per_call"], 2),
        "batched_same_sample_gcups": round(rates["batched"], 2),
        "batched_over_per_call": round(rates["batched"] / rates["per_call"], 1),
        "pairs": int(sum(s.count for s in subs)),
        "cells": int(cells),
        "launch": "one 256-thread workgroup per pair, one kernel launch per pair, same scorers; tapes in HBM",
    }


def config_line(szs, workload, args):
    """One configuration end to end: graph replay on HBM tapes, scoring-kernel roofline, per-call form, parity and
    the reference's CPU scorers."""
    import torch

    segments = workload.segments
    engines = [engine_for(s, szs) for s in segments]
    tapes = [(to_device_tape(s.firsts_data, s.firsts_offsets), to_device_tape(s.seconds_data, s.seconds_offsets))
             for s in segments]
    outputs = [torch.zeros(s.count, dtype=torch.int64, device="cuda") for s in segments]
    cells = workload.cells()
    batch = szs.Batch([(e, a, b) for e, (a, b) in zip(engines, tapes)], mode=szs.BATCH_GRAPH, outputs=outputs)
    steps = max(args.steps, 10)
    ms, _ = time_batch(batch, steps, max(args.warmup, 2))
    scoring = szs.Batch([(e, a, b) for e, (a, b) in zip(engines, tapes)], mode=szs.BATCH_PERSISTENT, outputs=outputs)
    _, kernel_ms = time_batch(scoring, steps, 1)
    nominal = sum(s.cells() * nominal_ops(s) for s in segments) / max(cells, 1)
    line = {
        "workload": workload.name,
        "pairs": workload.pairs,
        "cells": cells,
        "gcups": round(cells / (ms / 1e3) / 1e9, 2),
        "ms_per_step": round(ms, 3),
        "roofline": roofline_of(scoring.core_ops, cells, kernel_ms, nominal),
        "parity": parity_check([s for s in segments], outputs, 256),
    }
    del batch, scoring
    if not args.skip_per_call:
        line["per_call"] = per_call_line(szs, engines, segments, min(workload.pairs, args.per_call_sample))
    if not args.skip_cpu:
        line["cpu_baseline"] = cpu_baselines(segments, max(3.0, args.cpu_seconds / 3))
        line["over_cpu"] = round(line["gcups"] / line["cpu_baseline"]["value"], 1)
    return line


def allpairs_alone(szs, length: int, tile: int, calls: int, kind: str = "levenshtein") -> dict:
    """One all-pairs call (the reference-shaped `szs_*_u64tape` on a scope) over `tile` random ACGT strings of
    `length` bytes against a second tape of the same strings, timed over `calls` calls after one warm-up; for
    profiling the published all-pairs shape on its own. NW and SW score with the unit costs of the published shape
    (match 0, mismatch and gap -1)."""
    import torch
    from stringzillas_amd import _native

    rng = np.random.default_rng(0x535A0011)
    letters = np.frombuffer(b"acgt", dtype=np.uint8)
    data = letters[rng.integers(0, 4, size=tile * length)]
    offsets = np.arange(tile + 1, dtype=np.uint64) * length
    queries, candidates = to_device_tape(data, offsets), to_device_tape(data, offsets)
    results = torch.empty((tile, tile), dtype=torch.int64, device="cuda")
    if kind == "levenshtein":
        engine = szs.LevenshteinDistances()
    else:
        class_costs = np.full((32, 32), -1, dtype=np.int8)
        np.fill_diagonal(class_costs, 0)
        byte_to_class = np.zeros(256, dtype=np.uint8)
        byte_to_class[letters] = np.arange(1, 5, dtype=np.uint8)
        maker = szs.NeedlemanWunschScores if kind == "needleman_wunsch" else szs.SmithWatermanScores
        engine = maker(byte_to_class, class_costs, -1, -1)
    scope = szs._scope(None)
    call = getattr(_native.lib, engine._prefix + "_u64tape")
    error = ctypes.c_char_p()
    first, second = queries.tape(), candidates.tape()

    def once():
        status = call(engine.handle, scope.handle, ctypes.byref(first), ctypes.byref(second),
                      ctypes.c_void_p(results.data_ptr()), ctypes.c_size_t(tile), ctypes.byref(error))
        _native.check(status, error, "all-pairs")

    once()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(calls):
        once()
    torch.cuda.synchronize()
    seconds = (time.perf_counter() - start) / calls
    return {"allpairs": kind, "bytes": length, "tile": f"{tile}x{tile}", "calls": calls,
            "call_ms": round(seconds * 1e3, 3), "gcups": round(tile * tile * length * length / seconds / 1e9, 1)}


def cross_shapes(szs, args):
    """The reference's published shape: all-pairs cross-products of equal-length ACGT strings through the
    reference-shaped call (`szs_*_u64tape` with tapes and the result matrix in HBM), under unit costs
    (match/mismatch/gap 0/1/1, negated for the maximizing engines), at 100 and 1,000 bytes. Tiles: the
    reference's per-call budget rule (256 pairs per parallel unit: 256 x 256 CUs, a 256 x 256 tile here) and a
    large tile where the call is bound by the kernels."""
    import torch
    from oracle import oracle as O
    from stringzillas_amd import _native

    class_costs = np.full((32, 32), -1, dtype=np.int8)
    np.fill_diagonal(class_costs, 0)
    byte_to_class = np.zeros(256, dtype=np.uint8)
    byte_to_class[np.frombuffer(b"acgt", dtype=np.uint8)] = np.arange(1, 5, dtype=np.uint8)
    kinds = {
        "levenshtein": (szs.LevenshteinDistances(), O.Params("levenshtein")),
        "levenshtein_utf8": (szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8")),
        "needleman_wunsch": (szs.NeedlemanWunschScores(byte_to_class, class_costs, -1, -1),
                             O.Params("needleman_wunsch", open=-1, extend=-1, byte_to_class=byte_to_class,
                                      class_costs=class_costs)),
        "smith_waterman": (szs.SmithWatermanScores(byte_to_class, class_costs, -1, -1),
                           O.Params("smith_waterman", open=-1, extend=-1, byte_to_class=byte_to_class,
                                    class_costs=class_costs)),
    }
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    reference_tile = int(np.sqrt(256 * cus))
    scope = szs._scope(None)
    report = {"strings": "equal-length random 'acgt', all-pairs cross-product, unit costs",
              "reference_tile": f"{reference_tile}x{reference_tile} (256 pairs per CU, as the reference sizes its "
                                "H100 tile: 256 x 132 SMs -> 183x184)", "shapes": []}
    rng = np.random.default_rng(0x535A0010)
    letters = np.frombuffer(b"acgt", dtype=np.uint8)
    for length, large in ((100, 4096), (1000, 1024)):
        for tile in (reference_tile, large):
            data = letters[rng.integers(0, 4, size=tile * length)]
            offsets = np.arange(tile + 1, dtype=np.uint64) * length
            tape = to_device_tape(data, offsets)
            results = torch.empty((tile, tile), dtype=torch.int64, device="cuda")
            view = tape.tape()
            for name, (engine, params) in kinds.items():
                if tile == large and name == "levenshtein_utf8":
                    continue  # the large tiles time the four engines' kernels; UTF-8 only at the reference tile
                call = getattr(_native.lib, engine._prefix + "_u64tape")
                error = ctypes.c_char_p()

                def once():
                    status = call(engine.handle, scope.handle, ctypes.byref(view), ctypes.byref(view),
                                  ctypes.c_void_p(results.data_ptr()), ctypes.c_size_t(tile), ctypes.byref(error))
                    _native.check(status, error, name)

                once()
                torch.cuda.synchronize()
                calls, start = 0, time.perf_counter()
                while calls < 3 or time.perf_counter() - start < 0.3:
                    once()
                    calls += 1
                seconds = (time.perf_counter() - start) / calls
                cells = tile * tile * length * length
                picks = rng.integers(0, tile, size=(6, 2))
                got = results.cpu().numpy()
                strings = [data[i * length:(i + 1) * length].tobytes() for i in range(tile)]
                for q, c in picks:
                    if int(got[q, c]) != O.score(params, strings[q], strings[c]):
                        raise SystemExit(f"cross-product {name} {length} B differs from the oracle at ({q}, {c})")
                icelake, h100 = PUBLISHED[length][name]
                entry = {"kind": name, "bytes": length, "tile": f"{tile}x{tile}", "call_ms": round(seconds * 1e3, 3),
                         "gcups": round(cells / seconds / 1e9, 1), "reference_icelake_16core_gcups": icelake,
                         "reference_h100_gcups": h100, "checked_cells": len(picks)}
                if h100:
                    entry["over_h100"] = round(entry["gcups"] / h100, 2)
                log(f"cross {entry}")
                report["shapes"].append(entry)
            del tape, results
    if not args.skip_cpu:  # the reference's Ice Lake scorer on pairs of the same strings, this host's share
        threads = cpu_share()
        variant = "icelake" if O.reference_available("icelake") else "serial" if O.reference_available("serial") else None
        cpu = []
        for length, count in ((100, 1_000_000), (1000, 16_000)):
            data = letters[rng.integers(0, 4, size=2 * count * length)]
            a = (np.ascontiguousarray(data[:count * length]), np.arange(count + 1, dtype=np.uint64) * length)
            b = (np.ascontiguousarray(data[count * length:]), np.arange(count + 1, dtype=np.uint64) * length)
            start = time.perf_counter()
            if variant:
                O.reference_pairs(O.Params("levenshtein"), a, b, threads=threads, variant=variant)
            else:
                O.pairs(O.Params("levenshtein"), a, b, threads=threads)
            seconds = time.perf_counter() - start
            cpu.append({"kind": "levenshtein", "bytes": length, "pairs": count, "cores": threads,
                        "gcups": round(count * length * length / seconds / 1e9, 2),
                        "scorer": f"reference {variant}" if variant else "oracle"})
        report["cpu_baseline"] = cpu
        log(f"cross cpu {cpu}")
    return report


# --------------------------------------------------------------------------------------------------------------


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
    if args.only_allpairs:
        print(json.dumps(allpairs_alone(szs, *args.only_allpairs, max(1, args.steps), args.allpairs_kind)), flush=True)
        return
    device = torch.device("cuda", 0)
    t0 = time.time()
    workload = W.mixed_c3(args.pairs, args.long_pairs, device="cuda")
    torch.cuda.synchronize()
    log(f"generated {workload.name}: {workload.pairs} pairs in {time.time() - t0:.1f}s")
    cells = workload.cells()
    nominal = workload.weighted_ops_per_cell()

    engines = [engine_for(s, szs) for s in workload.segments]
    tapes, outputs = [], []
    for segment in workload.segments:
        tapes.append((TensorTape(segment.firsts_data, segment.firsts_offsets),
                      TensorTape(segment.seconds_data, segment.seconds_offsets)))
        outputs.append(torch.zeros(segment.count, dtype=torch.int64, device=device))
    if args.only_kind:
        index = next(i for i, s in enumerate(workload.segments) if s.kind == args.only_kind)
        alone = szs.Batch([(engines[index], *tapes[index])], mode=szs.BATCH_PERSISTENT, outputs=[outputs[index]])
        alone.run()
        _, alone_ms = time_batch(alone, max(1, args.steps), 0)
        print(json.dumps({"kind": args.only_kind, "pairs": workload.segments[index].count, "cells": alone.cells,
                          "core_ops": alone.core_ops, "kernel_ms": round(alone_ms, 3),
                          "gcups": round(alone.cells / (alone_ms / 1e3) / 1e9, 1)}), flush=True)
        return
    mode = szs.BATCH_GRAPH if args.mode == "graph" else szs.BATCH_PERSISTENT
    t0 = time.time()
    batch = szs.Batch([(e, a, b) for e, (a, b) in zip(engines, tapes)], mode=mode, outputs=outputs)
    log(f"prepared the batch in {time.time() - t0:.1f}s ({batch.cells / 1e12:.3f} T cells)")
    assert batch.cells == cells, (batch.cells, cells)

    ms_per_step, replay_ms = time_batch(batch, args.steps, args.warmup)
    gcups = cells / (ms_per_step / 1e3) / 1e9
    # The scoring kernel alone, from HIP events on its stream, in the non-graph form of the same batch.
    scoring = szs.Batch([(e, a, b) for e, (a, b) in zip(engines, tapes)], mode=szs.BATCH_PERSISTENT, outputs=outputs)
    scoring.run()
    _, kernel_ms = time_batch(scoring, max(1, min(args.steps, 5)), 0)
    roofline = roofline_of(scoring.core_ops, cells, kernel_ms, nominal)
    instructions = MEASURED_VALU_INSTRUCTIONS.get(workload.name)
    roofline["traffic"] = MEASURED_TRAFFIC_BYTES.get(workload.name)
    if instructions:  # from the committed rocprofv3 profile of this kernel on this workload
        roofline["valu_issue_frac"] = round(instructions / (kernel_ms / 1e3) / VALU_ISSUE_PEAK, 4)
        roofline["valu_per_cell"] = round(instructions * 64 / cells, 3)
    roofline["note"] = ("frac = core recurrence lane-ops (each cell weighted by the op count of its tier's update: "
                        "Myers words, packed 16-bit or int32 DP) over the scoring kernel's HIP-event time and the "
                        "INT32 VALU peak; valu_issue_frac = all issued VALU instructions (rocprofv3 SQ_INSTS_VALU) "
                        "over the same time and the nominal issue rate of one wave64 instruction per cycle per CU (packed and VOP3 forms "
                        "issue at most ~0.9 of it, simple VOP2 forms more), so frac <= valu_issue_frac")
    del scoring
    log(f"batched: {ms_per_step:.1f} ms/step ({replay_ms:.1f} ms replay), {gcups:.1f} GCUPS; scoring kernel "
        f"{kernel_ms:.1f} ms; roofline {roofline['frac']}")
    if args.only_headline:
        print(json.dumps({"value": round(gcups, 2), "ms_per_step": round(ms_per_step, 2), "roofline": roofline}),
              flush=True)
        return

    # Each kind alone (persistent form): where the scoring time goes, against each tier's ceiling.
    per_kind = {}
    if not args.skip_breakdown:
        for engine, (a, b), segment, output in zip(engines, tapes, workload.segments, outputs):
            alone = szs.Batch([(engine, a, b)], mode=szs.BATCH_PERSISTENT, outputs=[output])
            alone.run()
            _, alone_ms = time_batch(alone, 2, 0)
            seg_cells, core = alone.cells, alone.core_ops
            rate = seg_cells / (alone_ms / 1e3)
            ceiling = INT32_VALU_PEAK_TOPS * 1e12 / (core / seg_cells)  # cells/s if every lane-op were core work
            entry = {
                "pairs": segment.count, "cells": seg_cells, "kernel_ms": round(alone_ms, 2),
                "gcups": round(rate / 1e9, 1),
                "core_ops_per_cell": round(core / seg_cells, 4),
                "tier_ceiling_gcups": round(ceiling / 1e9, 1),
                "frac": round(rate / ceiling, 4),
                "nominal_ops_per_cell": nominal_ops(segment),
                "nominal_int32_frac": round(rate * nominal_ops(segment) / 1e12 / INT32_VALU_PEAK_TOPS, 3),
            }
            if segment.kind == "needleman_wunsch":  # packed 16-bit cells: two per lane-op at the nominal count
                entry["int16x2_frac"] = round(rate * nominal_ops(segment) / 2 / 1e12 / INT32_VALU_PEAK_TOPS, 3)
            per_kind[segment.kind] = entry
            del alone
        log(f"per kind: {per_kind}")

    host_segments = [s.host() for s in workload.segments]
    parity = parity_check(host_segments, outputs, args.check_pairs)
    log(f"parity: {parity}")
    if parity["mismatches"]:
        raise SystemExit("batched scores differ from the oracle")

    per_call = None
    if not args.skip_per_call:
        per_call = per_call_line(szs, engines, host_segments, args.per_call_sample)
        log(f"per-call: {per_call}")

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baselines(host_segments, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    del batch, tapes, workload, host_segments

    configs = None
    if not args.skip_configs:
        configs = {}
        for key, make in (("levenshtein_1024", W.levenshtein_c0), ("needleman_wunsch_16384", W.needleman_wunsch_c1),
                          ("smith_waterman_100000", W.smith_waterman_c2)):
            t0 = time.time()
            configs[key] = config_line(szs, make(), args)
            log(f"config {key}: {configs[key]} ({time.time() - t0:.1f}s)")
            if configs[key]["parity"]["mismatches"]:
                raise SystemExit(f"{key}: scores differ from the oracle")

    cross = None if args.skip_cross else cross_shapes(szs, args)

    serving = serving_run(args.serve_qps, args.serve_seconds, args.baseline_seconds) if args.serve_qps else None

    fingerprints = None
    if not args.skip_fingerprints:
        fingerprints = fingerprints_run(args.fingerprint_mb * 1_000_000, args.steps, not args.skip_cpu)

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
            "workload": MIXED if args.pairs == 1_000_000 and args.long_pairs == 32 else f"mixed_{args.pairs}",
            "pairs": args.pairs + args.long_pairs,
            "cells": cells,
            "mix": "50% byte Levenshtein ACGT, 25% UTF-8 Levenshtein, 25% NW affine -11/-1 (21-class matrix)",
            "launch": args.mode,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "per_call": per_call,
        # Same-sample ratio: the persistent launch over one launch per pair on identical pairs.
        "batched_over_per_call": per_call["batched_over_per_call"] if per_call else None,
        "full_batch_over_per_call": round(gcups / per_call["gcups"], 1) if per_call else None,
        "parity": parity,
        "per_kind": per_kind or None,
        "configs": configs,
        "cross_product": cross,
        "serving": serving,
        "fingerprints": fingerprints,
    }
    print(json.dumps(line), flush=True)


# --------------------------------------------------------------------------------------------------------------
# Serving


def serving_run(rates, seconds: float, baseline_seconds: float):
    """Poisson arrivals into the continuous-batching server at each rate, every completed request checked against
    the same request scored alone; then one plain call per request at the first rate.

    Requests follow the mixed batch's kind and length mix with 1..4096 pairs each, drawn from 512 HBM-resident
    templates; latency runs from a request's arrival to the completion of its round.
    """
    from stringzillas_amd import serving as S

    t0 = time.time()
    templates = S.make_templates(512)
    engines = S.engines_by_kind()
    log(f"serving: {len(templates)} templates in {time.time() - t0:.1f}s")
    S.run_server(templates[:4], engines, qps=50.0, duration=0.2, hbm_budget=16 << 30)  # warm-up
    report = {"requests": "log-uniform 1..4096 pairs of the mixed kinds, lengths log-uniform 16..16384",
              "templates": len(templates),
              "mean_cells_per_request": round(float(np.mean([t.cells for t in templates]))),
              "seconds_per_rate": seconds, "server": [], "per_request": None}
    for qps in rates:
        summary, results = S.run_server(templates, engines, qps, seconds, hbm_budget=16 << 30, drain_timeout=20.0)
        check = S.verify(templates, engines, results)
        if check["equal_to_alone"] != check["verified"]:
            raise SystemExit(f"served results differ from the same requests scored alone at {qps} req/s: {check}")
        summary.update(check)
        log(f"serving at {qps:g} req/s: {summary}")
        report["server"].append({k: summary[k] for k in ("offered_qps", "achieved_qps", "gcups", "p50_ms", "p99_ms",
                                                         "completed", "rounds", "verified", "equal_to_alone",
                                                         "distinct_requests")})
    baseline = S.run_baseline(templates, engines, rates[0], baseline_seconds, drain_timeout=5.0)
    log(f"one call per request at {rates[0]:g} req/s: {baseline}")
    report["per_request"] = {k: baseline[k] for k in ("offered_qps", "achieved_qps", "gcups", "p50_ms", "p99_ms",
                                                      "completed", "dropped_after_drain_timeout")}
    report["per_request"]["seconds"] = baseline_seconds
    return report


# --------------------------------------------------------------------------------------------------------------
# Fingerprints

# The reference's published fingerprint throughput on an H100 (include/stringzillas/fingerprints/README.md:15-42):
# equal-length ACGT documents, 64 dimensions, default window widths; MB there is 2^20 bytes.
FINGERPRINT_H100_MIB_S = {100: 438.7, 1000: 3400.0 * 1.024, 10000: 11200.0 * 1.024, 100000: 5600.0 * 1.024}
# Per (ndim, document bytes): VALU wave instructions per hash update (SQ_INSTS_VALU x 64 / updates) and HBM bytes
# per launch of the hashing kernel (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the factors of a plain device copy),
# recorded in profiles/fingerprint_kernels_acgt_documents.md.
FINGERPRINT_VALU_PER_UPDATE = {(64, 100): 13.6, (64, 1000): 11.3, (64, 10000): 11.6, (64, 100000): 11.4,
                               (1024, 10000): 10.4}
FINGERPRINT_TRAFFIC_BYTES = {(64, 10000): 673.7e6, (1024, 10000): 571.1e6}


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

    threads = cpu_share()
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


if __name__ == "__main__":
    main()
