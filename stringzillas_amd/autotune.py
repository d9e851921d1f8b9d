"""Autotuner for the kernel-shape ladders (lanes per pair, rows per lane, cooperative or not, DP or bit-parallel).

For each family slot and each band of row counts (the shorter side of a pair), it times every shape that can
score that band on a synthetic batch, through the same persistent launch the engine uses, and installs the
fastest shape per band via ``szs_tuning_set_ladder``. The result is cached as JSON (``tuned_shapes.json`` next to
this file) and applied when the package is imported on a device of the same architecture.

Run on the GPU: ``python -m stringzillas_amd.autotune``.
"""

from __future__ import annotations

import ctypes
import json
import os
import time
from typing import Dict, List, Tuple

import numpy as np

import stringzillas_amd as szs
from stringzillas_amd import _native
from stringzillas_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned_shapes.json")

FAMILIES = ["lev_linear", "lev_affine", "nw_linear", "nw_affine", "sw_linear", "sw_affine"]
# Slots of the unit-cost Levenshtein over bytes and over runes follow the family x {bytes, runes} slots; only their
# ladders may use bit-parallel shapes.
UNIT_BYTES_SLOT = 2 * len(FAMILIES)
UNIT_RUNES_SLOT = 2 * len(FAMILIES) + 1
UNIT_SLOTS = (UNIT_BYTES_SLOT, UNIT_RUNES_SLOT)
# Rows beyond which global families use the chained cooperative shape (the library's default ladders agree).
CHAINED_ROWS = 24576
BANDS = [16, 32, 64, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192, 16384]


def shapes() -> List[Tuple[int, int, int, int, int, int]]:
    """Every compiled shape as (lanes per pair, rows per lane, cooperative, bit-parallel, packed 16-bit, chained)."""
    out = []
    for index in range(_native.lib.szs_tuning_shapes()):
        values = [ctypes.c_int() for _ in range(6)]
        _native.lib.szs_tuning_shape(index, *[ctypes.byref(v) for v in values])
        out.append(tuple(v.value for v in values))
    return out


def wide_shapes() -> List[int]:
    """Shapes that keep int64 cells: the range-bound fallback, never a tuning candidate."""
    return [i for i in range(_native.lib.szs_tuning_shapes()) if _native.lib.szs_tuning_shape_wide(i)]


def get_ladder(slot: int):
    shapes_out = (ctypes.c_int * 32)()
    rows_out = (ctypes.c_int * 32)()
    count, beyond = ctypes.c_int(), ctypes.c_int()
    _native.lib.szs_tuning_get_ladder(slot, shapes_out, rows_out, 32, ctypes.byref(count), ctypes.byref(beyond))
    return [(shapes_out[i], rows_out[i]) for i in range(count.value)], beyond.value


def set_ladder(slot: int, ladder, beyond: int) -> None:
    count = len(ladder)
    shape_array = (ctypes.c_int * max(count, 1))(*[s for s, _ in ladder])
    rows_array = (ctypes.c_int * max(count, 1))(*[r for _, r in ladder])
    error = ctypes.c_char_p()
    status = _native.lib.szs_tuning_set_ladder(slot, shape_array, rows_array, count, beyond, ctypes.byref(error))
    _native.check(status, error, f"tuning slot {slot}")


def slot_of(family: str, runes: bool = False, unit: bool = False) -> int:
    if unit:
        return UNIT_RUNES_SLOT if runes else UNIT_BYTES_SLOT
    return FAMILIES.index(family) * 2 + (1 if runes else 0)


def _engine_for(slot: int):
    b2c, costs = W.protein_matrix(np.random.default_rng(W.SEED_BASE + 1))
    if slot == UNIT_BYTES_SLOT:
        return szs.LevenshteinDistances(), W.DNA
    if slot == UNIT_RUNES_SLOT:
        return szs.LevenshteinDistancesUTF8(), None  # codepoints of the mixed batch's four ranges
    family, runes = FAMILIES[slot // 2], slot % 2 == 1
    if family == "lev_linear":
        engine = (szs.LevenshteinDistancesUTF8 if runes else szs.LevenshteinDistances)(0, 1, 2, 2)
        return engine, W.DNA
    if family == "lev_affine":
        engine = (szs.LevenshteinDistancesUTF8 if runes else szs.LevenshteinDistances)(0, 4, 3, 2)
        return engine, W.DNA
    gaps = (-2, -2) if family.endswith("linear") else (-11, -1)
    cls = szs.NeedlemanWunschScores if family.startswith("nw") else szs.SmithWatermanScores
    return cls(b2c, costs, *gaps), W.AMINO


def _batch_for(rng, alphabet, low: int, high: int, target_cells: float, max_pairs: int = 200_000):
    mean_rows = (low + high) / 2
    mean_cells = mean_rows * max(mean_rows * 2.5, (16384 - mean_rows) / np.log(16384 / mean_rows))
    # Enough pairs to give every workgroup of the device (256 CUs x 4) two work items, as inside a large batch,
    # so shapes are compared at full occupancy: a 4-lane shape takes 64 pairs per item.
    floor = 131_072 if high <= 256 else 8192
    pairs = int(min(max_pairs, max(floor, target_cells / mean_cells)))
    rows = rng.integers(low, high + 1, size=pairs)
    # The longer side log-uniform from the rows up to 16,384 (4x the rows at least), like the mixed batch's pairs.
    top = np.maximum(rows * 4, 16384)
    columns = np.exp(rng.uniform(np.log(rows), np.log(top))).astype(np.int64)
    if alphabet is None:  # UTF-8 of the mixed batch: about 2.5 bytes per rune, rows counted in runes
        import torch

        generator = torch.Generator().manual_seed(int(rng.integers(1 << 62)))
        tapes = [W._torch_utf8_tape(generator, torch.from_numpy(lengths * 5 // 2)) for lengths in (rows, columns)]
        return tuple(szs.Strs.from_tape(data[:-1].numpy(), offsets.numpy()) for data, offsets in tapes)
    a = W._random_tape(rng, alphabet, rows)
    b = W._random_tape(rng, alphabet, columns)
    return szs.Strs.from_tape(*a), szs.Strs.from_tape(*b)


def _time(engine, firsts, seconds, repeats: int = 3) -> float:
    batch = szs.Batch([(engine, firsts, seconds)], mode=szs.BATCH_PERSISTENT)
    batch.run()
    best = float("inf")
    for _ in range(repeats):
        batch.run()
        best = min(best, batch.last_kernel_ms)
    return batch.cells / (best / 1e3) / 1e9


def candidates_for(slot: int, low: int, high: int, table) -> List[int]:
    family = "lev_linear" if slot in UNIT_SLOTS else FAMILIES[slot // 2]
    uniform = family.startswith("lev")
    wide = set(wide_shapes())
    out = []
    for index, (group, rows_per_lane, cooperative, bit_parallel, packed, chained) in enumerate(table):
        if index in wide:  # int64 cells are the range bound's fallback, slower than any int32 shape
            continue
        if bit_parallel and slot not in UNIT_SLOTS:
            continue
        if chained:  # for pairs far taller than the bands measured here
            continue
        if packed and (slot in UNIT_SLOTS or family.startswith("lev") or slot % 2 == 1):
            continue
        if not uniform and rows_per_lane > 12:
            continue
        capacity = group * rows_per_lane
        if cooperative:
            if high > 2 * capacity:  # needs stripes for several waves to pay off
                out.append(index)
            continue
        if group < 64:
            if high <= capacity:
                out.append(index)
        elif bit_parallel:
            if high > 256:  # a partly filled 64-word stripe still scores 64 cells per word operation
                out.append(index)
        elif high <= 16 * capacity and (high > capacity // 2 or capacity <= 128):
            out.append(index)  # up to 16 stripes swept one after another by one wave
    return out


def tune(slots: List[int] = None, target_cells: float = 4e9, verbose: bool = True, bands=None) -> Dict:
    import torch

    table = shapes()
    if slots is None:
        slots = [UNIT_BYTES_SLOT, UNIT_RUNES_SLOT, slot_of("lev_linear", runes=True), slot_of("nw_affine"),
                 slot_of("sw_linear")]
    result = {"arch": torch.cuda.get_device_properties(0).gcnArchName.split(":")[0], "slots": {}}
    for slot in slots:
        default_ladder, default_beyond = get_ladder(slot)
        engine, alphabet = _engine_for(slot)
        rng = np.random.default_rng(slot)
        measurements, chosen = {}, []
        low = 1
        for high in bands or BANDS:
            options = candidates_for(slot, low, high, table)
            if not options:
                low = high + 1
                continue
            firsts, seconds = _batch_for(rng, alphabet, low, high, target_cells)
            timings = {}
            for shape in options:
                group = table[shape][0]
                if table[shape][2] or group == 64:
                    set_ladder(slot, [], shape if group == 64 else default_beyond)
                else:
                    set_ladder(slot, [(shape, high)], default_beyond)
                timings[shape] = _time(engine, firsts, seconds)
            best = max(timings, key=timings.get)
            measurements[f"{low}-{high}"] = {str(k): round(v, 1) for k, v in timings.items()}
            chosen.append((best, high))
            if verbose:
                print(f"slot {slot} rows {low}-{high}: best shape {best} {table[best]} "
                      f"{timings[best]:.0f} GCUPS of {timings}", flush=True)
            low = high + 1
        ladder: List[Tuple[int, int]] = []
        # The last band's winner, if cooperative, becomes the shape for every taller pair; the ladder lists the
        # winners of the bands below it, cooperative ones included.
        if chosen and table[chosen[-1][0]][2]:
            tail = chosen[-1][0]
            chosen = chosen[:-1]
        else:
            tail = None
        for shape, high in chosen:
            earlier = [s for s, _ in ladder]
            if shape in earlier:
                # A shape that wins again after others takes over the bands in between (a ladder lists each
                # shape once, in increasing row limits).
                ladder = ladder[: earlier.index(shape)]
            ladder.append((shape, high))
        # Sub-wave shapes hold one stripe: clamp their limits to capacity.
        fixed = []
        for shape, high in ladder:
            group, rows_per_lane = table[shape][0], table[shape][1]
            if group < 64:
                high = min(high, group * rows_per_lane)
            if not fixed or high > fixed[-1][1]:
                fixed.append((shape, high))
        beyond = tail if tail is not None else default_beyond
        if beyond in [s for s, _ in fixed]:  # it also won a lower band: it takes over from there
            fixed = fixed[: [s for s, _ in fixed].index(beyond)]
        # Global DP families split pairs beyond the measured bands over chained workgroups (the tallest pairs
        # would otherwise set a batch's critical path on one workgroup each).
        family = "lev_linear" if slot in UNIT_SLOTS else FAMILIES[slot // 2]
        if slot not in UNIT_SLOTS and not family.startswith("sw"):
            chained = [i for i, s in enumerate(table) if s[5] and s[4] == table[beyond][4]]
            if chained:
                fixed.append((beyond, CHAINED_ROWS))
                beyond = chained[0]
        set_ladder(slot, fixed, beyond)
        result["slots"][str(slot)] = {"ladder": fixed, "beyond": beyond, "measurements": measurements}
    return result


class _DeviceTape:
    """A tape held by device tensors (8-byte offsets), handed to the C ABI in place."""

    def __init__(self, data, offsets):
        self.data, self.offsets = data.contiguous(), offsets.contiguous()

    def __len__(self):
        return self.offsets.numel() - 1

    def tape(self):
        return _native.U64Tape(self.data.data_ptr(), self.offsets.data_ptr(), len(self))


def _segment_ms(items, repeats: int = 3) -> float:
    """Best scoring-kernel time of `items` (a list of (engine, firsts, seconds)) under the installed ladders."""
    batch = szs.Batch(items, mode=szs.BATCH_PERSISTENT)
    batch.run()
    best = float("inf")
    for _ in range(repeats):
        batch.run()
        best = min(best, batch.last_kernel_ms)
    return best


def _moves(slot: int, ladder, beyond: int, table) -> List[Tuple[str, list, int]]:
    """Ladders one step away from `ladder`: another admissible shape for one entry or beyond it, one limit moved to
    the next band edge either way, or one entry dropped (the next one covers its rows)."""
    out = []
    used = {s for s, _ in ladder} | {beyond}
    edges = sorted(set(BANDS) | {g * r for g, r, *_ in table})
    lows = [1] + [limit + 1 for _, limit in ladder]
    for i, (shape, limit) in enumerate(ladder):
        for other in candidates_for(slot, lows[i], limit, table):
            group, rows_per_lane = table[other][0], table[other][1]
            if other in used or table[other][2] or (group < 64 and limit > group * rows_per_lane):
                continue
            # Bit-parallel words and DP rows are not close alternatives, nor int32 cells for packed ones.
            if table[other][3] != table[shape][3] or table[other][4] < table[shape][4]:
                continue
            out.append((f"entry {i}: shape {shape} -> {other}", ladder[:i] + [(other, limit)] + ladder[i + 1:], beyond))
        floor = ladder[i - 1][1] if i else 0
        ceiling = ladder[i + 1][1] if i + 1 < len(ladder) else None
        below = [e for e in edges if floor < e < limit]
        above = [e for e in edges if e > limit and (ceiling is None or e < ceiling)]
        group, rows_per_lane = table[shape][0], table[shape][1]
        if group < 64:
            above = [e for e in above if e <= group * rows_per_lane]
        for edge in ([below[-1]] if below else []) + ([above[0]] if above and ceiling is not None else []):
            out.append((f"entry {i}: limit {limit} -> {edge}", ladder[:i] + [(shape, edge)] + ladder[i + 1:], beyond))
        if i + 1 < len(ladder):
            out.append((f"entry {i}: drop shape {shape}", ladder[:i] + ladder[i + 1:], beyond))
        # An unused shape for the lower part of this entry's rows, up to a band edge inside them.
        for edge in below:
            for other in candidates_for(slot, lows[i], edge, table):
                group, rows_per_lane = table[other][0], table[other][1]
                if other in used or table[other][2] or table[other][3] != table[shape][3]:
                    continue
                if table[other][4] < table[shape][4]:
                    continue
                if group < 64 and edge > group * rows_per_lane:
                    continue
                out.append((f"entry {i}: shape {other} up to {edge}", ladder[:i] + [(other, edge)] + ladder[i:],
                            beyond))
    top = ladder[-1][1] if ladder else 0
    for other in candidates_for(slot, top + 1, max(top + 1, 16384), table):
        if table[other][2] and other not in used and table[other][3] == table[beyond][3]:
            out.append((f"beyond: shape {beyond} -> {other}", list(ladder), other))
    return out


def refine(slot: int, items, rounds: int = 2, gain: float = 0.01, verbose: bool = True) -> Dict:
    """Local search over one slot's ladder on a real workload rather than on bands in isolation.

    `items` are the (engine, firsts, seconds) lists whose pairs use `slot`. Every ladder one move away (see
    `_moves`) is installed and the whole workload timed; a move is kept when it shortens the scoring kernel by more
    than `gain`, and the search repeats from the new ladder for up to `rounds` passes. Shapes then compete inside a
    full, mixed-length launch, as they do in the engine's real batches.
    """
    table = shapes()
    ladder, beyond = get_ladder(slot)
    ladder = [tuple(x) for x in ladder]
    best_ms = _segment_ms(items)
    start_ms = best_ms
    history = []
    if verbose:
        print(f"slot {slot}: start {best_ms:.2f} ms with {ladder} beyond {beyond}", flush=True)
    for _ in range(rounds):
        improved = False
        for label, candidate, candidate_beyond in _moves(slot, ladder, beyond, table):
            try:
                set_ladder(slot, candidate, candidate_beyond)
            except Exception:  # an inadmissible ladder: the library rejects it
                continue
            ms = _segment_ms(items)
            if verbose:
                print(f"slot {slot}:   {label}: {ms:.2f} ms", flush=True)
            if ms < best_ms * (1.0 - gain):
                history.append({"move": label, "ms": round(ms, 3)})
                if verbose:
                    print(f"slot {slot}: {label}: {best_ms:.2f} -> {ms:.2f} ms", flush=True)
                ladder, beyond, best_ms, improved = [tuple(x) for x in candidate], candidate_beyond, ms, True
            set_ladder(slot, ladder, beyond)
        if not improved:
            break
    set_ladder(slot, ladder, beyond)
    return {"ladder": [list(x) for x in ladder], "beyond": beyond, "start_ms": round(start_ms, 3),
            "refined_ms": round(best_ms, 3), "moves": history}


def refine_on_mixed_batch(pairs: int = 1_000_000, long_pairs: int = 32, verbose: bool = True) -> Dict:
    """Refines the ladders of the mixed batch's three kinds (unit-cost bytes and runes, affine NW) on its segments,
    generated on the device from the benchmark's seed."""
    import torch

    workload = W.mixed_c3(pairs, long_pairs, device="cuda")
    torch.cuda.synchronize()
    b2c, costs = W.protein_matrix(np.random.default_rng(W.SEED_BASE + 1))
    result = {}
    for segment in workload.segments:
        if segment.kind == "levenshtein":
            engine, slot = szs.LevenshteinDistances(), UNIT_BYTES_SLOT
        elif segment.kind == "levenshtein_utf8":
            engine, slot = szs.LevenshteinDistancesUTF8(), UNIT_RUNES_SLOT
        else:
            engine = szs.NeedlemanWunschScores(segment.byte_to_class, segment.class_costs, segment.open,
                                               segment.extend)
            slot = slot_of("nw_affine")
        items = [(engine, _DeviceTape(segment.firsts_data, segment.firsts_offsets),
                  _DeviceTape(segment.seconds_data, segment.seconds_offsets))]
        result[str(slot)] = refine(slot, items, verbose=verbose)
    return result


def apply(path: str = CACHE) -> bool:
    """Installs a cached tuning if it was measured on the architecture the library is compiled for (gfx950).

    The check needs no GPU: the library's kernels exist only for ``build.ARCH``, so a tuning from another
    architecture is ignored rather than applied.
    """
    if not os.path.exists(path):
        return False
    from . import build

    with open(path) as f:
        tuned = json.load(f)
    if tuned.get("arch") != build.ARCH:
        return False
    for slot, entry in tuned["slots"].items():
        set_ladder(int(slot), [tuple(x) for x in entry["ladder"]], int(entry["beyond"]))
    return True


if __name__ == "__main__":
    import argparse

    parser = argparse.ArgumentParser(description="Measure and cache the fastest shape ladders.")
    parser.add_argument("slots", type=int, nargs="*", help="slots to tune (default: the benchmark's kinds)")
    parser.add_argument("--out", default=CACHE, help="where to write the merged tuning (default: the package cache)")
    parser.add_argument("--refine", action="store_true",
                        help="start from the installed ladders and refine those of the mixed batch's kinds on its "
                             "own segments instead of timing bands in isolation")
    arguments = parser.parse_args()
    start = time.time()
    if arguments.refine:
        import torch

        apply()
        refined = refine_on_mixed_batch()
        tuned = {"arch": torch.cuda.get_device_properties(0).gcnArchName.split(":")[0], "slots": {}}
        cached_slots = {}
        if os.path.exists(CACHE):
            with open(CACHE) as f:
                cached_slots = json.load(f).get("slots", {})
        for slot, entry in refined.items():
            # The band measurements stay; the ladder is the refined one, with the search's record beside it.
            tuned["slots"][slot] = {**cached_slots.get(slot, {}), "ladder": entry["ladder"], "beyond": entry["beyond"],
                                    "refined_on_mixed_batch": {k: entry[k] for k in ("start_ms", "refined_ms", "moves")}}
    else:
        tuned = tune(arguments.slots or None)
    # Slots not tuned in this run keep their cached ladders when the cache was measured on the same architecture.
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            cached = json.load(f)
        if cached.get("arch") == tuned["arch"]:
            tuned["slots"] = {**cached["slots"], **tuned["slots"]}
            if "note" in cached:
                tuned["note"] = cached["note"]
    with open(arguments.out, "w") as f:
        json.dump(tuned, f, indent=1)
    print(f"wrote {arguments.out} in {time.time() - start:.0f}s")
