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
    out = []
    for index, (group, rows_per_lane, cooperative, bit_parallel, packed, chained) in enumerate(table):
        if bit_parallel and slot not in UNIT_SLOTS:
            continue
        if chained:  # for pairs far taller than the bands measured here
            continue
        if packed and (slot in UNIT_SLOTS or family not in ("nw_linear", "nw_affine") or slot % 2 == 1):
            continue
        if not uniform and rows_per_lane > 8:
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
    arguments = parser.parse_args()
    start = time.time()
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
