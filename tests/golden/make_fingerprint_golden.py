"""Writes tests/golden/fingerprints.json: fingerprints computed by the reference's own engines.

The engines are compiled from the reference headers into oracle/_ref/ by oracle/Makefile (run ``make -C oracle``
first): the serial sliced `floating_rolling_hashers` when the layout allows it, else the per-dimension
`basic_rolling_hashers`, laid out as `szs_fingerprints_init` lays them out. The texts include the ones the
reference's own tests use (test/fingerprints.py:15-20, 62-76). Hashes and counts are stored as base64 of
little-endian u32 rows.

Run from the repository root: ``python tests/golden/make_fingerprint_golden.py``.
"""

from __future__ import annotations

import base64
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fingerprints.json")

CASES = [  # (ndim, window widths or None for the defaults, seed)
    (1, None, 0),
    (7, None, 0),
    (64, None, 0),
    (512, None, 0),
    (256, [4, 8], 7),
    (96, [3, 33], 123),
]


def texts():
    rng = np.random.default_rng(2024)
    fixed = [b"hello", b"world", b"hello", b"a", b"a" * 5, b"a" * 40, b"b", b"", b"the quick brown fox",
             b"the quick brown dog"]
    randoms = [bytes(rng.integers(0, 256, int(n), dtype=np.uint8)) for n in rng.integers(0, 120, 6)]
    dna = [bytes(rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), 300))]
    return fixed + randoms + dna


def main():
    if not O.reference_fingerprints_available("serial"):
        raise SystemExit("build oracle/_ref first: make -C oracle")
    strings = texts()
    tape = O.tape(strings)
    cases = []
    for ndim, widths, seed in CASES:
        hashes, counts = O.reference_fingerprints(tape, ndim, widths, seed=seed)
        cases.append({
            "ndim": ndim, "window_widths": widths, "seed": seed,
            "hashes": base64.b64encode(hashes.astype("<u4").tobytes()).decode(),
            "counts": base64.b64encode(counts.astype("<u4").tobytes()).decode(),
        })
    payload = {
        "generator": "tests/golden/make_fingerprint_golden.py",
        "engines": "oracle/ref_fingerprints.cpp over include/stringzillas/fingerprints/serial.hpp",
        "texts": [base64.b64encode(s).decode() for s in strings],
        "cases": cases,
    }
    with open(OUT, "w") as f:
        json.dump(payload, f, indent=1)
    print(f"wrote {OUT}: {len(strings)} texts, {len(cases)} cases")


if __name__ == "__main__":
    main()
