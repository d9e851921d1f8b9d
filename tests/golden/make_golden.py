"""Writes tests/golden/fixtures.json: inputs and expected outputs that pin the oracle.

Two kinds of vectors:
* ``known_answers``: the literal (input, expected) pairs asserted by the reference's own tests
  (test/similarities.py:134-143, 160-168, 188-197, 219-227), transcribed as data;
* ``reference_vectors``: seeded random pairs scored by the reference's own serial C++ scorers, compiled from the
  reference headers into oracle/_ref/ by oracle/Makefile (run ``make -C oracle`` first). Lengths cross the
  reference's 8/16/32-bit cell-width tiers (serial.hpp:293-356) and its Myers/DP split.

Run from the repository root: ``python tests/golden/make_golden.py``.
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")

KNOWN_ANSWERS = [
    # test/similarities.py:134-143 (unit costs)
    ("levenshtein", dict(), "hello", "hello", 0),
    ("levenshtein", dict(), "hello", "hell", 1),
    ("levenshtein", dict(), "", "", 0),
    ("levenshtein", dict(), "", "abc", 3),
    ("levenshtein", dict(), "abc", "", 3),
    ("levenshtein", dict(), "abc", "ac", 1),
    ("levenshtein", dict(), "abc", "a_bc", 1),
    ("levenshtein", dict(), "abc", "adc", 1),
    ("levenshtein", dict(), "ggbuzgjux{}l", "gbuzgjux{}l", 1),
    ("levenshtein", dict(), "abcdefgABCDEFG", "ABCDEFGabcdefg", 14),
    # test/similarities.py:160-168 (UTF-8, unit costs)
    ("levenshtein_utf8", dict(), "hello", "hell", 1),
    ("levenshtein_utf8", dict(), "\U0002070e \U00020731 \U00020779 \U00020c53",
     "\U0002070e\U00020731\U00020779\U00020c53", 3),
    ("levenshtein_utf8", dict(), "\U0001f496", "\U0001f497", 1),
    ("levenshtein_utf8", dict(), "αβγδ", "αγδ", 1),
    ("levenshtein_utf8", dict(), "école", "école", 2),
    ("levenshtein_utf8", dict(), "façade", "facade", 1),
    ("levenshtein_utf8", dict(), "Schön", "Schön", 2),
    ("levenshtein_utf8", dict(), "München", "Muenchen", 2),
    ("levenshtein_utf8", dict(), "こんにちは世界",
     "こんばんは世界", 2),
    # test/similarities.py:188-197 (mismatch 4, open 3, extend 2)
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "hello", "hello", 0),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "hello", "hell", 3),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "", "", 0),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "", "abc", 3 + 2 * 2),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "abc", "", 3 + 2 * 2),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "abc", "ac", 3),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "abc", "a_bc", 3),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "abc", "adc", 4),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "ggbuzgjux{}l", "gbuzgjux{}l", 3),
    ("levenshtein", dict(mismatch=4, open=3, extend=2), "abcdefgABCDEFG", "ABCDEFGabcdefg", min(14 * 4, 2 * 3 + 12 * 2)),
    # test/similarities.py:219-227 (UTF-8, mismatch 4, open = extend = 3)
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "hello", "hell", 3),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "\U0002070e \U00020731 \U00020779 \U00020c53",
     "\U0002070e\U00020731\U00020779\U00020c53", 9),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "\U0001f496", "\U0001f497", 4),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "αβγδ", "αγδ", 3),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "école", "école", 7),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "façade", "facade", 4),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "Schön", "Schön", 7),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "München", "Muenchen", 7),
    ("levenshtein_utf8", dict(mismatch=4, open=3, extend=3), "こんにちは世界",
     "こんばんは世界", min(2 * 4, 4 * 3)),
]


def protein_tables(rng, asymmetric=False):
    byte_to_class = np.zeros(256, dtype=np.uint8)
    for i, c in enumerate(b"ACDEFGHIKLMNPQRSTVWY"):
        byte_to_class[c] = i + 1
    costs = rng.integers(-4, 4, size=(32, 32)).astype(np.int8)
    if not asymmetric:
        costs = np.triu(costs) + np.triu(costs, 1).T
    for i in range(32):
        costs[i, i] = rng.integers(4, 12)
    return byte_to_class, np.ascontiguousarray(costs)


def b64(data: bytes) -> str:
    return base64.b64encode(data).decode()


def main():
    if not O.reference_available("serial"):
        raise SystemExit("oracle/_ref/libszref_serial.so is missing: run `make -C oracle` with the reference mounted")
    rng = np.random.default_rng(0x60D)
    b2c, costs = protein_tables(rng)
    b2c_a, costs_a = protein_tables(rng, asymmetric=True)
    utf8_alphabet = [chr(c) for c in list(range(0x61, 0x67)) + list(range(0x430, 0x436)) + [0x4E2D, 0x1F600]]
    configs = [
        ("levenshtein", dict(), b"ACGT"),
        ("levenshtein", dict(mismatch=4, open=3, extend=2), b"ACGT"),
        ("levenshtein", dict(match=1, mismatch=3, open=2, extend=2), b"ACGTN"),
        ("levenshtein", dict(mismatch=90, open=70, extend=60), b"ab"),  # crosses the 8/16-bit tiers early
        ("levenshtein_utf8", dict(), None),
        ("levenshtein_utf8", dict(mismatch=4, open=3, extend=2), None),
        ("needleman_wunsch", dict(open=-11, extend=-1, t=0), b"ACDEFGHIKLMNPQRSTVWY"),
        ("needleman_wunsch", dict(open=-2, extend=-2, t=0), b"ACDEFGHIKLMNPQRSTVWY"),
        ("needleman_wunsch", dict(open=-5, extend=-1, t=1), b"ACDEFGHIKLMNPQRSTVWY"),
        ("needleman_wunsch", dict(open=3, extend=1, t=0), b"ACDEFGHIKLMNPQRSTVWY"),
        ("smith_waterman", dict(open=-2, extend=-2, t=0), b"ACDEFGHIKLMNPQRSTVWY"),
        ("smith_waterman", dict(open=-5, extend=-1, t=1), b"ACDEFGHIKLMNPQRSTVWY"),
        ("smith_waterman", dict(open=3, extend=1, t=1), b"ACDEFGHIKLMNPQRSTVWY"),
    ]
    vectors = []
    for kind, kwargs, alphabet in configs:
        kwargs = dict(kwargs)
        table = kwargs.pop("t", None)
        params = O.Params(kind, **kwargs)
        if table is not None:
            params.byte_to_class, params.class_costs = (b2c, costs) if table == 0 else (b2c_a, costs_a)
        firsts, seconds = [], []
        for low, high in [(0, 6), (5, 40), (30, 140), (120, 260)]:
            for _ in range(8):
                if alphabet is None:
                    firsts.append("".join(rng.choice(utf8_alphabet, size=rng.integers(low, high + 1))).encode())
                    seconds.append("".join(rng.choice(utf8_alphabet, size=rng.integers(low, high + 1))).encode())
                else:
                    pool = np.frombuffer(alphabet, dtype=np.uint8)
                    firsts.append(bytes(rng.choice(pool, size=rng.integers(low, high + 1))))
                    seconds.append(bytes(rng.choice(pool, size=rng.integers(low, high + 1))))
        expected = O.reference_pairs(params, O.tape(firsts), O.tape(seconds), threads=4, variant="serial")
        vectors.append({
            "kind": kind,
            "match": params.match, "mismatch": params.mismatch, "open": params.open, "extend": params.extend,
            "table": table,
            "firsts": [b64(s) for s in firsts],
            "seconds": [b64(s) for s in seconds],
            "expected": [int(x) for x in expected],
        })
    known = [{"kind": kind, "params": params, "first": a, "second": b, "expected": expected}
             for kind, params, a, b, expected in KNOWN_ANSWERS]
    tables = {
        "0": {"byte_to_class": b64(b2c.tobytes()), "class_costs": b64(costs.tobytes())},
        "1": {"byte_to_class": b64(b2c_a.tobytes()), "class_costs": b64(costs_a.tobytes())},
    }
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "scorer": "reference serial (oracle/_ref)",
                   "tables": tables, "known_answers": known, "reference_vectors": vectors}, f, indent=0)
    print(f"wrote {OUT}: {len(known)} known answers, {sum(len(v['expected']) for v in vectors)} reference vectors")


if __name__ == "__main__":
    main()
