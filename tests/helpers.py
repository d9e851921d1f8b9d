"""Input generators shared by the test-suites (seeded, numpy only)."""

from __future__ import annotations

import numpy as np

PROTEIN = b"ACDEFGHIKLMNPQRSTVWY"
DNA = b"ACGT"
UTF8_ALPHABET = (
    [chr(c) for c in range(0x20, 0x7F)]
    + [chr(c) for c in range(0x400, 0x500)]
    + [chr(c) for c in range(0x4E00, 0x4E80)]
    + [chr(c) for c in range(0x1F300, 0x1F380)]
)


def random_bytes(rng, alphabet: bytes, low: int, high: int) -> bytes:
    length = int(rng.integers(low, high + 1))
    return bytes(rng.choice(np.frombuffer(alphabet, dtype=np.uint8), size=length))


def random_utf8(rng, low: int, high: int) -> bytes:
    length = int(rng.integers(low, high + 1))
    return "".join(rng.choice(UTF8_ALPHABET, size=length)).encode("utf-8")


def protein_tables(rng, asymmetric: bool = False):
    """A class table over the 20 amino acids (classes 1..20) and a seeded substitution matrix."""
    byte_to_class = np.zeros(256, dtype=np.uint8)
    for i, c in enumerate(PROTEIN):
        byte_to_class[c] = i + 1
    costs = rng.integers(-4, 4, size=(32, 32)).astype(np.int8)
    if not asymmetric:
        costs = np.triu(costs) + np.triu(costs, 1).T
    for i in range(32):
        costs[i, i] = rng.integers(4, 12)
    return byte_to_class, np.ascontiguousarray(costs, dtype=np.int8)


def dna_tables(match: int = 2, mismatch: int = -3):
    byte_to_class = np.zeros(256, dtype=np.uint8)
    for i, c in enumerate(DNA):
        byte_to_class[c] = i + 1
    costs = np.full((32, 32), mismatch, dtype=np.int8)
    np.fill_diagonal(costs, match)
    return byte_to_class, costs
