"""GPU parity of large all-pairs products, where a wave sweeps many consecutive items of one query.

The queue hands out up to 16 cheap items per draw at these sizes (8 per draw for the 100-byte 4,096^2 product).

At these sizes the queue hands out several cheap items per draw and the pairs of an item share their query, so the
bit-parallel groups keep the query's LDS tables from one item to the next instead of rebuilding them. Queries of
small alphabets (match-word tables) alternate with wide ones (bit-planes, no table to keep), so kept and rebuilt
tables follow each other on the same waves. The candidates are a second tape with the same strings, so entry
(q, c) and entry (c, q) are computed independently with different rows strings: the product must be symmetric with
a zero diagonal at full size, and whole query rows plus a random sample equal the oracle.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")

ACGT = np.frombuffer(b"acgt", dtype=np.uint8)
WIDE = np.frombuffer(bytes(range(48, 48 + 40)), dtype=np.uint8)


def _strings(rng, count, length):
    strings = []
    for i in range(count):
        alphabet = WIDE if i % 7 == 3 else ACGT
        strings.append(bytes(rng.choice(alphabet, size=length)))
    return strings


@pytest.mark.parametrize("length", [100, 128, 200])
def test_all_pairs_products_keep_query_tables(length):
    rng = np.random.default_rng(length)
    count = 1024
    strings = _strings(rng, count, length)
    engine = szs.LevenshteinDistances()
    got = engine(szs.Strs(strings), szs.Strs(list(strings))).astype(np.int64)
    assert got.shape == (count, count)
    np.testing.assert_array_equal(np.diagonal(got), 0)
    mismatched = np.argwhere(got != got.T)
    assert len(mismatched) == 0, f"asymmetric at {mismatched[:5].tolist()}"
    params = O.Params("levenshtein")
    rows = [0, 1, 3, 17, 511, 1023]
    expected = O.cross(params, O.tape([strings[q] for q in rows]), O.tape(strings), threads=16)
    np.testing.assert_array_equal(got[rows], expected)
    picks = rng.integers(0, count, size=(2000, 2))
    a = O.tape([strings[q] for q, _ in picks])
    b = O.tape([strings[c] for _, c in picks])
    expected = O.pairs(params, a, b, threads=16)
    np.testing.assert_array_equal(got[picks[:, 0], picks[:, 1]], expected)


@pytest.mark.parametrize("length", [60, 100])
def test_all_pairs_products_with_many_items_per_draw(length):
    """4,096^2 products: the queue hands out 16 (60-byte strings, DP groups) or 8 (100 bytes, Myers groups with
    kept query tables) cheap items per draw. Every entry must be written and equal its transposed twin, and whole
    query rows equal the oracle."""
    rng = np.random.default_rng(1000 + length)
    count = 4096
    strings = _strings(rng, count, length)
    engine = szs.LevenshteinDistances()
    got = engine(szs.Strs(strings), szs.Strs(list(strings))).astype(np.int64)
    assert got.shape == (count, count)
    np.testing.assert_array_equal(np.diagonal(got), 0)
    assert (got >= 0).all() and (got <= length).all()
    assert np.array_equal(got, got.T)
    rows = [0, 2049, 4095]
    expected = O.cross(O.Params("levenshtein"), O.tape([strings[q] for q in rows]), O.tape(strings), threads=16)
    np.testing.assert_array_equal(got[rows], expected)
