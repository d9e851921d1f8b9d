"""GPU parity of scores beyond 32 bits and of the int64-cell shape.

The reference picks its cell width per pair from the reach bound (include/stringzillas/similarities/serial.hpp:
117-140, 293-356) and moves to 8-byte cells when 32 bits could overflow. Here the int32 sweeps keep values relative
to their anti-diagonal (global families) or absolute local scores, which stay within a bound that grows with the
shorter side; only the final score is formed in 64 bits. Pairs beyond that bound run the int64 shape. The tests:
the int64 shape forced on every engine through the tuning ladders, then pairs whose scores leave int32 (a
Levenshtein distance in [2^31, 2^32), where the reference uses u32 cells, and an NW score below -2^31) and a pair
that the range bound sends to the int64 shape, each against the oracle (int64 cells throughout).
"""

import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, dna_tables, protein_tables, random_bytes, random_utf8

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import autotune  # noqa: E402

WIDE = autotune.wide_shapes()[0]


@pytest.fixture
def every_slot_wide():
    """Every shape-table slot sends every pair to the int64 shape; the tuned ladders come back afterwards."""
    slots = range(autotune._native.lib.szs_tuning_slots())
    saved = {slot: autotune.get_ladder(slot) for slot in slots}
    for slot in slots:
        autotune.set_ladder(slot, [], WIDE)
    yield
    for slot, (ladder, beyond) in saved.items():
        autotune.set_ladder(slot, ladder, beyond)


def _engines(rng):
    b2c, costs = protein_tables(rng)
    b2c_a, costs_a = protein_tables(rng, asymmetric=True)
    dna_b2c, dna_costs = dna_tables()
    wide_costs = np.ascontiguousarray(rng.integers(-100, 101, size=(32, 32)).astype(np.int8))
    nw = lambda b, c, o, e: (szs.NeedlemanWunschScores(b, c, o, e),  # noqa: E731
                             O.Params("needleman_wunsch", open=o, extend=e, byte_to_class=b, class_costs=c))
    sw = lambda b, c, o, e: (szs.SmithWatermanScores(b, c, o, e),  # noqa: E731
                             O.Params("smith_waterman", open=o, extend=e, byte_to_class=b, class_costs=c))
    return {
        "lev_unit": ((szs.LevenshteinDistances(), O.Params("levenshtein")), DNA),
        "lev_affine": ((szs.LevenshteinDistances(0, 4, 3, 2), O.Params("levenshtein", 0, 4, 3, 2)), DNA),
        "lev_linear": ((szs.LevenshteinDistances(1, 3, 2, 2), O.Params("levenshtein", 1, 3, 2, 2)), b"ACGTN"),
        "utf8_unit": ((szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8")), None),
        "utf8_affine": ((szs.LevenshteinDistancesUTF8(0, 4, 3, 2), O.Params("levenshtein_utf8", 0, 4, 3, 2)), None),
        "nw_affine": (nw(b2c, costs, -11, -1), PROTEIN),
        "nw_linear": (nw(b2c, costs, -2, -2), PROTEIN),
        "nw_asymmetric": (nw(b2c_a, costs_a, -5, -1), PROTEIN),
        "nw_odd_signs": (nw(b2c, costs, 3, 1), PROTEIN),
        "nw_wide_affine": (nw(b2c, wide_costs, -100, -40), PROTEIN),
        "sw_linear": (sw(dna_b2c, dna_costs, -2, -2), DNA),
        "sw_affine": (sw(b2c, costs, -5, -1), PROTEIN),
        "sw_odd_signs": (sw(b2c_a, costs_a, 3, 1), PROTEIN),
        "sw_wide_linear": (sw(b2c, wide_costs, -90, -90), PROTEIN),
    }


NAMES = list(_engines(np.random.default_rng(0)))


@pytest.mark.parametrize("name", NAMES)
def test_int64_shape_matches_oracle(name, every_slot_wide):
    """One stripe, several stripes (512-row stripes), ragged edges, empty sides, either side longer."""
    rng = np.random.default_rng(zlib.crc32(f"wide:{name}".encode()))
    (engine, params), alphabet = _engines(rng)[name]

    def make(low, high):
        return random_utf8(rng, low, high) if alphabet is None else random_bytes(rng, alphabet, low, high)

    sizes = [(0, 5), (1, 1), (1, 700), (63, 64), (64, 65), (511, 513), (512, 512), (513, 700), (900, 1300),
             (1500, 1100), (2100, 2500)]
    a = [make(n, n) for n, _ in sizes] + [make(0, 800) for _ in range(30)]
    b = [make(m, m) for _, m in sizes] + [make(0, 800) for _ in range(30)]
    expected = O.pairs(params, O.tape(a), O.tape(b), threads=16)
    np.testing.assert_array_equal(engine.pairs(szs.Strs(a), szs.Strs(b)).astype(np.int64), expected)
    # A cross-product and the per-pair mode through the same shape.
    np.testing.assert_array_equal(engine(szs.Strs(a[:7]), szs.Strs(b[:5])).astype(np.int64),
                                  O.cross(params, O.tape(a[:7]), O.tape(b[:5]), threads=16))
    batch = szs.Batch([(engine, szs.Strs(a), szs.Strs(b))], mode=szs.BATCH_PER_PAIR)
    np.testing.assert_array_equal(batch.run()[0].astype(np.int64), expected)


def test_levenshtein_distance_in_the_unsigned_32_bit_band():
    """match 0, mismatch/gap 127: 17 Mi 'A' against 'C' costs 127 x 17 Mi = 2,264,924,160, past 2^31 (the reference
    keeps such distances in u32 cells)."""
    engine = szs.LevenshteinDistances(0, 127, 127, 127)
    long_side = b"A" * (17 << 20)
    got = engine.pairs(szs.Strs([long_side, b"C"]), szs.Strs([b"C", long_side]))
    assert got.tolist() == [127 * (17 << 20)] * 2
    assert got[0] > 2**31
    assert int(got[0]) == O.score(O.Params("levenshtein", 0, 127, 127, 127), long_side, b"C")


def test_needleman_wunsch_score_below_minus_2_31():
    """Linear gaps of -127 over a 100 x 17 Mi pair: the score is near -127 x 17 Mi, below int32, while every
    stored value stays near its anti-diagonal."""
    rng = np.random.default_rng(127)
    b2c, costs = protein_tables(rng)
    costs = costs.copy()
    costs[1, 2] = costs[2, 1] = -127  # magnitude 127 over the matrix as well
    engine = szs.NeedlemanWunschScores(b2c, costs, -127, -127)
    params = O.Params("needleman_wunsch", open=-127, extend=-127, byte_to_class=b2c, class_costs=costs)
    short, long_side = random_bytes(rng, PROTEIN, 100, 100), random_bytes(rng, PROTEIN, 17 << 20, 17 << 20)
    want = O.score(params, short, long_side)
    assert want < -(2**31)
    got = engine.pairs(szs.Strs([short, long_side]), szs.Strs([long_side, short])).astype(np.int64)
    assert got.tolist() == [want, want]


def test_wide_costs_over_millions_of_columns_take_the_int64_shape():
    """Costs whose bias leaves int8 run the plain recurrences, whose scores grow with n + m: a 64 x 4 Mi pair at
    magnitude 100 passes the int32 bound and is scored with int64 cells, like the oracle."""
    rng = np.random.default_rng(400)
    b2c, _ = protein_tables(rng)
    costs = np.ascontiguousarray(rng.integers(-100, 101, size=(32, 32)).astype(np.int8))
    engine = szs.NeedlemanWunschScores(b2c, costs, -100, -40)
    params = O.Params("needleman_wunsch", open=-100, extend=-40, byte_to_class=b2c, class_costs=costs)
    short, long_side = random_bytes(rng, PROTEIN, 64, 64), random_bytes(rng, PROTEIN, 4 << 20, 4 << 20)
    batch = szs.Batch([(engine, szs.Strs([short]), szs.Strs([long_side]))])
    assert batch.core_ops == 20 * batch.cells  # the int64 shape's accounting (affine)
    got = batch.run()[0].astype(np.int64)
    assert got.tolist() == [O.score(params, short, long_side)]


def test_pairs_of_millions_of_columns_keep_stripe_scratch_bounded():
    """A multi-stripe pair of 1.2 M columns: stripe-boundary rows for every resident workgroup would take ~29 GB, so
    the launch runs fewer workgroups and the batch stays within a few GB; the score equals the oracle."""
    rng = np.random.default_rng(1200)
    b2c, costs = protein_tables(rng)
    engine = szs.NeedlemanWunschScores(b2c, costs, -11, -1)
    params = O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs)
    short, long_side = random_bytes(rng, PROTEIN, 600, 600), random_bytes(rng, PROTEIN, 1_200_000, 1_200_000)
    batch = szs.Batch([(engine, szs.Strs([short, b"ACD"]), szs.Strs([long_side, b"ACE"]))])
    assert batch.device_bytes < (5 << 30)
    got = batch.run()[0].astype(np.int64)
    assert got.tolist() == [O.score(params, short, long_side), O.score(params, b"ACD", b"ACE")]
