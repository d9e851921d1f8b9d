"""GPU parity of the packed 16-bit Needleman-Wunsch sweeps (two cells per register, values around a moving base).

The default ladders are replaced by ladders that send every pair to the packed shapes (one stripe, several stripes
on one wave, the cooperative workgroup), and every score is compared bit-exactly with the CPU oracle. Engines on
both sides of the range bound run: those admitting 8 rows per lane, one admitting only 4 (its 8-row shapes run
the int32 twin), and one admitting none. Identical and near-identical long strings drive the absolute scores far
beyond 16 bits, which only the per-chunk rebasing keeps exact.
"""

import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import PROTEIN, protein_tables, random_bytes

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import _native, autotune  # noqa: E402

TABLE = autotune.shapes()
PACKED = {(s[1], s[2]): i for i, s in enumerate(TABLE) if s[4] and not s[5]}  # (rows per lane, cooperative) -> id
LADDERS = {
    # Every pair on 512-row packed stripes swept by one wave (taller pairs: several stripes through HBM).
    "wave8": ([(PACKED[(8, 0)], 1 << 20)], PACKED[(8, 1)]),
    # 256- and 384-row packed stripes for short pairs, 512-row ones up to 1536 rows, cooperative beyond.
    "mixed": ([(PACKED[(4, 0)], 256), (PACKED[(6, 0)], 384), (PACKED[(8, 0)], 1536)], PACKED[(8, 1)]),
    # Every pair on the whole workgroup, 512-row stripes pipelined over its four waves.
    "cooperative": ([], PACKED[(8, 1)]),
}
CHAINED = {s[4]: i for i, s in enumerate(TABLE) if s[5]}  # packed -> chained cooperative shape id
# Every pair split over 16 workgroups that hand bottom rows over through HBM (engines outside the packed range
# bound run the int32 twin of the packed shape).
CHAINED_LADDERS = {"chained": ([], CHAINED[1]), "chained_int32": ([], CHAINED[0])}
# Edge lengths: tiny, around the 64-column chunk and 128-step ramp, around the 256/384/512-row stripes.
EDGES = [1, 2, 3, 5, 63, 64, 65, 127, 128, 129, 191, 255, 256, 257, 383, 384, 385, 511, 512, 513, 1023, 1025]


def _engine(name, rng):
    b2c, costs = protein_tables(rng)
    if name == "nw_affine":
        open_, extend = -11, -1
    elif name == "nw_linear":
        open_, extend = -2, -2
    elif name == "nw_asymmetric":
        b2c, costs = protein_tables(rng, asymmetric=True)
        open_, extend = -5, -1
    elif name == "nw_wide_gap":  # delta = 52
        open_, extend = -40, -1
    elif name == "nw_big_diagonal":  # delta = 112
        costs = costs.copy()
        np.fill_diagonal(costs, 100)
        open_, extend = -11, -1
    elif name == "nw_far_gap":  # delta = 200: admits 4 rows per lane only (its 6- and 8-row shapes run int32)
        costs = (costs.astype(np.int64) - 85).astype(np.int8)
        open_, extend = -128, -72
    elif name == "nw_over_byte":  # 120 - (o + e) does not fit a profile byte: everything runs int32
        costs = costs.copy()
        np.fill_diagonal(costs, 120)
        open_, extend = -11, -1
    else:
        raise KeyError(name)
    engine = szs.NeedlemanWunschScores(b2c, costs, open_, extend)
    params = O.Params("needleman_wunsch", open=open_, extend=extend, byte_to_class=b2c, class_costs=costs)
    return engine, params


EXPECTED_ROWS = {"nw_affine": 8, "nw_linear": 8, "nw_asymmetric": 8, "nw_wide_gap": 8, "nw_big_diagonal": 8,
                 "nw_far_gap": 4, "nw_over_byte": 0}


@pytest.fixture
def packed_ladders():
    """Installs a ladder on both global class-cost byte slots, restoring the tuned ones afterwards."""
    slots = [autotune.slot_of("nw_affine"), autotune.slot_of("nw_linear")]
    saved = {slot: autotune.get_ladder(slot) for slot in slots}

    def install(name):
        ladder, beyond = {**LADDERS, **CHAINED_LADDERS}[name]
        for slot in slots:
            autotune.set_ladder(slot, ladder, beyond)

    yield install
    for slot, (ladder, beyond) in saved.items():
        autotune.set_ladder(slot, ladder, beyond)


def _mutated(rng, text: bytes, substitutions: float, indels: float) -> bytes:
    out = bytearray()
    for byte in text:
        roll = rng.random()
        if roll < indels / 2:
            continue  # deletion
        if roll < indels:
            out += bytes(rng.choice(list(PROTEIN), size=int(rng.integers(1, 4))))  # insertion
        out.append(int(rng.choice(list(PROTEIN))) if rng.random() < substitutions else byte)
    return bytes(out)


def _check(engine, params, a, b, label):
    expected = O.pairs(params, O.tape(a), O.tape(b), threads=16)
    got = engine.pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected, err_msg=label)


@pytest.mark.parametrize("ladder", list(LADDERS))
@pytest.mark.parametrize("name", list(EXPECTED_ROWS))
def test_packed_shapes_match_oracle(name, ladder, packed_ladders):
    rng = np.random.default_rng(zlib.crc32(f"{name}:{ladder}".encode()))
    engine, params = _engine(name, rng)
    assert _native.lib.szs_tuning_packed_rows(engine.handle) == EXPECTED_ROWS[name]
    packed_ladders(ladder)
    # Every pair of edge lengths once (rows the shorter side, either side first), plus random ones.
    a = [random_bytes(rng, PROTEIN, n, n) for n in EDGES for _ in EDGES]
    b = [random_bytes(rng, PROTEIN, m, m) for _ in EDGES for m in EDGES]
    a += [random_bytes(rng, PROTEIN, 1, 1400) for _ in range(40)]
    b += [random_bytes(rng, PROTEIN, 1, 1400) for _ in range(40)]
    _check(engine, params, a, b, f"{name} {ladder}")


@pytest.mark.parametrize("ladder", ["wave8", "cooperative"])
@pytest.mark.parametrize("name", ["nw_affine", "nw_linear", "nw_big_diagonal", "nw_far_gap"])
def test_packed_shapes_keep_exact_scores_far_beyond_16_bits(name, ladder, packed_ladders):
    """Identical and lightly mutated long strings: scores grow by ~10 per row, far past any fixed 16-bit window."""
    rng = np.random.default_rng(zlib.crc32(f"drift:{name}:{ladder}".encode()))
    engine, params = _engine(name, rng)
    packed_ladders(ladder)
    a, b = [], []
    for length in (2900, 4100, 7000):
        text = random_bytes(rng, PROTEIN, length, length)
        a += [text, text, text]
        b += [text, _mutated(rng, text, 0.05, 0.01), _mutated(rng, text, 0.3, 0.05)]
    _check(engine, params, a, b, f"{name} {ladder} long similar pairs")
    expected = O.pairs(params, O.tape(a[6:7]), O.tape(b[6:7]), threads=1)
    assert abs(int(expected[0])) > 32767  # a 7,000-residue self-alignment scores far beyond int16


# Rows whose last 512-row stripe keeps at most 256 rows, 257-384 or more: it runs 4, 6 or 8 rows per lane.
TAIL_ROWS = [1, 127, 128, 129, 255, 256, 257, 383, 384, 385, 512, 513, 640, 641, 768, 769, 896, 897, 1024, 1152,
             1153, 1280, 1281, 1408, 1409, 2000]


@pytest.mark.parametrize("name", ["nw_affine", "nw_linear", "nw_asymmetric", "nw_far_gap"])
def test_packed_tail_stripes(name, packed_ladders):
    """Every pair on 512-row stripes: the last stripe alone or after full ones, on both sides of each tail size
    (2, 4 and 6 packed rows per lane)."""
    rng = np.random.default_rng(zlib.crc32(f"tails:{name}".encode()))
    engine, params = _engine(name, rng)
    packed_ladders("wave8")
    a, b = [], []
    for n in TAIL_ROWS:
        for m in (n, n + 1, n + 130, 3000):
            rows, columns = random_bytes(rng, PROTEIN, n, n), random_bytes(rng, PROTEIN, m, m)
            # Either side first: the shorter side is the rows either way.
            a.append(rows if m % 2 else columns)
            b.append(columns if m % 2 else rows)
    text = random_bytes(rng, PROTEIN, 4400, 4400)  # 8 full stripes and a 304-row tail, scores beyond 16 bits
    a += [text, text]
    b += [text, _mutated(rng, text, 0.05, 0.01)]
    _check(engine, params, a, b, f"{name} tail stripes")


def test_packed_batches_and_graph_replays(packed_ladders):
    """The packed sweep inside a mixed batch, as a persistent launch, one launch per pair and graph replays."""
    rng = np.random.default_rng(29)
    packed_ladders("mixed")
    engine, params = _engine("nw_affine", rng)
    lev = szs.LevenshteinDistances()
    a = [random_bytes(rng, PROTEIN, 1, 3000) for _ in range(60)]
    b = [random_bytes(rng, PROTEIN, 1, 3000) for _ in range(60)]
    dna_a = [random_bytes(rng, b"ACGT", 1, 2000) for _ in range(40)]
    dna_b = [random_bytes(rng, b"ACGT", 1, 2000) for _ in range(40)]
    expected = [O.pairs(params, O.tape(a), O.tape(b), threads=16),
                O.pairs(O.Params("levenshtein"), O.tape(dna_a), O.tape(dna_b), threads=16)]
    segments = [(engine, szs.Strs(a), szs.Strs(b)), (lev, szs.Strs(dna_a), szs.Strs(dna_b))]
    for mode in (szs.BATCH_PERSISTENT, szs.BATCH_PER_PAIR, szs.BATCH_GRAPH):
        batch = szs.Batch(segments, mode=mode)
        for replay in range(3 if mode == szs.BATCH_GRAPH else 1):
            for out in batch.outputs:
                out.fill(777)
            for got, want in zip(batch.run(), expected):
                np.testing.assert_array_equal(got.astype(np.int64), want, err_msg=f"mode {mode} replay {replay}")


@pytest.mark.parametrize("ladder", list(CHAINED_LADDERS))
@pytest.mark.parametrize("name", ["nw_affine", "nw_linear", "nw_far_gap", "nw_big_diagonal"])
def test_chained_groups_match_oracle(name, ladder, packed_ladders):
    """One-stripe pairs (15 empty groups), one stripe per group, and pairs spanning all 16 groups."""
    rng = np.random.default_rng(zlib.crc32(f"chained:{name}:{ladder}".encode()))
    engine, params = _engine(name, rng)
    packed_ladders(ladder)
    sizes = [(1, 1), (5, 700), (511, 513), (513, 2000), (700, 700), (3000, 3100), (4100, 4096), (8193, 8300),
             (9100, 9000), (2, 9500)]
    a = [random_bytes(rng, PROTEIN, n, n) for n, _ in sizes]
    b = [random_bytes(rng, PROTEIN, m, m) for _, m in sizes]
    text = random_bytes(rng, PROTEIN, 8800, 8800)  # similar long strings: scores far beyond 16 bits
    a += [text, text]
    b += [text, _mutated(rng, text, 0.05, 0.01)]
    _check(engine, params, a, b, f"{name} {ladder}")


def test_chained_levenshtein_dp_and_batch_modes():
    """Chained int32 stripes for a non-unit Levenshtein, in a persistent launch, per pair and graph replays."""
    rng = np.random.default_rng(31)
    slot = autotune.slot_of("lev_affine")
    saved = autotune.get_ladder(slot)
    try:
        autotune.set_ladder(slot, [], CHAINED[0])
        engine = szs.LevenshteinDistances(0, 4, 3, 2)
        params = O.Params("levenshtein", 0, 4, 3, 2)
        a = [random_bytes(rng, b"ACGT", n, n) for n in (1, 600, 2500, 8400, 9000)]
        b = [random_bytes(rng, b"ACGT", m, m) for m in (700, 1, 2600, 8500, 8999)]
        expected = O.pairs(params, O.tape(a), O.tape(b), threads=16)
        segments = [(engine, szs.Strs(a), szs.Strs(b))]
        for mode in (szs.BATCH_PERSISTENT, szs.BATCH_PER_PAIR, szs.BATCH_GRAPH):
            batch = szs.Batch(segments, mode=mode)
            for replay in range(3 if mode == szs.BATCH_GRAPH else 1):
                batch.outputs[0].fill(99)
                got = batch.run()[0]
                np.testing.assert_array_equal(got.astype(np.int64), expected, err_msg=f"mode {mode} replay {replay}")
    finally:
        autotune.set_ladder(slot, *saved)


def test_chained_pairs_of_more_than_256_stripes(packed_ladders):
    """A pair over 4 x 64 stripes (131,072 rows) puts more than four stripes in each of its chained groups, which a
    group then sweeps four at a time, so each group waits about as long as all the groups before it take; the
    wait bound grows with that work, and the scores equal the closed form and the one-workgroup sweep."""
    rng = np.random.default_rng(140000)
    engine, params = _engine("nw_affine", rng)
    text = random_bytes(rng, PROTEIN, 140_000, 140_000)
    other = _mutated(rng, text, 0.05, 0.01)
    assert min(len(text), len(other)) > 4 * 64 * 512
    # NW(a, a) = sum of C[a_i][a_i] when no off-diagonal cost reaches a diagonal one and gaps cost (as here).
    costs, classes = params.class_costs.astype(np.int64), params.byte_to_class
    codes = classes[np.frombuffer(text, dtype=np.uint8)]
    self_score = int(costs[codes, codes].sum())
    packed_ladders("chained")
    chained = engine.pairs(szs.Strs([text, text, other]), szs.Strs([text, other, text])).astype(np.int64)
    assert chained[0] == self_score
    assert chained[1] == chained[2]  # a symmetric matrix scores both orders alike
    packed_ladders("cooperative")  # the same pairs on one workgroup, no chain
    single = engine.pairs(szs.Strs([text, text]), szs.Strs([text, other])).astype(np.int64)
    np.testing.assert_array_equal(chained[:2], single)
