"""GPU parity of the packed 16-bit Smith-Waterman sweeps, linear and affine gaps (two cells per register, values
around a moving base, the zero floor one constant per lane and chunk, the best folded chunk by chunk).

The local slots' ladders are replaced by ladders that send every pair to the packed shapes (one stripe, several
stripes on one wave with 2-, 4- or 6-row tails, the cooperative workgroup), and every score is compared bit-exactly
with the CPU oracle. Engines on both sides of the range bound run: profile bytes that need no correction (the open
cost <= every cost) and ones that do (a mismatch below it, as in the usual DNA scoring), ones admitting only 4 rows
per lane (their 6- and 8-row shapes run the int32 twin), and ones admitting none. Long similar strings drive the
best score far beyond 16 bits, and local matches embedded in random flanks put it in the middle of the matrix.
"""

import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, dna_tables, protein_tables, random_bytes

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import _native, autotune  # noqa: E402

TABLE = autotune.shapes()
PACKED = {(s[1], s[2]): i for i, s in enumerate(TABLE) if s[4] and not s[5]}  # (rows per lane, cooperative) -> id
LADDERS = {
    "wave8": ([(PACKED[(8, 0)], 1 << 20)], PACKED[(8, 1)]),
    "mixed": ([(PACKED[(4, 0)], 256), (PACKED[(6, 0)], 384), (PACKED[(8, 0)], 1536)], PACKED[(8, 1)]),
    "cooperative": ([], PACKED[(8, 1)]),
}
EDGES = [1, 2, 3, 5, 63, 64, 65, 127, 128, 129, 191, 255, 256, 257, 383, 384, 385, 511, 512, 513, 1023, 1025]


AFFINE = {  # name -> (alphabet, open, extend)
    "sw_affine_protein": ("protein", -11, -1),  # costs + 11 >= 0: no correction
    "sw_affine_small_open": ("protein", -3, -1),  # costs below the open cost: biased profile
    "sw_affine_dna": ("dna", -5, -2),
    "sw_affine_dna_biased": ("dna", -2, -1),  # mismatch -3 below the open cost
    "sw_affine_far": ("protein", -110, -80),  # delta = 201: admits 4 rows per lane only
}


def _engine(name, rng):
    if name in AFFINE:
        kind, open_, extend = AFFINE[name]
        b2c, costs = dna_tables(2, -3) if kind == "dna" else protein_tables(rng)
        engine = szs.SmithWatermanScores(b2c, costs, open_, extend)
        params = O.Params("smith_waterman", open=open_, extend=extend, byte_to_class=b2c, class_costs=costs)
        return engine, params, DNA if kind == "dna" else PROTEIN
    if name.startswith("sw_dna"):
        b2c, costs = dna_tables(2, -3)
        gap = {"sw_dna": -2, "sw_dna_wide_gap": -5}[name]
        alphabet = DNA
    elif name == "sw_zero_match":  # match 0, mismatch -1, gap -1: every best score is 0
        b2c = np.zeros(256, dtype=np.uint8)
        b2c[np.frombuffer(DNA, dtype=np.uint8)] = np.arange(1, 5, dtype=np.uint8)
        costs = np.full((32, 32), -1, dtype=np.int8)
        np.fill_diagonal(costs, 0)
        gap, alphabet = -1, DNA
    else:
        b2c, costs = protein_tables(rng)
        alphabet = PROTEIN
        if name == "sw_protein":  # gap <= every cost: no correction of the profile bytes
            gap = -4
        elif name == "sw_protein_small_gap":  # costs below the gap: profile bytes biased by the smallest cost
            gap = -2
        elif name == "sw_big_diagonal":  # delta = 122
            costs = costs.copy()
            np.fill_diagonal(costs, 100)
            gap = -11
        elif name == "sw_far_gap":  # delta = 201: admits 4 rows per lane only
            gap = -95
        elif name == "sw_wide_range":  # costs from -40 to 100: `cost - gap` fits int8, `cost + 40` no byte
            costs = costs.copy()
            np.fill_diagonal(costs, 100)
            costs[1, 2] = costs[2, 1] = -40
            gap = -20
        elif name == "sw_over_byte":  # 120 - gap does not fit int8: the plain recurrences run
            costs = costs.copy()
            np.fill_diagonal(costs, 120)
            gap = -11
        else:
            raise KeyError(name)
    engine = szs.SmithWatermanScores(b2c, costs, gap, gap)
    params = O.Params("smith_waterman", open=gap, extend=gap, byte_to_class=b2c, class_costs=costs)
    return engine, params, alphabet


EXPECTED_ROWS = {"sw_protein": 8, "sw_protein_small_gap": 8, "sw_dna": 8, "sw_dna_wide_gap": 8, "sw_zero_match": 8,
                 "sw_big_diagonal": 8, "sw_far_gap": 4, "sw_wide_range": 0, "sw_over_byte": 0,
                 "sw_affine_protein": 8, "sw_affine_small_open": 8, "sw_affine_dna": 8, "sw_affine_dna_biased": 8,
                 "sw_affine_far": 4}


@pytest.fixture
def packed_ladders():
    """Installs a ladder on both local byte slots, restoring the tuned ones afterwards."""
    slots = [autotune.slot_of("sw_linear"), autotune.slot_of("sw_affine")]
    saved = {slot: autotune.get_ladder(slot) for slot in slots}

    def install(name):
        ladder, beyond = LADDERS[name]
        for slot in slots:
            autotune.set_ladder(slot, ladder, beyond)

    yield install
    for slot, (ladder, beyond) in saved.items():
        autotune.set_ladder(slot, ladder, beyond)


def _mutated(rng, alphabet, text: bytes, substitutions: float, indels: float) -> bytes:
    out = bytearray()
    for byte in text:
        roll = rng.random()
        if roll < indels / 2:
            continue  # deletion
        if roll < indels:
            out += bytes(rng.choice(list(alphabet), size=int(rng.integers(1, 4))))  # insertion
        out.append(int(rng.choice(list(alphabet))) if rng.random() < substitutions else byte)
    return bytes(out)


def _check(engine, params, a, b, label):
    expected = O.pairs(params, O.tape(a), O.tape(b), threads=16)
    got = engine.pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected, err_msg=label)
    return expected


def test_engines_admit_the_expected_packed_rows():
    rng = np.random.default_rng(5)
    for name, rows in EXPECTED_ROWS.items():
        engine, _, _ = _engine(name, rng)
        assert _native.lib.szs_tuning_packed_rows(engine.handle) == rows, name


@pytest.mark.parametrize("ladder", list(LADDERS))
@pytest.mark.parametrize("name", list(EXPECTED_ROWS))
def test_packed_local_shapes_match_oracle(name, ladder, packed_ladders):
    rng = np.random.default_rng(zlib.crc32(f"{name}:{ladder}".encode()))
    engine, params, alphabet = _engine(name, rng)
    packed_ladders(ladder)
    a = [random_bytes(rng, alphabet, n, n) for n in EDGES for _ in EDGES]
    b = [random_bytes(rng, alphabet, m, m) for _ in EDGES for m in EDGES]
    a += [random_bytes(rng, alphabet, 1, 1400) for _ in range(40)]
    b += [random_bytes(rng, alphabet, 1, 1400) for _ in range(40)]
    expected = _check(engine, params, a, b, f"{name} {ladder}")
    if name == "sw_zero_match":
        assert not expected.any()
    else:
        assert expected.max() > 0


@pytest.mark.parametrize("ladder", ["wave8", "cooperative"])
@pytest.mark.parametrize("name", ["sw_protein", "sw_protein_small_gap", "sw_dna", "sw_big_diagonal", "sw_far_gap",
                                  "sw_affine_protein", "sw_affine_small_open", "sw_affine_dna_biased", "sw_affine_far"])
def test_packed_local_exact_scores_far_beyond_16_bits(name, ladder, packed_ladders):
    """Identical and mutated long strings (the best grows along the whole diagonal, far past a 16-bit window), and
    local matches embedded at random offsets in random flanks (the best sits in the middle of the matrix)."""
    rng = np.random.default_rng(zlib.crc32(f"drift:{name}:{ladder}".encode()))
    engine, params, alphabet = _engine(name, rng)
    packed_ladders(ladder)
    a, b = [], []
    for length in (2900, 4100, 7000):
        text = random_bytes(rng, alphabet, length, length)
        a += [text, text, text]
        b += [text, _mutated(rng, alphabet, text, 0.05, 0.01), _mutated(rng, alphabet, text, 0.3, 0.05)]
    for core, flank in ((150, 1200), (600, 700), (1500, 3000)):
        text = random_bytes(rng, alphabet, core, core)
        left, right = int(rng.integers(0, flank)), int(rng.integers(0, flank))
        a.append(random_bytes(rng, alphabet, left, left) + text + random_bytes(rng, alphabet, right, right))
        b.append(_mutated(rng, alphabet, text, 0.1, 0.02))
    expected = _check(engine, params, a, b, f"{name} {ladder} long similar pairs")
    if alphabet == PROTEIN:
        assert int(expected[6]) > 32767  # a 7,000-residue self-alignment scores far beyond int16


TAIL_ROWS = [1, 127, 128, 129, 255, 256, 257, 383, 384, 385, 512, 513, 640, 641, 768, 769, 896, 897, 1024, 1152,
             1153, 1280, 1281, 1408, 1409, 2000]


@pytest.mark.parametrize("name", ["sw_protein_small_gap", "sw_dna", "sw_far_gap", "sw_affine_protein",
                                  "sw_affine_dna_biased", "sw_affine_far"])
def test_packed_local_tail_stripes(name, packed_ladders):
    """Every pair on 512-row stripes: the last stripe alone or after full ones, on both sides of each tail size
    (2, 4 and 6 packed rows per lane); the best of the full stripes and the tail's are combined."""
    rng = np.random.default_rng(zlib.crc32(f"tails:{name}".encode()))
    engine, params, alphabet = _engine(name, rng)
    packed_ladders("wave8")
    a, b = [], []
    for n in TAIL_ROWS:
        for m in (n, n + 1, n + 130, 3000):
            rows, columns = random_bytes(rng, alphabet, n, n), random_bytes(rng, alphabet, m, m)
            a.append(rows if m % 2 else columns)
            b.append(columns if m % 2 else rows)
    # The best in the tail only, and in the full stripes only.
    text = random_bytes(rng, alphabet, 300, 300)
    noise = random_bytes(rng, alphabet, 1800, 1800)
    a += [noise[:1500] + text, text + noise[:1500]]
    b += [text, text]
    _check(engine, params, a, b, f"{name} tail stripes")


@pytest.mark.parametrize("name", ["sw_dna", "sw_protein", "sw_affine_protein", "sw_affine_dna_biased"])
def test_packed_local_cross_products(name, packed_ladders):
    """A symmetric all-pairs product (each score also written to its mirrored slot) and a rectangular one."""
    rng = np.random.default_rng(zlib.crc32(f"cross:{name}".encode()))
    engine, params, alphabet = _engine(name, rng)
    packed_ladders("mixed")
    strings = [random_bytes(rng, alphabet, 200, 1100) for _ in range(48)]
    tape = szs.Strs(strings)
    got = engine(tape, tape).astype(np.int64)
    expected = O.cross(params, O.tape(strings), O.tape(strings), threads=16)
    np.testing.assert_array_equal(got, expected)
    others = [random_bytes(rng, alphabet, 1, 1500) for _ in range(20)]
    got = engine(szs.Strs(strings), szs.Strs(others)).astype(np.int64)
    np.testing.assert_array_equal(got, O.cross(params, O.tape(strings), O.tape(others), threads=16))


def test_packed_local_batches_and_graph_replays(packed_ladders):
    """The packed local sweeps inside a mixed batch (linear and affine gaps, the profile corrected or not, beside
    Levenshtein pairs), as a persistent launch, one launch per pair and graph replays."""
    rng = np.random.default_rng(31)
    packed_ladders("mixed")
    engine, params, _ = _engine("sw_dna", rng)
    affine, affine_params, _ = _engine("sw_affine_protein", rng)
    lev = szs.LevenshteinDistances()
    a = [random_bytes(rng, DNA, 1, 3000) for _ in range(60)]
    b = [random_bytes(rng, DNA, 1, 3000) for _ in range(60)]
    c = [random_bytes(rng, PROTEIN, 1, 3000) for _ in range(60)]
    d = [random_bytes(rng, PROTEIN, 1, 3000) for _ in range(60)]
    dna_a = [random_bytes(rng, b"ACGT", 1, 2000) for _ in range(40)]
    dna_b = [random_bytes(rng, b"ACGT", 1, 2000) for _ in range(40)]
    expected = [O.pairs(params, O.tape(a), O.tape(b), threads=16),
                O.pairs(affine_params, O.tape(c), O.tape(d), threads=16),
                O.pairs(O.Params("levenshtein"), O.tape(dna_a), O.tape(dna_b), threads=16)]
    segments = [(engine, szs.Strs(a), szs.Strs(b)), (affine, szs.Strs(c), szs.Strs(d)),
                (lev, szs.Strs(dna_a), szs.Strs(dna_b))]
    for mode in (szs.BATCH_PERSISTENT, szs.BATCH_PER_PAIR, szs.BATCH_GRAPH):
        batch = szs.Batch(segments, mode=mode)
        for replay in range(3 if mode == szs.BATCH_GRAPH else 1):
            for out in batch.outputs:
                out.fill(777)
            for got, want in zip(batch.run(), expected):
                np.testing.assert_array_equal(got.astype(np.int64), want, err_msg=f"mode {mode} replay {replay}")


@pytest.mark.parametrize("name", ["sw_dna", "sw_affine_dna_biased"])
@pytest.mark.parametrize("ladder", ["wave8", "mixed"])
def test_packed_local_rows_past_the_pair_never_raise_the_best(ladder, name, packed_ladders):
    """Columns with long runs of 'N' (class 0, which scores +2 against itself under the DNA matrix) against ACGT rows
    that end inside their stripe: a row past the pair given a real class-0 profile would collect those matches, so
    these pairs pin that such rows never raise the best."""
    rng = np.random.default_rng(zlib.crc32(f"past:{ladder}:{name}".encode()))
    engine, params, _ = _engine(name, rng)
    packed_ladders(ladder)
    a, b = [], []
    for n in (1, 40, 130, 300, 450, 513, 700, 1100):
        rows = random_bytes(rng, DNA, n, n)
        columns = b"N" * 400 + random_bytes(rng, DNA, 300, 900) + b"N" * 600
        a.append(rows if n % 2 else columns)
        b.append(columns if n % 2 else rows)
    expected = _check(engine, params, a, b, f"{ladder} rows past the pair")
    # A best of at most 2 per row: for the short rows far below what the 600-byte 'N' run would give such a row.
    assert all(score <= 2 * min(len(x), len(y)) for x, y, score in zip(a, b, expected))
