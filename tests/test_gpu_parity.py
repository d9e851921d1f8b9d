"""GPU parity: every engine, shape and mode scored by the HIP kernels (through the C ABI) against the CPU oracle.

Bit-exact integer equality is required. Lengths are chosen so every compiled kernel shape runs, including
multi-pass stripes (rows > 1024) and 16-lane groups holding four pairs per wave.
"""

import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, dna_tables, protein_tables, random_bytes, random_utf8

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")

# The last two ranges reach the cooperative workgroup tier (rows > 2048 for class costs, > 4096 for uniform).
RANGES = [(0, 20), (10, 70), (60, 300), (200, 600), (500, 1100), (1000, 2600), (2100, 3600), (4200, 6000)]


def _wide_tables(rng):
    """Costs whose difference to the gap-open cost leaves int8: these engines run the unshifted-profile sweep."""
    byte_to_class, _ = protein_tables(rng)
    costs = rng.integers(-100, 101, size=(32, 32)).astype(np.int8)
    return byte_to_class, np.ascontiguousarray(costs)


def _engines(rng):
    b2c, costs = protein_tables(rng)
    b2c_a, costs_a = protein_tables(rng, asymmetric=True)
    dna_b2c, dna_costs = dna_tables()
    b2c_w, costs_w = _wide_tables(rng)
    return {
        "lev_unit": (szs.LevenshteinDistances(), O.Params("levenshtein"), DNA, False),
        "lev_affine": (szs.LevenshteinDistances(0, 4, 3, 2), O.Params("levenshtein", 0, 4, 3, 2), DNA, False),
        "lev_linear": (szs.LevenshteinDistances(1, 3, 2, 2), O.Params("levenshtein", 1, 3, 2, 2), b"ACGTN", False),
        "utf8_unit": (szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8"), None, True),
        "utf8_affine": (szs.LevenshteinDistancesUTF8(0, 4, 3, 2), O.Params("levenshtein_utf8", 0, 4, 3, 2), None, True),
        "nw_affine": (szs.NeedlemanWunschScores(b2c, costs, -11, -1),
                      O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs), PROTEIN, False),
        "nw_linear": (szs.NeedlemanWunschScores(b2c, costs, -2, -2),
                      O.Params("needleman_wunsch", open=-2, extend=-2, byte_to_class=b2c, class_costs=costs), PROTEIN, False),
        "nw_asymmetric": (szs.NeedlemanWunschScores(b2c_a, costs_a, -5, -1),
                          O.Params("needleman_wunsch", open=-5, extend=-1, byte_to_class=b2c_a, class_costs=costs_a), PROTEIN, False),
        "nw_odd_signs": (szs.NeedlemanWunschScores(b2c, costs, 3, 1),
                         O.Params("needleman_wunsch", open=3, extend=1, byte_to_class=b2c, class_costs=costs), PROTEIN, False),
        "sw_linear": (szs.SmithWatermanScores(dna_b2c, dna_costs, -2, -2),
                      O.Params("smith_waterman", open=-2, extend=-2, byte_to_class=dna_b2c, class_costs=dna_costs), DNA, False),
        "sw_affine": (szs.SmithWatermanScores(b2c, costs, -5, -1),
                      O.Params("smith_waterman", open=-5, extend=-1, byte_to_class=b2c, class_costs=costs), PROTEIN, False),
        "sw_odd_signs": (szs.SmithWatermanScores(b2c_a, costs_a, 3, 1),
                         O.Params("smith_waterman", open=3, extend=1, byte_to_class=b2c_a, class_costs=costs_a), PROTEIN, False),
        "nw_wide_affine": (szs.NeedlemanWunschScores(b2c_w, costs_w, -100, -40),
                           O.Params("needleman_wunsch", open=-100, extend=-40, byte_to_class=b2c_w, class_costs=costs_w),
                           PROTEIN, False),
        "sw_wide_linear": (szs.SmithWatermanScores(b2c_w, costs_w, -90, -90),
                           O.Params("smith_waterman", open=-90, extend=-90, byte_to_class=b2c_w, class_costs=costs_w),
                           PROTEIN, False),
    }


ENGINE_NAMES = list(_engines(np.random.default_rng(0)).keys())


def _strings(rng, alphabet, utf8, count, low, high):
    if utf8:
        return [random_utf8(rng, low, high) for _ in range(count)]
    return [random_bytes(rng, alphabet, low, high) for _ in range(count)]


@pytest.mark.parametrize("low,high", RANGES)
@pytest.mark.parametrize("name", ENGINE_NAMES)
def test_pairs_match_oracle(name, low, high):
    rng = np.random.default_rng(zlib.crc32(f"{name}:{low}:{high}".encode()))
    engine, params, alphabet, utf8 = _engines(rng)[name]
    count = 96 if high <= 600 else 24 if high <= 1100 else 8 if high <= 2600 else 3
    a = _strings(rng, alphabet, utf8, count, low, high)
    b = _strings(rng, alphabet, utf8, count, low, high)
    expected = O.pairs(params, O.tape(a), O.tape(b), threads=16)
    got = engine.pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected, err_msg=f"{name} lengths {low}-{high}")


@pytest.mark.parametrize("name", ENGINE_NAMES)
def test_cross_product_and_symmetric(name):
    rng = np.random.default_rng(7)
    engine, params, alphabet, utf8 = _engines(rng)[name]
    queries = _strings(rng, alphabet, utf8, 9, 0, 90)
    candidates = _strings(rng, alphabet, utf8, 13, 0, 300)
    expected = O.cross(params, O.tape(queries), O.tape(candidates), threads=16)
    got = engine(szs.Strs(queries), szs.Strs(candidates))
    assert got.shape == (9, 13)
    np.testing.assert_array_equal(got.astype(np.int64), expected)
    expected_sym = O.cross(params, O.tape(queries), None, threads=16)
    got_sym = engine(szs.Strs(queries))
    np.testing.assert_array_equal(got_sym.astype(np.int64), expected_sym)


def test_strided_out_buffer_and_untouched_padding():
    engine = szs.LevenshteinDistances()
    queries, candidates = szs.Strs(["abc", "hello", ""]), szs.Strs(["abd", "", "help", "xyz"])
    out = np.full((3, 6), 777, dtype=np.uint64)
    result = engine(queries, candidates, out=out)
    assert result is out
    expected = O.cross(O.Params("levenshtein"), O.tape(["abc", "hello", ""]), O.tape(["abd", "", "help", "xyz"]))
    np.testing.assert_array_equal(out[:, :4].astype(np.int64), expected)
    assert (out[:, 4:] == 777).all()


def test_known_answers_reference_cases():
    """Known answers from the reference's test/similarities.py:134-143 and 160-168."""
    lev = szs.LevenshteinDistances()
    cases = [("hello", "hello", 0), ("hello", "hell", 1), ("", "", 0), ("", "abc", 3), ("abc", "", 3),
             ("abc", "ac", 1), ("abc", "a_bc", 1), ("abc", "adc", 1), ("ggbuzgjux{}l", "gbuzgjux{}l", 1),
             ("abcdefgABCDEFG", "ABCDEFGabcdefg", 14)]
    got = lev.pairs(szs.Strs([c[0] for c in cases]), szs.Strs([c[1] for c in cases]))
    assert got.tolist() == [c[2] for c in cases]
    utf8 = szs.LevenshteinDistancesUTF8()
    cases = [("hello", "hell", 1), ("𠜎 𠜱 𠝹 𠱓", "𠜎𠜱𠝹𠱓", 3), ("💖", "💗", 1), ("αβγδ", "αγδ", 1),
             ("\u00e9cole", "e\u0301cole", 2), ("façade", "facade", 1), ("Sch\u00f6n", "Scho\u0308n", 2),
             ("München", "Muenchen", 2), ("こんにちは世界", "こんばんは世界", 2)]
    got = utf8.pairs(szs.Strs([c[0] for c in cases]), szs.Strs([c[1] for c in cases]))
    assert got.tolist() == [c[2] for c in cases]


def test_cpu_scope_is_rejected():
    """A GPU engine on a CPU scope fails with device_code_mismatch (reference test/stringzillas.py:65-73)."""
    engine = szs.LevenshteinDistances()
    with pytest.raises(RuntimeError):
        engine(szs.Strs(["a"]), szs.Strs(["b"]), device=szs.DeviceScope(cpu_cores=2))


def test_batch_modes_agree_and_graph_replays():
    rng = np.random.default_rng(11)
    b2c, costs = protein_tables(rng)
    lev = szs.LevenshteinDistances()
    utf8 = szs.LevenshteinDistancesUTF8()
    nw = szs.NeedlemanWunschScores(b2c, costs, -11, -1)
    segments = [
        (lev, szs.Strs([random_bytes(rng, DNA, 1, 1500) for _ in range(150)]),
         szs.Strs([random_bytes(rng, DNA, 1, 1500) for _ in range(150)])),
        (utf8, szs.Strs([random_utf8(rng, 0, 400) for _ in range(100)]),
         szs.Strs([random_utf8(rng, 0, 400) for _ in range(100)])),
        (nw, szs.Strs([random_bytes(rng, PROTEIN, 50, 700) for _ in range(80)]),
         szs.Strs([random_bytes(rng, PROTEIN, 50, 700) for _ in range(80)])),
    ]
    params = [O.Params("levenshtein"), O.Params("levenshtein_utf8"),
              O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs)]
    expected = [O.pairs(p, O.tape([s[1][i] for i in range(len(s[1]))]), O.tape([s[2][i] for i in range(len(s[2]))]),
                        threads=16) for p, s in zip(params, segments)]
    for mode in (szs.BATCH_PERSISTENT, szs.BATCH_PER_PAIR, szs.BATCH_GRAPH):
        batch = szs.Batch(segments, mode=mode)
        for replay in range(3 if mode == szs.BATCH_GRAPH else 1):
            for out in batch.outputs:
                out.fill(12345)
            outputs = batch.run()
            for got, want in zip(outputs, expected):
                np.testing.assert_array_equal(got.astype(np.int64), want, err_msg=f"mode {mode} replay {replay}")
        assert batch.cells > 0


@pytest.mark.parametrize("alphabet_name", ["binary", "dna", "thirteen", "fourteen", "printable", "all_bytes"])
def test_unit_levenshtein_bit_parallel_alphabets(alphabet_name):
    """Unit costs take the bit-parallel tier above 64 rows; its symbol renumbering must not depend on the alphabet.
    Rows of at most 13 distinct bytes take the match-word table, 14 and more the bit-planes."""
    alphabets = {"binary": b"ab", "dna": DNA, "thirteen": b"abcdefghijklm", "fourteen": b"abcdefghijklmn",
                 "printable": bytes(range(32, 127)), "all_bytes": bytes(range(256))}
    rng = np.random.default_rng(zlib.crc32(alphabet_name.encode()))
    alphabet = alphabets[alphabet_name]
    lengths = [(50, 300), (250, 1100), (1000, 4200), (4000, 9000)]
    a, b = [], []
    for low, high in lengths:
        for _ in range(6 if high < 5000 else 2):
            a.append(random_bytes(rng, alphabet, low, high))
            b.append(random_bytes(rng, alphabet, low, high))
    expected = O.pairs(O.Params("levenshtein"), O.tape(a), O.tape(b), threads=16)
    got = szs.LevenshteinDistances().pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected)


def test_unit_levenshtein_cooperative_tall_pair():
    """Rows above 16,384 run on all four waves of a workgroup (cooperative bit-parallel stripes)."""
    rng = np.random.default_rng(5)
    a = [random_bytes(rng, DNA, 17000, 18000), random_bytes(rng, b"ACGTN", 16500, 17500)]
    b = [random_bytes(rng, DNA, 17000, 19000), random_bytes(rng, b"ACGTN", 16500, 17000)]
    b[0] = a[0][:9000] + b[0][:3000] + a[0][9000:]  # a related pair too
    expected = O.pairs(O.Params("levenshtein"), O.tape(a), O.tape(b), threads=16)
    got = szs.LevenshteinDistances().pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected)


@pytest.mark.parametrize("low,high", [(1, 40), (200, 400), (900, 3000)])
def test_utf8_malformed_and_truncated_input(low, high):
    """Stray continuation bytes, leads without their continuations and sequences cut by the end of a string:
    the unchecked decode (utf8_runes/serial.h:101-114) fixes the runes, on the one-lane and whole-wave paths."""
    rng = np.random.default_rng(low + high)
    pieces = [b"a", b"Z", b"\xd0\x96", b"\xe4\xb8\x80", b"\xf0\x9f\x8c\x80", b"\x80", b"\xbf", b"\xc3", b"\xe2\x82",
              b"\xf0\x9f", b"\xff", b"\xfe"]

    def make(clean: bool):
        size = int(rng.integers(low, high + 1))
        out = bytearray()
        while len(out) < size:
            out += pieces[int(rng.integers(0, 5 if clean else len(pieces)))]
        return bytes(out[:size])  # the cut may split a sequence at the end

    firsts = [make(i % 3 == 0) for i in range(60)]
    seconds = [make(i % 4 == 0) for i in range(60)]
    for params, engine in ((O.Params("levenshtein_utf8"), szs.LevenshteinDistancesUTF8()),
                           (O.Params("levenshtein_utf8", 0, 4, 3, 2), szs.LevenshteinDistancesUTF8(0, 4, 3, 2))):
        a, b = szs.Strs(firsts), szs.Strs(seconds)
        expected = O.pairs(params, (a.data, a.offsets), (b.data, b.offsets), threads=8)
        np.testing.assert_array_equal(engine.pairs(a, b), expected)
        np.testing.assert_array_equal(engine.pairs(szs.to_device(a), szs.to_device(b)), expected)


@pytest.mark.parametrize("rows_per_lane", [8, 10])
@pytest.mark.parametrize("family", ["nw_affine", "nw_linear", "sw_affine", "sw_linear", "lev_affine"])
def test_sixteen_lane_groups(family, rows_per_lane):
    """Four pairs per wave on 16 lanes of 8 or 10 rows (up to 128 or 160 rows; class-cost profiles of 10 rows per
    lane take 12 bytes per lane and class), forced through the tuning ladder."""
    from stringzillas_amd import autotune

    table = autotune.shapes()
    shape = next(i for i, s in enumerate(table) if tuple(s) == (16, rows_per_lane, 0, 0, 0, 0))
    capacity = 16 * rows_per_lane
    slot = autotune.slot_of(family)
    saved = autotune.get_ladder(slot)
    rng = np.random.default_rng(zlib.crc32(f"g16r{rows_per_lane}:{family}".encode()))
    if family != "lev_affine":
        b2c, costs = protein_tables(rng)
        kind, engine_type = {"nw": ("needleman_wunsch", szs.NeedlemanWunschScores),
                             "sw": ("smith_waterman", szs.SmithWatermanScores)}[family[:2]]
        gaps = (-11, -1) if family.endswith("affine") else (-4, -4)
        engine = engine_type(b2c, costs, *gaps)
        params = O.Params(kind, open=gaps[0], extend=gaps[1], byte_to_class=b2c, class_costs=costs)
        alphabet = PROTEIN
    else:
        engine = szs.LevenshteinDistances(0, 4, 3, 2)
        params = O.Params("levenshtein", 0, 4, 3, 2)
        alphabet = DNA
    edges = [n for n in (1, 2, 15, 16, 17, 63, 64, 65, 127, 128, 129, 150, 159, 160) if n <= capacity]
    a = [random_bytes(rng, alphabet, n, n) for n in edges for _ in range(3)]
    b = [random_bytes(rng, alphabet, 1, 3000) for _ in edges for _ in range(3)]
    a += [random_bytes(rng, alphabet, 1, capacity) for _ in range(60)]
    b += [random_bytes(rng, alphabet, 1, 2000) for _ in range(60)]
    try:
        autotune.set_ladder(slot, [(shape, capacity)], saved[1])
        expected = O.pairs(params, O.tape(a), O.tape(b), threads=16)
        got = engine.pairs(szs.Strs(a), szs.Strs(b))
        np.testing.assert_array_equal(got.astype(np.int64), expected)
    finally:
        autotune.set_ladder(slot, *saved)
