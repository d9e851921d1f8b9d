"""GPU parity of the bit-parallel unit-cost Levenshtein over UTF-8 runes (Myers over raw rune bit-planes).

Every bit-parallel shape (4-, 16- and 64-lane groups, several stripes on one wave, the cooperative workgroup) is
forced through the tuning ladder of the unit-cost rune slot and compared bit-exactly with the CPU oracle. The
alphabets set the number of planes a wave needs (ASCII 7 bits, Cyrillic 11, CJK 16, emoji 17, up to the 21 bits
a malformed 4-byte sequence decodes to), and pairs whose columns hold runes wider than every row rune check that
those match nothing.
"""

import zlib

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import autotune  # noqa: E402

TABLE = autotune.shapes()
# (lanes per pair, rows per lane, cooperative) -> shape id of the bit-parallel shapes
MYERS = {(s[0], s[1], s[2]): i for i, s in enumerate(TABLE) if s[3] and not s[5]}
LADDERS = {
    "small_groups": ([(MYERS[(4, 64, 0)], 256), (MYERS[(16, 64, 0)], 1024)], MYERS[(64, 64, 1)]),
    "wave": ([(MYERS[(64, 64, 0)], 1 << 20)], MYERS[(64, 64, 1)]),
    "cooperative": ([], MYERS[(64, 64, 1)]),
    # 32-row words: 16-lane groups up to 512 rows, then 2048-row stripes on one wave for every taller pair.
    "half_words": ([(MYERS[(16, 32, 0)], 512), (MYERS[(64, 32, 0)], 1 << 20)], MYERS[(64, 64, 1)]),
    # 4-lane groups of 32-row words up to 128 rows, then 4-lane groups of 64-row words.
    "small_half_words": ([(MYERS[(4, 32, 0)], 128), (MYERS[(4, 64, 0)], 256), (MYERS[(16, 64, 0)], 1024)],
                         MYERS[(64, 64, 1)]),
}
ALPHABETS = {
    "ascii": [chr(c) for c in range(0x20, 0x7F)],
    "cyrillic": [chr(c) for c in range(0x400, 0x500)] + list("abc"),
    "cjk": [chr(c) for c in range(0x4E00, 0x4E00 + 300)] + [chr(0x9FFF)],
    "mixed": [chr(c) for c in range(0x20, 0x7F)] + [chr(c) for c in range(0x400, 0x500)]
    + [chr(c) for c in range(0x4E00, 0x4E00 + 500)] + [chr(c) for c in range(0x1F300, 0x1F3FF)],
    "nul_and_astral": ["\x00", "a", "\U0010FFFF", "\U00010000", "z", "߿", "ࠀ"],
}
EDGES = [1, 2, 63, 64, 65, 255, 256, 257, 1023, 1025, 4095, 4097]


def _text(rng, alphabet, length):
    return "".join(rng.choice(alphabet, size=length)).encode("utf-8")


@pytest.fixture
def rune_ladder():
    slot = autotune.slot_of("lev_linear", runes=True, unit=True)
    saved = autotune.get_ladder(slot)

    def install(name):
        ladder, beyond = LADDERS[name]
        autotune.set_ladder(slot, ladder, beyond)

    yield install
    autotune.set_ladder(slot, *saved)


@pytest.fixture
def byte_ladder():
    slot = autotune.slot_of("lev_linear", unit=True)
    saved = autotune.get_ladder(slot)

    def install(name):
        ladder, beyond = LADDERS[name]
        autotune.set_ladder(slot, ladder, beyond)

    yield install
    autotune.set_ladder(slot, *saved)


def _check(a, b, label):
    engine = szs.LevenshteinDistancesUTF8()
    expected = O.pairs(O.Params("levenshtein_utf8"), O.tape(a), O.tape(b), threads=16)
    got = engine.pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected, err_msg=label)


@pytest.mark.parametrize("ladder", list(LADDERS))
@pytest.mark.parametrize("alphabet", list(ALPHABETS))
def test_rune_myers_matches_oracle(alphabet, ladder, rune_ladder):
    rng = np.random.default_rng(zlib.crc32(f"{alphabet}:{ladder}".encode()))
    rune_ladder(ladder)
    letters = ALPHABETS[alphabet]
    edges = [e for e in EDGES if ladder not in ("small_groups", "small_half_words") or e <= 1025]
    if ladder == "half_words":
        edges += [511, 512, 513, 2047, 2048, 2049]
    a = [_text(rng, letters, n) for n in edges for _ in edges]
    b = [_text(rng, letters, m) for _ in edges for m in edges]
    a += [_text(rng, letters, int(rng.integers(0, 3000))) for _ in range(30)]
    b += [_text(rng, letters, int(rng.integers(0, 3000))) for _ in range(30)]
    _check(a, b, f"{alphabet} {ladder}")


@pytest.mark.parametrize("ladder", list(LADDERS))
def test_rune_myers_wider_columns_and_malformed_bytes(ladder, rune_ladder):
    """Rows of narrow runes against columns of wider ones (no match across widths), and raw malformed bytes."""
    rng = np.random.default_rng(zlib.crc32(f"wide:{ladder}".encode()))
    rune_ladder(ladder)
    ascii_, cjk, mixed = ALPHABETS["ascii"], ALPHABETS["cjk"], ALPHABETS["mixed"]
    a, b = [], []
    for n in (5, 70, 300, 1100, 2100):
        narrow = _text(rng, ascii_, n)
        a += [narrow, narrow, _text(rng, cjk, n)]
        # Columns: the same ASCII text with every other rune replaced by a wide one whose low bits equal it.
        wide = "".join(chr(0x1F300 | (ord(c) & 0xFF)) if i % 2 else c for i, c in enumerate(narrow.decode()))
        b += [wide.encode(), _text(rng, mixed, n + 7), _text(rng, ascii_, n // 2 + 1)]
    # Malformed input: stray continuation bytes, truncated sequences and 4-byte leads that decode past U+10FFFF.
    for n in (40, 500, 1500):
        raw = bytes(rng.choice([0x41, 0x80, 0xBF, 0xC3, 0xE2, 0xF4, 0xF7, 0x9F, 0xA4], size=n).tolist())
        a.append(raw)
        b.append(bytes(rng.choice([0x41, 0x80, 0xF7, 0xBF, 0x9F], size=n + 3).tolist()))
    _check(a, b, f"wide columns and malformed bytes, {ladder}")


@pytest.mark.parametrize("ladder", ["half_words", "small_half_words"])
@pytest.mark.parametrize("alphabet", [b"AC", b"ACGT", bytes(range(32, 127)), bytes(range(256))])
def test_half_word_myers_over_bytes(alphabet, ladder, byte_ladder):
    """32-row words over bytes: table mode (few symbols) and bit-planes (up to 256), one and several stripes, in
    16-lane and 64-lane groups and in 4-lane groups (up to 128 rows, 16 pairs per wave)."""
    rng = np.random.default_rng(zlib.crc32(alphabet + ladder.encode()))
    byte_ladder(ladder)
    table = np.frombuffer(alphabet, dtype=np.uint8)
    lengths = [1, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 300, 511, 512, 513, 2047, 2048, 2049, 5000]
    a = [bytes(rng.choice(table, size=n)) for n in lengths for _ in lengths]
    b = [bytes(rng.choice(table, size=m)) for _ in lengths for m in lengths]
    engine = szs.LevenshteinDistances()
    expected = O.pairs(O.Params("levenshtein"), O.tape(a), O.tape(b), threads=16)
    got = engine.pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected)


# Rows whose last 4096-row stripe keeps at most 2048 rows (32-row words) or more (64-row words).
MYERS_TAIL_ROWS = [1, 2047, 2048, 2049, 4096, 4097, 6144, 6145, 8192, 8193, 11000]


@pytest.mark.parametrize("alphabet", [b"ACGT", bytes(range(256))])
def test_myers_tail_stripes_over_bytes(alphabet, byte_ladder):
    """Every pair on 4096-row stripes: the last stripe runs 32-row words when its rows fit (table and planes)."""
    rng = np.random.default_rng(zlib.crc32(b"tails:" + alphabet))
    byte_ladder("wave")
    table = np.frombuffer(alphabet, dtype=np.uint8)
    a = [bytes(rng.choice(table, size=n)) for n in MYERS_TAIL_ROWS for _ in range(2)]
    b = [bytes(rng.choice(table, size=n + extra)) for n in MYERS_TAIL_ROWS for extra in (0, 70)]
    engine = szs.LevenshteinDistances()
    expected = O.pairs(O.Params("levenshtein"), O.tape(a), O.tape(b), threads=16)
    got = engine.pairs(szs.Strs(a), szs.Strs(b))
    np.testing.assert_array_equal(got.astype(np.int64), expected)


@pytest.mark.parametrize("alphabet", ["cjk", "mixed"])
def test_myers_tail_stripes_over_runes(alphabet, rune_ladder):
    rng = np.random.default_rng(zlib.crc32(f"tails:{alphabet}".encode()))
    rune_ladder("wave")
    letters = ALPHABETS[alphabet]
    a = [_text(rng, letters, n) for n in MYERS_TAIL_ROWS for _ in range(2)]
    b = [_text(rng, letters, n + extra) for n in MYERS_TAIL_ROWS for extra in (1, 90)]
    _check(a, b, f"{alphabet} tail stripes")


@pytest.mark.parametrize("alphabet", ["cyrillic", "cjk", "mixed"])
def test_wave_decode_across_lane_and_pass_boundaries(alphabet):
    """Strings decode in passes of 1 KB from the 16-byte boundary at or before their first byte, as rows of 64 byte
    positions: sequences straddle rows and passes, strings start at every offset within a 16-byte block, and strings
    cut at byte lengths around those edges end inside a sequence."""
    rng = np.random.default_rng(zlib.crc32(f"decode:{alphabet}".encode()))
    letters = ALPHABETS[alphabet]
    sizes = [1, 2, 3, 15, 16, 17, 63, 64, 65, 127, 256, 257, 271, 1008, 1023, 1024, 1025, 1039, 2047, 2048, 2049,
             3073, 5000]
    a = [_text(rng, letters, size)[:size] for size in sizes for _ in range(3)]
    b = [_text(rng, letters, size + 5)[: size + int(rng.integers(0, 40))] for size in sizes for _ in range(3)]
    _check(a, b, f"{alphabet} decode boundaries")


@pytest.mark.parametrize("position", [0, 1, 62, 63, 64, 1021, 1022, 1023, 1024, 1025, 2047, 2048])
def test_wave_decode_rejects_malformed_bytes_at_row_and_pass_edges(position):
    """A stray continuation byte, a lead cut short by an ASCII byte, or a lead 0xF8-0xFF placed at a row or pass
    edge of an otherwise well-formed string sends the string to the one-lane walk, whose runes the oracle fixes."""
    rng = np.random.default_rng(position)
    letters = ALPHABETS["cjk"]
    a, b = [], []
    for flaw in (b"\x80", b"\xe4a", b"\xf0\x9fa", b"\xf9\x80\x80\x80", b"\xbf\xbf"):
        for skew in (0, 5, 15):
            clean = _text(rng, letters, position + 64)
            text = (b"x" * skew + clean)[: skew + position] + flaw + clean[: 40]
            a.append(text)
            b.append(_text(rng, letters, 80))
    _check(a, b, f"malformed at {position}")
