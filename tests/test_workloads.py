"""CPU tests of the seeded benchmark inputs: sizes, length bounds, alphabets, UTF-8 validity and the matrix."""

import numpy as np
import pytest

from oracle import oracle as O
from stringzillas_amd import workloads as W


def test_levenshtein_config():
    w = W.levenshtein_c0()
    (s,) = w.segments
    assert s.count == 1024 and s.kind == "levenshtein"
    for which in ("firsts", "seconds"):
        lengths = s.lengths(which)
        assert lengths.min() >= 64 and lengths.max() <= 256
    assert set(np.unique(s.firsts_data[:-1]).tobytes()) <= set(b"ACGT")
    assert w.cells() == int(np.dot(s.lengths("firsts"), s.lengths("seconds")))
    assert W.levenshtein_c0().cells() == w.cells()  # seeded


def test_needleman_wunsch_config_and_matrix():
    w = W.needleman_wunsch_c1(2048)
    (s,) = w.segments
    lengths = np.concatenate([s.lengths("firsts"), s.lengths("seconds")])
    assert lengths.min() >= 50 and lengths.max() <= 2000
    assert 250 < np.median(lengths) < 450
    assert (s.open, s.extend) == (-11, -1)
    dense = W.dense_matrix_256(s.byte_to_class, s.class_costs)
    assert dense.shape == (256, 256) and (dense == dense.T).all()
    letters = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)
    diagonal = dense[letters, letters]
    assert diagonal.min() >= 4 and diagonal.max() <= 11
    off = dense[np.ix_(letters, letters)][~np.eye(20, dtype=bool)]
    assert off.min() >= -4 and off.max() <= 3


def test_smith_waterman_config_embeds_the_query():
    w = W.smith_waterman_c2(40)
    (s,) = w.segments
    assert (s.lengths("firsts") == 150).all()
    assert s.lengths("seconds").min() >= 150
    params = O.Params("smith_waterman", open=-2, extend=-2, byte_to_class=s.byte_to_class, class_costs=s.class_costs)
    scores = O.pairs(params, (s.firsts_data, s.firsts_offsets), (s.seconds_data, s.seconds_offsets), threads=4)
    # A copy with ~5% substitutions and ~1% short indels scores far above random DNA (~60 for these lengths).
    assert np.median(scores) > 150


def test_mixed_config_split_lengths_and_utf8():
    pytest.importorskip("torch")
    w = W.mixed_c3(4000, 4, 3000)
    kinds = [s.kind for s in w.segments]
    assert kinds == ["levenshtein", "levenshtein_utf8", "needleman_wunsch"]
    counts = np.array([s.count for s in w.segments]) - np.array([2, 1, 1])
    assert abs(counts[0] / 4000 - 0.5) < 0.05 and abs(counts[1] / 4000 - 0.25) < 0.05
    for s in w.segments:
        host = s.host()
        byte_lengths = np.diff(host.firsts_offsets.astype(np.int64))
        assert byte_lengths[:-2].max() <= 16384 and (byte_lengths[-1] == 3000 or s.kind == "levenshtein_utf8")
        if s.kind != "levenshtein_utf8":
            assert byte_lengths[:-2].min() >= 16
    utf8 = w.segments[1].host()
    text = bytes(utf8.firsts_data[: int(utf8.firsts_offsets[-1])])
    decoded = text.decode("utf-8")  # valid UTF-8
    ranges = W.UTF8_RANGES
    assert all(any(lo <= ord(ch) <= hi for lo, hi in ranges) for ch in decoded[:5000])
    runes = utf8.lengths("firsts")
    assert (runes <= np.diff(utf8.firsts_offsets.astype(np.int64))).all()
