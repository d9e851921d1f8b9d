"""Pins the fingerprint oracle (oracle/sz_fingerprints.c).

- splitmix64 against published outputs of Vigna's generator;
- every window hash against the Rabin-Karp polynomial evaluated with exact Python integers, with a brute-force
  minimum and count, so a dropped term, a wrong sign or a wrong reduction in the rolling f64 steps shows;
- the whole fingerprint against the reference's own engines compiled into oracle/_ref/ (serial sliced, AVX2 and
  AVX-512 sliced, and the per-dimension engine), including texts long enough for the reference's chunked path
  (serial.hpp:817-860, texts of 256 KiB and more on one thread).
"""

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, random_bytes

MASK64 = (1 << 64) - 1


def _window_hashes(text: bytes, multiplier: int, modulo: int, window: int):
    """Every window's hash as the textbook polynomial sum((b + 1) * multiplier^(window - 1 - k)) mod modulo."""
    powers = [pow(multiplier, window - 1 - k, modulo) for k in range(window)]
    return [sum((text[j + k] + 1) * powers[k] for k in range(window)) % modulo for j in range(len(text) - window + 1)]


def _closed_form(text: bytes, hashers, ndim: int):
    hashes, counts = [], []
    for d in range(ndim):
        h = hashers[d]
        values = _window_hashes(text, int(h.multiplier), int(h.modulo), int(h.window_width))
        if not values:
            hashes.append(0xFFFFFFFF), counts.append(0)
            continue
        smallest = min(values)
        hashes.append(smallest & 0xFFFFFFFF), counts.append(values.count(smallest))
    return np.array(hashes, dtype=np.uint32), np.array(counts, dtype=np.uint32)


def test_splitmix64_published_outputs():
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    state, golden = 1234567, 0x9E3779B97F4A7C15
    expected = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                16408922859458223821]
    for value in expected:
        assert O.splitmix64(state) == value
        state = (state + golden) & MASK64


def test_hasher_parameters_follow_their_definitions():
    hashers, sliced = O.fingerprint_hashers(1024, seed=42)
    assert sliced
    for d in (0, 1, 63, 64, 500, 1023):
        h = hashers[d]
        mixed = O.splitmix64(42 + d)
        assert h.multiplier == 256 + mixed % 384
        assert h.modulo == 4503599626977 - O.splitmix64(mixed) % (1 << 20)
        assert h.window_width == O.DEFAULT_WINDOW_WIDTHS[(d // 64) % 8]
        m, mod, w = int(h.multiplier), int(h.modulo), int(h.window_width)
        assert -h.negative_discarding_multiplier == pow(m, w - 1, mod)
        assert h.discarding_multiplier == (-pow(m, w, mod)) % mod
    hashers, sliced = O.fingerprint_hashers(7, seed=0)
    assert not sliced
    assert [hashers[d].window_width for d in range(7)] == list(O.DEFAULT_WINDOW_WIDTHS[:7])


@pytest.mark.parametrize("ndim,widths", [(7, None), (64, None), (24, [2, 3, 8]), (128, [1, 5])])
def test_oracle_matches_the_polynomial_closed_form(ndim, widths):
    rng = np.random.default_rng(ndim)
    hashers, _ = O.fingerprint_hashers(ndim, widths, seed=5)
    texts = [random_bytes(rng, DNA, 0, 40) for _ in range(6)] + [bytes(rng.integers(0, 256, 60, dtype=np.uint8))]
    texts += [b"aaaaaaaaaaaaaaaaaaaa", b"ab" * 15, b"x"]  # repeated windows: counts above 1
    hashes, counts = O.fingerprints(O.tape(texts), ndim, widths, seed=5)
    for i, text in enumerate(texts):
        expected_hashes, expected_counts = _closed_form(text, hashers, ndim)
        np.testing.assert_array_equal(hashes[i], expected_hashes)
        np.testing.assert_array_equal(counts[i], expected_counts)


def test_degenerate_texts():
    hashes, counts = O.fingerprints(O.tape([b"", b"ab", b"a" * 40, b"a" * 40]), 64)
    assert (hashes[0] == 0xFFFFFFFF).all() and (counts[0] == 0).all()
    short = O.fingerprint_hashers(64)[0]
    for d in range(64):  # shorter than the window: skipped; otherwise exactly one window
        if short[d].window_width > 2:
            assert hashes[1, d] == 0xFFFFFFFF and counts[1, d] == 0
    # a run of one byte repeats every window: the minimum is that one hash, counted once per window
    for d in range(64):
        assert counts[2, d] == 40 - short[d].window_width + 1
    np.testing.assert_array_equal(hashes[2], hashes[3])


@pytest.mark.skipif(not O.reference_fingerprints_available("serial"),
                    reason="oracle/_ref is built only where the reference is mounted")
@pytest.mark.parametrize("variant", ["serial", "haswell", "skylake"])
@pytest.mark.parametrize("ndim,widths", [(1, None), (7, None), (64, None), (1024, None), (256, [4, 8]),
                                         (192, [3, 33, 64]), (100, [2, 6, 9])])
def test_live_differential_against_reference(variant, ndim, widths):
    if not O.reference_fingerprints_available(variant):
        pytest.skip(f"{variant} reference build needs the matching CPU")
    rng = np.random.default_rng(ndim * 7 + len(variant))
    texts = [bytes(rng.integers(0, 256, int(rng.integers(0, 400)), dtype=np.uint8)) for _ in range(150)]
    texts += [random_bytes(rng, DNA, 0, 200) for _ in range(50)] + [b"", b"a", b"hello", b"hello", b"a" * 70]
    tape = O.tape(texts)
    hashes, counts = O.fingerprints(tape, ndim, widths, seed=ndim)
    ref_hashes, ref_counts = O.reference_fingerprints(tape, ndim, widths, seed=ndim, variant=variant)
    np.testing.assert_array_equal(hashes, ref_hashes)
    np.testing.assert_array_equal(counts, ref_counts)


@pytest.mark.skipif(not O.reference_fingerprints_available("serial"),
                    reason="oracle/_ref is built only where the reference is mounted")
def test_per_dimension_engine_and_long_texts_against_reference():
    rng = np.random.default_rng(77)
    # 300 KiB of DNA: above the reference's 256 KiB single-thread threshold, so its chunked path runs.
    texts = [random_bytes(rng, DNA, 300_000, 300_000), bytes(rng.integers(0, 256, 270_000, dtype=np.uint8))]
    tape = O.tape(texts)
    for ndim, per_dimension in ((512, False), (7, True)):  # the sliced engines, then the per-dimension one
        hashes, counts = O.fingerprints(tape, ndim, seed=3)
        ref_hashes, ref_counts = O.reference_fingerprints(tape, ndim, seed=3, per_dimension=per_dimension)
        np.testing.assert_array_equal(hashes, ref_hashes)
        np.testing.assert_array_equal(counts, ref_counts)


def load_golden():
    import base64
    import json
    import os

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fingerprints.json")) as f:
        payload = json.load(f)
    texts = [base64.b64decode(s) for s in payload["texts"]]
    cases = []
    for case in payload["cases"]:
        shape = (len(texts), case["ndim"])
        hashes = np.frombuffer(base64.b64decode(case["hashes"]), dtype="<u4").reshape(shape)
        counts = np.frombuffer(base64.b64decode(case["counts"]), dtype="<u4").reshape(shape)
        cases.append((case["ndim"], case["window_widths"], case["seed"], hashes, counts))
    return texts, cases


def test_oracle_matches_the_committed_reference_vectors():
    texts, cases = load_golden()
    tape = O.tape(texts)
    for ndim, widths, seed, hashes, counts in cases:
        got_hashes, got_counts = O.fingerprints(tape, ndim, widths, seed=seed)
        np.testing.assert_array_equal(got_hashes, hashes)
        np.testing.assert_array_equal(got_counts, counts)
