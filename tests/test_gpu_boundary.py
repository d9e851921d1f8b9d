"""GPU parity through every reference-shaped entry point of include/stringzillas/stringzillas.h.

For the four engines: the ``sz_sequence_t`` callback forms (``szs_E``, stringzillas.h:184, 253, 325, 395), the
u32-tape forms (``szs_E_u32tape``, :199, 268, 340, 410) and the u64-tape forms, cross-product and symmetric,
with host and HBM-resident inputs and padded result rows, each bit-exact against the oracle. Also the scope's
grow-only workspace across calls of growing and shrinking sizes, and tape offsets in HBM that decrease or run past
the tape, which must fail the call (or only their own request when serving) without scoring anything.
"""

import ctypes
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, dna_tables, protein_tables, random_bytes, random_utf8

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import _native  # noqa: E402

KINDS = ["levenshtein", "levenshtein_affine", "levenshtein_utf8", "needleman_wunsch", "smith_waterman"]


def _setup(kind, rng):
    """(engine, oracle params, string maker) for one engine kind."""
    b2c, costs = protein_tables(rng)
    dna_b2c, dna_costs = dna_tables()
    if kind == "levenshtein":
        return szs.LevenshteinDistances(), O.Params("levenshtein"), lambda lo, hi: random_bytes(rng, DNA, lo, hi)
    if kind == "levenshtein_affine":
        return (szs.LevenshteinDistances(0, 3, 4, 1), O.Params("levenshtein", 0, 3, 4, 1),
                lambda lo, hi: random_bytes(rng, DNA, lo, hi))
    if kind == "levenshtein_utf8":
        return szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8"), lambda lo, hi: random_utf8(rng, lo, hi)
    if kind == "needleman_wunsch":
        return (szs.NeedlemanWunschScores(b2c, costs, -11, -1),
                O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs),
                lambda lo, hi: random_bytes(rng, PROTEIN, lo, hi))
    return (szs.SmithWatermanScores(dna_b2c, dna_costs, -2, -2),
            O.Params("smith_waterman", open=-2, extend=-2, byte_to_class=dna_b2c, class_costs=dna_costs),
            lambda lo, hi: random_bytes(rng, DNA, lo, hi))


def _u32(strings):
    """A host tape of the strings with 32-bit offsets."""
    wide = szs.Strs(strings)
    return szs.Strs.from_tape(wide.data, wide.offsets.astype(np.uint32))


class _Sequence:
    """An ``sz_sequence_t`` whose callbacks hand out the strings of a Python list (kept alive here)."""

    def __init__(self, strings):
        self.buffers = [ctypes.create_string_buffer(s, max(len(s), 1)) for s in strings]
        self.lengths = [len(s) for s in strings]
        self.start = _native.SequenceStart(lambda handle, i: ctypes.addressof(self.buffers[i]))
        self.length = _native.SequenceLength(lambda handle, i: self.lengths[i])
        self.struct = _native.Sequence(None, len(strings), self.start, self.length)


def _call_sequence(engine, queries, candidates, padding):
    q = _Sequence(queries)
    c = _Sequence(candidates) if candidates is not None else None
    columns = len(candidates) if candidates is not None else len(queries)
    out = np.full((len(queries), columns + padding), 777, dtype=engine._dtype)
    error = ctypes.c_char_p()
    status = getattr(_native.lib, engine._prefix)(
        engine.handle, szs._scope(None).handle, ctypes.byref(q.struct), ctypes.byref(c.struct) if c else None,
        ctypes.c_void_p(out.ctypes.data), ctypes.c_size_t(columns + padding), ctypes.byref(error))
    _native.check(status, error, engine._prefix)
    assert (out[:, columns:] == 777).all(), "cells beyond the matrix must stay untouched"
    return out[:, :columns]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("symmetric", [False, True])
def test_sequence_callback_forms(kind, symmetric):
    rng = np.random.default_rng(zlib.crc32(f"seq:{kind}:{symmetric}".encode()))
    engine, params, make = _setup(kind, rng)
    queries = [make(0, 400) for _ in range(11)]
    candidates = None if symmetric else [make(0, 1300) for _ in range(7)]
    expected = O.cross(params, O.tape(queries), O.tape(candidates) if candidates else None, threads=16)
    got = _call_sequence(engine, queries, candidates, padding=3)
    np.testing.assert_array_equal(got.astype(np.int64), expected)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("resident", ["host", "hbm"])
def test_u32_tape_forms(kind, resident):
    rng = np.random.default_rng(zlib.crc32(f"u32:{kind}:{resident}".encode()))
    engine, params, make = _setup(kind, rng)
    queries = [make(0, 700) for _ in range(13)]
    candidates = [make(0, 2500) for _ in range(9)]
    q, c = _u32(queries), _u32(candidates)
    if resident == "hbm":
        q, c = szs.to_device(q), szs.to_device(c)
    assert q.u32 and isinstance(q.tape(), _native.U32Tape)
    expected = O.cross(params, O.tape(queries), O.tape(candidates), threads=16)
    out = np.full((13, 12), 555, dtype=engine._dtype)
    got = engine(q, c, out=out)
    assert got is out
    np.testing.assert_array_equal(out[:, :9].astype(np.int64), expected)
    assert (out[:, 9:] == 555).all()
    expected_sym = O.cross(params, O.tape(queries), None, threads=16)
    np.testing.assert_array_equal(engine(q).astype(np.int64), expected_sym)
    # The pair-list extension takes the same strings (widened to u64 offsets).
    firsts, seconds = queries[:9], candidates
    np.testing.assert_array_equal(engine.pairs(_u32(firsts), _u32(seconds)).astype(np.int64),
                                  O.pairs(params, O.tape(firsts), O.tape(seconds), threads=16))


def test_arrow_string_input_scores_like_list_input():
    """test/similarities.py:701-722: pyarrow string and large_string arrays score as the same strings in a list."""
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(701)
    texts = [random_utf8(rng, 0, 300).decode() for _ in range(40)]
    others = [random_utf8(rng, 0, 900).decode() for _ in range(25)]
    for engine, params in ((szs.LevenshteinDistances(), O.Params("levenshtein")),
                           (szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8"))):
        baseline = engine(szs.Strs(texts), szs.Strs(others))
        expected = O.cross(params, O.tape([t.encode() for t in texts]), O.tape([t.encode() for t in others]),
                           threads=16)
        np.testing.assert_array_equal(baseline.astype(np.int64), expected)
        for kind in (pa.string(), pa.large_string()):
            got = engine(pa.array(texts, type=kind), pa.array(others, type=kind))
            np.testing.assert_array_equal(got, baseline, err_msg=str(kind))
        sliced = pa.array(["skip"] + texts, type=pa.string()).slice(1)
        np.testing.assert_array_equal(engine(sliced, pa.array(others)), baseline)


def test_workspace_reuse_across_growing_and_shrinking_calls():
    """One scope's buffers only grow: every call, larger or smaller than the one before, scores like the oracle."""
    rng = np.random.default_rng(31337)
    scope = szs.DeviceScope(gpu_device=0)
    plans = [("levenshtein", 5, 0, 40), ("needleman_wunsch", 40, 500, 3000), ("levenshtein_utf8", 30, 0, 1500),
             ("smith_waterman", 3, 0, 20), ("levenshtein", 60, 2000, 9000), ("levenshtein_affine", 8, 0, 100),
             ("needleman_wunsch", 2, 1, 30), ("levenshtein_utf8", 4, 0, 10)]
    for step, (kind, count, low, high) in enumerate(plans * 2):
        engine, params, make = _setup(kind, rng)
        a = [make(low, high) for _ in range(count)]
        b = [make(low, high) for _ in range(count)]
        if step % 3 == 2:  # a symmetric cross-product now and then
            got = engine(szs.Strs(a), device=scope)
            want = O.cross(params, O.tape(a), None, threads=16)
        else:
            got = engine.pairs(szs.Strs(a), szs.Strs(b), device=scope)
            want = O.pairs(params, O.tape(a), O.tape(b), threads=16)
        np.testing.assert_array_equal(got.astype(np.int64), want, err_msg=f"step {step}: {kind} x{count}")


def _bad_device_tape():
    """An HBM tape whose second offset runs past the data and whose third decreases."""
    data = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)
    host = szs.Strs.from_tape(data, np.array([0, 2, 4], dtype=np.uint64))
    tape = szs.to_device(host)
    tape._offsets[1] = 15  # past offsets[count] = 4
    return tape


@pytest.mark.parametrize("kind", KINDS)
def test_device_offsets_outside_the_tape_fail_the_call(kind):
    rng = np.random.default_rng(zlib.crc32(f"bad:{kind}".encode()))
    engine, params, make = _setup(kind, rng)
    bad = _bad_device_tape()
    good = szs.to_device(szs.Strs([make(1, 30), make(1, 30)]))
    for call in (lambda: engine.pairs(bad, good), lambda: engine.pairs(good, bad), lambda: engine(bad, good),
                 lambda: engine(bad), lambda: szs.Batch([(engine, good, bad)]).run()):
        with pytest.raises(ValueError, match="offsets"):
            call()
    # The scope still scores afterwards.
    a, b = [make(0, 50) for _ in range(5)], [make(0, 50) for _ in range(5)]
    np.testing.assert_array_equal(engine.pairs(szs.Strs(a), szs.Strs(b)).astype(np.int64),
                                  O.pairs(params, O.tape(a), O.tape(b)))


def test_server_fails_only_requests_with_bad_offsets():
    rng = np.random.default_rng(4242)
    engine = szs.LevenshteinDistances()
    a = [random_bytes(rng, DNA, 0, 200) for _ in range(20)]
    b = [random_bytes(rng, DNA, 0, 200) for _ in range(20)]
    host_bad = szs.Strs.from_tape(np.frombuffer(b"ACGTACGT", dtype=np.uint8), np.array([0, 6, 3], dtype=np.uint64))
    server = szs.Server(hbm_budget_bytes=256 << 20)
    try:
        with pytest.raises(ValueError, match="decrease"):  # host offsets are checked when submitted
            server.submit(engine, host_bad, szs.Strs(["A", "C"]))
        bad_ticket = server.submit(engine, _bad_device_tape(), szs.to_device(szs.Strs(["A", "C"])))
        good_ticket = server.submit(engine, szs.Strs(a), szs.Strs(b))
        with pytest.raises(ValueError, match="offsets"):
            server.wait(bad_ticket)
        np.testing.assert_array_equal(server.wait(good_ticket).astype(np.int64),
                                      O.pairs(O.Params("levenshtein"), O.tape(a), O.tape(b)))
    finally:
        server.close()
