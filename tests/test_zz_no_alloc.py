"""The allocation-free call ``szs_batch_run_noalloc`` against the call it varies (``szs_batch_init`` in persistent
mode + ``szs_batch_run``), through the C ABI: two calls in a row, the second smaller, each bit-exact against the
existing entry point and the oracle; host memory and inputs larger than the static scratch are refused.
"""

import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, dna_tables, protein_tables, random_bytes, random_utf8

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
torch = pytest.importorskip("torch")
from stringzillas_amd import _native  # noqa: E402


def _workload(rng, pairs, longest):
    """Four segments (byte and UTF-8 Levenshtein, affine NW, linear SW) of `pairs` pairs up to `longest` bytes."""
    b2c, costs = protein_tables(rng)
    dna_b2c, dna_costs = dna_tables()
    kinds = [
        (szs.LevenshteinDistances(), O.Params("levenshtein"), lambda: random_bytes(rng, DNA, 0, longest)),
        (szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8"), lambda: random_utf8(rng, 0, longest // 3)),
        (szs.NeedlemanWunschScores(b2c, costs, -11, -1),
         O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs),
         lambda: random_bytes(rng, PROTEIN, 1, longest)),
        (szs.SmithWatermanScores(dna_b2c, dna_costs, -2, -2),
         O.Params("smith_waterman", open=-2, extend=-2, byte_to_class=dna_b2c, class_costs=dna_costs),
         lambda: random_bytes(rng, DNA, 1, longest)),
    ]
    segments = []
    for engine, params, make in kinds:
        firsts = [make() for _ in range(pairs)]
        seconds = [make() for _ in range(pairs)]
        segments.append((engine, params, firsts, seconds))
    return segments


def _run_noalloc(scope, segments):
    """Scores device-resident copies of the segments with the allocation-free call; returns host results."""
    keep, outputs = [], []
    array = (_native.Segment * len(segments))()
    for i, (engine, _, firsts, seconds) in enumerate(segments):
        a, b = szs.to_device(szs.Strs(firsts)), szs.to_device(szs.Strs(seconds))
        out = torch.full((len(firsts),), -7, dtype=torch.int64, device="cuda")
        ta, tb = a.tape(), b.tape()
        keep += [a, b, ta, tb]
        array[i].kind = engine._kind
        array[i].engine = engine.handle
        array[i].firsts = ctypes.pointer(ta)
        array[i].seconds = ctypes.pointer(tb)
        array[i].results = ctypes.c_void_p(out.data_ptr())
        array[i].firsts_bytes = a.nbytes
        array[i].seconds_bytes = b.nbytes
        outputs.append(out)
    error = ctypes.c_char_p()
    status = _native.lib.szs_batch_run_noalloc(scope.handle, array, ctypes.c_size_t(len(segments)),
                                               ctypes.byref(error))
    _native.check(status, error, "szs_batch_run_noalloc")
    return [out.cpu().numpy() for out in outputs]


def test_noalloc_call_matches_the_batch_call_twice():
    rng = np.random.default_rng(20260)
    scope = szs.DeviceScope(gpu_device=0)
    # The first call has multi-stripe pairs (boundary-row scratch); the second is smaller and reuses the array.
    for pairs, longest in ((160, 2600), (24, 300)):
        segments = _workload(rng, pairs, longest)
        got = _run_noalloc(scope, segments)
        batch = szs.Batch([(engine, szs.Strs(f), szs.Strs(s)) for engine, _, f, s in segments], device=scope)
        want = batch.run()
        for (engine, params, firsts, seconds), g, w in zip(segments, got, want):
            name = type(engine).__name__
            np.testing.assert_array_equal(g, w.astype(np.int64), err_msg=f"{name}, {pairs} pairs: vs szs_batch_run")
            oracle = O.pairs(params, O.tape(firsts), O.tape(seconds), threads=16)
            np.testing.assert_array_equal(g, oracle, err_msg=f"{name}, {pairs} pairs: vs the oracle")


def test_noalloc_call_refuses_host_memory_and_oversized_inputs():
    scope = szs.DeviceScope(gpu_device=0)
    engine = szs.LevenshteinDistances()
    host_a, host_b = szs.Strs(["ACGT", "GG"]), szs.Strs(["AGT", "G"])
    out = np.zeros(2, dtype=np.uint64)
    ta, tb = host_a.tape(), host_b.tape()
    segment = _native.Segment(0, engine.handle, ctypes.pointer(ta), ctypes.pointer(tb),
                              ctypes.c_void_p(out.ctypes.data), 0, 0)
    error = ctypes.c_char_p()
    status = _native.lib.szs_batch_run_noalloc(scope.handle, ctypes.byref(segment), ctypes.c_size_t(1),
                                               ctypes.byref(error))
    assert status == -18  # sz_device_memory_mismatch_k
    # Five million short pairs need ~300 MB of task records alone: more than the static array holds.
    count = 5_000_000
    lengths = torch.randint(0, 4, (count,), device="cuda", dtype=torch.int64)
    offsets = torch.zeros(count + 1, dtype=torch.int64, device="cuda")
    offsets[1:] = torch.cumsum(lengths, 0)
    data = torch.randint(65, 69, (int(offsets[-1].item()) + 1,), device="cuda", dtype=torch.uint8)
    results = torch.zeros(count, dtype=torch.int64, device="cuda")
    tape = _native.U64Tape(data.data_ptr(), offsets.data_ptr(), count)
    segment = _native.Segment(0, engine.handle, ctypes.pointer(tape), ctypes.pointer(tape),
                              ctypes.c_void_p(results.data_ptr()), int(offsets[-1].item()), int(offsets[-1].item()))
    status = _native.lib.szs_batch_run_noalloc(scope.handle, ctypes.byref(segment), ctypes.c_size_t(1),
                                               ctypes.byref(error))
    assert status == -10 and b"SZS_NOALLOC_SCRATCH_BYTES" in error.value  # sz_bad_alloc_k, stated clearly
    # The array is intact for the next call.
    rng = np.random.default_rng(3)
    segments = _workload(rng, 8, 200)[:1]
    np.testing.assert_array_equal(_run_noalloc(scope, segments)[0],
                                  O.pairs(segments[0][1], O.tape(segments[0][2]), O.tape(segments[0][3])))
