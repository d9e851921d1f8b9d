"""GPU parity at the benchmark configurations' real sizes.

The Levenshtein (1,024 pairs), Needleman-Wunsch (16,384 protein pairs) and Smith-Waterman (100,000 DNA queries
embedded in longer targets) configurations are scored in full and compared pair by pair with the oracle. The
mixed batch (1,000,000 pairs plus 32 of 65,536 bytes) is too large for the oracle, so it is checked through
properties that hold at any size: every kind is symmetric in its two sides here (unit costs, a symmetric matrix),
so scoring the swapped batch must give the same scores; a seeded sample, including the longest pairs, is
compared with the oracle; and a reduced mixed batch is compared in full.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import workloads as W  # noqa: E402

THREADS = min(16, os.cpu_count() or 1)


def _engine(segment):
    if segment.kind == "levenshtein":
        return szs.LevenshteinDistances(segment.match, segment.mismatch, segment.open, segment.extend)
    if segment.kind == "levenshtein_utf8":
        return szs.LevenshteinDistancesUTF8(segment.match, segment.mismatch, segment.open, segment.extend)
    cls = szs.NeedlemanWunschScores if segment.kind == "needleman_wunsch" else szs.SmithWatermanScores
    return cls(segment.byte_to_class, segment.class_costs, segment.open, segment.extend)


def _params(segment):
    kwargs = dict(match=segment.match, mismatch=segment.mismatch, open=segment.open, extend=segment.extend)
    if segment.byte_to_class is not None:
        kwargs.update(byte_to_class=segment.byte_to_class, class_costs=segment.class_costs)
    return O.Params(segment.kind, **kwargs)


def _oracle(segment):
    return O.pairs(_params(segment), (segment.firsts_data, segment.firsts_offsets),
                   (segment.seconds_data, segment.seconds_offsets), threads=THREADS)


def _score(segment):
    engine = _engine(segment)
    firsts = szs.Strs.from_tape(segment.firsts_data, segment.firsts_offsets)
    seconds = szs.Strs.from_tape(segment.seconds_data, segment.seconds_offsets)
    return engine.pairs(firsts, seconds)


@pytest.mark.parametrize("make", [W.levenshtein_c0, W.needleman_wunsch_c1, W.smith_waterman_c2],
                         ids=["levenshtein_1024", "needleman_wunsch_16384", "smith_waterman_100000"])
def test_configuration_in_full(make):
    (segment,) = make().segments
    got = _score(segment)
    expected = _oracle(segment)
    mismatches = np.flatnonzero(got != expected)
    assert mismatches.size == 0, f"{mismatches.size} of {segment.count} pairs differ, first at {mismatches[:5]}"


def test_reduced_mixed_batch_in_full():
    workload = W.mixed_c3(2000, 0)
    batch_inputs, expected = [], []
    for segment in workload.segments:
        host = segment.host()
        batch_inputs.append((_engine(host), szs.Strs.from_tape(host.firsts_data, host.firsts_offsets),
                             szs.Strs.from_tape(host.seconds_data, host.seconds_offsets)))
        expected.append(_oracle(host))
    batch = szs.Batch(batch_inputs)
    batch.run()
    for out, want, segment in zip(batch.outputs, expected, workload.segments):
        np.testing.assert_array_equal(out, want, err_msg=segment.kind)


def test_full_mixed_batch_symmetry_and_sample():
    import torch

    workload = W.mixed_c3(1_000_000, 32, device="cuda")
    engines = [_engine(segment) for segment in workload.segments]

    class Tape:
        def __init__(self, data, offsets):
            self.data, self.offsets = data, offsets

        def __len__(self):
            return self.offsets.numel() - 1

        def tape(self):
            from stringzillas_amd import _native

            return _native.U64Tape(self.data.data_ptr(), self.offsets.data_ptr(), len(self))

    def run(swapped: bool):
        segments = []
        for engine, segment in zip(engines, workload.segments):
            a = Tape(segment.firsts_data, segment.firsts_offsets)
            b = Tape(segment.seconds_data, segment.seconds_offsets)
            segments.append((engine, b, a) if swapped else (engine, a, b))
        batch = szs.Batch(segments)
        batch.run()
        return [out.copy() for out in batch.outputs]

    straight, swapped = run(False), run(True)
    for a, b, segment in zip(straight, swapped, workload.segments):
        assert segment.kind in ("levenshtein", "levenshtein_utf8") or (
            segment.class_costs == segment.class_costs.T).all()
        np.testing.assert_array_equal(a, b, err_msg=f"{segment.kind}: scores change when the sides swap")
    rng = np.random.default_rng(11)
    for out, segment in zip(straight, workload.segments):
        host = segment.host()
        picks = np.sort(rng.choice(host.count, size=40, replace=False))
        longest = np.argsort(host.lengths("firsts") * host.lengths("seconds"))[-3:]
        picks = np.unique(np.concatenate([picks, longest]))
        sub = host.subset(picks)
        np.testing.assert_array_equal(out[picks], _oracle(sub), err_msg=segment.kind)
    del workload
    torch.cuda.empty_cache()
