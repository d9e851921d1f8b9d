"""GPU parity of the fingerprint kernels through the C ABI: bit-exact against the committed reference vectors and
the oracle, on the reference's own test shapes (test/fingerprints.py), on random texts across every run shape
(1, 2 and 4 slots per lane, per-dimension and sliced layouts), on texts long enough to be split into segments, and
with device-resident inputs and outputs."""

import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, random_bytes
from tests.test_oracle_fingerprints import load_golden

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")


def _random_texts(rng, count, low, high, alphabet=None):
    if alphabet is None:
        return [bytes(rng.integers(0, 256, int(rng.integers(low, high + 1)), dtype=np.uint8)) for _ in range(count)]
    return [random_bytes(rng, alphabet, low, high) for _ in range(count)]


KERNELS = [(0, 0), (1, 0), (2, 0), (1, 100), (2, 64)]  # (kernel, segment bytes): auto, wave, lane, tiny segments


def _check(texts, ndim, widths=None, seed=0, kernels=KERNELS):
    """Every kernel choice and segment length must produce the oracle's fingerprints."""
    widths_array = None if widths is None else np.array(widths, dtype=np.uint64)
    engine = szs.Fingerprints(ndim=ndim, window_widths=widths_array, seed=seed)
    expected_hashes, expected_counts = O.fingerprints(O.tape(texts), ndim, widths, seed=seed, threads=8)
    for kernel, segment in kernels:
        engine.tune(kernel, segment)
        hashes, counts = engine(szs.Strs(texts))
        np.testing.assert_array_equal(hashes, expected_hashes, err_msg=f"kernel {kernel}, segment {segment}")
        np.testing.assert_array_equal(counts, expected_counts, err_msg=f"kernel {kernel}, segment {segment}")
    return hashes, counts


def test_committed_reference_vectors():
    texts, cases = load_golden()
    for ndim, widths, seed, hashes, counts in cases:
        print(f"golden case ndim={ndim} widths={widths} seed={seed}", flush=True)
        engine = szs.Fingerprints(ndim=ndim, window_widths=None if widths is None else np.array(widths, np.uint64),
                                  seed=seed)
        for kernel, segment in KERNELS:
            engine.tune(kernel, segment)
            got_hashes, got_counts = engine(szs.Strs(texts))
            np.testing.assert_array_equal(got_hashes, hashes, err_msg=f"kernel {kernel}, segment {segment}")
            np.testing.assert_array_equal(got_counts, counts, err_msg=f"kernel {kernel}, segment {segment}")


@pytest.mark.parametrize("ndim", [1, 7, 64, 1024])
def test_reference_test_shapes(ndim):
    """test/fingerprints.py:10-76: shapes, dtypes, identical and different strings, degenerate inputs."""
    engine = szs.Fingerprints(ndim=ndim)
    hashes, counts = engine(szs.Strs([]))
    assert hashes.shape == (0, ndim) and counts.shape == (0, ndim)
    assert hashes.dtype == np.uint32 and counts.dtype == np.uint32
    hashes, counts = _check([b"hello", b"world", b"hello"], ndim)
    assert np.array_equal(hashes[0], hashes[2]) and np.array_equal(counts[0], counts[2])
    assert not np.array_equal(hashes[0], hashes[1])
    _check([b"a"], ndim)
    _check([b"a" * 5] * 3, ndim)
    first = _check([b"a", b"a" * 40, b"b", b""], ndim)
    second = szs.Fingerprints(ndim=ndim)(szs.Strs([b"a", b"a" * 40, b"b", b""]))
    np.testing.assert_array_equal(first[0], second[0])
    np.testing.assert_array_equal(first[1], second[1])


@pytest.mark.parametrize("ndim,widths,seed", [
    (64, None, 1),        # per-dimension layout, 8 dimensions per width: partly filled runs
    (512, None, 2),       # sliced, 64 per width: one slot per lane
    (1024, None, 3),      # sliced, 128 per width: two slots
    (2048, None, 4),      # sliced, 256 per width: four slots
    (200, [2, 3, 5], 5),  # per-dimension, uneven widths
    (192, [1, 47, 300], 6),  # width 1, and widths longer than many texts
])
def test_random_texts(ndim, widths, seed):
    rng = np.random.default_rng(seed)
    texts = _random_texts(rng, 300, 0, 1500) + _random_texts(rng, 100, 0, 3000, DNA)
    texts += [b"", b"x", b"xy" * 100, bytes(range(256)) * 3]
    print(f"random texts ndim={ndim} widths={widths}", flush=True)
    _check(texts, ndim, widths, seed)


def test_long_texts_are_split_into_segments():
    rng = np.random.default_rng(11)
    # 16,384 bytes is the last whole-text length; longer texts take the segmented pass and the merge.
    texts = [random_bytes(rng, DNA, n, n) for n in (16383, 16384, 16385, 16384 + 30, 40000, 100000)]
    texts += [b"A" * 50000, b"", bytes(rng.integers(0, 256, 70000, dtype=np.uint8))]
    texts += _random_texts(rng, 20, 0, 500)
    print("long texts", flush=True)
    hashes, counts = _check(texts, 1024, kernels=[(0, 0), (2, 0), (2, 4096)])
    # A run of one byte: every window of a dimension hashes alike, so the count is the number of windows.
    widths = [O.fingerprint_hashers(1024)[0][d].window_width for d in range(1024)]
    np.testing.assert_array_equal(counts[6], 50000 - np.array(widths, dtype=np.int64) + 1)
    _check(texts[:8], 7, kernels=[(0, 0), (1, 0), (2, 1000)])


def test_device_inputs_and_strided_device_outputs():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(12)
    texts = _random_texts(rng, 500, 0, 800) + [bytes(rng.integers(0, 256, 20000, dtype=np.uint8))]
    strs = szs.Strs(texts)
    engine = szs.Fingerprints(ndim=1024, seed=9)
    expected = engine(strs)
    device_texts = szs.to_device(strs)
    wide_hashes = torch.zeros((len(texts), 1024 + 16), dtype=torch.int32, device="cuda")
    wide_counts = torch.zeros((len(texts), 1024 + 32), dtype=torch.int32, device="cuda")
    engine(device_texts, out=(wide_hashes[:, :1024], wide_counts[:, 1024 + 32 - 1024:]))
    torch.cuda.synchronize()
    got_hashes = wide_hashes[:, :1024].cpu().numpy().view(np.uint32)
    got_counts = wide_counts[:, 32:].cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got_hashes, expected[0])
    np.testing.assert_array_equal(got_counts, expected[1])
    assert (wide_hashes[:, 1024:] == 0).all() and (wide_counts[:, :32] == 0).all()  # nothing written outside rows


def test_u32_tapes_strided_host_outputs_and_utf8_entry_points():
    lib = szs._native.lib
    rng = np.random.default_rng(13)
    texts = _random_texts(rng, 100, 0, 600)
    data, offsets = O.tape(texts)
    offsets32 = offsets.astype(np.uint32)
    ndim = 256
    expected_hashes, expected_counts = O.fingerprints((data, offsets), ndim, seed=21)
    scope = szs.DeviceScope()
    for init, call, free in (("szs_fingerprints_init", "szs_fingerprints_u32tape", "szs_fingerprints_free"),
                             ("szs_fingerprints_utf8_init", "szs_fingerprints_utf8_u32tape",
                              "szs_fingerprints_utf8_free")):
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        status = getattr(lib, init)(ctypes.c_size_t(ndim), ctypes.c_size_t(256), None, ctypes.c_size_t(0),
                                    ctypes.c_uint64(21), None, ctypes.c_int(0), ctypes.byref(handle), ctypes.byref(error))
        assert status == 0
        tape = szs._native.U64Tape(data.ctypes.data, offsets32.ctypes.data, len(texts))  # same layout, u32 offsets
        hashes = np.full((len(texts), ndim + 5), 7, dtype=np.uint32)  # row stride of ndim + 5 elements
        counts = np.full((len(texts), ndim + 3), 7, dtype=np.uint32)
        status = getattr(lib, call)(handle, scope.handle, ctypes.byref(tape), ctypes.c_void_p(hashes.ctypes.data),
                                    ctypes.c_size_t(hashes.strides[0]), ctypes.c_void_p(counts.ctypes.data),
                                    ctypes.c_size_t(counts.strides[0]), ctypes.byref(error))
        assert status == 0, error.value
        np.testing.assert_array_equal(hashes[:, :ndim], expected_hashes)
        np.testing.assert_array_equal(counts[:, :ndim], expected_counts)
        assert (hashes[:, ndim:] == 7).all() and (counts[:, ndim:] == 7).all()
        getattr(lib, free)(handle)


def test_inputs_written_on_the_default_stream_are_complete_before_hashing():
    """Offsets and bytes produced by torch on its default stream, handed over with no synchronization: the
    library's own stream waits for that work (a read racing it once saw stale offsets)."""
    torch = pytest.importorskip("torch")
    letters = torch.tensor(list(b"ACGT"), dtype=torch.uint8, device="cuda")
    engine = szs.Fingerprints(ndim=64)
    for length, count in ((100000, 40), (700, 3000)):
        scratch = torch.full((count + 1,), 0x4147, dtype=torch.int64, device="cuda")  # stale-looking contents
        del scratch
        data = letters[torch.randint(0, 4, (count * length,), device="cuda")]
        offsets = torch.arange(count + 1, dtype=torch.int64, device="cuda") * length
        tape = szs._native.U64Tape(data.data_ptr(), offsets.data_ptr(), count)

        class Texts:
            def __len__(self):
                return count

            def tape(self):
                return tape

        hashes, counts = engine(Texts())
        host = data.view(count, length)[:3].cpu().numpy()
        expected_hashes, expected_counts = O.fingerprints(O.tape([row.tobytes() for row in host]), 64)
        np.testing.assert_array_equal(hashes[:3], expected_hashes)
        np.testing.assert_array_equal(counts[:3], expected_counts)


def test_decreasing_offsets_are_rejected():
    data = np.frombuffer(b"abcdefgh", dtype=np.uint8)
    for offsets in (np.array([0, 5, 3, 8], dtype=np.uint64), np.array([4, 0], dtype=np.uint64)):
        strs = szs.Strs.from_tape(data, offsets)
        with pytest.raises(ValueError, match="decrease"):
            szs.Fingerprints(ndim=64)(strs)
