"""GPU tests of on-disk tapes: files streamed into HBM score exactly like the host strings they were saved from,
in direct calls, batches and the server; damaged files are rejected before anything reaches the device."""

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, random_bytes, random_utf8

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")


def test_loaded_tapes_score_like_host_strings(tmp_path):
    rng = np.random.default_rng(21)
    firsts = szs.Strs([random_bytes(rng, DNA, 0, 900) for _ in range(300)])
    seconds = szs.Strs([random_bytes(rng, DNA, 0, 900) for _ in range(300)])
    szs.save_tape(firsts, str(tmp_path / "a.tape"))
    szs.save_tape(seconds, str(tmp_path / "b.tape"))
    a, b = szs.load_tape(str(tmp_path / "a.tape")), szs.load_tape(str(tmp_path / "b.tape"))
    assert len(a) == 300 and a.nbytes == firsts.nbytes
    engine = szs.LevenshteinDistances()
    expected = O.pairs(O.Params("levenshtein"), (firsts.data, firsts.offsets), (seconds.data, seconds.offsets))
    np.testing.assert_array_equal(engine.pairs(a, b), expected)
    batch = szs.Batch([(engine, a, b)], mode=szs.BATCH_GRAPH)
    batch.run()
    np.testing.assert_array_equal(batch.outputs[0], expected)
    server = szs.Server(hbm_budget_bytes=64 << 20)
    try:
        np.testing.assert_array_equal(server.wait(server.submit(engine, a, b)), expected)
    finally:
        server.close()


def test_large_utf8_tape_streams_through_staging(tmp_path):
    rng = np.random.default_rng(22)
    # Long enough (about 80 MB) to cycle the two 32 MB staging buffers several times.
    chunk = "".join(rng.choice(["a", "Ж", "中", "😀"], size=20_000)).encode()
    strings = [chunk * 40 for _ in range(25)] + [random_utf8(rng, 0, 300) for _ in range(50)]
    path = str(tmp_path / "big.tape")
    szs.save_tape(szs.Strs(strings), path)
    loaded = szs.load_tape(path)
    assert loaded.nbytes == sum(len(s) for s in strings)
    small_firsts = szs.Strs(strings[25:50])
    small_seconds = szs.Strs(strings[50:75])
    szs.save_tape(small_firsts, str(tmp_path / "f.tape"))
    szs.save_tape(small_seconds, str(tmp_path / "s.tape"))
    engine = szs.LevenshteinDistancesUTF8()
    got = engine.pairs(szs.load_tape(str(tmp_path / "f.tape")), szs.load_tape(str(tmp_path / "s.tape")))
    expected = O.pairs(O.Params("levenshtein_utf8"), (small_firsts.data, small_firsts.offsets),
                       (small_seconds.data, small_seconds.offsets))
    np.testing.assert_array_equal(got, expected)
    # The big tape's strings compared with themselves (distance 0) exercise the whole-wave transcoding path.
    np.testing.assert_array_equal(engine.pairs(loaded, loaded), np.zeros(len(strings), dtype=np.uint64))


def test_damaged_tape_files_are_rejected(tmp_path):
    path = str(tmp_path / "t.tape")
    szs.save_tape(szs.Strs([b"abc", b"de"]), path)
    raw = bytearray(open(path, "rb").read())
    for position, value in ((0, ord("X")), (8, 200), (64 + 8, 9)):  # magic, count, an offset past the data
        damaged = bytearray(raw)
        damaged[position] = value
        open(path, "wb").write(bytes(damaged))
        with pytest.raises(ValueError):
            szs.load_tape(path)
