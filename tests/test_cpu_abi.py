"""CPU-only checks of the C ABI library and its Python mirror: the library loads, exports every symbol the
public header declares, validates arguments like the reference binding (python/stringzillas/similarities.c), and
fails loudly (never silently computes on the CPU) when no GPU is visible. No kernel runs here.
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import dna_tables

szs = pytest.importorskip("stringzillas_amd")
from stringzillas_amd import _native  # noqa: E402


def test_library_exports_every_declared_symbol():
    declared = _native.declared_symbols()
    assert len(declared) >= 50
    missing = [name for name in declared if not hasattr(_native.lib, name)]
    assert not missing, missing
    for name in ["szs_levenshtein_distances_u64tape", "szs_needleman_wunsch_scores_u32tape",
                 "szs_smith_waterman_scores", "szs_device_scope_init_gpu_device", "szs_batch_init", "szs_server_init"]:
        assert name in declared


def test_version_and_capabilities():
    assert szs.__version__ == "5.3.2"
    comptime = _native.lib.szs_capabilities_comptime()
    assert comptime & (1 << 25)  # sz_cap_rocm_k
    if not gpu_available():
        assert szs.__capabilities__ == ()


def test_engine_argument_validation_mirrors_the_reference():
    with pytest.raises(ValueError):
        szs.LevenshteinDistances(match=200)
    with pytest.raises(OverflowError):  # negative costs would wrap the reference's unsigned cells
        szs.LevenshteinDistances(mismatch=-1)
    b2c = np.zeros(256, dtype=np.uint8)
    costs = np.zeros((32, 32), dtype=np.int8)
    with pytest.raises(TypeError):
        szs.NeedlemanWunschScores(b2c.astype(np.int32), costs)
    with pytest.raises(ValueError):
        szs.NeedlemanWunschScores(b2c[:100], costs)
    with pytest.raises(TypeError):
        szs.NeedlemanWunschScores(b2c, costs.astype(np.int16))
    with pytest.raises(ValueError):
        szs.SmithWatermanScores(b2c, np.zeros((16, 16), dtype=np.int8))
    with pytest.raises(TypeError):
        szs.NeedlemanWunschScores()
    bad = b2c.copy()
    bad[7] = 40
    with pytest.raises(ValueError):
        szs.NeedlemanWunschScores(bad, costs)
    engine = szs.NeedlemanWunschScores(b2c, costs, open=300, extend=-1)  # silently truncated to int8
    assert "open=44" in repr(engine)


def test_inputs_must_be_tapes():
    engine = szs.LevenshteinDistances()
    with pytest.raises(TypeError):
        engine(["a", "b"], ["c"])


def test_out_argument_validation():
    engine = szs.LevenshteinDistances()
    q, c = szs.Strs(["ab", "c"]), szs.Strs(["a"])
    with pytest.raises(TypeError):
        engine(q, c, out=[[0], [0]])
    with pytest.raises(ValueError):
        engine(q, c, out=np.zeros(2, dtype=np.uint64))
    with pytest.raises(ValueError):
        engine(q, c, out=np.zeros((1, 1), dtype=np.uint64))
    with pytest.raises(TypeError):
        engine(q, c, out=np.zeros((2, 1), dtype=np.int64))
    with pytest.raises(ValueError):
        engine(q, c, out=np.zeros((2, 4), dtype=np.uint64)[:, ::2])


def test_pair_outputs_are_checked_before_the_library_writes():
    engine = szs.LevenshteinDistances()
    a, b = szs.Strs(["ab", "c", "d"]), szs.Strs(["a", "b", "e"])
    with pytest.raises(TypeError):
        engine.pairs(a, b, out=np.zeros(3, dtype=np.int32))  # 4-byte elements would overrun
    with pytest.raises(TypeError):
        engine.pairs(a, b, out=np.zeros(3, dtype=np.int64))  # Levenshtein results are uint64
    with pytest.raises(ValueError):
        engine.pairs(a, b, out=np.zeros(2, dtype=np.uint64))
    with pytest.raises(ValueError):
        engine.pairs(a, b, out=np.zeros((3, 1), dtype=np.uint64))
    with pytest.raises(ValueError):
        engine.pairs(a, b, out=np.zeros(6, dtype=np.uint64)[::2])
    frozen = np.zeros(3, dtype=np.uint64)
    frozen.flags.writeable = False
    with pytest.raises(ValueError):
        engine.pairs(a, b, out=frozen)
    with pytest.raises(TypeError):
        engine.pairs(a, b, out=[0, 0, 0])
    with pytest.raises(TypeError):
        szs.Batch([(engine, a, b)], outputs=[np.zeros(3, dtype=np.float64)])
    with pytest.raises(ValueError):
        szs.Batch([(engine, a, b)], outputs=[np.zeros(1, dtype=np.uint64)])
    with pytest.raises(ValueError):
        szs.Batch([(engine, a, szs.Strs(["x"]))])


def test_decreasing_offsets_are_rejected_before_any_device_work():
    # An intermediate offset past the last one would have the kernels read outside the tape.
    data = np.frombuffer(b"ACGTACGTACGTACGTACGT", dtype=np.uint8)
    bad = szs.Strs.from_tape(data, np.array([0, 15, 5], dtype=np.uint64))
    good = szs.Strs(["ACG", "TT"])
    nw = szs.NeedlemanWunschScores(*dna_tables(), -2, -1)
    for engine in (szs.LevenshteinDistances(), szs.LevenshteinDistancesUTF8(), nw):
        with pytest.raises(ValueError, match="decrease"):
            engine.pairs(bad, good)
        with pytest.raises(ValueError, match="decrease"):
            engine.pairs(good, bad)
        with pytest.raises(ValueError, match="decrease"):
            engine(bad, good)
        with pytest.raises(ValueError, match="decrease"):
            engine(bad)  # symmetric
        with pytest.raises(ValueError, match="decrease"):
            szs.Batch([(engine, good, bad)])


def test_device_scopes_without_scoring():
    cpu = szs.DeviceScope(cpu_cores=3)
    cores = ctypes.c_size_t()
    error = ctypes.c_char_p()
    assert _native.lib.szs_device_scope_get_cpu_cores(cpu.handle, ctypes.byref(cores), ctypes.byref(error)) == 0
    assert cores.value == 3
    gpu_index = ctypes.c_size_t()
    status = _native.lib.szs_device_scope_get_gpu_device(cpu.handle, ctypes.byref(gpu_index), ctypes.byref(error))
    assert status == -17  # sz_device_code_mismatch_k
    with pytest.raises(ValueError):
        szs.DeviceScope(cpu_cores=1, gpu_device=0)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_scoring_without_a_gpu_fails_loudly():
    engine = szs.LevenshteinDistances()
    with pytest.raises(RuntimeError, match="GPU"):
        engine(szs.Strs(["abc"]), szs.Strs(["abd"]))
    with pytest.raises(RuntimeError, match="GPU"):
        engine.pairs(szs.Strs(["abc"]), szs.Strs(["abd"]))
    with pytest.raises(RuntimeError):
        szs.DeviceScope(gpu_device=0)
    with pytest.raises(RuntimeError):
        szs.Batch([(engine, szs.Strs(["a"]), szs.Strs(["b"]))])
    with pytest.raises(RuntimeError):
        szs.Server(hbm_budget_bytes=1 << 20)


def test_cpu_scope_rejects_gpu_engines_before_touching_a_device():
    engine = szs.LevenshteinDistances()
    with pytest.raises(RuntimeError, match="CPU"):
        engine(szs.Strs(["a"]), szs.Strs(["b"]), device=szs.DeviceScope(cpu_cores=2))


def test_arrow_inputs_are_viewed_as_tapes():
    pa = pytest.importorskip("pyarrow")
    strings = ["alpha", "", "βeta", "gamma" * 10]
    for array in (pa.array(strings, type=pa.string()), pa.array(strings, type=pa.large_string()),
                  pa.array(strings, type=pa.string()).slice(1, 3)):
        tape = szs._as_tape(array, "queries")
        expected = strings[1:4] if len(array) == 3 else strings
        assert [tape[i].decode() for i in range(len(tape))] == expected
        # Arrow's 32-bit offsets are viewed, not widened: the engines then take their u32-tape entry points.
        assert tape.u32 == (array.type == pa.string())
        assert np.shares_memory(tape.offsets, np.frombuffer(array.buffers()[1], dtype=tape.offsets.dtype))
        assert isinstance(tape.tape(), _native.U32Tape if tape.u32 else _native.U64Tape)
        wide = tape.as_u64()
        assert wide.offsets.dtype == np.uint64 and [wide[i] for i in range(len(wide))] == [tape[i] for i in range(len(tape))]


def test_reference_import_name_is_this_package():
    import stringzillas

    assert stringzillas is szs
    assert stringzillas.LevenshteinDistances is szs.LevenshteinDistances


def test_u32_and_u64_tapes_do_not_mix_in_one_call():
    """python/stringzillas/similarities.c:283-303 rejects a cross-product of a 32- and a 64-bit offset tape."""
    narrow = szs.Strs.from_tape(np.frombuffer(b"ACGTT", dtype=np.uint8), np.array([0, 2, 5], dtype=np.uint32))
    wide = szs.Strs(["AC", "GTT"])
    assert narrow.u32 and not wide.u32
    with pytest.raises(TypeError):
        szs.LevenshteinDistances()(narrow, wide)


def test_tuning_ladders_validate_and_round_trip():
    from stringzillas_amd import autotune

    table = autotune.shapes()
    assert len(table) == _native.lib.szs_tuning_shapes()
    unit = autotune.slot_of("lev_linear", unit=True)
    nw = autotune.slot_of("nw_affine")
    plain = autotune.slot_of("nw_affine", runes=True)  # engines on the plain recurrences: no packed cells
    sw_linear, sw_affine = autotune.slot_of("sw_linear"), autotune.slot_of("sw_affine")
    lev_affine = autotune.slot_of("lev_affine")
    saved = {slot: autotune.get_ladder(slot) for slot in (unit, nw, plain, sw_linear, sw_affine, lev_affine)}
    try:
        ladder, beyond = saved[nw]
        assert ladder and all(b[1] > a[1] for a, b in zip(ladder, ladder[1:]))
        autotune.set_ladder(nw, ladder[:2], beyond)
        assert autotune.get_ladder(nw) == (ladder[:2], beyond)
        myers = next(i for i, s in enumerate(table) if s[3] and s[0] == 64 and not s[2])
        with pytest.raises(ValueError):
            autotune.set_ladder(nw, [(myers, 64)], beyond)  # bit-parallel shapes score unit costs only
        with pytest.raises(ValueError):
            autotune.set_ladder(nw, [(0, 10_000)], beyond)  # a sub-wave shape holds one stripe
        with pytest.raises(ValueError):
            autotune.set_ladder(nw, [(3, 100), (4, 50)], beyond)  # limits must grow
        packed = next(i for i, s in enumerate(table) if s[4])
        with pytest.raises(ValueError):
            autotune.set_ladder(plain, [(packed, 512)], beyond)
        autotune.set_ladder(plain, [], autotune.wide_shapes()[0])  # the int64 shape scores any slot
        assert autotune.get_ladder(nw) == (ladder[:2], beyond)  # rejected ladders change nothing
        # Local alignment takes packed shapes, never chained groups (its score needs every stripe's best);
        # uniform costs never take packed shapes.
        cooperative_packed = next(i for i, s in enumerate(table) if s[4] and s[2] and not s[5])
        chained_packed = next(i for i, s in enumerate(table) if s[4] and s[5])
        autotune.set_ladder(sw_linear, [(packed, 512)], cooperative_packed)
        assert autotune.get_ladder(sw_linear) == ([(packed, 512)], cooperative_packed)
        with pytest.raises(ValueError):
            autotune.set_ladder(sw_linear, [(packed, 512)], chained_packed)
        autotune.set_ladder(sw_affine, [(packed, 512)], cooperative_packed)
        with pytest.raises(ValueError):
            autotune.set_ladder(lev_affine, [(packed, 512)], saved[lev_affine][1])
    finally:
        for slot, (ladder, beyond) in saved.items():
            autotune.set_ladder(slot, ladder, beyond)


def test_built_in_class_cost_ladders_match_the_tuned_ones():
    """A C caller that never imports the package gets the library's built-in ladders: for the byte slots of affine
    NW and of SW (packed shapes where they measured faster) they equal the measured ladders of `tuned_shapes.json`.
    A fresh process loads the library alone, since importing the package installs the tuned file."""
    import json

    from stringzillas_amd import autotune

    script = (
        "import ctypes, json, sys\n"
        "lib = ctypes.CDLL(sys.argv[1])\n"
        "out = {}\n"
        "for slot in map(int, sys.argv[2:]):\n"
        "    shapes, rows = (ctypes.c_int * 64)(), (ctypes.c_int * 64)()\n"
        "    count, beyond = ctypes.c_int(), ctypes.c_int()\n"
        "    assert lib.szs_tuning_get_ladder(slot, shapes, rows, 64, ctypes.byref(count), ctypes.byref(beyond)) == 0\n"
        "    out[slot] = [[[shapes[i], rows[i]] for i in range(count.value)], beyond.value]\n"
        "print(json.dumps(out))\n"
    )
    slots = [autotune.slot_of(family) for family in ("nw_affine", "sw_linear", "sw_affine")]
    done = subprocess.run(["python3", "-c", script, _native.LIBRARY_PATH, *map(str, slots)], capture_output=True,
                          text=True, timeout=120)
    assert done.returncode == 0, done.stderr
    built_in = json.loads(done.stdout)
    with open(autotune.CACHE) as f:
        tuned = json.load(f)["slots"]
    for slot in slots:
        assert built_in[str(slot)] == [tuned[str(slot)]["ladder"], tuned[str(slot)]["beyond"]], slot


def test_refinement_moves_are_admissible_neighbours():
    """Every ladder the workload-level search would try is one the library accepts, and differs from the start."""
    from stringzillas_amd import autotune

    table = autotune.shapes()
    for slot in (autotune.UNIT_BYTES_SLOT, autotune.UNIT_RUNES_SLOT, autotune.slot_of("nw_affine")):
        saved = autotune.get_ladder(slot)
        ladder, beyond = [tuple(x) for x in saved[0]], saved[1]
        moves = autotune._moves(slot, ladder, beyond, table)
        assert moves, slot
        try:
            for label, candidate, candidate_beyond in moves:
                assert (candidate, candidate_beyond) != (ladder, beyond), label
                limits = [limit for _, limit in candidate]
                assert limits == sorted(set(limits)), label
                assert len({s for s, _ in candidate} | {candidate_beyond}) == len(candidate) + 1, label
                for shape, limit in candidate:
                    group, rows_per_lane = table[shape][0], table[shape][1]
                    assert group == 64 or limit <= group * rows_per_lane, label
                    assert table[shape][3] == (slot in autotune.UNIT_SLOTS) or limit <= 64, label
                autotune.set_ladder(slot, candidate, candidate_beyond)  # the library validates it
        finally:
            autotune.set_ladder(slot, *saved)


def test_tape_files_round_trip_and_reject_damage(tmp_path):
    pa = pytest.importorskip("pyarrow")
    strings = [b"", b"ACGT" * 40, "Ж中😀".encode(), b"x"]
    path = str(tmp_path / "strings.tape")
    szs.save_tape(szs.Strs(strings), path)
    back = szs.read_tape(path)
    assert [back[i] for i in range(len(back))] == strings
    # The sections are an Arrow large_binary array's buffers, and a slice saves as a standalone tape.
    arrow = pa.array(strings, type=pa.large_binary()).slice(1, 2)
    szs.save_tape(arrow, path)
    assert [szs.read_tape(path)[i] for i in range(2)] == strings[1:3]
    raw = bytearray(open(path, "rb").read())
    raw[0] ^= 1
    open(path, "wb").write(bytes(raw))
    with pytest.raises(ValueError):
        szs.read_tape(path)
    if not gpu_available():
        with pytest.raises(ValueError):  # the native loader validates before it touches a device... or fails on it
            try:
                szs.load_tape(path)
            except RuntimeError as error:  # no GPU in this container
                raise ValueError(str(error))


def test_fingerprints_argument_validation_mirrors_the_reference():
    """python/stringzillas/fingerprints.c:34-83: widths must be a contiguous 1-D uint64 array."""
    engine = szs.Fingerprints(ndim=1024)
    assert engine.ndim == 1024 and "ndim=1024" in repr(engine)
    assert _native.lib.szs_fingerprints_dimensions(engine.handle) == 1024
    szs.Fingerprints(ndim=7, window_widths=np.array([3, 5], dtype=np.uint64), alphabet_size=4, seed=9)
    with pytest.raises(TypeError):
        szs.Fingerprints(ndim=64, window_widths=[3, 5])
    with pytest.raises(TypeError):
        szs.Fingerprints(ndim=64, window_widths=np.array([3, 5], dtype=np.int64))
    with pytest.raises(ValueError):
        szs.Fingerprints(ndim=64, window_widths=np.array([[3, 5]], dtype=np.uint64))
    with pytest.raises(ValueError):
        szs.Fingerprints(ndim=64, window_widths=np.array([3, 0, 5, 0], dtype=np.uint64)[::2])
    with pytest.raises(ValueError):  # a zero-width window hashes nothing
        szs.Fingerprints(ndim=64, window_widths=np.array([3, 0], dtype=np.uint64))
    hashes, counts = engine(szs.Strs([]))  # empty input: no device work at all
    assert hashes.shape == (0, 1024) and counts.dtype == np.uint32
    if not gpu_available():
        with pytest.raises(RuntimeError, match="GPU"):
            engine(szs.Strs(["hello"]))
    with pytest.raises(RuntimeError, match="CPU"):
        engine(szs.Strs(["hello"]), device=szs.DeviceScope(cpu_cores=2))


def test_packed_rows_follow_the_range_bound():
    """Rows per lane of the packed 16-bit sweep an engine admits (no GPU needed: engines are host objects).

    Worked values for the seeded protein matrix (costs in [-4, 11]): adjacent cells of the anti-diagonal domain
    differ by at most delta = max(|o|, s_max - min(o, e) + max(0, o - e)) + |e|; R rows per lane need
    delta * (R + 70) plus a slack of 2 delta + |s_min| + |o| + |e| + |o - e| + 64 to stay within 15871 of the
    lane's base.
    """
    from stringzillas_amd import _native
    from tests.helpers import dna_tables, protein_tables

    b2c, costs = protein_tables(np.random.default_rng(3))
    rows = lambda engine: _native.lib.szs_tuning_packed_rows(engine.handle)  # noqa: E731
    assert rows(szs.NeedlemanWunschScores(b2c, costs, -11, -1)) == 8  # delta = 23: 23 * 78 = 1,794
    assert rows(szs.NeedlemanWunschScores(b2c, costs, -2, -2)) == 8  # linear, delta = 15
    assert rows(szs.NeedlemanWunschScores(b2c, costs, -40, -1)) == 8  # delta = 52
    big = costs.copy()
    np.fill_diagonal(big, 100)
    assert rows(szs.NeedlemanWunschScores(b2c, big, -11, -1)) == 8  # delta = 112: 112 * 78 + 314 = 9,050
    np.fill_diagonal(big, 120)
    assert rows(szs.NeedlemanWunschScores(b2c, big, -11, -1)) == 0  # 120 - (-12) = 132: not an unsigned byte
    # o = -128, e = -72 with every cost 85 lower (in [-89, -74], cost - (o + e) in [111, 126]): delta = 200, slack
    # 809; 4 rows: 200 * 74 + 809 = 15,609 fits, 6 rows: 200 * 76 + 809 = 16,009 does not.
    assert rows(szs.NeedlemanWunschScores(b2c, (costs.astype(np.int64) - 85).astype(np.int8), -128, -72)) == 4
    assert rows(szs.NeedlemanWunschScores(b2c, costs, 3, 1)) == 0  # odd signs: the bound does not apply
    # Local alignment: profile bytes `cost - min(o, s_min)`, the same delta as the global families.
    assert rows(szs.SmithWatermanScores(b2c, costs, -11, -1)) == 8  # bytes cost + 11 in [7, 22], delta = 23
    assert rows(szs.SmithWatermanScores(b2c, costs, -110, -80)) == 4  # delta = 201: 201 * 74 + 690 = 15,564
    assert rows(szs.SmithWatermanScores(b2c, costs, -4, -4)) == 8  # bytes cost + 4 in [0, 15], delta = 19
    assert rows(szs.SmithWatermanScores(b2c, costs, -2, -2)) == 8  # bias -4 below g: the correction add runs
    assert rows(szs.SmithWatermanScores(b2c, costs, -95, -95)) == 4  # delta = 201: 201 * 74 + 660 = 15,534
    assert rows(szs.SmithWatermanScores(b2c, costs, 1, 1)) == 0  # a positive gap: the bound does not apply
    wide = costs.copy()
    np.fill_diagonal(wide, 100)
    wide[1, 2] = wide[2, 1] = -40
    assert rows(szs.SmithWatermanScores(b2c, wide, -20, -20)) == 0  # 100 - (-40) = 140: not an unsigned byte
    assert rows(szs.LevenshteinDistances()) == 0
    dna_b2c, dna_costs = dna_tables()
    assert rows(szs.NeedlemanWunschScores(dna_b2c, dna_costs, -1, -1)) == 0  # -3 - (-2) < 0: no unsigned profile
    assert rows(szs.NeedlemanWunschScores(dna_b2c, dna_costs, -2, -2)) == 8  # -3 - (-4) = 1: fits
    assert rows(szs.SmithWatermanScores(dna_b2c, dna_costs, -2, -2)) == 8  # bias -3: cost + 3 in [0, 5]


REFERENCE_INCLUDE = "/root/reference/include"

CALLER = r"""
#include <stdio.h>
#include <stringzillas/stringzillas.h>
static sz_cptr_t start(void const *handle, sz_sorted_idx_t i) { return ((char const *const *)handle)[i]; }
static sz_size_t length(void const *handle, sz_sorted_idx_t i) {
    sz_size_t n = 0;
    while (((char const *const *)handle)[i][n]) ++n;
    return n;
}
int main(void) {
    char const *error = NULL;
    szs_device_scope_t scope = NULL;
    if (szs_device_scope_init_default(&scope, &error) != sz_success_k) return 2;
    szs_levenshtein_distances_t lev = NULL;
    if (szs_levenshtein_distances_init(0, 1, 1, 1, NULL, sz_cap_serial_k, &lev, &error) != sz_success_k) return 3;
    char const data[] = "kittensitting";
    sz_u64_t wide[3] = {0, 6, 13};
    sz_u32_t narrow[3] = {0, 6, 13};
    sz_sequence_u64tape_t tape64 = {data, wide, 2};
    sz_sequence_u32tape_t tape32 = {data, narrow, 2};
    char const *words[2] = {"kitten", "sitting"};
    sz_sequence_t sequence = {(void const *)words, 2, start, length};
    sz_size_t results[3][4] = {{0}};
    sz_status_t statuses[3];
    statuses[0] = szs_levenshtein_distances_u64tape(lev, scope, &tape64, NULL, results[0], 2, &error);
    statuses[1] = szs_levenshtein_distances_u32tape(lev, scope, &tape32, &tape32, results[1], 2, &error);
    statuses[2] = szs_levenshtein_distances(lev, scope, &sequence, NULL, results[2], 2, &error);
    sz_u8_t classes[256] = {0};
    sz_error_cost_t costs[32 * 32] = {0};
    szs_needleman_wunsch_scores_t nw = NULL;
    if (szs_needleman_wunsch_scores_init(classes, costs, -1, -1, NULL, sz_cap_serial_k, &nw, &error) != sz_success_k)
        return 4;
    sz_ssize_t scores[4] = {0};
    sz_status_t const nw_status = szs_needleman_wunsch_scores_u64tape(nw, scope, &tape64, NULL, scores, 2, &error);
    for (int k = 0; k < 3; ++k)
        printf("%d %zu %zu %zu %zu\n", (int)statuses[k], results[k][0], results[k][1], results[k][2], results[k][3]);
    printf("%d %td %td %td %td\n", (int)nw_status, scores[0], scores[1], scores[2], scores[3]);
    szs_needleman_wunsch_scores_free(nw);
    szs_levenshtein_distances_free(lev);
    szs_device_scope_free(scope);
    return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE_INCLUDE, "stringzillas")),
                    reason="needs the reference's own include/stringzillas/stringzillas.h")
def test_a_caller_built_against_the_reference_header_links_unchanged(tmp_path):
    """A C program written against include/stringzillas/stringzillas.h (with the dynamic-dispatch declarations its
    shared-library users compile with) links against this library unchanged and gets a score, or
    sz_missing_gpu_k where no GPU is visible."""
    source = tmp_path / "caller.c"
    source.write_text(CALLER)
    binary = tmp_path / "caller"
    library_dir = os.path.dirname(_native.LIBRARY_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-Wno-unknown-pragmas", "-DSZ_DYNAMIC_DISPATCH=1", f"-I{REFERENCE_INCLUDE}",
                    str(source), f"-L{library_dir}", "-lstringzillas_amd", f"-Wl,-rpath,{library_dir}", "-o",
                    str(binary)], check=True, capture_output=True, text=True)
    lines = subprocess.run([str(binary)], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [[int(x) for x in line.split()] for line in lines if line.strip()]
    if gpu_available():
        for row in rows[:3]:
            assert row == [0, 0, 3, 3, 0]  # Lev(kitten, sitting) = 3
        assert rows[3] == [0, 0, -3, -3, 0]  # NW with 0 / -1 costs is minus Levenshtein
    else:
        assert all(row[0] == -16 for row in rows)  # sz_missing_gpu_k
