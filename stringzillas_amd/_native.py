"""ctypes binding of ``libstringzillas_amd.so``: loading, struct layouts and status-to-exception mapping.

The library is built in-tree by ``stringzillas_amd/build.py``; importing without it raises ``ImportError``.
"""

from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(HERE, "libstringzillas_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "stringzillas_amd.h")

try:
    # PyTorch's ROCm wheels ship their own HIP runtime under the same soname as /opt/rocm's. Whichever loads first
    # serves the whole process, and torch cannot start on the other one; loading torch first keeps one runtime, so
    # device tensors and this library's device pointers are interchangeable.
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for host-only use
    pass

if not os.path.exists(LIBRARY_PATH):
    # A fresh tree: compile it in place (hipcc is part of the toolchain); anything else is a hard error.
    try:
        from . import build as _build

        _build.build(verbose=False)
    except Exception as error:  # pragma: no cover - exercised only on trees without a built library
        raise ImportError(f"{LIBRARY_PATH} is missing and building it failed: {error}") from error
lib = ctypes.CDLL(LIBRARY_PATH)


class U64Tape(ctypes.Structure):
    """``sz_sequence_u64tape_t`` (include/stringzillas/stringzillas.h:78-82)."""

    _fields_ = [("data", ctypes.c_void_p), ("offsets", ctypes.c_void_p), ("count", ctypes.c_size_t)]


class U32Tape(ctypes.Structure):
    """``sz_sequence_u32tape_t`` (include/stringzillas/stringzillas.h:68-72)."""

    _fields_ = [("data", ctypes.c_void_p), ("offsets", ctypes.c_void_p), ("count", ctypes.c_size_t)]


SequenceStart = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64)
SequenceLength = ctypes.CFUNCTYPE(ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint64)


class Sequence(ctypes.Structure):
    """``sz_sequence_t`` (include/stringzilla/types.h:1177-1182): strings behind two callbacks."""

    _fields_ = [("handle", ctypes.c_void_p), ("count", ctypes.c_size_t), ("get_start", SequenceStart),
                ("get_length", SequenceLength)]


class Segment(ctypes.Structure):
    """``szs_pairs_segment_t`` (include/stringzillas_amd.h)."""

    _fields_ = [
        ("kind", ctypes.c_int),
        ("engine", ctypes.c_void_p),
        ("firsts", ctypes.POINTER(U64Tape)),
        ("seconds", ctypes.POINTER(U64Tape)),
        ("results", ctypes.c_void_p),
        ("firsts_bytes", ctypes.c_size_t),
        ("seconds_bytes", ctypes.c_size_t),
    ]


class ServerStats(ctypes.Structure):
    """``szs_server_stats_t`` (include/stringzillas_amd.h)."""

    _fields_ = [
        ("rounds", ctypes.c_uint64), ("requests", ctypes.c_uint64), ("cells", ctypes.c_uint64),
        ("largest_round", ctypes.c_uint64), ("pending", ctypes.c_uint64), ("inflight", ctypes.c_uint64),
        ("prepare_seconds", ctypes.c_double), ("device_seconds", ctypes.c_double),
    ]


STATUS = {
    0: "success", -10: "bad_alloc", -12: "invalid_utf8", -13: "contains_duplicates", -14: "overflow_risk",
    -15: "unexpected_dimensions", -16: "missing_gpu", -17: "device_code_mismatch", -18: "device_memory_mismatch",
    -19: "authentication_failed", -1: "unknown",
}


class SzsError(RuntimeError):
    """Raised for statuses the reference maps to ``RuntimeError`` (GPU missing, device mismatches, unknown)."""

    def __init__(self, message: str, status: int):
        super().__init__(message)
        self.status = status


def check(status: int, error: ctypes.c_char_p, context: str) -> None:
    """Maps ``sz_status_t`` to exceptions like python/stringzillas/stringzillas.c:21-32."""
    if status == 0:
        return
    detail = error.value.decode() if error is not None and error.value else STATUS.get(status, str(status))
    message = f"{context}: {detail}"
    if status == -10:
        raise MemoryError(message)
    if status in (-12, -15):
        raise ValueError(message)
    if status == -14:
        raise OverflowError(message)
    raise SzsError(message, status)


def declared_symbols() -> list:
    """Every function the public header declares (used by tests to check the exports)."""
    with open(HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:sz_status_t|void|int|void \*|sz_capability_t|sz_size_t|double)\s*\*?\s*(\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


_c = ctypes
for name, restype in {
    "szs_version_major": _c.c_int, "szs_version_minor": _c.c_int, "szs_version_patch": _c.c_int,
    "szs_capabilities": _c.c_int, "szs_capabilities_comptime": _c.c_int, "szs_capabilities_runtime": _c.c_int,
    "szs_batch_cells": _c.c_size_t, "szs_batch_core_ops": _c.c_size_t, "szs_batch_device_bytes": _c.c_size_t, "szs_batch_last_kernel_ms": _c.c_double, "szs_server_poll": _c.c_int,
    "szs_server_bytes_in_use": _c.c_size_t, "szs_unified_alloc": _c.c_void_p, "szs_tape_bytes": _c.c_size_t,
    "szs_fingerprints_dimensions": _c.c_size_t, "szs_fingerprints_last_kernel_ms": _c.c_double,
}.items():
    getattr(lib, name).restype = restype
lib.szs_tuning_packed_rows.restype = _c.c_int
lib.szs_tuning_packed_rows.argtypes = [_c.c_void_p]
for name in ("szs_device_scope_free", "szs_levenshtein_distances_free", "szs_levenshtein_distances_utf8_free",
             "szs_needleman_wunsch_scores_free", "szs_smith_waterman_scores_free", "szs_batch_free",
             "szs_server_free", "szs_unified_free", "szs_tape_free", "szs_fingerprints_free",
             "szs_fingerprints_utf8_free"):
    getattr(lib, name).restype = None
    getattr(lib, name).argtypes = [_c.c_void_p] if name != "szs_unified_free" else [_c.c_void_p, _c.c_size_t]
