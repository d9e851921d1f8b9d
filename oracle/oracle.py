"""TEST INFRASTRUCTURE ONLY: ctypes bindings to the CPU oracle (``liboracle.so``) and to the reference shim
(``_ref/libszref_*.so``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this module, and only
as the checker. The product package never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KINDS = {"levenshtein": 0, "levenshtein_utf8": 1, "needleman_wunsch": 2, "smith_waterman": 3}


class _Params(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("match", ctypes.c_int8),
        ("mismatch", ctypes.c_int8),
        ("open", ctypes.c_int8),
        ("extend", ctypes.c_int8),
        ("byte_to_class", ctypes.c_uint8 * 256),
        ("class_costs", ctypes.c_int8 * 1024),
    ]


class _RefParams(ctypes.Structure):
    _fields_ = [
        ("match", ctypes.c_int8),
        ("mismatch", ctypes.c_int8),
        ("open", ctypes.c_int8),
        ("extend", ctypes.c_int8),
        ("byte_to_class", ctypes.c_uint8 * 256),
        ("class_costs", ctypes.c_int8 * 1024),
    ]


@dataclass
class Params:
    """Scoring parameters of one engine; mirrors the arguments of the four ``szs_*_init`` calls."""

    kind: str
    match: int = 0
    mismatch: int = 1
    open: int = 1
    extend: int = 1
    byte_to_class: np.ndarray = field(default_factory=lambda: np.zeros(256, dtype=np.uint8))
    class_costs: np.ndarray = field(default_factory=lambda: np.zeros((32, 32), dtype=np.int8))

    def _fill(self, p):
        p.match, p.mismatch, p.open, p.extend = self.match, self.mismatch, self.open, self.extend
        ctypes.memmove(p.byte_to_class, np.ascontiguousarray(self.byte_to_class, dtype=np.uint8).ctypes.data, 256)
        ctypes.memmove(p.class_costs, np.ascontiguousarray(self.class_costs, dtype=np.int8).ctypes.data, 1024)
        return p

    def c_struct(self) -> _Params:
        p = _Params()
        p.kind = KINDS[self.kind]
        return self._fill(p)

    def ref_struct(self) -> _RefParams:
        return self._fill(_RefParams())


def tape(strings: Sequence[bytes]):
    """Arrow-style tape: concatenated bytes plus ``count + 1`` u64 offsets."""
    strings = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
    offsets = np.zeros(len(strings) + 1, dtype=np.uint64)
    if strings:
        offsets[1:] = np.cumsum([len(s) for s in strings], dtype=np.uint64)
    data = np.frombuffer(b"".join(strings) + b"\0", dtype=np.uint8)
    return data, offsets


def _ptr(array: np.ndarray):
    return ctypes.c_void_p(array.ctypes.data)


_lib = None


def _oracle():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, os.path.join(HERE, "liboracle.so")], check=True)
        _lib = ctypes.CDLL(path)
        _lib.szo_score.restype = ctypes.c_int64
        _lib.szo_pairs.restype = ctypes.c_int
        _lib.szo_cross.restype = ctypes.c_int
        _lib.szo_utf8_decode.restype = ctypes.c_size_t
        _lib.szo_splitmix64.restype = ctypes.c_uint64
        _lib.szo_fingerprint_hashers.restype = ctypes.c_int
        _lib.szo_fingerprints.restype = ctypes.c_int
    return _lib


def score(params: Params, a: bytes, b: bytes) -> int:
    a = a.encode() if isinstance(a, str) else bytes(a)
    b = b.encode() if isinstance(b, str) else bytes(b)
    return int(_oracle().szo_score(ctypes.byref(params.c_struct()), a, ctypes.c_size_t(len(a)), b, ctypes.c_size_t(len(b))))


def utf8_decode(text: bytes) -> np.ndarray:
    runes = np.zeros(len(text) + 1, dtype=np.uint32)
    count = _oracle().szo_utf8_decode(bytes(text), ctypes.c_size_t(len(text)), _ptr(runes))
    return runes[:count]


def pairs(params: Params, a_tape, b_tape, threads: int = 1) -> np.ndarray:
    (ad, ao), (bd, bo) = a_tape, b_tape
    count = len(ao) - 1
    out = np.zeros(count, dtype=np.int64)
    _oracle().szo_pairs(ctypes.byref(params.c_struct()), _ptr(ad), _ptr(ao), _ptr(bd), _ptr(bo),
                        ctypes.c_uint64(count), _ptr(out), ctypes.c_int(threads))
    return out


def cross(params: Params, q_tape, c_tape=None, threads: int = 1) -> np.ndarray:
    qd, qo = q_tape
    qn = len(qo) - 1
    if c_tape is None:
        cd, co, cn = None, None, qn
    else:
        cd, co = c_tape
        cn = len(co) - 1
    out = np.zeros((qn, cn), dtype=np.int64)
    _oracle().szo_cross(ctypes.byref(params.c_struct()), _ptr(qd), _ptr(qo), ctypes.c_uint64(qn),
                        None if cd is None else _ptr(cd), None if co is None else _ptr(co), ctypes.c_uint64(cn),
                        _ptr(out), ctypes.c_uint64(cn), ctypes.c_int(threads))
    return out


# --- the reference's own scorers, compiled into oracle/_ref/ -------------------------------------------------

_refs = {}


def _cpu_has_icelake_features() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
    except OSError:
        return False
    needed = ["avx512f", "avx512bw", "avx512vl", "avx512vbmi", "avx512_vbmi2", "avx512_vnni", "avx512_bitalg",
              "avx512_vpopcntdq", "gfni", "vaes", "vpclmulqdq"]
    return all(f" {flag}" in flags for flag in needed)


def reference_available(variant: str = "serial") -> bool:
    if variant == "icelake" and not _cpu_has_icelake_features():
        return False
    return os.path.exists(os.path.join(HERE, "_ref", f"libszref_{variant}.so"))


def _reference(variant: str):
    if variant not in _refs:
        lib = ctypes.CDLL(os.path.join(HERE, "_ref", f"libszref_{variant}.so"))
        lib.szref_pairs.restype = ctypes.c_int
        _refs[variant] = lib
    return _refs[variant]


def reference_pairs(params: Params, a_tape, b_tape, threads: int = 1, variant: str = "serial") -> np.ndarray:
    """Scores pairs with the reference's own CPU scorer (serial or Ice Lake), driven by OpenMP threads."""
    (ad, ao), (bd, bo) = a_tape, b_tape
    count = len(ao) - 1
    out = np.zeros(count, dtype=np.int64)
    status = _reference(variant).szref_pairs(KINDS[params.kind], ctypes.byref(params.ref_struct()), _ptr(ad), _ptr(ao),
                                             _ptr(bd), _ptr(bo), ctypes.c_uint64(count), _ptr(out),
                                             ctypes.c_int(threads))
    if status != 0:
        raise RuntimeError(f"reference scorer failed with status {status}")
    return out


# --- rolling fingerprints -------------------------------------------------------------------------------------

DEFAULT_WINDOW_WIDTHS = (3, 4, 5, 7, 9, 11, 15, 31)


class Hasher(ctypes.Structure):
    """One dimension of the reference's f64 rolling hasher (fingerprints/serial.hpp:364-431)."""

    _fields_ = [
        ("multiplier", ctypes.c_double),
        ("modulo", ctypes.c_double),
        ("inverse_modulo", ctypes.c_double),
        ("negative_discarding_multiplier", ctypes.c_double),
        ("discarding_multiplier", ctypes.c_double),
        ("window_width", ctypes.c_uint64),
    ]


def splitmix64(state: int) -> int:
    return int(_oracle().szo_splitmix64(ctypes.c_uint64(state)))


def fingerprint_hashers(ndim: int, window_widths=None, seed: int = 0):
    """Per-dimension hashers in output order and whether the sliced layout applies."""
    hashers = (Hasher * max(ndim, 1))()
    widths = None if window_widths is None else np.ascontiguousarray(window_widths, dtype=np.uint64)
    sliced = _oracle().szo_fingerprint_hashers(ctypes.c_uint64(ndim), None if widths is None else _ptr(widths),
                                               ctypes.c_uint64(0 if widths is None else len(widths)),
                                               ctypes.c_uint64(seed), hashers)
    return hashers, bool(sliced)


def fingerprints(texts_tape, ndim: int, window_widths=None, seed: int = 0, threads: int = 1):
    """(hashes, counts), two uint32 matrices of shape (texts, ndim), from the oracle restatement."""
    data, offsets = texts_tape
    count = len(offsets) - 1
    hashers, _ = fingerprint_hashers(ndim, window_widths, seed)
    hashes = np.zeros((count, ndim), dtype=np.uint32)
    counts = np.zeros((count, ndim), dtype=np.uint32)
    if count and ndim:
        _oracle().szo_fingerprints(hashers, ctypes.c_uint64(ndim), _ptr(data), _ptr(offsets), ctypes.c_uint64(count),
                                   _ptr(hashes), ctypes.c_uint64(ndim * 4), _ptr(counts), ctypes.c_uint64(ndim * 4),
                                   ctypes.c_int(threads))
    return hashes, counts


def reference_fingerprints_available(variant: str = "serial") -> bool:
    if variant == "skylake" and not _cpu_has_icelake_features():
        return False
    return os.path.exists(os.path.join(HERE, "_ref", f"libszref_fp_{variant}.so"))


def reference_fingerprints(texts_tape, ndim: int, window_widths=None, seed: int = 0, alphabet_size: int = 256,
                           per_dimension: bool = False, threads: int = 1, variant: str = "serial"):
    """(hashes, counts) from the reference's own engines (sliced or per-dimension), compiled into oracle/_ref/."""
    key = f"fp_{variant}"
    if key not in _refs:
        lib = ctypes.CDLL(os.path.join(HERE, "_ref", f"libszref_fp_{variant}.so"))
        lib.szref_fingerprints.restype = ctypes.c_int
        _refs[key] = lib
    data, offsets = texts_tape
    count = len(offsets) - 1
    widths = None if window_widths is None else np.ascontiguousarray(window_widths, dtype=np.uint64)
    hashes = np.zeros((count, ndim), dtype=np.uint32)
    counts = np.zeros((count, ndim), dtype=np.uint32)
    if count and ndim:
        status = _refs[key].szref_fingerprints(
            ctypes.c_uint64(ndim), ctypes.c_uint64(alphabet_size), None if widths is None else _ptr(widths),
            ctypes.c_uint64(0 if widths is None else len(widths)), ctypes.c_uint64(seed), _ptr(data), _ptr(offsets),
            ctypes.c_uint64(count), _ptr(hashes), _ptr(counts), ctypes.c_uint64(ndim), ctypes.c_int(int(per_dimension)),
            ctypes.c_int(threads))
        if status != 0:
            raise RuntimeError(f"reference fingerprints failed with status {status}")
    return hashes, counts
