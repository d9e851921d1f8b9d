"""MI355X-native drop-in for the ``stringzillas`` similarity and fingerprint classes.

Mirrors the reference's Python surface (python/stringzillas/similarities.c, fingerprints.c, device_scope.c,
stringzillas.c): ``LevenshteinDistances``, ``LevenshteinDistancesUTF8``, ``NeedlemanWunschScores``,
``SmithWatermanScores``, ``Fingerprints``, ``DeviceScope``, ``Strs``, ``__capabilities__``, ``__version__``,
``reset_capabilities`` and ``to_device``, with the same argument names, defaults, dtypes and exception types.
Every score and fingerprint is computed by the HIP kernels in ``libstringzillas_amd.so``; importing this package
fails loudly when that library is missing.

Extensions beyond the reference: ``engine.pairs(firsts, seconds)`` scores independent pairs, ``Batch`` scores a
mixed list of (engine, firsts, seconds) segments in one persistent launch, ``Server`` is the continuous-batching
request loop, and ``DeviceStrs`` keeps a tape resident in HBM.
"""

from __future__ import annotations

import ctypes
import os
from typing import Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _native
from ._native import SzsError, check as _check, lib as _lib, U32Tape as _U32Tape, U64Tape as _U64Tape

__version__ = "%d.%d.%d" % (_lib.szs_version_major(), _lib.szs_version_minor(), _lib.szs_version_patch())

_CAPABILITY_NAMES = {1: "serial", 1 << 2: "parallel", 1 << 21: "cuda", 1 << 25: "rocm"}


def _caps_to_tuple(mask: int) -> Tuple[str, ...]:
    return tuple(name for bit, name in sorted(_CAPABILITY_NAMES.items()) if mask & bit)


__capabilities__ = _caps_to_tuple(_lib.szs_capabilities())
_active_capabilities = _lib.szs_capabilities()


def reset_capabilities(capabilities: Sequence[str]) -> None:
    """Restricts the capabilities later engines are created with (python/stringzillas/stringzillas.c:111-147)."""
    global _active_capabilities, __capabilities__
    mask = 0
    names = {v: k for k, v in _CAPABILITY_NAMES.items()}
    for name in capabilities:
        if name not in names:
            raise ValueError(f"Unknown capability: {name}")
        mask |= names[name]
    _active_capabilities = mask & _lib.szs_capabilities()
    __capabilities__ = _caps_to_tuple(_active_capabilities)


#  Strings -----------------------------------------------------------------------------------------------------


def _offsets_array(offsets) -> np.ndarray:
    """Tape offsets as ``uint32`` (from 32-bit input) or ``uint64`` (anything wider), C-contiguous."""
    offsets = np.ascontiguousarray(offsets)
    if offsets.dtype in (np.int32, np.uint32):
        return offsets.view(np.uint32)
    if offsets.dtype in (np.int64, np.uint64):
        return offsets.view(np.uint64)
    return offsets.astype(np.uint64)


class Strs:
    """An Arrow-style tape of byte strings (``data`` + ``count + 1`` offsets), the input type the engines take.

    Plays the role of ``stringzilla.Strs`` (python/stringzilla/strs.c), which the reference's engines require.
    Offsets are ``uint64`` (built from a list, or an Arrow ``large_string``/``large_binary`` array) or ``uint32``
    (32-bit offsets given to ``from_tape``, or an Arrow ``string``/``binary`` array): the engines then call the
    ``*_u64tape`` or ``*_u32tape`` entry point, as python/stringzillas/similarities.c:243-270 does.
    """

    __slots__ = ("data", "offsets", "_keep")

    def __init__(self, strings: Union[Iterable[Union[str, bytes]], "Strs", None] = None):
        if isinstance(strings, Strs):
            self.data, self.offsets, self._keep = strings.data, strings.offsets, strings._keep
            return
        encoded = [s.encode("utf-8") if isinstance(s, str) else bytes(s) for s in (strings or [])]
        offsets = np.zeros(len(encoded) + 1, dtype=np.uint64)
        if encoded:
            offsets[1:] = np.cumsum([len(s) for s in encoded], dtype=np.uint64)
        self.data = np.frombuffer(b"".join(encoded) + b"\0", dtype=np.uint8)
        self.offsets = offsets
        self._keep = None

    @classmethod
    def from_tape(cls, data: np.ndarray, offsets: np.ndarray) -> "Strs":
        """A tape over existing buffers: 32-bit offsets stay 32-bit (a u32 tape), wider ones become u64."""
        strs = cls.__new__(cls)
        strs.data = np.ascontiguousarray(data, dtype=np.uint8)
        strs.offsets = _offsets_array(offsets)
        strs._keep = None
        return strs

    @classmethod
    def from_arrow(cls, array) -> "Strs":
        """Zero-copy view of a pyarrow ``string``/``binary`` (u32 tape) or ``large_string``/``large_binary``
        (u64 tape) array without nulls."""
        import pyarrow as pa

        if array.null_count:
            raise ValueError("Strs cannot hold null strings")
        if pa.types.is_string(array.type) or pa.types.is_binary(array.type):
            width = np.int32
        elif pa.types.is_large_string(array.type) or pa.types.is_large_binary(array.type):
            width = np.int64
        else:
            raise TypeError(f"Strs views Arrow string or binary arrays, not {array.type}")
        buffers = array.buffers()
        offsets = np.frombuffer(buffers[1], dtype=width)[array.offset: array.offset + len(array) + 1]
        data = np.frombuffer(buffers[2], dtype=np.uint8) if buffers[2] is not None else np.zeros(1, np.uint8)
        strs = cls.from_tape(data, offsets)
        strs._keep = array
        return strs

    @property
    def u32(self) -> bool:
        """Whether the offsets are 32-bit (the engines then call their ``*_u32tape`` entry points)."""
        return self.offsets.dtype == np.uint32

    def as_u64(self) -> "Strs":
        """The same strings with 64-bit offsets (the data is shared, not copied)."""
        if not self.u32:
            return self
        strs = Strs.from_tape(self.data, self.offsets.astype(np.uint64))
        strs._keep = self
        return strs

    def __len__(self) -> int:
        return len(self.offsets) - 1

    @property
    def nbytes(self) -> int:
        """Bytes of string data, ``offsets[count]``."""
        return int(self.offsets[-1]) if len(self.offsets) else 0

    def __getitem__(self, index: int) -> bytes:
        begin, end = int(self.offsets[index]), int(self.offsets[index + 1])
        return self.data[begin:end].tobytes()

    def tape(self):
        """``sz_sequence_u32tape_t`` or ``sz_sequence_u64tape_t`` over these buffers."""
        kind = _U32Tape if self.u32 else _U64Tape
        return kind(self.data.ctypes.data, self.offsets.ctypes.data, len(self))


class DeviceStrs:
    """A tape resident in HBM (torch device tensors), used in place by every call.

    ``to_device`` (python/stringzillas/stringzillas.c:148-162) returns one of these; 32-bit offsets stay 32-bit.
    """

    def __init__(self, strs: Strs, device: int = 0):
        import torch

        self.device = torch.device("cuda", device)
        self._data = torch.from_numpy(np.array(strs.data, dtype=np.uint8)).to(self.device)
        self._offsets = torch.from_numpy(strs.offsets.astype(np.int32 if strs.u32 else np.int64)).to(self.device)
        self.count = len(strs)
        self.host = strs

    def __len__(self) -> int:
        return self.count

    @property
    def u32(self) -> bool:
        return self.host.u32

    @property
    def nbytes(self) -> int:
        return self.host.nbytes

    def as_u64(self) -> "DeviceStrs":
        """The same strings with 64-bit offsets (widened in HBM; the data is shared)."""
        if not self.u32:
            return self
        wide = DeviceStrs.__new__(DeviceStrs)
        wide.device, wide._data, wide.count = self.device, self._data, self.count
        import torch

        wide._offsets = self._offsets.to(torch.int64)
        wide.host = self.host.as_u64()
        return wide

    def tape(self):
        kind = _U32Tape if self.u32 else _U64Tape
        return kind(self._data.data_ptr(), self._offsets.data_ptr(), self.count)


def to_device(strs: Strs, device: int = 0) -> DeviceStrs:
    return DeviceStrs(strs if isinstance(strs, Strs) else Strs(strs), device)


TAPE_MAGIC = b"SZSTAPE1"


def save_tape(strs, path: str) -> None:
    """Writes a tape file: 64-byte header, u64 offsets, u8 data (the layout of an Arrow ``large_binary`` array)."""
    tape = _as_tape(strs, "strs")
    if not isinstance(tape, Strs):
        raise TypeError("save_tape takes host strings (Strs, a pyarrow array or a list)")
    error = ctypes.c_char_p()
    tape = tape.as_u64()  # held until the call returns: the view points into it
    view = tape.tape()
    _check(_lib.szs_tape_save(path.encode(), ctypes.byref(view), ctypes.byref(error)), error, "Saving a tape")


def read_tape(path: str) -> Strs:
    """Maps a tape file into host memory (no copy): the same layout `save_tape` and `load_tape` use."""
    raw = np.memmap(path, dtype=np.uint8, mode="r")
    if raw.size < 72 or raw[:8].tobytes() != TAPE_MAGIC:
        raise ValueError(f"{path} is not a tape file")
    count, size = (int(x) for x in raw[8:24].view(np.uint64))
    data_start = (64 + (count + 1) * 8 + 63) // 64 * 64
    if data_start + size > raw.size:
        raise ValueError(f"{path} is shorter than its header says")
    offsets = raw[64:64 + (count + 1) * 8].view(np.uint64)
    if int(offsets[0]) != 0 or int(offsets[-1]) != size or (np.diff(offsets.astype(np.int64)) < 0).any():
        raise ValueError(f"{path} has offsets that do not span its data")
    strs = Strs.__new__(Strs)
    strs.data, strs.offsets, strs._keep = raw[data_start:data_start + size], offsets, raw
    return strs


class DeviceTape:
    """A tape file streamed into HBM by the library (``load_tape``); scored in place like ``DeviceStrs``."""

    def __init__(self, path: str, device: Optional[DeviceScope] = None):
        self.scope = _scope(device)
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        self._view = _U64Tape()
        _check(_lib.szs_tape_load(self.scope.handle, path.encode(), ctypes.byref(handle), ctypes.byref(self._view),
                                  ctypes.byref(error)), error, "Loading a tape")
        self.handle = handle

    def __len__(self) -> int:
        return int(self._view.count)

    @property
    def nbytes(self) -> int:
        return int(_lib.szs_tape_bytes(self.handle))

    def tape(self) -> _U64Tape:
        return self._view

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.szs_tape_free(self.handle)
            self.handle = None


def load_tape(path: str, device: Optional[DeviceScope] = None) -> DeviceTape:
    """Streams a tape file into HBM without reading it whole into host memory."""
    return DeviceTape(path, device)


def _is_u32(tape) -> bool:
    return bool(getattr(tape, "u32", False))


def _u64(tape):
    """The pair-list, batch and serving entry points take u64 tapes: 32-bit offsets are widened for them."""
    return tape.as_u64() if _is_u32(tape) else tape


def _as_tape(obj, name: str):
    if isinstance(obj, (Strs, DeviceStrs)) or (hasattr(obj, "tape") and hasattr(obj, "__len__")):
        return obj
    try:
        import pyarrow as pa

        if isinstance(obj, (pa.Array, pa.ChunkedArray)):
            if isinstance(obj, pa.ChunkedArray):
                obj = obj.combine_chunks()
            return Strs.from_arrow(obj)
    except ImportError:  # pragma: no cover - pyarrow is optional
        pass
    raise TypeError(
        f"Expected Strs objects, got {type(obj).__name__} for '{name}'. Convert using: Strs(your_string_list)"
    )


#  Device scopes -----------------------------------------------------------------------------------------------


class DeviceScope:
    """``DeviceScope(cpu_cores=None, gpu_device=None)`` (python/stringzillas/device_scope.c:17-61)."""

    def __init__(self, cpu_cores: Optional[int] = None, gpu_device: Optional[int] = None):
        if cpu_cores is not None and gpu_device is not None:
            raise ValueError("Specify either cpu_cores or gpu_device, not both")
        handle = ctypes.c_void_p()
        error = ctypes.c_char_p()
        if gpu_device is not None:
            status = _lib.szs_device_scope_init_gpu_device(ctypes.c_size_t(gpu_device), ctypes.byref(handle), ctypes.byref(error))
            self.description = f"gpu_device={gpu_device}"
        elif cpu_cores is not None:
            status = _lib.szs_device_scope_init_cpu_cores(ctypes.c_size_t(cpu_cores), ctypes.byref(handle), ctypes.byref(error))
            self.description = f"cpu_cores={cpu_cores}"
        else:
            status = _lib.szs_device_scope_init_default(ctypes.byref(handle), ctypes.byref(error))
            self.description = "default"
        _check(status, error, "Device scope initialization")
        self.handle = handle

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:  # module globals are gone at interpreter exit
            _lib.szs_device_scope_free(self.handle)
            self.handle = None

    def __repr__(self) -> str:
        return f"DeviceScope({self.description})"


_default_scope: Optional[DeviceScope] = None


def _scope(device: Optional[DeviceScope]) -> DeviceScope:
    global _default_scope
    if device is None:
        if _default_scope is None:
            _default_scope = DeviceScope()
        return _default_scope
    if not isinstance(device, DeviceScope):
        raise TypeError("device must be a DeviceScope instance")
    return device


#  Engines -----------------------------------------------------------------------------------------------------


def _int8(value, name: str) -> int:
    value = int(value)
    if value < -128 or value > 127:
        raise ValueError(f"{name} cost must fit in 8-bit signed integer")
    return value


def _pair_output(out, count: int, dtype) -> int:
    """Checks a caller's pair-list output and returns its address: the library writes ``count`` 8-byte results
    there, so it must be a 1-D, C-contiguous, writable array (or device tensor) of the engine's dtype holding at
    least ``count`` elements."""
    if hasattr(out, "data_ptr"):  # a torch tensor, typically in HBM
        if out.dim() != 1 or out.shape[0] < count:
            raise ValueError(f"out must be a 1-D tensor of at least {count} elements")
        if out.element_size() != 8 or out.is_floating_point() or out.is_complex():
            raise TypeError(f"out must hold 64-bit integers ({np.dtype(dtype).name})")
        if not out.is_contiguous():
            raise ValueError("out must be contiguous")
        return out.data_ptr()
    if not isinstance(out, np.ndarray):
        raise TypeError("out must be a NumPy array or a device tensor")
    if out.ndim != 1 or out.shape[0] < count:
        raise ValueError(f"out must be a 1-D array of at least {count} elements")
    if out.dtype != np.dtype(dtype):
        raise TypeError(f"out array must have {np.dtype(dtype).name} dtype")
    if not out.flags.c_contiguous:
        raise ValueError("out array must be C-contiguous")
    if not out.flags.writeable:
        raise ValueError("out array must be writable")
    return out.ctypes.data


class _Engine:
    _kind = 0
    _dtype = np.uint64
    _prefix = ""

    def _finish_init(self, status, handle, error, description):
        _check(status, error, f"{type(self).__name__} initialization")
        self.handle = handle
        self.description = description

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            getattr(_lib, f"{self._prefix}_free")(self.handle)
            self.handle = None

    def __repr__(self) -> str:
        return f"{type(self).__name__}({self.description})"

    def __call__(self, queries, candidates=None, device: Optional[DeviceScope] = None, out=None) -> np.ndarray:
        """Cross-product ``out[q, c] = score(queries[q], candidates[c])``; symmetric when ``candidates`` is None."""
        queries = _as_tape(queries, "queries")
        if candidates is not None:
            candidates = _as_tape(candidates, "candidates")
        scope = _scope(device)
        rows = len(queries)
        columns = len(candidates) if candidates is not None else rows
        if out is None:
            results = np.empty((rows, columns), dtype=self._dtype)
        else:
            if not isinstance(out, np.ndarray):
                raise TypeError("out argument must be a NumPy array")
            if out.ndim != 2:
                raise ValueError("out array must be 2-dimensional")
            if out.shape[0] < rows or out.shape[1] < columns:
                raise ValueError("out array is too small for results")
            if out.dtype != self._dtype:
                raise TypeError(f"out array must have {np.dtype(self._dtype).name} dtype")
            if out.strides[1] != 8:
                raise ValueError("out array rows must be contiguous (unit stride along columns)")
            results = out
        stride = results.strides[0] // 8 if rows else columns
        narrow = _is_u32(queries)
        if candidates is not None and _is_u32(candidates) != narrow:
            if len(queries) and len(candidates):  # python/stringzillas/similarities.c:283-303
                raise TypeError(f"Expected Strs objects of one offset width, got {type(queries).__name__} and "
                                f"{type(candidates).__name__} with 32- and 64-bit offsets")
            queries, candidates, narrow = _u64(queries), _u64(candidates), False
        q_tape = queries.tape()
        c_tape = candidates.tape() if candidates is not None else None
        error = ctypes.c_char_p()
        status = getattr(_lib, f"{self._prefix}_u32tape" if narrow else f"{self._prefix}_u64tape")(
            self.handle, scope.handle, ctypes.byref(q_tape), ctypes.byref(c_tape) if c_tape is not None else None,
            ctypes.c_void_p(results.ctypes.data), ctypes.c_size_t(stride), ctypes.byref(error))
        _check(status, error, type(self).__name__)
        return results

    def pairs(self, firsts, seconds, device: Optional[DeviceScope] = None, out=None):
        """Independent pairs: ``out[i] = score(firsts[i], seconds[i])`` (an extension of the reference API)."""
        firsts = _as_tape(firsts, "firsts")
        seconds = _as_tape(seconds, "seconds")
        if len(firsts) != len(seconds):
            raise ValueError("firsts and seconds must hold the same number of strings")
        scope = _scope(device)
        if out is None:
            out = np.empty(len(firsts), dtype=self._dtype)
        pointer = _pair_output(out, len(firsts), self._dtype)
        firsts, seconds = _u64(firsts), _u64(seconds)  # held until the call returns: the tapes point into them
        a, b = firsts.tape(), seconds.tape()
        error = ctypes.c_char_p()
        status = getattr(_lib, f"{self._prefix}_pairs_u64tape")(
            self.handle, scope.handle, ctypes.byref(a), ctypes.byref(b), ctypes.c_void_p(pointer), ctypes.byref(error))
        _check(status, error, type(self).__name__)
        return out


class LevenshteinDistances(_Engine):
    """Byte-level Levenshtein distances (python/stringzillas/similarities.c:36-140); ``uint64`` results."""

    _kind = 0
    _dtype = np.uint64
    _prefix = "szs_levenshtein_distances"

    def __init__(self, match: int = 0, mismatch: int = 1, open: int = 1, extend: int = 1, capabilities=None):
        match, mismatch = _int8(match, "match"), _int8(mismatch, "mismatch")
        open, extend = _int8(open, "open"), _int8(extend, "extend")
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        status = getattr(_lib, f"{self._prefix}_init")(
            ctypes.c_int8(match), ctypes.c_int8(mismatch), ctypes.c_int8(open), ctypes.c_int8(extend), None,
            ctypes.c_int(_active_capabilities), ctypes.byref(handle), ctypes.byref(error))
        self._finish_init(status, handle, error, f"{match},{mismatch},{open},{extend}")


class LevenshteinDistancesUTF8(LevenshteinDistances):
    """Codepoint-level Levenshtein distances over UTF-8 (python/stringzillas/similarities.c:379-560)."""

    _kind = 1
    _prefix = "szs_levenshtein_distances_utf8"


class NeedlemanWunschScores(_Engine):
    """Global alignment scores (python/stringzillas/similarities.c:793-926); ``int64`` results."""

    _kind = 2
    _dtype = np.int64
    _prefix = "szs_needleman_wunsch_scores"

    def __init__(self, byte_to_class=None, class_substitution_costs=None, open: int = -1, extend: int = -1,
                 capabilities=None):
        if byte_to_class is None or class_substitution_costs is None:
            raise TypeError(f"{type(self).__name__} requires 'byte_to_class' and 'class_substitution_costs'")
        if not isinstance(byte_to_class, np.ndarray):
            raise TypeError("byte_to_class must be a NumPy array")
        if byte_to_class.ndim != 1 or byte_to_class.shape[0] != 256:
            raise ValueError("byte_to_class must be a 256-element array")
        if byte_to_class.dtype != np.uint8:
            raise TypeError("byte_to_class must have uint8 dtype")
        if not byte_to_class.flags.c_contiguous:
            raise ValueError("byte_to_class must be a C-contiguous array. Use np.ascontiguousarray() to convert.")
        if not isinstance(class_substitution_costs, np.ndarray):
            raise TypeError("class_substitution_costs must be a NumPy array")
        if class_substitution_costs.shape != (32, 32):
            raise ValueError("class_substitution_costs must be a 32x32 array")
        if class_substitution_costs.dtype != np.int8:
            raise TypeError("class_substitution_costs must have int8 dtype")
        if not class_substitution_costs.flags.c_contiguous:
            raise ValueError("class_substitution_costs must be a C-contiguous array. Use np.ascontiguousarray() to convert.")
        # The reference truncates open/extend to int8 silently (python/stringzillas/similarities.c:856-865).
        open8 = int(np.int64(open).astype(np.int8))
        extend8 = int(np.int64(extend).astype(np.int8))
        self._tables = (byte_to_class, class_substitution_costs)
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        status = getattr(_lib, f"{self._prefix}_init")(
            ctypes.c_void_p(byte_to_class.ctypes.data), ctypes.c_void_p(class_substitution_costs.ctypes.data),
            ctypes.c_int8(open8), ctypes.c_int8(extend8), None, ctypes.c_int(_active_capabilities),
            ctypes.byref(handle), ctypes.byref(error))
        self._finish_init(status, handle, error, f"open={open8},extend={extend8}")


class SmithWatermanScores(NeedlemanWunschScores):
    """Local alignment scores (python/stringzillas/similarities.c:1090-1520); ``int64`` results."""

    _kind = 3
    _prefix = "szs_smith_waterman_scores"


#  Fingerprints -----------------------------------------------------------------------------------------------


class Fingerprints:
    """``Fingerprints(ndim, window_widths=None, alphabet_size=256, seed=0, capabilities=None)``
    (python/stringzillas/fingerprints.c:34-83, 212-245): rolling MinHash + Count-Min sketches.

    ``engine(texts, device=None, out=None)`` returns ``(hashes, counts)``, two ``uint32`` matrices of shape
    ``(len(texts), ndim)``: per dimension the low 32 bits of the smallest window hash (0xFFFFFFFF when the text is
    shorter than that dimension's window) and how many windows reach it. ``out`` (an extension: the reference
    accepts and ignores it) may be a pair of ``uint32`` arrays or device tensors of that shape to write into.
    """

    def __init__(self, ndim: int, window_widths=None, alphabet_size: int = 256, seed: int = 0, capabilities=None):
        ndim = int(ndim)
        if ndim < 0:
            raise ValueError("ndim must be non-negative")
        widths_pointer, widths_count = None, 0
        if window_widths is not None:
            if not isinstance(window_widths, np.ndarray):
                raise TypeError("window_widths must be a numpy array of uint64")
            if window_widths.dtype != np.uint64:
                raise TypeError("window_widths must have dtype uint64")
            if window_widths.ndim != 1:
                raise ValueError("window_widths must be a 1D array")
            if not window_widths.flags.c_contiguous:
                raise ValueError("window_widths must be a contiguous C-style array (no strides)")
            widths_pointer, widths_count = ctypes.c_void_p(window_widths.ctypes.data), window_widths.size
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        status = _lib.szs_fingerprints_init(
            ctypes.c_size_t(ndim), ctypes.c_size_t(int(alphabet_size)), widths_pointer, ctypes.c_size_t(widths_count),
            ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), None, ctypes.c_int(_active_capabilities),
            ctypes.byref(handle), ctypes.byref(error))
        _check(status, error, "Fingerprints initialization")
        self.handle = handle
        self.ndim = ndim
        self.description = f"ndim={ndim},window_widths={widths_count},alphabet_size={alphabet_size},seed={seed}"

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.szs_fingerprints_free(self.handle)
            self.handle = None

    def __repr__(self) -> str:
        return f"Fingerprints({self.description})"

    @property
    def capabilities(self) -> Tuple[str, ...]:
        return _caps_to_tuple(_active_capabilities)

    @property
    def last_kernel_ms(self) -> float:
        """Device time of the last call's hashing kernels (HIP events), in milliseconds."""
        return float(_lib.szs_fingerprints_last_kernel_ms(self.handle))

    KERNEL_AUTO, KERNEL_WAVE, KERNEL_LANE = 0, 1, 2

    def tune(self, kernel: int = 0, segment_bytes: int = 0) -> None:
        """Tuning (results never change): force the wave-per-text (1) or lane-per-text (2) kernel, or pick
        automatically (0); cap the window starts of one work unit (0: automatic)."""
        error = ctypes.c_char_p()
        _check(_lib.szs_fingerprints_tune(self.handle, ctypes.c_int(kernel), ctypes.c_size_t(segment_bytes),
                                          ctypes.byref(error)), error, "Fingerprints tuning")

    @staticmethod
    def _output(array, rows: int, columns: int, name: str):
        if hasattr(array, "data_ptr"):  # a torch tensor, typically in HBM
            if tuple(array.shape) != (rows, columns) or array.element_size() != 4 or array.stride(1) != 1:
                raise ValueError(f"{name} must be a ({rows}, {columns}) 32-bit tensor with contiguous rows")
            return array.data_ptr(), array.stride(0) * 4
        if not isinstance(array, np.ndarray) or array.dtype != np.uint32:
            raise TypeError(f"{name} must be a uint32 NumPy array or a device tensor")
        if array.shape != (rows, columns) or (columns and array.strides[1] != 4):
            raise ValueError(f"{name} must have shape ({rows}, {columns}) with contiguous rows")
        return array.ctypes.data, array.strides[0] if rows > 1 else columns * 4

    def __call__(self, texts, device: Optional[DeviceScope] = None, out=None):
        if isinstance(texts, (list, tuple)) and len(texts) == 0:
            texts = Strs([])
        texts = _as_tape(texts, "texts")
        scope = _scope(device)
        rows = len(texts)
        if out is None:
            hashes = np.empty((rows, self.ndim), dtype=np.uint32)
            counts = np.empty((rows, self.ndim), dtype=np.uint32)
        else:
            hashes, counts = out
        if rows == 0 or self.ndim == 0:
            return hashes, counts
        hashes_pointer, hashes_stride = self._output(hashes, rows, self.ndim, "hashes")
        counts_pointer, counts_stride = self._output(counts, rows, self.ndim, "counts")
        tape = texts.tape()
        error = ctypes.c_char_p()
        status = (_lib.szs_fingerprints_u32tape if _is_u32(texts) else _lib.szs_fingerprints_u64tape)(
            self.handle, scope.handle, ctypes.byref(tape), ctypes.c_void_p(hashes_pointer),
            ctypes.c_size_t(max(hashes_stride, self.ndim * 4)), ctypes.c_void_p(counts_pointer),
            ctypes.c_size_t(max(counts_stride, self.ndim * 4)), ctypes.byref(error))
        _check(status, error, "Fingerprints computation")
        return hashes, counts


#  Batches and serving (extensions) ----------------------------------------------------------------------------

BATCH_PERSISTENT, BATCH_GRAPH, BATCH_PER_PAIR = 0, 1, 2


class Batch:
    """Several (engine, firsts, seconds) pair lists scored together in one persistent, length-binned launch.

    ``mode``: ``BATCH_PERSISTENT`` (default), ``BATCH_GRAPH`` (captured once into a HIP graph, replayed by
    ``run``) or ``BATCH_PER_PAIR`` (one kernel launch per pair: the per-call baseline). ``outputs`` may be NumPy
    arrays or device tensors with 8-byte elements.
    """

    def __init__(self, segments, mode: int = BATCH_PERSISTENT, device: Optional[DeviceScope] = None, outputs=None):
        self.scope = _scope(device)
        self._keep = []
        array = (_native.Segment * max(len(segments), 1))()
        self.outputs = []
        for i, (engine, firsts, seconds) in enumerate(segments):
            firsts, seconds = _u64(_as_tape(firsts, "firsts")), _u64(_as_tape(seconds, "seconds"))
            if len(firsts) != len(seconds):
                raise ValueError("firsts and seconds must hold the same number of strings")
            if outputs is not None:
                out = outputs[i]
            else:
                out = np.zeros(len(firsts), dtype=engine._dtype)
            pointer = _pair_output(out, len(firsts), engine._dtype)
            a, b = firsts.tape(), seconds.tape()
            self._keep += [firsts, seconds, a, b, out]
            array[i].kind = engine._kind
            array[i].engine = engine.handle
            array[i].firsts = ctypes.pointer(a)
            array[i].seconds = ctypes.pointer(b)
            array[i].results = ctypes.c_void_p(pointer)
            array[i].firsts_bytes = int(getattr(firsts, "nbytes", 0) or 0)
            array[i].seconds_bytes = int(getattr(seconds, "nbytes", 0) or 0)
            self.outputs.append(out)
        self._engines = [s[0] for s in segments]
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        status = _lib.szs_batch_init(self.scope.handle, array, ctypes.c_size_t(len(segments)), ctypes.c_int(mode),
                                     ctypes.byref(handle), ctypes.byref(error))
        _check(status, error, "Batch preparation")
        self.handle = handle

    @property
    def cells(self) -> int:
        return int(_lib.szs_batch_cells(self.handle))

    @property
    def device_bytes(self) -> int:
        """Device memory the prepared batch holds."""
        return int(_lib.szs_batch_device_bytes(self.handle))

    @property
    def core_ops(self) -> int:
        """Core recurrence lane-ops per run (each cell weighted by its tier's cell-update op count)."""
        return int(_lib.szs_batch_core_ops(self.handle))

    @property
    def last_kernel_ms(self) -> float:
        return float(_lib.szs_batch_last_kernel_ms(self.handle))

    def launch(self) -> None:
        error = ctypes.c_char_p()
        _check(_lib.szs_batch_launch(self.handle, ctypes.byref(error)), error, "Batch launch")

    def wait(self) -> None:
        error = ctypes.c_char_p()
        _check(_lib.szs_batch_wait(self.handle, ctypes.byref(error)), error, "Batch")

    def run(self):
        self.launch()
        self.wait()
        return self.outputs

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.szs_batch_free(self.handle)
            self.handle = None


class Server:
    """Continuous-batching request loop (native, in the library): ``submit`` returns a ticket, ``wait`` joins it."""

    def __init__(self, hbm_budget_bytes: int = 64 << 30, max_inflight_rounds: int = 2,
                 device: Optional[DeviceScope] = None):
        self.scope = _scope(device)
        handle, error = ctypes.c_void_p(), ctypes.c_char_p()
        status = _lib.szs_server_init(self.scope.handle, ctypes.c_size_t(hbm_budget_bytes),
                                      ctypes.c_size_t(max_inflight_rounds), ctypes.byref(handle), ctypes.byref(error))
        _check(status, error, "Server initialization")
        self.handle = handle
        self._requests = {}

    def submit(self, engine: _Engine, firsts, seconds) -> int:
        firsts, seconds = _u64(_as_tape(firsts, "firsts")), _u64(_as_tape(seconds, "seconds"))
        out = np.zeros(len(firsts), dtype=engine._dtype)
        a, b = firsts.tape(), seconds.tape()
        segment = _native.Segment()
        segment.kind = engine._kind
        segment.engine = engine.handle
        segment.firsts = ctypes.pointer(a)
        segment.seconds = ctypes.pointer(b)
        segment.results = ctypes.c_void_p(out.ctypes.data)
        segment.firsts_bytes = int(getattr(firsts, "nbytes", 0) or 0)
        segment.seconds_bytes = int(getattr(seconds, "nbytes", 0) or 0)
        ticket = ctypes.c_uint64()
        error = ctypes.c_char_p()
        _check(_lib.szs_server_submit(self.handle, ctypes.byref(segment), ctypes.byref(ticket), ctypes.byref(error)),
               error, "Server submission")
        self._requests[ticket.value] = (firsts, seconds, a, b, out, engine)
        return ticket.value

    def wait(self, ticket: int) -> np.ndarray:
        return self.wait_timed(ticket)[0]

    def wait_timed(self, ticket: int) -> Tuple[np.ndarray, float]:
        """Results and the time the request's round finished, on the clock of ``time.perf_counter``."""
        error = ctypes.c_char_p()
        completed = ctypes.c_double()
        status = _lib.szs_server_wait_timed(self.handle, ctypes.c_uint64(ticket), ctypes.byref(completed),
                                            ctypes.byref(error))
        kept = self._requests.pop(ticket)
        _check(status, error, "Server request")
        return kept[4], completed.value

    def stats(self) -> dict:
        stats = _native.ServerStats()
        _lib.szs_server_get_stats(self.handle, ctypes.byref(stats))
        return {name: getattr(stats, name) for name, _ in stats._fields_}

    def poll(self, ticket: int) -> bool:
        return bool(_lib.szs_server_poll(self.handle, ctypes.c_uint64(ticket)))

    @property
    def bytes_in_use(self) -> int:
        return int(_lib.szs_server_bytes_in_use(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None) and _lib is not None:
            _lib.szs_server_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def _apply_cached_tuning() -> None:
    """Installs the autotuned shape ladders when they were measured on this GPU architecture."""
    if os.path.exists(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_shapes.json")):
        from . import autotune

        autotune.apply()


_apply_cached_tuning()

__all__ = [
    "Strs", "DeviceStrs", "to_device", "DeviceScope", "LevenshteinDistances", "LevenshteinDistancesUTF8",
    "NeedlemanWunschScores", "SmithWatermanScores", "Batch", "Server", "reset_capabilities", "SzsError",
    "save_tape", "read_tape", "load_tape", "DeviceTape",
    "BATCH_PERSISTENT", "BATCH_GRAPH", "BATCH_PER_PAIR", "__capabilities__", "__version__",
]
