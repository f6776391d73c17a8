"""Python mirror of the reference's codec surface over the HIP C ABI.

Names, argument meanings and errors follow the reference crate:
``Profile::resolve`` / ``Profile::recommended`` (scrs/src/selector.rs:40,67),
the ``BatchEncoder`` / ``BatchDecoder`` / ``Decoder`` / ``IncrementalEncoder``
traits (scrs/src/codec.rs:122-201) and the error enums
(scrs/src/error.rs:12-126).  Every method calls into ``libscrs_hip.so``;
nothing here computes payload bytes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import IntEnum
from typing import Iterable, Sequence

import numpy as np

from . import _lib


class Field(IntEnum):
    """codec.rs:20-26."""

    GF256 = 0
    GF65536 = 1


class Engine(IntEnum):
    """codec.rs:30-40."""

    STANDARD_CAUCHY = 0
    GOOD_CAUCHY = 1
    TOWER = 2
    AFFT = 3

    def field(self) -> Field:
        return Field(_lib.load().scrs_engine_field(int(self)))


class Kernel(IntEnum):
    """Multiply-XOR kernel of the GF(256) batch codecs (no reference
    counterpart): split-nibble LDS tables, or the GF(2) bit-matrix product on
    int8 MFMA or on e2m1 (fp4) MFMA.  All give the same bytes."""

    TABLE = 0
    BITMATRIX = 1
    BITMATRIX_FP4 = 2


# ---------------------------------------------------------------------------
# Errors
# ---------------------------------------------------------------------------


class ScrsError(Exception):
    """Base of every error raised by the codec surface."""

    def __init__(self, variant: str, **fields):
        self.variant = variant
        self.fields = fields
        body = ", ".join(f"{k}: {v}" for k, v in fields.items())
        super().__init__(f"{type(self).__name__}::{variant}" + (f" {{ {body} }}" if body else ""))

    def __eq__(self, other):  # compare like the reference's PartialEq errors
        return type(self) is type(other) and self.variant == other.variant and self.fields == other.fields

    def __hash__(self):
        return hash((type(self).__name__, self.variant, tuple(sorted(self.fields.items()))))


class ConfigError(ScrsError):
    """error.rs:12."""


class EncodeError(ScrsError):
    """error.rs:35."""


class DecodeError(ScrsError):
    """error.rs:75."""


class TransformLengthError(ScrsError):
    """afft/transform.rs:16-22 (a struct there; variant name kept as its own)."""


class DeviceError(ScrsError):
    """A HIP runtime failure or C-ABI misuse (no reference counterpart)."""


# code -> (class, variant, field names for a and b)
_CODES = {
    10: (ConfigError, "ZeroDimension", ()),
    11: (ConfigError, "TooManySymbols", ("cap",)),
    12: (ConfigError, "ZeroSymbolLen", ()),
    13: (ConfigError, "OddSymbolLen", ()),
    14: (ConfigError, "UnsupportedMode", ("engine",)),
    20: (EncodeError, "WrongInputLen", ("expected", "got")),
    21: (EncodeError, "IndexOutOfRange", ("index", "n")),
    22: (EncodeError, "WrongPayloadLen", ("expected", "got")),
    23: (EncodeError, "DuplicateData", ("index",)),
    24: (EncodeError, "WrongOutputLen", ("expected", "got")),
    25: (EncodeError, "ScratchMismatch", ()),
    30: (DecodeError, "WrongCount", ("expected", "got")),
    31: (DecodeError, "IndexOutOfRange", ("index", "n")),
    32: (DecodeError, "DuplicateIndex", ("index",)),
    33: (DecodeError, "WrongPayloadLen", ("expected", "got")),
    34: (DecodeError, "TooManySymbols", ("cap", "received")),
    35: (DecodeError, "InsufficientRank", ("rank", "k")),
    36: (DecodeError, "WrongOutputLen", ("expected", "got")),
    37: (DecodeError, "ScratchMismatch", ()),
    40: (TransformLengthError, "TransformLengthError", ("expected", "got")),
    90: (DeviceError, "InvalidArgument", ("position",)),
    91: (DeviceError, "Device", ("hip_error",)),
    92: (DeviceError, "OutOfMemory", ("bytes",)),
    93: (DeviceError, "CapacityExceeded", ("limit",)),
}


def error_from_status(code: int, a: int, b: int) -> ScrsError:
    cls, variant, names = _CODES.get(code, (DeviceError, f"Unknown{code}", ()))
    vals = (a, b)
    fields = {name: vals[i] for i, name in enumerate(names)}
    if cls is ConfigError and variant == "UnsupportedMode":
        fields["engine"] = Engine(a)
    return cls(variant, **fields)


def check(status: _lib.Status) -> None:
    if status.code != 0:
        raise error_from_status(status.code, status.a, status.b)


# ---------------------------------------------------------------------------
# Buffers
# ---------------------------------------------------------------------------


def _in_buf(buf) -> tuple[int, int, object]:
    """(address, length, keepalive) of a readable host byte buffer."""
    if isinstance(buf, np.ndarray):
        arr = np.ascontiguousarray(buf).view(np.uint8).reshape(-1)
    else:
        arr = np.frombuffer(memoryview(buf).cast("B"), dtype=np.uint8) if len(memoryview(buf)) else np.zeros(0, np.uint8)
    if arr.size == 0:
        return 0, 0, arr
    return arr.ctypes.data, arr.size, arr


def _out_buf(buf) -> tuple[int, int, object]:
    """(address, length, keepalive) of a writable host byte buffer."""
    if isinstance(buf, np.ndarray):
        if not buf.flags.c_contiguous or not buf.flags.writeable:
            raise ValueError("output array must be C-contiguous and writable")
        arr = buf.view(np.uint8).reshape(-1)
    else:
        mv = memoryview(buf)
        if mv.readonly:
            raise ValueError("output buffer must be writable")
        arr = np.frombuffer(mv.cast("B"), dtype=np.uint8) if len(mv) else np.zeros(0, np.uint8)
    if arr.size == 0:
        return 0, 0, arr
    return arr.ctypes.data, arr.size, arr


def _dev_ptr(t) -> int:
    """Device address of a torch tensor (or a raw integer address)."""
    return t if isinstance(t, int) else int(t.data_ptr())


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


# ---------------------------------------------------------------------------
# Profile
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class Profile:
    """A validated geometry (codec.rs:54-103)."""

    engine: Engine
    k: int
    m: int
    symbol_len: int

    @property
    def n(self) -> int:
        return self.k + self.m

    @property
    def field(self) -> Field:
        return self.engine.field()

    def _struct(self) -> _lib.ProfileStruct:
        return _lib.ProfileStruct(int(self.engine), self.k, self.m, self.symbol_len)

    @staticmethod
    def resolve(engine: Engine, k: int, m: int, symbol_len: int) -> "Profile":
        """Profile::resolve (selector.rs:40-61)."""
        out = _lib.ProfileStruct()
        check(_lib.load().scrs_profile_resolve(int(engine), k, m, symbol_len, C.byref(out)))
        return Profile(Engine(out.engine), out.k, out.m, out.symbol_len)

    @staticmethod
    def recommended(field: Field, k: int, m: int, symbol_len: int) -> "Profile":
        """Profile::recommended (selector.rs:67-87)."""
        out = _lib.ProfileStruct()
        check(_lib.load().scrs_profile_recommended(int(field), k, m, symbol_len, C.byref(out)))
        return Profile(Engine(out.engine), out.k, out.m, out.symbol_len)


def recommended_gf16_engine(k: int, m: int) -> Engine:
    """lib.rs:85-93."""
    return Engine(_lib.load().scrs_recommended_gf16_engine(k, m))


class _Handle:
    _destroy_name = ""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def close(self) -> None:
        if self._h is not None and self._h.value:
            getattr(_lib.load(), self._destroy_name)(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class EncodeScratch(_Handle):
    _destroy_name = "scrs_encode_scratch_destroy"


class DecodeScratch(_Handle):
    _destroy_name = "scrs_decode_scratch_destroy"


class _Coded(_Handle):
    profile: Profile

    @property
    def k(self) -> int:
        return self.profile.k

    @property
    def m(self) -> int:
        return self.profile.m

    @property
    def n(self) -> int:
        return self.profile.n

    @property
    def symbol_len(self) -> int:
        return self.profile.symbol_len


# ---------------------------------------------------------------------------
# BatchEncoder
# ---------------------------------------------------------------------------


class BatchEncoder(_Coded):
    """BatchEncoder (codec.rs:136-151) on the device."""

    _destroy_name = "scrs_batch_encoder_destroy"

    def __init__(self, profile: Profile):
        h = C.c_void_p()
        check(_lib.load().scrs_batch_encoder_create(C.byref(profile._struct()), C.byref(h)))
        super().__init__(h.value)
        self.profile = profile

    def set_kernel(self, kernel: Kernel) -> None:
        """Select the multiply-XOR kernel (GF(256) engines only)."""
        check(_lib.load().scrs_batch_encoder_set_kernel(self._h, int(kernel)))

    def scratch(self) -> EncodeScratch:
        h = C.c_void_p()
        check(_lib.load().scrs_encode_scratch_create(self._h, C.byref(h)))
        return EncodeScratch(h.value)

    def encode_into(self, data, repairs) -> None:
        dp, dn, _k1 = _in_buf(data)
        rp, rn, _k2 = _out_buf(repairs)
        check(_lib.load().scrs_encode_into(self._h, dp, dn, rp, rn))

    def encode_into_with(self, data, repairs, scratch: EncodeScratch) -> None:
        dp, dn, _k1 = _in_buf(data)
        rp, rn, _k2 = _out_buf(repairs)
        check(_lib.load().scrs_encode_into_with(self._h, dp, dn, rp, rn, scratch._h))

    def encode(self, data) -> np.ndarray:
        """Convenience: return the m*symbol_len repair bytes."""
        out = np.empty(self.m * self.symbol_len, np.uint8)
        self.encode_into(data, out)
        return out

    def encode_batch_dev(self, data_dev, repairs_dev, count: int, stream=None) -> None:
        check(_lib.load().scrs_encode_batch_dev(self._h, _dev_ptr(data_dev), _dev_ptr(repairs_dev), count,
                                                _stream_ptr(stream)))


def batch_encoder(profile: Profile) -> BatchEncoder:
    """batch_encoder (selector.rs:457-467)."""
    return BatchEncoder(profile)


# ---------------------------------------------------------------------------
# BatchDecoder
# ---------------------------------------------------------------------------


def _symbols_arrays(symbols: Sequence[tuple[int, object]]):
    n = len(symbols)
    idx = (C.c_uint64 * max(n, 1))()
    ptrs = (C.c_void_p * max(n, 1))()
    lens = (C.c_uint64 * max(n, 1))()
    keep = []
    for i, (index, payload) in enumerate(symbols):
        p, ln, k = _in_buf(payload)
        keep.append(k)
        idx[i] = index
        # A zero-length payload still needs a non-null address.
        ptrs[i] = p if p else C.addressof(lens)
        lens[i] = ln
    return idx, ptrs, lens, keep


class BatchDecoder(_Coded):
    """BatchDecoder (codec.rs:157-175) on the device."""

    _destroy_name = "scrs_batch_decoder_destroy"

    def __init__(self, profile: Profile):
        h = C.c_void_p()
        check(_lib.load().scrs_batch_decoder_create(C.byref(profile._struct()), C.byref(h)))
        super().__init__(h.value)
        self.profile = profile

    def set_kernel(self, kernel: Kernel) -> None:
        """Select the multiply-XOR kernel (GF(256) engines only)."""
        check(_lib.load().scrs_batch_decoder_set_kernel(self._h, int(kernel)))

    def scratch(self) -> DecodeScratch:
        h = C.c_void_p()
        check(_lib.load().scrs_decode_scratch_create(self._h, C.byref(h)))
        return DecodeScratch(h.value)

    def decode_into(self, symbols: Sequence[tuple[int, object]], out) -> None:
        idx, ptrs, lens, _keep = _symbols_arrays(symbols)
        op, on, _k = _out_buf(out)
        check(_lib.load().scrs_decode_into(self._h, idx, ptrs, lens, len(symbols), op, on))

    def decode_into_with(self, symbols: Sequence[tuple[int, object]], out, scratch: DecodeScratch) -> None:
        idx, ptrs, lens, _keep = _symbols_arrays(symbols)
        op, on, _k = _out_buf(out)
        check(_lib.load().scrs_decode_into_with(self._h, idx, ptrs, lens, len(symbols), op, on, scratch._h))

    def decode(self, symbols: Sequence[tuple[int, object]]) -> np.ndarray:
        out = np.empty(self.k * self.symbol_len, np.uint8)
        self.decode_into(symbols, out)
        return out

    def decode_batch_dev(self, indices: np.ndarray, data_dev, repairs_dev, out_dev, count: int,
                         stream=None) -> None:
        idx = np.ascontiguousarray(indices, dtype=np.uint16)
        if idx.size != count * self.k:
            raise ValueError("indices must hold count * k entries")
        check(_lib.load().scrs_decode_batch_dev(self._h, idx.ctypes.data_as(_lib.u16p), _dev_ptr(data_dev),
                                                _dev_ptr(repairs_dev), _dev_ptr(out_dev), count,
                                                _stream_ptr(stream)))

    def decode_batch_dev_noalloc(self, indices: np.ndarray, data_dev, repairs_dev, out_dev, count: int,
                                 stream=None) -> None:
        """decode_batch_dev with its launch workspace in the library's fixed
        per-device array: the call allocates nothing (GF(256) engines)."""
        idx = np.ascontiguousarray(indices, dtype=np.uint16)
        if idx.size != count * self.k:
            raise ValueError("indices must hold count * k entries")
        check(_lib.load().scrs_decode_batch_dev_noalloc(self._h, idx.ctypes.data_as(_lib.u16p), _dev_ptr(data_dev),
                                                        _dev_ptr(repairs_dev), _dev_ptr(out_dev), count,
                                                        _stream_ptr(stream)))

    def noalloc_bytes(self, count: int) -> int:
        """Workspace bytes of a decode_batch_dev_noalloc call of `count` codewords."""
        return int(_lib.load().scrs_decode_batch_dev_noalloc_bytes(self._h, count))

    def decode_batch_ptrs_dev(self, indices: np.ndarray, payload_ptrs_dev, out_dev, count: int,
                              stream=None) -> None:
        idx = np.ascontiguousarray(indices, dtype=np.uint16)
        if idx.size != count * self.k:
            raise ValueError("indices must hold count * k entries")
        check(_lib.load().scrs_decode_batch_ptrs_dev(self._h, idx.ctypes.data_as(_lib.u16p),
                                                     _dev_ptr(payload_ptrs_dev), _dev_ptr(out_dev), count,
                                                     _stream_ptr(stream)))


def batch_decoder(profile: Profile) -> BatchDecoder:
    """batch_decoder (selector.rs:270-279)."""
    return BatchDecoder(profile)


# ---------------------------------------------------------------------------
# Decoder (streaming)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class PushOutcome:
    """PushOutcome (transport/symbol_sink.rs:8-28): ``Advanced`` carries the
    rank and received count, ``Complete`` and ``Dependent`` carry nothing."""

    kind: str
    rank: int | None = None
    received: int | None = None

    @staticmethod
    def advanced(rank: int, received: int) -> "PushOutcome":
        return PushOutcome("Advanced", rank, received)


PushOutcome.COMPLETE = PushOutcome("Complete")
PushOutcome.DEPENDENT = PushOutcome("Dependent")
_PUSH_KINDS = {0: "Advanced", 1: "Complete", 2: "Dependent"}


def _push_outcome(o: _lib.PushOutcomeStruct) -> PushOutcome:
    kind = _PUSH_KINDS[o.kind]
    if kind == "Advanced":
        return PushOutcome.advanced(int(o.rank), int(o.received))
    return PushOutcome(kind)


class DecoderScratch(_Handle):
    """Decoder::Scratch: a recipe cache (GF(2^8)) or a geometry-tagged
    workspace (Tower)."""

    _destroy_name = "scrs_decoder_scratch_destroy"

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        check(_lib.load().scrs_decoder_scratch_stats(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("hits", "misses", "len", "recipe_bytes"), (int(x.value) for x in v)))


class Decoder(_Coded):
    """Decoder + SymbolSink (codec.rs:180-201, transport/symbol_sink.rs:36-54)
    on the device: pushes land in device slots, finalize reconstructs there."""

    _destroy_name = "scrs_decoder_destroy"

    def __init__(self, profile: Profile):
        h = C.c_void_p()
        check(_lib.load().scrs_decoder_create(C.byref(profile._struct()), C.byref(h)))
        super().__init__(h.value)
        self.profile = profile

    def scratch(self) -> DecoderScratch:
        h = C.c_void_p()
        check(_lib.load().scrs_decoder_scratch_create(self._h, C.byref(h)))
        return DecoderScratch(h.value)

    def push(self, index: int, payload) -> PushOutcome:
        if index < 0:
            raise DecodeError("IndexOutOfRange", index=index, n=self.n)
        p, n, _k = _in_buf(payload)
        o = _lib.PushOutcomeStruct()
        check(_lib.load().scrs_decoder_push(self._h, index, p, n, C.byref(o)))
        return _push_outcome(o)

    def push_dev(self, index: int, payload_dev, length: int, stream=None) -> PushOutcome:
        o = _lib.PushOutcomeStruct()
        check(_lib.load().scrs_decoder_push_dev(self._h, index, _dev_ptr(payload_dev), length,
                                                _stream_ptr(stream), C.byref(o)))
        return _push_outcome(o)

    def is_complete(self) -> bool:
        return bool(_lib.load().scrs_decoder_is_complete(self._h))

    def rank(self) -> int:
        return int(_lib.load().scrs_decoder_rank(self._h))

    def received(self) -> int:
        return int(_lib.load().scrs_decoder_received(self._h))

    def reset(self) -> None:
        check(_lib.load().scrs_decoder_reset(self._h))

    def finalize_into(self, out) -> None:
        op, on, _k = _out_buf(out)
        check(_lib.load().scrs_decoder_finalize_into(self._h, op, on))

    def finalize_into_with(self, out, scratch: DecoderScratch) -> None:
        op, on, _k = _out_buf(out)
        check(_lib.load().scrs_decoder_finalize_into_with(self._h, op, on, scratch._h))

    def finalize_into_dev(self, out_dev, length: int, stream=None) -> None:
        check(_lib.load().scrs_decoder_finalize_into_dev(self._h, _dev_ptr(out_dev), length, _stream_ptr(stream)))

    def finalize(self) -> np.ndarray:
        """SymbolSink::finalize (returns the data; the decoder stays usable)."""
        out = np.empty(self.k * self.symbol_len, np.uint8)
        self.finalize_into(out)
        return out

    # BatchDecoder for every Decoder (codec.rs:203-240): WrongCount, then
    # reset, push each symbol (a Dependent push is DuplicateIndex), finalize.
    def decode_into_with(self, symbols: Sequence[tuple[int, object]], out, scratch: DecoderScratch) -> None:
        if len(symbols) != self.k:
            raise DecodeError("WrongCount", expected=self.k, got=len(symbols))
        self.reset()
        for index, payload in symbols:
            if self.push(index, payload) == PushOutcome.DEPENDENT:
                raise DecodeError("DuplicateIndex", index=index)
        self.finalize_into_with(out, scratch)

    def decode_into(self, symbols: Sequence[tuple[int, object]], out) -> None:
        with self.scratch() as scratch:
            self.decode_into_with(symbols, out, scratch)


def decoder(profile: Profile) -> Decoder:
    """decoder (selector.rs:117-125)."""
    return Decoder(profile)


# ---------------------------------------------------------------------------
# IncrementalEncoder
# ---------------------------------------------------------------------------


class IncrementalEncoder(_Coded):
    """IncrementalEncoder (codec.rs:122-131) on the device: fed symbols are
    folded into device-resident repair accumulators."""

    _destroy_name = "scrs_incr_encoder_destroy"

    def __init__(self, profile: Profile):
        h = C.c_void_p()
        check(_lib.load().scrs_incr_encoder_create(C.byref(profile._struct()), C.byref(h)))
        super().__init__(h.value)
        self.profile = profile

    def feed(self, index: int, data) -> None:
        if index < 0:
            raise EncodeError("IndexOutOfRange", index=index, n=self.n)
        p, n, _k = _in_buf(data)
        check(_lib.load().scrs_incr_feed(self._h, index, p, n))

    def feed_dev(self, index: int, data_dev, length: int, stream=None) -> None:
        check(_lib.load().scrs_incr_feed_dev(self._h, index, _dev_ptr(data_dev), length, _stream_ptr(stream)))

    def repair(self, index: int) -> bytes:
        """Repair ``index`` as the partial sum of the data fed so far (a copy
        of the library's view)."""
        if index < 0:
            raise EncodeError("IndexOutOfRange", index=index, n=self.n)
        ptr = C.c_void_p()
        ln = C.c_uint64()
        check(_lib.load().scrs_incr_repair(self._h, index, C.byref(ptr), C.byref(ln)))
        return C.string_at(ptr.value, ln.value)

    def repairs_dev(self, stream=None) -> int:
        """Device address of the m x symbol_len repair accumulators."""
        ptr = C.c_void_p()
        check(_lib.load().scrs_incr_repairs_dev(self._h, C.byref(ptr), _stream_ptr(stream)))
        return int(ptr.value)

    def fed_count(self) -> int:
        return int(_lib.load().scrs_incr_fed_count(self._h))

    def reset(self) -> None:
        check(_lib.load().scrs_incr_reset(self._h))

    def encode_batch_dev(self, data_dev, repairs_dev, count: int, stream=None) -> None:
        check(_lib.load().scrs_incr_encode_batch_dev(self._h, _dev_ptr(data_dev), _dev_ptr(repairs_dev), count,
                                                     _stream_ptr(stream)))


def incremental_encoder(profile: Profile) -> IncrementalEncoder:
    """incremental_encoder (selector.rs:370-381)."""
    return IncrementalEncoder(profile)


# ---------------------------------------------------------------------------
# Additive-FFT transform plan
# ---------------------------------------------------------------------------


class TransformPlan(_Handle):
    """TransformPlan (afft/transform.rs:24-95): the novel-basis additive FFT
    over the points 0..size-1 of GF(2^16), size a power of two up to 65536."""

    _destroy_name = "scrs_transform_plan_destroy"

    def __init__(self, size: int):
        h = C.c_void_p()
        check(_lib.load().scrs_transform_plan_create(size, C.byref(h)))
        super().__init__(h.value)

    @staticmethod
    def new(size: int) -> "TransformPlan | None":
        """TransformPlan::new: None for a size that is not a power of two in 1..=65536."""
        if size <= 0 or size > 65536 or size & (size - 1):
            return None
        return TransformPlan(size)

    @property
    def size(self) -> int:
        return int(_lib.load().scrs_transform_plan_size(self._h))

    @property
    def log_size(self) -> int:
        return int(_lib.load().scrs_transform_plan_log_size(self._h))

    def _host(self, fn, values: np.ndarray) -> None:
        if not (isinstance(values, np.ndarray) and values.dtype == np.uint16 and values.flags.c_contiguous):
            raise TypeError("values must be a contiguous uint16 array (modified in place)")
        check(fn(self._h, values.ctypes.data_as(_lib.u16p), values.size))

    def forward(self, values: np.ndarray) -> None:
        """Evaluate novel-basis coefficients at the points, in place."""
        self._host(_lib.load().scrs_transform_forward, values)

    def inverse(self, values: np.ndarray) -> None:
        """Evaluations back to novel-basis coefficients, in place."""
        self._host(_lib.load().scrs_transform_inverse, values)

    def forward_dev(self, rows_dev, symbol_len: int, count: int = 1, stream=None) -> None:
        check(_lib.load().scrs_transform_forward_dev(self._h, _dev_ptr(rows_dev), symbol_len, count,
                                                     _stream_ptr(stream)))

    def inverse_dev(self, rows_dev, symbol_len: int, count: int = 1, stream=None) -> None:
        check(_lib.load().scrs_transform_inverse_dev(self._h, _dev_ptr(rows_dev), symbol_len, count,
                                                     _stream_ptr(stream)))


# ---------------------------------------------------------------------------
# Fused uneven batches and the serving loop
# ---------------------------------------------------------------------------


@dataclass
class Job:
    """One codeword's work for a BatchEngine or Server (device payloads).

    Encode: ``src`` = k*symbol_len data bytes, ``dst`` = m*symbol_len repairs.
    Decode: ``indices`` = k received wire indices, ``payloads`` = their device
    addresses, ``dst`` = k*symbol_len data bytes.
    """

    op: str
    profile: Profile
    src: object = None
    dst: object = None
    indices: Sequence[int] | None = None
    payloads: Sequence[int] | None = None

    @staticmethod
    def encode(profile: Profile, src_dev, dst_dev) -> "Job":
        return Job("encode", profile, src_dev, dst_dev)

    @staticmethod
    def decode(profile: Profile, indices: Sequence[int], payloads: Sequence[int], dst_dev) -> "Job":
        return Job("decode", profile, None, dst_dev, list(indices), list(payloads))


def _job_array(jobs: Sequence[Job]):
    arr = (_lib.JobStruct * max(len(jobs), 1))()
    keep = []
    for i, j in enumerate(jobs):
        s = arr[i]
        s.op = 0 if j.op == "encode" else 1
        s.profile = j.profile._struct()
        s.src = _dev_ptr(j.src) if j.src is not None else None
        s.dst = _dev_ptr(j.dst) if j.dst is not None else None
        if j.op == "decode":
            idx = np.ascontiguousarray(j.indices, dtype=np.uint16)
            pay = (C.c_void_p * max(len(j.payloads), 1))(*[int(x) for x in j.payloads])
            keep += [idx, pay]
            s.indices = idx.ctypes.data_as(_lib.u16p)
            s.payloads = C.cast(pay, C.POINTER(C.c_void_p))
    return arr, keep


def _batch_config(tile_bytes: int, cw_per_unit: int, use_graph: bool, cache_bytes: int) -> _lib.BatchConfigStruct:
    return _lib.BatchConfigStruct(tile_bytes, cw_per_unit, 1 if use_graph else 0, cache_bytes)


class BatchEngine(_Handle):
    """Runs a batch of encode / decode jobs of uneven geometry as one fused,
    persistent launch chain, with resident tables and decode recipes."""

    _destroy_name = "scrs_batch_engine_destroy"

    def __init__(self, tile_bytes: int = 0, cw_per_unit: int = 0, use_graph: bool = True, cache_bytes: int = 0):
        h = C.c_void_p()
        cfg = _batch_config(tile_bytes, cw_per_unit, use_graph, cache_bytes)
        check(_lib.load().scrs_batch_engine_create(C.byref(cfg), C.byref(h)))
        super().__init__(h.value)

    def submit(self, jobs: Sequence[Job], stream=None) -> None:
        arr, _keep = _job_array(jobs)
        failed = C.c_uint64(0)
        st = _lib.load().scrs_batch_engine_submit(self._h, arr, len(jobs), _stream_ptr(stream), C.byref(failed))
        if st.code != 0:
            err = error_from_status(st.code, st.a, st.b)
            err.job = int(failed.value)
            raise err

    def stats(self) -> dict:
        s = _lib.BatchStatsStruct()
        check(_lib.load().scrs_batch_engine_stats(self._h, C.byref(s)))
        return {name: int(getattr(s, name)) for name, _ in s._fields_}


class Server(_Handle):
    """Serving loop: admission by payload footprint, continuous batching
    (``batched=True``) or one codec call per job."""

    _destroy_name = "scrs_server_destroy"

    def __init__(self, batched: bool = True, max_batch_jobs: int = 0, hbm_budget: int = 0, tile_bytes: int = 0,
                 cw_per_unit: int = 0, use_graph: bool = True, cache_bytes: int = 0):
        h = C.c_void_p()
        cfg = _lib.ServerConfigStruct(1 if batched else 0, max_batch_jobs, hbm_budget,
                                      _batch_config(tile_bytes, cw_per_unit, use_graph, cache_bytes))
        check(_lib.load().scrs_server_create(C.byref(cfg), C.byref(h)))
        super().__init__(h.value)

    def submit(self, job: Job, tag: int) -> None:
        arr, _keep = _job_array([job])
        check(_lib.load().scrs_server_submit(self._h, arr, tag))

    def poll(self, cap: int = 1024, wait: bool = False) -> list[dict]:
        buf = (_lib.CompletionStruct * cap)()
        n = C.c_uint64(0)
        check(_lib.load().scrs_server_poll(self._h, buf, cap, C.byref(n), 1 if wait else 0))
        return [{"tag": int(c.tag), "code": int(c.status.code), "t_submit_ns": int(c.t_submit_ns),
                 "t_start_ns": int(c.t_start_ns), "t_done_ns": int(c.t_done_ns)} for c in buf[:n.value]]

    def stats(self) -> dict:
        s = _lib.ServerStatsStruct()
        check(_lib.load().scrs_server_get_stats(self._h, C.byref(s)))
        return {name: int(getattr(s, name)) for name, _ in s._fields_}

    @staticmethod
    def prepare(jobs: Sequence[Job]):
        """Marshal a job list once for repeated replays."""
        return _job_array(jobs) + (len(jobs),)

    def replay(self, jobs, arrival_ns: Sequence[int]) -> np.ndarray:
        """Submit jobs[i] at arrival_ns[i] after the start and wait for all;
        returns per-job (code, t_submit_ns, t_start_ns, t_done_ns).  `jobs`
        is a job list or the result of prepare()."""
        arr, _keep, n = jobs if isinstance(jobs, tuple) else Server.prepare(jobs)
        jobs = range(n)
        arrivals = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        out = (_lib.CompletionStruct * max(len(jobs), 1))()
        check(_lib.load().scrs_server_replay(self._h, arr, arrivals.ctypes.data_as(_lib.u64p), len(jobs), out))
        res = np.zeros(len(jobs), dtype=[("code", "i4"), ("t_submit_ns", "u8"), ("t_start_ns", "u8"),
                                         ("t_done_ns", "u8")])
        for i in range(len(jobs)):
            c = out[i]
            res[i] = (c.status.code, c.t_submit_ns, c.t_start_ns, c.t_done_ns)
        return res


def noalloc_scratch_bytes() -> int:
    """Size of the per-device workspace array of the no-allocation decode."""
    return int(_lib.load().scrs_noalloc_scratch_bytes())


def device_alloc_count() -> int:
    """Device and pinned-host allocations the library has made so far."""
    return int(_lib.load().scrs_device_alloc_count())


def host_alloc_count() -> int:
    """Heap allocations (operator new) the library's own code has made so far."""
    return int(_lib.load().scrs_host_alloc_count())


def set_term_split(enable: bool) -> bool:
    """Let codec launches too small to fill the device split each codeword's
    terms across workgroups (process-wide, off by default).  Returns the
    previous setting."""
    return bool(_lib.load().scrs_set_term_split(1 if enable else 0))


def afft_set_targeted_max(max_missing: int) -> int:
    """Largest number of missing data symbols for which an additive-FFT
    decode takes the targeted path instead of the transforms (process-wide,
    default 5 as TARGETED_MAX_MISSING in scrs/src/afft/decoder.rs:578,
    clamped to [0, 64]).  Returns the previous setting."""
    return int(_lib.load().scrs_afft_set_targeted_max(int(max_missing)))


def now_ns() -> int:
    return int(_lib.load().scrs_now_ns())


def fill_random_dev(dst_dev, length: int, seed: int, byte_offset: int = 0, stream=None) -> None:
    """Seeded splitmix64 counter-stream bytes written on the device."""
    check(_lib.load().scrs_fill_random_dev(_dev_ptr(dst_dev), length, seed, byte_offset, _stream_ptr(stream)))
