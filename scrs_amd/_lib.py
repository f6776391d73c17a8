"""ctypes binding of the C ABI declared in ``include/scrs.h``.

The library is the in-tree ``libscrs_hip.so`` built by ``scrs_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libscrs_hip.so"


class Status(C.Structure):
    _fields_ = [("code", C.c_int32), ("a", C.c_uint64), ("b", C.c_uint64)]


class ProfileStruct(C.Structure):
    _fields_ = [("engine", C.c_int32), ("k", C.c_uint64), ("m", C.c_uint64), ("symbol_len", C.c_uint64)]


class PushOutcomeStruct(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rank", C.c_uint64), ("received", C.c_uint64)]


u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p
H = C.c_void_p  # opaque handle
HP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes).  Every symbol declared in include/scrs.h.
SIGNATURES: dict[str, tuple] = {
    "scrs_abi_version": (C.c_uint32, []),
    "scrs_status_name": (C.c_char_p, [C.c_int32]),
    "scrs_engine_field": (C.c_int32, [C.c_int32]),
    "scrs_recommended_gf16_engine": (C.c_int32, [C.c_uint64, C.c_uint64]),
    "scrs_profile_resolve": (Status, [C.c_int32, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(ProfileStruct)]),
    "scrs_profile_recommended": (Status, [C.c_int32, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(ProfileStruct)]),
    # batch encoder
    "scrs_batch_encoder_create": (Status, [C.POINTER(ProfileStruct), HP]),
    "scrs_batch_encoder_destroy": (None, [H]),
    "scrs_batch_encoder_profile": (Status, [H, C.POINTER(ProfileStruct)]),
    "scrs_encode_scratch_create": (Status, [H, HP]),
    "scrs_encode_scratch_destroy": (None, [H]),
    "scrs_encode_into": (Status, [H, vp, C.c_uint64, vp, C.c_uint64]),
    "scrs_encode_into_with": (Status, [H, vp, C.c_uint64, vp, C.c_uint64, H]),
    "scrs_encode_batch_dev": (Status, [H, vp, vp, C.c_uint64, vp]),
    # batch decoder
    "scrs_batch_decoder_create": (Status, [C.POINTER(ProfileStruct), HP]),
    "scrs_batch_decoder_destroy": (None, [H]),
    "scrs_batch_decoder_profile": (Status, [H, C.POINTER(ProfileStruct)]),
    "scrs_decode_scratch_create": (Status, [H, HP]),
    "scrs_decode_scratch_destroy": (None, [H]),
    "scrs_decode_into": (Status, [H, u64p, C.POINTER(vp), u64p, C.c_uint64, vp, C.c_uint64]),
    "scrs_decode_into_with": (Status, [H, u64p, C.POINTER(vp), u64p, C.c_uint64, vp, C.c_uint64, H]),
    "scrs_decode_batch_dev": (Status, [H, u16p, vp, vp, vp, C.c_uint64, vp]),
    "scrs_decode_batch_ptrs_dev": (Status, [H, u16p, vp, vp, C.c_uint64, vp]),
    # streaming decoder
    "scrs_decoder_create": (Status, [C.POINTER(ProfileStruct), HP]),
    "scrs_decoder_destroy": (None, [H]),
    "scrs_decoder_profile": (Status, [H, C.POINTER(ProfileStruct)]),
    "scrs_decoder_push": (Status, [H, C.c_uint64, vp, C.c_uint64, C.POINTER(PushOutcomeStruct)]),
    "scrs_decoder_push_dev": (Status, [H, C.c_uint64, vp, C.c_uint64, vp, C.POINTER(PushOutcomeStruct)]),
    "scrs_decoder_rank": (C.c_uint64, [H]),
    "scrs_decoder_received": (C.c_uint64, [H]),
    "scrs_decoder_is_complete": (C.c_int32, [H]),
    "scrs_decoder_reset": (Status, [H]),
    "scrs_decoder_scratch_create": (Status, [H, HP]),
    "scrs_decoder_scratch_destroy": (None, [H]),
    "scrs_decoder_scratch_stats": (Status, [H, u64p, u64p, u64p, u64p]),
    "scrs_decoder_finalize_into": (Status, [H, vp, C.c_uint64]),
    "scrs_decoder_finalize_into_with": (Status, [H, vp, C.c_uint64, H]),
    "scrs_decoder_finalize_into_dev": (Status, [H, vp, C.c_uint64, vp]),
    # incremental encoder
    "scrs_incr_encoder_create": (Status, [C.POINTER(ProfileStruct), HP]),
    "scrs_incr_encoder_destroy": (None, [H]),
    "scrs_incr_encoder_profile": (Status, [H, C.POINTER(ProfileStruct)]),
    "scrs_incr_feed": (Status, [H, C.c_uint64, vp, C.c_uint64]),
    "scrs_incr_feed_dev": (Status, [H, C.c_uint64, vp, C.c_uint64, vp]),
    "scrs_incr_repair": (Status, [H, C.c_uint64, C.POINTER(vp), u64p]),
    "scrs_incr_repairs_dev": (Status, [H, C.POINTER(vp), vp]),
    "scrs_incr_fed_count": (C.c_uint64, [H]),
    "scrs_incr_reset": (Status, [H]),
    "scrs_incr_encode_batch_dev": (Status, [H, vp, vp, C.c_uint64, vp]),
    # utilities
    "scrs_fill_random_dev": (Status, [vp, C.c_uint64, C.c_uint64, C.c_uint64, vp]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the HIP library and bind every C-ABI symbol.

    Raises ``OSError`` when the library has not been built: the codec has no
    other implementation to fall back to.
    """
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise OSError(f"{LIB_PATH} is missing; build it with `python -m scrs_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
