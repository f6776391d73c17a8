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


class JobStruct(C.Structure):
    _fields_ = [("op", C.c_int32), ("profile", ProfileStruct), ("src", C.c_void_p), ("dst", C.c_void_p),
                ("indices", C.POINTER(C.c_uint16)), ("payloads", C.POINTER(C.c_void_p))]


class BatchConfigStruct(C.Structure):
    _fields_ = [("tile_bytes", C.c_uint32), ("cw_per_unit", C.c_uint32), ("use_graph", C.c_uint32),
                ("cache_bytes", C.c_uint64)]


class BatchStatsStruct(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("jobs", C.c_uint64), ("recipe_hits", C.c_uint64),
                ("recipe_misses", C.c_uint64), ("table_hits", C.c_uint64), ("table_misses", C.c_uint64),
                ("cache_bytes", C.c_uint64), ("tile_bytes", C.c_uint32), ("cw_per_unit", C.c_uint32),
                ("tuned_signatures", C.c_uint32)]


class ServerConfigStruct(C.Structure):
    _fields_ = [("batched", C.c_int32), ("max_batch_jobs", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("batch", BatchConfigStruct)]


class ServerStatsStruct(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("jobs", C.c_uint64), ("max_batch_jobs", C.c_uint64),
                ("host_ns", C.c_uint64), ("gpu_ns", C.c_uint64), ("max_gpu_ns", C.c_uint64)]


class CompletionStruct(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", Status), ("t_submit_ns", C.c_uint64), ("t_start_ns", C.c_uint64),
                ("t_done_ns", C.c_uint64)]


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
    "scrs_device_alloc_count": (C.c_uint64, []),
    "scrs_host_alloc_count": (C.c_uint64, []),
    "scrs_set_term_split": (C.c_int32, [C.c_int32]),
    "scrs_afft_set_targeted_max": (C.c_int32, [C.c_int32]),
    "scrs_engine_field": (C.c_int32, [C.c_int32]),
    "scrs_recommended_gf16_engine": (C.c_int32, [C.c_uint64, C.c_uint64]),
    "scrs_profile_resolve": (Status, [C.c_int32, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(ProfileStruct)]),
    "scrs_profile_recommended": (Status, [C.c_int32, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(ProfileStruct)]),
    # batch encoder
    "scrs_batch_encoder_create": (Status, [C.POINTER(ProfileStruct), HP]),
    "scrs_batch_encoder_destroy": (None, [H]),
    "scrs_batch_encoder_profile": (Status, [H, C.POINTER(ProfileStruct)]),
    "scrs_batch_encoder_set_kernel": (Status, [H, C.c_int32]),
    "scrs_batch_decoder_set_kernel": (Status, [H, C.c_int32]),
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
    "scrs_decode_batch_dev_noalloc": (Status, [H, u16p, vp, vp, vp, C.c_uint64, vp]),
    "scrs_noalloc_scratch_bytes": (C.c_uint64, []),
    "scrs_decode_batch_dev_noalloc_bytes": (C.c_uint64, [H, C.c_uint64]),
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
    # additive-FFT transform plan
    "scrs_transform_plan_create": (Status, [C.c_uint64, HP]),
    "scrs_transform_plan_destroy": (None, [H]),
    "scrs_transform_plan_size": (C.c_uint64, [H]),
    "scrs_transform_plan_log_size": (C.c_uint32, [H]),
    "scrs_transform_forward": (Status, [H, u16p, C.c_uint64]),
    "scrs_transform_inverse": (Status, [H, u16p, C.c_uint64]),
    "scrs_transform_forward_dev": (Status, [H, vp, C.c_uint64, C.c_uint64, vp]),
    "scrs_transform_inverse_dev": (Status, [H, vp, C.c_uint64, C.c_uint64, vp]),
    # fused uneven batches
    "scrs_batch_engine_create": (Status, [C.POINTER(BatchConfigStruct), HP]),
    "scrs_batch_engine_destroy": (None, [H]),
    "scrs_batch_engine_submit": (Status, [H, C.POINTER(JobStruct), C.c_uint64, vp, u64p]),
    "scrs_batch_engine_stats": (Status, [H, C.POINTER(BatchStatsStruct)]),
    # serving loop
    "scrs_server_create": (Status, [C.POINTER(ServerConfigStruct), HP]),
    "scrs_server_destroy": (None, [H]),
    "scrs_server_submit": (Status, [H, C.POINTER(JobStruct), C.c_uint64]),
    "scrs_server_poll": (Status, [H, C.POINTER(CompletionStruct), C.c_uint64, u64p, C.c_int32]),
    "scrs_server_replay": (Status, [H, C.POINTER(JobStruct), u64p, C.c_uint64, C.POINTER(CompletionStruct)]),
    "scrs_now_ns": (C.c_uint64, []),
    "scrs_server_get_stats": (Status, [H, C.POINTER(ServerStatsStruct)]),
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
