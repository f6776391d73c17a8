"""CPU oracle: a C restatement of the reference's payload arithmetic.

TEST INFRASTRUCTURE ONLY.  Imported by the test-suite, the smoke check and
the CPU-baseline leg of ``bench.py`` -- never by the product package.  The C
sources cite the reference file:line each function restates; the header of
``scrs_oracle.h`` lists what pins each one.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_szp = C.POINTER(C.c_size_t)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.or_gf256_mul.restype = C.c_uint8
        L.or_gf256_mul.argtypes = [C.c_uint8, C.c_uint8]
        L.or_gf256_inv.restype = C.c_uint8
        L.or_gf256_inv.argtypes = [C.c_uint8]
        L.or_gf256_mul_gfni.restype = C.c_int
        L.or_gf256_mul_gfni.argtypes = [C.c_uint8, C.c_uint8]
        L.or_gf256_exp_log.argtypes = [_u8p, _u8p]
        for name in ("or_gf16_mul", "or_gf16_mul_schoolbook"):
            getattr(L, name).restype = C.c_uint16
            getattr(L, name).argtypes = [C.c_uint16, C.c_uint16]
        L.or_gf16_square.restype = C.c_uint16
        L.or_gf16_square.argtypes = [C.c_uint16]
        L.or_gf16_inv.restype = C.c_uint16
        L.or_gf16_inv.argtypes = [C.c_uint16]
        L.or_gf16_pow.restype = C.c_uint16
        L.or_gf16_pow.argtypes = [C.c_uint16, C.c_uint32]
        for name in ("or_std_cauchy_matrix", "or_good_cauchy_matrix", "or_good_cauchy_matrix_factored"):
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [C.c_size_t, C.c_size_t, _u8p]
        L.or_tower_cauchy_matrix.restype = C.c_size_t
        L.or_tower_cauchy_matrix.argtypes = [C.c_size_t, C.c_size_t, _u16p]
        L.or_gf256_encode.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p]
        L.or_gf256_decode.restype = C.c_int
        L.or_gf256_decode.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _szp, C.POINTER(C.c_void_p),
                                      _szp, C.c_size_t, _u8p, C.c_size_t, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.or_gf256_decode_rref.restype = C.c_int
        L.or_gf256_decode_rref.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _szp, C.POINTER(C.c_void_p),
                                           _u8p]
        L.or_gf256_rows_terms.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p, C.POINTER(C.c_void_p)]
        L.or_gf256_rows_terms_fast.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p,
                                               C.POINTER(C.c_void_p)]
        L.or_tower_encode.argtypes = [_u16p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p]
        L.or_tower_decode_rref.restype = C.c_int
        L.or_tower_decode_rref.argtypes = [_u16p, C.c_size_t, C.c_size_t, C.c_size_t, _szp, C.POINTER(C.c_void_p),
                                           _u8p]
        L.or_cpu_plan.restype = C.c_int
        L.or_gf256_encode_many.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p, C.c_size_t,
                                           C.c_int]
        L.or_gf256_decode_many.restype = C.c_int
        L.or_gf256_decode_many.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, _u16p, _u8p, _u8p, _u8p,
                                           C.c_size_t, C.c_int]
        L.or_fill_bytes.argtypes = [C.c_uint64, C.c_uint64, _u8p, C.c_size_t]
        L.or_rs16_interp.restype = C.c_int
        L.or_rs16_interp.argtypes = [C.c_size_t, C.c_size_t, _szp, C.POINTER(C.c_void_p), C.c_size_t, _szp, _u8p]
        L.or_afft_encode.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p]
        L.or_cpu16_plan.restype = C.c_int
        L.or_tower_encode_port.argtypes = [_u16p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p]
        L.or_tower_finalize_port.restype = C.c_int
        L.or_tower_finalize_port.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, _szp, C.POINTER(C.c_void_p), _u8p]
        L.or_tower_encode_port_many.argtypes = [_u16p, C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p,
                                                C.c_size_t, C.c_int]
        L.or_tower_finalize_port_many.restype = C.c_int
        L.or_tower_finalize_port_many.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, _u16p, _u8p, _u8p, _u8p,
                                                  C.c_size_t, C.c_int]
        L.or_afft_encode_port.restype = C.c_int
        L.or_afft_encode_port.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, _u8p, _u8p, C.c_int, C.c_size_t]
        L.or_afft_novel_eval.argtypes = [_u16p, C.c_size_t, _u16p, C.c_size_t, _u16p]
        L.or_afft_decode_port.restype = C.c_int
        L.or_afft_decode_port.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, _szp, C.POINTER(C.c_void_p), _u8p,
                                          C.c_int]
        _lib = L
    return _lib


def _p8(a: np.ndarray):
    return a.ctypes.data_as(_u8p)


def _p16(a: np.ndarray):
    return a.ctypes.data_as(_u16p)


# ---- field ----------------------------------------------------------------

def gf256_mul(a: int, b: int) -> int:
    return lib().or_gf256_mul(a, b)


def gf256_inv(a: int) -> int:
    return lib().or_gf256_inv(a)


def gf256_mul_gfni(a: int, b: int) -> int:
    return lib().or_gf256_mul_gfni(a, b)


def gf256_exp_log() -> tuple[np.ndarray, np.ndarray]:
    e = np.zeros(255, np.uint8)
    lg = np.zeros(256, np.uint8)
    lib().or_gf256_exp_log(_p8(e), _p8(lg))
    return e, lg


def gf256_mul_table() -> np.ndarray:
    t = np.zeros((256, 256), np.uint8)
    for a in range(256):
        for b in range(256):
            t[a, b] = lib().or_gf256_mul(a, b)
    return t


def gf16_mul(a: int, b: int) -> int:
    return lib().or_gf16_mul(a, b)


def gf16_mul_schoolbook(a: int, b: int) -> int:
    return lib().or_gf16_mul_schoolbook(a, b)


def gf16_square(a: int) -> int:
    return lib().or_gf16_square(a)


def gf16_inv(a: int) -> int:
    return lib().or_gf16_inv(a)


def gf16_pow(a: int, e: int) -> int:
    return lib().or_gf16_pow(a, e)


# ---- matrices -------------------------------------------------------------

def std_cauchy_matrix(k: int, m: int) -> np.ndarray:
    out = np.zeros(k * m, np.uint8)
    cap = lib().or_std_cauchy_matrix(k, m, _p8(out))
    if cap:
        raise ValueError(f"k+m exceeds capacity {cap}")
    return out.reshape(k, m)


def good_cauchy_matrix(k: int, m: int, factored: bool = False) -> np.ndarray:
    out = np.zeros(k * m, np.uint8)
    fn = lib().or_good_cauchy_matrix_factored if factored else lib().or_good_cauchy_matrix
    cap = fn(k, m, _p8(out))
    if cap:
        raise ValueError(f"k+m exceeds capacity {cap}")
    return out.reshape(k, m)


def tower_cauchy_matrix(k: int, m: int) -> np.ndarray:
    out = np.zeros(k * m, np.uint16)
    cap = lib().or_tower_cauchy_matrix(k, m, _p16(out))
    if cap:
        raise ValueError(f"k+m exceeds capacity {cap}")
    return out.reshape(k, m)


def gf256_matrix(engine: int, k: int, m: int) -> np.ndarray:
    """engine 0 = Standard Cauchy, 1 = Good Cauchy."""
    return std_cauchy_matrix(k, m) if engine == 0 else good_cauchy_matrix(k, m)


# ---- GF(256) codec --------------------------------------------------------

def gf256_encode(coeffs: np.ndarray, k: int, m: int, L: int, data: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(coeffs, np.uint8).reshape(-1)
    d = np.ascontiguousarray(data, np.uint8).reshape(-1)
    assert d.size == k * L
    out = np.zeros(m * L, np.uint8)
    lib().or_gf256_encode(_p8(c), k, m, L, _p8(d), _p8(out))
    return out


def _sym_arrays(symbols):
    n = len(symbols)
    idx = (C.c_size_t * max(n, 1))(*[s[0] for s in symbols])
    keep = [np.ascontiguousarray(np.frombuffer(bytes(s[1]), np.uint8)) if not isinstance(s[1], np.ndarray)
            else np.ascontiguousarray(s[1], np.uint8).reshape(-1) for s in symbols]
    ptrs = (C.c_void_p * max(n, 1))(*[k.ctypes.data if k.size else 0 for k in keep])
    lens = (C.c_size_t * max(n, 1))(*[k.size for k in keep])
    return idx, ptrs, lens, keep


def gf256_decode(coeffs: np.ndarray, k: int, m: int, L: int, symbols, out_len: int | None = None):
    """Returns (code, a, b, out)."""
    c = np.ascontiguousarray(coeffs, np.uint8).reshape(-1)
    idx, ptrs, lens, _keep = _sym_arrays(symbols)
    out_len = k * L if out_len is None else out_len
    out = np.zeros(max(out_len, 1), np.uint8)
    a = C.c_uint64()
    b = C.c_uint64()
    code = lib().or_gf256_decode(_p8(c), k, m, L, idx, ptrs, lens, len(symbols), _p8(out), out_len, C.byref(a),
                                 C.byref(b))
    return code, a.value, b.value, out[:out_len]


def gf256_decode_rref(coeffs: np.ndarray, k: int, m: int, L: int, symbols) -> np.ndarray | None:
    c = np.ascontiguousarray(coeffs, np.uint8).reshape(-1)
    idx, ptrs, _lens, _keep = _sym_arrays(symbols)
    out = np.zeros(k * L, np.uint8)
    rank = lib().or_gf256_decode_rref(_p8(c), k, m, L, idx, ptrs, _p8(out))
    return out if rank == k else None


def gf256_rows_terms(dst: np.ndarray, coef_tr: np.ndarray, srcs: list[np.ndarray], fast: bool = False) -> None:
    """dst (rows x L, modified in place) ^= coef_tr[t][r] * srcs[t]."""
    rows, L = dst.shape
    keep = [np.ascontiguousarray(s, np.uint8) for s in srcs]
    ptrs = (C.c_void_p * max(len(keep), 1))(*[k.ctypes.data for k in keep])
    c = np.ascontiguousarray(coef_tr, np.uint8).reshape(-1)
    fn = lib().or_gf256_rows_terms_fast if fast else lib().or_gf256_rows_terms
    fn(_p8(dst), L, rows, len(keep), _p8(c), ptrs)


def cpu_plan() -> int:
    return lib().or_cpu_plan()


def gf256_encode_many(coeffs: np.ndarray, k: int, m: int, L: int, data: np.ndarray, count: int,
                      threads: int = 1) -> np.ndarray:
    c = np.ascontiguousarray(coeffs, np.uint8).reshape(-1)
    d = np.ascontiguousarray(data, np.uint8).reshape(-1)
    assert d.size == count * k * L
    out = np.empty(count * m * L, np.uint8)
    lib().or_gf256_encode_many(_p8(c), k, m, L, _p8(d), _p8(out), count, threads)
    return out


def gf256_decode_many(coeffs: np.ndarray, k: int, m: int, L: int, idx: np.ndarray, data: np.ndarray,
                      repairs: np.ndarray, count: int, threads: int = 1) -> np.ndarray:
    """Fast CPU decode of `count` back-to-back codewords (idx: count x k, validated)."""
    c = np.ascontiguousarray(coeffs, np.uint8).reshape(-1)
    ix = np.ascontiguousarray(idx, np.uint16).reshape(-1)
    d = np.ascontiguousarray(data, np.uint8).reshape(-1)
    r = np.ascontiguousarray(repairs, np.uint8).reshape(-1)
    assert ix.size == count * k and d.size >= count * k * L and r.size >= count * m * L
    out = np.empty(count * k * L, np.uint8)
    if not lib().or_gf256_decode_many(_p8(c), k, m, L, _p16(ix), _p8(d), _p8(r), _p8(out), count, threads):
        raise ArithmeticError("singular submatrix")
    return out


# ---- Tower codec ----------------------------------------------------------

def tower_encode(coeffs: np.ndarray, k: int, m: int, L: int, data: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(coeffs, np.uint16).reshape(-1)
    d = np.ascontiguousarray(data, np.uint8).reshape(-1)
    out = np.zeros(m * L, np.uint8)
    lib().or_tower_encode(_p16(c), k, m, L, _p8(d), _p8(out))
    return out


def tower_decode_rref(coeffs: np.ndarray, k: int, m: int, L: int, symbols) -> np.ndarray | None:
    c = np.ascontiguousarray(coeffs, np.uint16).reshape(-1)
    idx, ptrs, _lens, _keep = _sym_arrays(symbols)
    out = np.zeros(k * L, np.uint8)
    rank = lib().or_tower_decode_rref(_p16(c), k, m, L, idx, ptrs, _p8(out))
    return out if rank == k else None


# ---- additive-FFT codec (Reed-Solomon at points = wire index) ---------------

def rs16_interp(k: int, L: int, points, values, outs) -> np.ndarray:
    """Evaluate the degree < k polynomial through (points[t], values[t]) at `outs`."""
    keep = [np.ascontiguousarray(np.frombuffer(bytes(v), np.uint8)) if not isinstance(v, np.ndarray)
            else np.ascontiguousarray(v, np.uint8).reshape(-1) for v in values]
    assert len(keep) == k and all(v.size == L for v in keep)
    pts = (C.c_size_t * max(k, 1))(*[int(p) for p in points])
    vals = (C.c_void_p * max(k, 1))(*[v.ctypes.data for v in keep])
    outs = list(outs)
    o = (C.c_size_t * max(len(outs), 1))(*[int(x) for x in outs])
    out = np.zeros(max(len(outs) * L, 1), np.uint8)
    if not lib().or_rs16_interp(k, L, pts, vals, len(outs), o, _p8(out)):
        raise ValueError("points are not distinct")
    return out[:len(outs) * L]


def afft_encode(k: int, m: int, L: int, data: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(data, np.uint8).reshape(-1)
    assert d.size == k * L
    out = np.zeros(m * L, np.uint8)
    lib().or_afft_encode(k, m, L, _p8(d), _p8(out))
    return out


def afft_decode(k: int, L: int, symbols) -> np.ndarray:
    """Data symbols from any k received (index, payload) pairs."""
    return rs16_interp(k, L, [s[0] for s in symbols], [s[1] for s in symbols], range(k))


# ---- seeded inputs --------------------------------------------------------

def fill_bytes(seed: int, byte_offset: int, length: int) -> np.ndarray:
    out = np.zeros(length, np.uint8)
    lib().or_fill_bytes(seed, byte_offset, _p8(out), length)
    return out


def splitmix_bytes_np(seed: int, byte_offset: int, length: int) -> np.ndarray:
    """Vectorised numpy form of fill_bytes for large samples (byte_offset % 8 == 0)."""
    assert byte_offset % 8 == 0
    words = (length + 7) // 8
    w = np.arange(byte_offset // 8, byte_offset // 8 + words, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (w + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z.astype("<u8").view(np.uint8)[:length].copy()


# ---------------------------------------------------------------------------
# CPU ports of the reference's GF(2^16) paths (the CPU baseline)
# ---------------------------------------------------------------------------


def cpu16_plan() -> int:
    """2 when the AVX2+GFNI kernels run, 0 for the scalar fallback."""
    return lib().or_cpu16_plan()


def tower_encode_port(coeffs: np.ndarray, k: int, m: int, L: int, data: np.ndarray, count: int = 1,
                      threads: int = 1) -> np.ndarray:
    """Tower incremental encode of `count` codewords (the reference's GFNI
    byte-swap rows kernel), one codeword per thread at a time."""
    c = np.ascontiguousarray(coeffs, dtype=np.uint16)
    d = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.empty(count * m * L, np.uint8)
    lib().or_tower_encode_port_many(_p16(c), k, m, L, _p8(d), _p8(out), count, threads)
    return out


def tower_finalize_port(k: int, m: int, L: int, idx: np.ndarray, data: np.ndarray, repairs: np.ndarray,
                        count: int = 1, threads: int = 1) -> np.ndarray | None:
    """Tower finalize of `count` codewords from idx (count x k received
    indices; data symbols from `data`, repairs from `repairs`): the
    rational-Lagrange recipe and output-major apply.  None on failure."""
    ix = np.ascontiguousarray(idx, dtype=np.uint16)
    d = np.ascontiguousarray(data, dtype=np.uint8)
    r = np.ascontiguousarray(repairs, dtype=np.uint8)
    out = np.empty(count * k * L, np.uint8)
    ok = lib().or_tower_finalize_port_many(k, m, L, _p16(ix), _p8(d), _p8(r), _p8(out), count, threads)
    return out if ok else None


def afft_encode_port(k: int, m: int, L: int, data: np.ndarray, threads: int = 1, max_strips: int = 0,
                     out: np.ndarray | None = None) -> tuple[np.ndarray, int]:
    """Strip-blocked fused additive-FFT encode (k a power of two, m <= k).
    Returns (repairs, strip width); with max_strips only the first strips'
    columns are computed."""
    d = np.ascontiguousarray(data, dtype=np.uint8)
    if out is None:
        out = np.zeros(m * L, np.uint8)
    w = lib().or_afft_encode_port(k, m, L, _p8(d), _p8(out), threads, max_strips)
    if w < 0:
        raise ValueError("geometry is off the fused encode path")
    return out, w


def afft_decode_port(k: int, m: int, L: int, symbols, threads: int = 1) -> np.ndarray:
    """Complete erasure-locator additive-FFT decode of k (index, payload)
    symbols, column slices shared out over `threads`."""
    idx, ptrs, lens, keep = _sym_arrays(symbols)
    if len(symbols) != k or any(n != L for n in lens[:len(symbols)]) or any(i >= k + m for i in idx[:k]):
        raise ValueError("need k symbols of symbol_len bytes with indices below k + m")
    out = np.empty(k * L, np.uint8)
    if lib().or_afft_decode_port(k, m, L, idx, ptrs, _p8(out), threads) != 0:
        raise ValueError("bad geometry")
    return out


def afft_novel_eval(coef: np.ndarray, points) -> np.ndarray:
    """Direct evaluation of novel-basis coefficients (len a power of two) at
    the given points: sum_i coef[i] X_i(x)."""
    c = np.ascontiguousarray(coef, np.uint16)
    pts = np.ascontiguousarray(points, np.uint16)
    out = np.empty(pts.size, np.uint16)
    lib().or_afft_novel_eval(_p16(c), c.size, _p16(pts), pts.size, _p16(out))
    return out
