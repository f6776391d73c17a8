"""The CPU ports timed as the GF(2^16) baselines equal the scalar oracle.

The ports follow the reference's fast CPU paths (GFNI byte-swap
multiply-XOR and rational-Lagrange recipes for Tower; strip-blocked fused
encode and the complete erasure-locator decode for the additive FFT); the
oracle they are checked against is the plain per-element restatement
(schoolbook products, Gauss-Jordan, Lagrange interpolation), so a wrong
factor word, node factor, strip bound or locator term fails here.
"""
import numpy as np
import pytest

import oracle


def rand(n, seed):
    return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8)


@pytest.mark.parametrize("k,m,L", [(1, 1, 2), (3, 2, 30), (8, 4, 64), (20, 10, 98), (50, 16, 1024), (200, 50, 96)])
def test_tower_encode_port_equals_oracle(k, m, L):
    c = oracle.tower_cauchy_matrix(k, m)
    data = rand(k * L, k * 31 + L)
    got = oracle.tower_encode_port(c, k, m, L, data)
    assert np.array_equal(got, oracle.tower_encode(c, k, m, L, data))


def test_tower_encode_port_many_threads():
    k, m, L, count = 12, 5, 160, 7
    c = oracle.tower_cauchy_matrix(k, m)
    data = rand(count * k * L, 3)
    got = oracle.tower_encode_port(c, k, m, L, data, count=count, threads=3)
    for i in range(count):
        want = oracle.tower_encode(c, k, m, L, data[i * k * L:(i + 1) * k * L])
        assert np.array_equal(got[i * m * L:(i + 1) * m * L], want)


@pytest.mark.parametrize("k,m,L,seed", [(4, 4, 32, 1), (10, 6, 66, 2), (40, 20, 64, 3), (200, 50, 64, 4)])
def test_tower_finalize_port_round_trips(k, m, L, seed):
    rng = np.random.default_rng(seed)
    c = oracle.tower_cauchy_matrix(k, m)
    count = 3
    data = rand(count * k * L, seed)
    rep = oracle.tower_encode_port(c, k, m, L, data, count=count)
    idx = np.empty((count, k), np.uint16)
    for cw in range(count):
        r = int(rng.integers(0, min(k, m) + 1))
        erased = set(rng.choice(k, size=r, replace=False).tolist())
        keep = [i for i in range(k + m) if i not in erased]
        idx[cw] = rng.permutation(keep)[:k]
    out = oracle.tower_finalize_port(k, m, L, idx, data, rep, count=count, threads=2)
    assert out is not None and np.array_equal(out, data)
    # The same symbols through the oracle's Gauss-Jordan decode.
    sym = [(int(x), data[x * L:(x + 1) * L] if x < k else rep[(x - k) * L:(x - k + 1) * L]) for x in idx[0]]
    assert np.array_equal(oracle.tower_decode_rref(c, k, m, L, sym), data[:k * L])


def test_tower_finalize_port_reports_too_few_repairs():
    k, m, L = 6, 3, 32
    data = rand(k * L, 9)
    rep = oracle.tower_encode_port(oracle.tower_cauchy_matrix(k, m), k, m, L, data)
    idx = np.array([[1, 2, 3, 4, 5, 5]], np.uint16)  # one data symbol missing, no repair
    assert oracle.tower_finalize_port(k, m, L, idx, data, rep) is None


@pytest.mark.parametrize("k,m,L", [(2, 1, 2), (4, 4, 64), (16, 5, 40), (64, 64, 34), (256, 100, 100),
                                   (8192, 3, 200)])
def test_afft_encode_port_equals_lagrange_oracle(k, m, L):
    data = rand(k * L, k + m + L)
    got, width = oracle.afft_encode_port(k, m, L, data, threads=2)
    assert width == min(L, max(2, (786432 // k) & ~1))
    assert np.array_equal(got, oracle.afft_encode(k, m, L, data))


def test_afft_encode_port_bounded_sample_covers_leading_strips():
    k, m, L = 8192, 4, 400  # strips of 96 bytes
    data = rand(k * L, 5)
    full, w = oracle.afft_encode_port(k, m, L, data)
    part, _ = oracle.afft_encode_port(k, m, L, data, max_strips=2)
    assert w == 96
    rows_full, rows_part = full.reshape(m, L), part.reshape(m, L)
    assert np.array_equal(rows_full[:, :2 * w], rows_part[:, :2 * w])
    assert not rows_part[:, 2 * w:].any()


def test_afft_encode_port_rejects_unfused_geometry():
    with pytest.raises(ValueError):
        oracle.afft_encode_port(6, 2, 8, rand(48, 1))


@pytest.mark.parametrize("k,m,L,seed", [(4, 4, 8, 1), (8, 3, 34, 2), (64, 64, 40, 3), (100, 28, 66, 4),
                                        (512, 512, 16, 5)])
def test_afft_decode_port_round_trips(k, m, L, seed):
    rng = np.random.default_rng(seed)
    data = rand(k * L, seed)
    rep = oracle.afft_encode(k, m, L, data)
    word = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    for threads in (1, 3):
        idx = rng.permutation(k + m)[:k]
        out = oracle.afft_decode_port(k, m, L, [(int(i), word[i]) for i in idx], threads=threads)
        assert np.array_equal(out, data)


def test_afft_decode_port_with_every_data_symbol_missing():
    k, m, L = 16, 16, 24
    data = rand(k * L, 11)
    rep = oracle.afft_encode(k, m, L, data)
    sym = [(k + j, rep[j * L:(j + 1) * L]) for j in range(m)]
    assert np.array_equal(oracle.afft_decode_port(k, m, L, sym), data)
