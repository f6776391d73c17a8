"""The bit-matrix GF(2^8) kernels (int8 MFMA and e2m1 / fp4 MFMA) against
the CPU oracle and the table kernel, bit for bit, through the C ABI."""
import itertools

import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchDecoder, BatchEncoder, ConfigError, DeviceError, Engine, Kernel, Profile,
                            fill_random_dev)

pytestmark = pytest.mark.gpu

ENGINES = [Engine.STANDARD_CAUCHY, Engine.GOOD_CAUCHY]
KINDS = [Kernel.BITMATRIX, Kernel.BITMATRIX_FP4]
# Source counts per kernel: int8 MFMA chunks of 8 up to 64, fp4 chunks of 16
# up to every GF(256) geometry (k + m <= 255 for Good Cauchy), across 1 to 16
# chunks.
FP4_WIDE = [(65, 4, 64), (100, 20, 1028), (129, 17, 260), (200, 32, 516), (254, 1, 64), (240, 15, 8200)]


def bm_codec(engine, k, m, L, kind=Kernel.BITMATRIX):
    p = Profile.resolve(engine, k, m, L)
    enc, dec = BatchEncoder(p), BatchDecoder(p)
    enc.set_kernel(kind)
    dec.set_kernel(kind)
    return enc, dec, oracle.gf256_matrix(int(engine), k, m)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k,m,L", [
    (1, 1, 4), (2, 1, 8), (3, 2, 12), (8, 4, 64), (9, 5, 100), (13, 16, 1500), (16, 17, 260),
    (32, 16, 4096), (32, 16, 4100), (40, 32, 8192), (57, 3, 516), (64, 1, 2048), (64, 40, 1024),
    (10, 4, 4096),
])
def test_encode_matches_oracle(torch_cuda, engine, k, m, L, kind):
    rng = np.random.default_rng(k * 1000 + m * 10 + L)
    enc, _, c = bm_codec(engine, k, m, L, kind)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    want = oracle.gf256_encode(c, k, m, L, data)
    got = np.full(m * L, 0xAB, np.uint8)
    enc.encode_into(data, got)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k,m,L", FP4_WIDE)
def test_fp4_wide_encode_decode_match_oracle(torch_cuda, engine, k, m, L):
    rng = np.random.default_rng(k * 7 + m + L)
    enc, dec, c = bm_codec(engine, k, m, L, Kernel.BITMATRIX_FP4)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    want = oracle.gf256_encode(c, k, m, L, data)
    got = np.full(m * L, 0xAB, np.uint8)
    enc.encode_into(data, got)
    assert np.array_equal(got, want)
    syms = [data[i * L:(i + 1) * L] for i in range(k)] + [want[j * L:(j + 1) * L] for j in range(m)]
    lost = set(rng.choice(k, size=min(m, k), replace=False).tolist())
    order = [i for i in range(k + m) if i not in lost][:k]
    out = np.full(k * L, 0xCD, np.uint8)
    dec.decode_into([(i, syms[i]) for i in rng.permutation(order)], out)
    assert np.array_equal(out, data)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("engine", ENGINES)
def test_decode_every_subset_small(torch_cuda, engine, kind):
    k, m, L = 4, 3, 36
    rng = np.random.default_rng(7)
    enc, dec, c = bm_codec(engine, k, m, L, kind)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    rep = enc.encode(data)
    syms = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    for subset in itertools.combinations(range(k + m), k):
        for order in (subset, subset[::-1]):
            out = np.full(k * L, 0xCD, np.uint8)
            dec.decode_into([(i, syms[i]) for i in order], out)
            assert np.array_equal(out, data), order


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k,m,L", [(32, 16, 4096), (64, 32, 1028), (17, 9, 8)])
def test_batch_dev_matches_table_kernel(torch_cuda, engine, k, m, L, kind):
    torch = torch_cuda
    count = 24
    p = Profile.resolve(engine, k, m, L)
    table_enc, table_dec = BatchEncoder(p), BatchDecoder(p)
    enc, dec, _ = bm_codec(engine, k, m, L, kind)
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 99 + k)
    rep_t = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    rep_b = torch.full_like(rep_t, 0xAB)
    table_enc.encode_batch_dev(data, rep_t, count)
    enc.encode_batch_dev(data, rep_b, count)
    torch.cuda.synchronize()
    assert torch.equal(rep_b, rep_t)
    rng = np.random.default_rng(k + m)
    idx = np.empty((count, k), np.uint16)
    for cw in range(count):
        r = int(rng.integers(1, m + 1))
        keep = np.ones(k + m, bool)
        keep[rng.permutation(k + m)[:r]] = False
        survivors = np.flatnonzero(keep)
        idx[cw] = rng.permutation(survivors)[:k]
    out_t = torch.full_like(data, 0x11)
    out_b = torch.full_like(data, 0x22)
    table_dec.decode_batch_dev(idx, data, rep_t, out_t, count)
    dec.decode_batch_dev(idx, data, rep_b, out_b, count)
    torch.cuda.synchronize()
    assert torch.equal(out_t, data)
    assert torch.equal(out_b, data)


@pytest.mark.parametrize("kind", KINDS)
def test_bench_geometry_sample(torch_cuda, kind):
    # k=32, m=16, L=65536 on a handful of codewords, with repeats against the
    # oracle for one codeword.
    torch = torch_cuda
    k, m, L, count = 32, 16, 65536, 8
    enc, dec, c = bm_codec(Engine.GOOD_CAUCHY, k, m, L, kind)
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 5150)
    rep = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(data, rep, count)
    torch.cuda.synchronize()
    d0 = data[: k * L].cpu().numpy()
    assert np.array_equal(rep[: m * L].cpu().numpy(), oracle.gf256_encode(c, k, m, L, d0))
    idx = np.array([[i for i in range(k + m) if i not in (0, 5, 31, 40)][:k]] * count, np.uint16)
    out = torch.empty_like(data)
    dec.decode_batch_dev(idx, data, rep, out, count)
    torch.cuda.synchronize()
    assert torch.equal(out, data)


@pytest.mark.parametrize("kind", KINDS)
def test_unaligned_buffers_use_the_table_kernel(torch_cuda, kind):
    torch = torch_cuda
    k, m, L = 8, 4, 64
    enc, _, c = bm_codec(Engine.GOOD_CAUCHY, k, m, L, kind)
    buf = torch.empty(k * L + 1, dtype=torch.uint8, device="cuda")
    fill_random_dev(buf, buf.numel(), 3)
    rep = torch.empty(m * L + 1, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(buf[1:], rep[1:], 1)
    torch.cuda.synchronize()
    d = buf[1:].cpu().numpy()
    assert np.array_equal(rep[1:].cpu().numpy(), oracle.gf256_encode(c, k, m, L, d))


def test_kernel_selection_errors(torch_cuda):
    with pytest.raises(DeviceError) as e:  # k beyond the bit-matrix limit
        BatchEncoder(Profile.resolve(Engine.GOOD_CAUCHY, 65, 4, 64)).set_kernel(Kernel.BITMATRIX)
    assert e.value == DeviceError("CapacityExceeded", limit=64)
    with pytest.raises(DeviceError) as e:  # symbol_len not a multiple of 4
        BatchDecoder(Profile.resolve(Engine.STANDARD_CAUCHY, 8, 4, 66)).set_kernel(Kernel.BITMATRIX)
    assert e.value == DeviceError("InvalidArgument", position=2)
    with pytest.raises(ConfigError) as e:
        BatchDecoder(Profile.resolve(Engine.TOWER, 8, 4, 64)).set_kernel(Kernel.BITMATRIX)
    assert e.value == ConfigError("UnsupportedMode", engine=Engine.TOWER)
    # The table kernel is always accepted on GF(256), and switching back works.
    enc = BatchEncoder(Profile.resolve(Engine.GOOD_CAUCHY, 65, 4, 64))
    enc.set_kernel(Kernel.TABLE)
    # The fp4 kernel covers every GF(256) geometry; symbol_len must still be
    # a multiple of 4, and switching between the two MFMA kernels rebuilds
    # the operands.
    enc.set_kernel(Kernel.BITMATRIX_FP4)
    with pytest.raises(DeviceError) as e:
        BatchEncoder(Profile.resolve(Engine.GOOD_CAUCHY, 8, 4, 66)).set_kernel(Kernel.BITMATRIX_FP4)
    assert e.value == DeviceError("InvalidArgument", position=2)
    k, m, L = 8, 4, 64
    rng = np.random.default_rng(11)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    enc = BatchEncoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    want = oracle.gf256_encode(oracle.gf256_matrix(int(Engine.GOOD_CAUCHY), k, m), k, m, L, data)
    for kind in (Kernel.BITMATRIX, Kernel.BITMATRIX_FP4, Kernel.BITMATRIX, Kernel.TABLE):
        enc.set_kernel(kind)
        assert np.array_equal(enc.encode(data), want), kind
