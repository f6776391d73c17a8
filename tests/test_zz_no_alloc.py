"""The no-allocation batched decode (scrs_decode_batch_dev_noalloc) against
scrs_decode_batch_dev: same bytes on two calls in a row, the second with a
smaller input, and no device allocation by either call.  Its workspace is a
fixed per-device array, so consecutive calls reuse the same memory."""
import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchDecoder, BatchEncoder, ConfigError, DeviceError, Engine, Profile,
                            device_alloc_count, fill_random_dev, noalloc_scratch_bytes)

pytestmark = pytest.mark.gpu


def erasures(count, k, m, seed):
    rng = np.random.default_rng(seed)
    idx = np.empty((count, k), np.uint16)
    for c in range(count):
        r = int(rng.integers(1, m + 1))
        keep = np.ones(k + m, bool)
        keep[rng.permutation(k + m)[:r]] = False
        idx[c] = rng.permutation(np.flatnonzero(keep)[:k])  # any received order
    return idx


@pytest.mark.parametrize("engine,k,m,L,counts", [
    (Engine.GOOD_CAUCHY, 32, 16, 4096, (300, 37)),
    (Engine.STANDARD_CAUCHY, 10, 4, 1500, (64, 5)),
    (Engine.GOOD_CAUCHY, 200, 50, 512, (40, 3)),
])
def test_noalloc_decode_matches_decode_batch_dev(torch_cuda, engine, k, m, L, counts):
    torch = torch_cuda
    p = Profile.resolve(engine, k, m, L)
    enc, dec = BatchEncoder(p), BatchDecoder(p)
    n0 = counts[0]
    data = torch.empty(n0 * k * L, dtype=torch.uint8, device="cuda")
    rep = torch.empty(n0 * m * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 0x0A11 + k, 0)
    enc.encode_batch_dev(data, rep, n0)
    torch.cuda.synchronize()
    # The repairs are the oracle's on a sampled codeword.
    c = oracle.gf256_matrix(int(engine), k, m)
    d0 = data[:k * L].cpu().numpy()
    assert np.array_equal(rep[:m * L].cpu().numpy(), oracle.gf256_encode(c, k, m, L, d0))
    for call, count in enumerate(counts):
        assert dec.noalloc_bytes(count) <= noalloc_scratch_bytes()
        idx = erasures(count, k, m, 0x5EED + 17 * call + k)
        want = torch.full((count * k * L,), 0xAB, dtype=torch.uint8, device="cuda")
        got = torch.full((count * k * L,), 0xCD, dtype=torch.uint8, device="cuda")
        dec.decode_batch_dev(idx, data, rep, want, count)
        torch.cuda.synchronize()
        before = device_alloc_count()
        dec.decode_batch_dev_noalloc(idx, data, rep, got, count)
        torch.cuda.synchronize()
        assert device_alloc_count() == before, "the no-allocation decode allocated device memory"
        assert torch.equal(got, want), f"call {call}: no-allocation decode differs from decode_batch_dev"
        assert torch.equal(got, data[:count * k * L]), f"call {call}: decode did not reconstruct the data"


def test_noalloc_decode_rejects_what_does_not_fit(torch_cuda):
    torch = torch_cuda
    k, m, L = 200, 55, 64
    dec = BatchDecoder(Profile.resolve(Engine.GOOD_CAUCHY, k, m, L))
    buf = torch.zeros(1024, dtype=torch.uint8, device="cuda")
    count = 2000
    assert dec.noalloc_bytes(count) > noalloc_scratch_bytes()
    idx = np.tile(np.arange(m, k + m, dtype=np.uint16), (count, 1))
    with pytest.raises(DeviceError) as e:
        dec.decode_batch_dev_noalloc(idx, buf, buf, buf, count)
    assert e.value.variant == "CapacityExceeded"
    assert e.value.fields["limit"] == noalloc_scratch_bytes()
    # A bad index is reported before the size.
    idx[5, 3] = k + m
    with pytest.raises(Exception) as e:
        dec.decode_batch_dev_noalloc(idx, buf, buf, buf, count)
    assert e.value.variant == "IndexOutOfRange"


def test_noalloc_decode_is_gf256_only(torch_cuda):
    torch = torch_cuda
    dec = BatchDecoder(Profile.resolve(Engine.TOWER, 8, 4, 64))
    buf = torch.zeros(1024, dtype=torch.uint8, device="cuda")
    with pytest.raises(ConfigError) as e:
        dec.decode_batch_dev_noalloc(np.arange(8, dtype=np.uint16), buf, buf, buf, 1)
    assert e.value.variant == "UnsupportedMode"
