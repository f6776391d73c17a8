"""GF(2^8) parity: the HIP library against the CPU oracle, bit for bit.

Every case runs through the C ABI (scrs_amd.codec is a ctypes binding).
Shapes cover the reference's own test shapes (batch/codec.rs:596-1108),
vector-boundary lengths (simd/dispatch.rs:476) and the two GF(2^8) bench
geometries.
"""
import itertools

import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchDecoder, BatchEncoder, DecodeError, EncodeError, Engine, Profile,
                            fill_random_dev)

pytestmark = pytest.mark.gpu

ENGINES = [Engine.STANDARD_CAUCHY, Engine.GOOD_CAUCHY]


def codec(engine, k, m, L):
    p = Profile.resolve(engine, k, m, L)
    return BatchEncoder(p), BatchDecoder(p), oracle.gf256_matrix(int(engine), k, m)


def symbols_of(data, rep, k, m, L):
    return [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k,m,L", [
    (1, 1, 1), (1, 1, 2), (3, 2, 4), (4, 3, 17), (10, 4, 4096), (8, 4, 64), (16, 8, 1024),
    (32, 16, 1024), (32, 16, 65536), (128, 64, 1024), (200, 50, 256), (1, 254, 64), (254, 1, 64),
    (33, 17, 4099), (20, 20, 16 * 257), (64, 32, 1400),
])
def test_encode_into_matches_oracle(torch_cuda, engine, k, m, L):
    if engine == Engine.GOOD_CAUCHY and k + m > 255:
        pytest.skip("beyond Good Cauchy capacity")
    rng = np.random.default_rng(k * 1000 + m * 10 + L)
    enc, _, c = codec(engine, k, m, L)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    want = oracle.gf256_encode(c, k, m, L, data)
    got = np.full(m * L, 0xAB, np.uint8)  # garbage must be overwritten
    enc.encode_into(data, got)
    assert np.array_equal(got, want)
    got2 = np.zeros(m * L, np.uint8)
    enc.encode_into_with(data, got2, enc.scratch())
    assert np.array_equal(got2, want)


@pytest.mark.parametrize("L", [0 + 1, 7, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 1400, 4099, 4096 * 3 + 5])
def test_encode_vector_boundary_lengths(torch_cuda, L):
    k, m = 6, 5
    rng = np.random.default_rng(L)
    enc, _, c = codec(Engine.GOOD_CAUCHY, k, m, L)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    assert np.array_equal(enc.encode(data), oracle.gf256_encode(c, k, m, L, data))


def test_encode_errors(torch_cuda):
    enc, _, _ = codec(Engine.STANDARD_CAUCHY, 3, 2, 4)
    with pytest.raises(EncodeError) as e:
        enc.encode_into(np.zeros(11, np.uint8), np.zeros(8, np.uint8))
    assert e.value == EncodeError("WrongInputLen", expected=12, got=11)
    with pytest.raises(EncodeError) as e:
        enc.encode_into(np.zeros(12, np.uint8), np.zeros(7, np.uint8))
    assert e.value == EncodeError("WrongOutputLen", expected=8, got=7)
    # input length is checked first
    with pytest.raises(EncodeError) as e:
        enc.encode_into(np.zeros(11, np.uint8), np.zeros(7, np.uint8))
    assert e.value.variant == "WrongInputLen"
    # a GF(256) unit scratch of another geometry is accepted (selector.rs:516-527)
    other, _, _ = codec(Engine.GOOD_CAUCHY, 5, 5, 9)
    out = np.zeros(8, np.uint8)
    enc.encode_into_with(np.arange(12, dtype=np.uint8), out, other.scratch())
    assert np.array_equal(out, enc.encode(np.arange(12, dtype=np.uint8)))


def test_encode_batch_dev_matches_oracle(torch_cuda):
    torch = torch_cuda
    k, m, L, count = 32, 16, 65536, 6
    enc, _, c = codec(Engine.GOOD_CAUCHY, k, m, L)
    rng = np.random.default_rng(9)
    data = rng.integers(0, 256, count * k * L, dtype=np.uint8)
    d = torch.from_numpy(data).cuda()
    r = torch.full((count * m * L,), 0x5A, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(d, r, count, torch.cuda.current_stream())
    torch.cuda.synchronize()
    want = oracle.gf256_encode_many(c, k, m, L, data, count, threads=4)
    assert np.array_equal(r.cpu().numpy(), want)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k,m,L", [(3, 2, 4), (4, 3, 8), (3, 3, 4), (5, 4, 33), (1, 1, 2)])
def test_decode_every_subset(torch_cuda, engine, k, m, L):
    # batch/codec.rs:709-803: exhaustive k-of-n enumeration
    rng = np.random.default_rng(k * 7 + m)
    enc, dec, c = codec(engine, k, m, L)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    rep = enc.encode(data)
    sym = symbols_of(data, rep, k, m, L)
    scratch = dec.scratch()
    for subset in itertools.combinations(range(k + m), k):
        received = [(i, sym[i]) for i in subset]
        out = np.full(k * L, 0xAB, np.uint8)
        dec.decode_into(received, out)
        assert np.array_equal(out, data), subset
        out2 = np.zeros(k * L, np.uint8)
        dec.decode_into_with(received[::-1], out2, scratch)
        assert np.array_equal(out2, data), subset


def test_decode_matches_oracle_at_reference_shapes(torch_cuda):
    # batch/codec.rs:936-1013 (optimized_decode_matches_reference shapes)
    for (k, m, L) in [(1, 1, 2), (4, 2, 16), (8, 4, 64), (6, 6, 33), (16, 8, 100), (32, 16, 64)]:
        for engine in ENGINES:
            enc, dec, c = codec(engine, k, m, L)
            data = ((np.arange(k * L) * 131 + 7) % 256).astype(np.uint8)
            rep = enc.encode(data)
            sym = symbols_of(data, rep, k, m, L)
            rng = np.random.default_rng(k + m)
            for _ in range(12):
                subset = sorted(rng.choice(k + m, size=k, replace=False).tolist())
                received = [(i, sym[i]) for i in subset]
                got = dec.decode(received)
                code, _, _, want = oracle.gf256_decode(c, k, m, L, received)
                assert code == 0
                assert np.array_equal(got, want)
                assert np.array_equal(got, data)


def test_decode_repair_only_recovery(torch_cuda):
    # batch/codec.rs:758-774
    k, m, L = 3, 3, 4
    enc, dec, _ = codec(Engine.STANDARD_CAUCHY, k, m, L)
    data = np.array([0xDE, 0xAD, 0xBE, 0xEF, 0x12, 0x34, 0x56, 0x78, 0x9A, 0xBC, 0xDE, 0xF0], np.uint8)
    rep = enc.encode(data)
    sym = symbols_of(data, rep, k, m, L)
    assert np.array_equal(dec.decode([(3, sym[3]), (4, sym[4]), (5, sym[5])]), data)


def test_decode_errors_follow_reference_order(torch_cuda):
    # batch/codec.rs:805-866, 1015-1050
    k, m, L = 3, 2, 4
    enc, dec, _ = codec(Engine.STANDARD_CAUCHY, k, m, L)
    sym = symbols_of(np.zeros(12, np.uint8), enc.encode(np.zeros(12, np.uint8)), k, m, L)
    with pytest.raises(DecodeError) as e:
        dec.decode([(0, sym[0]), (1, sym[1])])
    assert e.value == DecodeError("WrongCount", expected=3, got=2)
    with pytest.raises(DecodeError) as e:
        dec.decode([(0, sym[0]), (1, sym[1]), (5, np.zeros(4, np.uint8))])
    assert e.value == DecodeError("IndexOutOfRange", index=5, n=5)
    with pytest.raises(DecodeError) as e:
        dec.decode([(0, sym[0]), (1, sym[1]), (0, sym[0])])
    assert e.value == DecodeError("DuplicateIndex", index=0)
    with pytest.raises(DecodeError) as e:
        dec.decode([(0, sym[0]), (1, sym[1]), (2, np.zeros(3, np.uint8))])
    assert e.value == DecodeError("WrongPayloadLen", expected=4, got=3)
    with pytest.raises(DecodeError) as e:
        dec.decode_into([(0, sym[0]), (1, sym[1]), (2, sym[2])], np.zeros(11, np.uint8))
    assert e.value == DecodeError("WrongOutputLen", expected=12, got=11)
    _, other, _ = codec(Engine.STANDARD_CAUCHY, 4, 2, 4)
    with pytest.raises(DecodeError) as e:
        dec.decode_into_with([(i, sym[i]) for i in range(3)], np.zeros(12, np.uint8), other.scratch())
    assert e.value == DecodeError("ScratchMismatch")
    # scratch mismatch wins over a wrong output length
    with pytest.raises(DecodeError) as e:
        dec.decode_into_with([(i, sym[i]) for i in range(3)], np.zeros(11, np.uint8), other.scratch())
    assert e.value.variant == "ScratchMismatch"


def test_config1_standard_cauchy_10_4(torch_cuda):
    # Standard Cauchy k=10, m=4, L=4096: encode, erase 4 of 14, decode the
    # 10 survivors in ascending order.
    k, m, L = 10, 4, 4096
    enc, dec, c = codec(Engine.STANDARD_CAUCHY, k, m, L)
    data = oracle.splitmix_bytes_np(0x5C25_0001, 0, k * L)
    rep = enc.encode(data)
    assert np.array_equal(rep, oracle.gf256_encode(c, k, m, L, data))
    sym = symbols_of(data, rep, k, m, L)
    rng = np.random.default_rng(0x5C25_0001)
    erased = set(rng.choice(k + m, size=4, replace=False).tolist())
    received = [(i, sym[i]) for i in range(k + m) if i not in erased]
    assert len(received) == k
    assert np.array_equal(dec.decode(received), data)


def _erasure_indices(rng, count, k, m, max_r):
    n = k + m
    idx = np.zeros((count, k), np.uint16)
    for c in range(count):
        r = int(rng.integers(1, max_r + 1))
        erased = set(rng.choice(n, size=r, replace=False).tolist())
        surv = [i for i in range(n) if i not in erased][:k]
        idx[c] = surv
    return idx


@pytest.mark.parametrize("engine", ENGINES)
def test_decode_batch_dev_config2_geometry(torch_cuda, engine):
    torch = torch_cuda
    k, m, L, count = 32, 16, 65536, 8
    enc, dec, c = codec(engine, k, m, L)
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    from scrs_amd.codec import fill_random_dev
    fill_random_dev(data, data.numel(), seed=1234)
    rep = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(data, rep, count)
    rng = np.random.default_rng(77)
    idx = _erasure_indices(rng, count, k, m, 16)
    out = torch.full((count * k * L,), 0xAB, dtype=torch.uint8, device="cuda")
    dec.decode_batch_dev(idx, data, rep, out, count)
    torch.cuda.synchronize()
    assert torch.equal(out, data)
    # The bytes on the device are the seeded stream.
    host = data[: 3 * L].cpu().numpy()
    assert np.array_equal(host, oracle.splitmix_bytes_np(1234, 0, 3 * L))
    # Repairs of the first codeword against the oracle.
    want = oracle.gf256_encode(c, k, m, L, data[: k * L].cpu().numpy())
    assert np.array_equal(rep[: m * L].cpu().numpy(), want)


def test_decode_batch_ptrs_dev(torch_cuda):
    torch = torch_cuda
    k, m, L, count = 8, 4, 1000, 5
    enc, dec, c = codec(Engine.STANDARD_CAUCHY, k, m, L)
    rng = np.random.default_rng(5)
    data = rng.integers(0, 256, count * k * L, dtype=np.uint8)
    d = torch.from_numpy(data).cuda()
    r = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(d, r, count)
    idx = _erasure_indices(rng, count, k, m, m)
    ptrs = np.zeros(count * k, np.uint64)
    for cw in range(count):
        for t in range(k):
            i = int(idx[cw, t])
            base = d.data_ptr() + (cw * k + i) * L if i < k else r.data_ptr() + (cw * m + i - k) * L
            ptrs[cw * k + t] = base
    p = torch.from_numpy(ptrs.view(np.int64)).cuda()
    out = torch.zeros(count * k * L, dtype=torch.uint8, device="cuda")
    dec.decode_batch_ptrs_dev(idx, p, out, count)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), data)


def test_decode_batch_dev_rejects_bad_indices_before_work(torch_cuda):
    torch = torch_cuda
    k, m, L = 4, 2, 16
    _, dec, _ = codec(Engine.GOOD_CAUCHY, k, m, L)
    d = torch.zeros(2 * k * L, dtype=torch.uint8, device="cuda")
    r = torch.zeros(2 * m * L, dtype=torch.uint8, device="cuda")
    out = torch.zeros(2 * k * L, dtype=torch.uint8, device="cuda")
    idx = np.array([[0, 1, 2, 3], [0, 1, 1, 4]], np.uint16)
    with pytest.raises(DecodeError) as e:
        dec.decode_batch_dev(idx, d, r, out, 2)
    assert e.value == DecodeError("DuplicateIndex", index=1)
    idx = np.array([[0, 1, 2, 3], [0, 1, 6, 4]], np.uint16)
    with pytest.raises(DecodeError) as e:
        dec.decode_batch_dev(idx, d, r, out, 2)
    assert e.value == DecodeError("IndexOutOfRange", index=6, n=6)


@pytest.mark.parametrize("engine", [Engine.GOOD_CAUCHY, Engine.AFFT])
def test_encoder_shared_across_threads(torch_cuda, engine):
    # Encoders take &self in the reference, so one handle may serve several
    # threads at once, each with its own scratch (or none).
    import threading

    k, m, L = 12, 6, 4096
    p = Profile.resolve(engine, k, m, L)
    enc = BatchEncoder(p)
    if engine == Engine.AFFT:
        want_of = lambda d: oracle.afft_encode(k, m, L, d)  # noqa: E731
    else:
        c = oracle.gf256_matrix(int(engine), k, m)
        want_of = lambda d: oracle.gf256_encode(c, k, m, L, d)  # noqa: E731
    rng = np.random.default_rng(77)
    inputs = [rng.integers(0, 256, k * L, dtype=np.uint8) for _ in range(6)]
    wants = [want_of(d) for d in inputs]
    errors = []

    def worker(tid):
        try:
            scratch = enc.scratch() if tid % 2 else None
            for it in range(12):
                i = (tid + it) % len(inputs)
                got = np.empty(m * L, np.uint8)
                if scratch is None:
                    enc.encode_into(inputs[i], got)
                else:
                    enc.encode_into_with(inputs[i], got, scratch)
                if not np.array_equal(got, wants[i]):
                    errors.append((tid, it))
        except Exception as e:  # noqa: BLE001
            errors.append((tid, repr(e)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert errors == []


@pytest.mark.parametrize("engine", [Engine.STANDARD_CAUCHY, Engine.GOOD_CAUCHY])
@pytest.mark.parametrize("k,m,L", [(100, 20, 65536), (40, 7, 4100), (17, 3, 1028), (200, 33, 2048)])
@pytest.mark.parametrize("split", [False, True])
def test_lone_codeword_term_split(torch_cuda, engine, k, m, L, split, term_split):
    """One codeword fills too few workgroups: with the term split enabled its
    terms run as segments whose partial rows are XORed into cleared
    destinations; encode and decode against the oracle either way."""
    torch = torch_cuda
    term_split(split)
    if engine == Engine.GOOD_CAUCHY and k + m > 255:
        k = 255 - m
    p = Profile.resolve(engine, k, m, L)
    data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), k * 31 + m)
    rep = torch.full((m * L,), 0xAB, dtype=torch.uint8, device="cuda")
    BatchEncoder(p).encode_batch_dev(data, rep, 1)
    torch.cuda.synchronize()
    d = data.cpu().numpy()
    assert np.array_equal(rep.cpu().numpy(), oracle.gf256_encode(oracle.gf256_matrix(int(engine), k, m), k, m, L, d))
    rng = np.random.default_rng(k + m)
    erased = set(rng.choice(k, size=min(m, k), replace=False).tolist())
    idx = np.array([[i for i in range(k + m) if i not in erased][:k]], np.uint16)
    out = torch.full_like(data, 0xCD)
    BatchDecoder(p).decode_batch_dev(idx, data, rep, out, 1)
    torch.cuda.synchronize()
    assert torch.equal(out, data)
