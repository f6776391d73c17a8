"""GF(2^16) Tower parity: the HIP library against the CPU oracle, bit for bit.

The Tower engine has an incremental encoder and a streaming/batch decoder
(selector.rs:117-125, 270-279, 370-381); its batch decode is the blanket
BatchDecoder over the streaming decoder (codec.rs:203-240).  Shapes follow
tower/encoder.rs:184-230 and tower/decoder.rs:600-675, plus the
k=200, m=50, symbol_len=8192 streaming geometry.
"""
import itertools

import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchDecoder, BatchEncoder, ConfigError, DecodeError, Decoder, EncodeError, Engine,
                            IncrementalEncoder, Profile, PushOutcome)

pytestmark = pytest.mark.gpu

TOWER = Engine.TOWER


def encode_with(k, m, L, data, order=None):
    enc = IncrementalEncoder(Profile.resolve(TOWER, k, m, L))
    order = range(k) if order is None else order
    for i in order:
        enc.feed(i, data[i * L:(i + 1) * L])
    return enc, np.frombuffer(b"".join(enc.repair(j) for j in range(m)), np.uint8)


def symbols_of(data, rep, k, m, L):
    return [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]


def test_batch_encoder_is_unsupported(torch_cuda):
    # batch_encoder (selector.rs:457-467): Tower is incremental-only
    with pytest.raises(ConfigError) as e:
        BatchEncoder(Profile.resolve(TOWER, 4, 2, 8))
    assert e.value == ConfigError("UnsupportedMode", engine=TOWER)


def test_profile_validation(torch_cuda):
    # tower/encoder.rs:185-191
    IncrementalEncoder(Profile.resolve(TOWER, 256, 2, 64))
    with pytest.raises(ConfigError):
        Profile.resolve(TOWER, 1, 1, 0)
    with pytest.raises(ConfigError):
        Profile.resolve(TOWER, 1, 1, 3)
    with pytest.raises(ConfigError) as e:
        Profile.resolve(TOWER, 32768, 32768, 2)
    assert e.value == ConfigError("TooManySymbols", cap=65535)


def test_incremental_repair_matches_direct_field_sum(torch_cuda):
    # tower/encoder.rs:193-217
    k, m, L = 4, 3, 10
    data = np.array([(row * 31 + byte * 7) % 256 for row in range(k) for byte in range(L)], np.uint8)
    _, rep = encode_with(k, m, L, data)
    want = oracle.tower_encode(oracle.tower_cauchy_matrix(k, m), k, m, L, data)
    assert np.array_equal(rep, want)


def test_reset_retains_configuration_and_clears_state(torch_cuda):
    # tower/encoder.rs:219-228
    enc = IncrementalEncoder(Profile.resolve(TOWER, 3, 2, 8))
    enc.feed(0, bytes([7] * 8))
    assert enc.repair(0) != bytes(8)
    enc.reset()
    assert enc.fed_count() == 0
    assert enc.repair(0) == bytes(8)


@pytest.mark.parametrize("k,m,L", [
    (1, 1, 2), (3, 2, 4), (4, 3, 18), (10, 4, 4096), (16, 16, 34), (33, 17, 4098), (64, 32, 66),
    (200, 50, 8192), (256, 3, 4), (300, 20, 1400), (40, 100, 64),
])
def test_incremental_encode_matches_oracle(torch_cuda, k, m, L):
    rng = np.random.default_rng(k * 7919 + m * 31 + L)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    order = rng.permutation(k).tolist()
    enc, rep = encode_with(k, m, L, data, order)
    assert enc.fed_count() == k
    want = oracle.tower_encode(oracle.tower_cauchy_matrix(k, m), k, m, L, data)
    assert np.array_equal(rep, want)


def test_partial_sums_before_all_data_is_fed(torch_cuda):
    # A repair read early is the sum over the data fed so far
    # (tower/encoder.rs:118-121).
    k, m, L = 70, 9, 130
    rng = np.random.default_rng(5)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    c = oracle.tower_cauchy_matrix(k, m)
    enc = IncrementalEncoder(Profile.resolve(TOWER, k, m, L))
    order = rng.permutation(k).tolist()
    fed = np.zeros(k * L, np.uint8)
    for step, i in enumerate(order):
        enc.feed(i, data[i * L:(i + 1) * L])
        fed[i * L:(i + 1) * L] = data[i * L:(i + 1) * L]
        if step in (0, 1, 31, 32, 33, 50, k - 1):
            got = np.frombuffer(b"".join(enc.repair(j) for j in range(m)), np.uint8)
            assert np.array_equal(got, oracle.tower_encode(c, k, m, L, fed)), step


def test_feed_errors_follow_reference_order(torch_cuda):
    # tower/encoder.rs:89-131: IndexOutOfRange{n=k+m} -> WrongPayloadLen -> DuplicateData
    enc = IncrementalEncoder(Profile.resolve(TOWER, 3, 2, 4))
    with pytest.raises(EncodeError) as e:
        enc.feed(3, bytes(3))
    assert e.value == EncodeError("IndexOutOfRange", index=3, n=5)
    with pytest.raises(EncodeError) as e:
        enc.feed(0, bytes(3))
    assert e.value == EncodeError("WrongPayloadLen", expected=4, got=3)
    enc.feed(0, bytes(4))
    with pytest.raises(EncodeError) as e:
        enc.feed(0, bytes(3))
    assert e.value.variant == "WrongPayloadLen"
    with pytest.raises(EncodeError) as e:
        enc.feed(0, bytes(4))
    assert e.value == EncodeError("DuplicateData", index=0)
    with pytest.raises(EncodeError) as e:
        enc.repair(2)
    assert e.value == EncodeError("IndexOutOfRange", index=2, n=5)


def test_exhaustive_small_k_of_n_recovery(torch_cuda):
    # tower/decoder.rs:600-614
    k, m, L = 4, 3, 18
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    _, rep = encode_with(k, m, L, data)
    sym = symbols_of(data, rep, k, m, L)
    dec = Decoder(Profile.resolve(TOWER, k, m, L))
    bdec = BatchDecoder(Profile.resolve(TOWER, k, m, L))
    for subset in itertools.combinations(range(k + m), k):
        dec.reset()
        for i in subset:
            dec.push(i, sym[i])
        assert np.array_equal(dec.finalize(), data), subset
        assert np.array_equal(bdec.decode([(i, sym[i]) for i in subset][::-1]), data), subset


def test_recovers_configuration_beyond_gf256_capacity(torch_cuda):
    # tower/decoder.rs:616-633
    k, m, L = 256, 3, 4
    rng = np.random.default_rng(2)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    _, rep = encode_with(k, m, L, data)
    sym = symbols_of(data, rep, k, m, L)
    dec = Decoder(Profile.resolve(TOWER, k, m, L))
    for i in range(k + m):
        if i not in (17, 255, k + 2):
            dec.push(i, sym[i])
        if dec.is_complete():
            break
    assert np.array_equal(dec.finalize(), data)


def test_randomized_larger_erasure_patterns_match_oracle(torch_cuda):
    # tower/decoder.rs:635-654 (LCG shuffles of 0..k+m, first k pushed)
    k, m, L = 64, 32, 66
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    c = oracle.tower_cauchy_matrix(k, m)
    rep = oracle.tower_encode(c, k, m, L, data)
    sym = symbols_of(data, rep, k, m, L)
    dec = Decoder(Profile.resolve(TOWER, k, m, L))
    scratch = dec.scratch()
    state = 0x1234_5678_9ABC_DEF0
    mask = (1 << 64) - 1
    for _ in range(20):
        indices = list(range(k + m))
        for pos in range(len(indices) - 1, 0, -1):
            state = (state * 6364136223846793005 + 1) & mask
            j = state % (pos + 1)
            indices[pos], indices[j] = indices[j], indices[pos]
        dec.reset()
        for i in indices[:k]:
            dec.push(i, sym[i])
        out = np.full(k * L, 0xAB, np.uint8)
        dec.finalize_into_with(out, scratch)
        want = oracle.tower_decode_rref(c, k, m, L, [(i, sym[i]) for i in indices[:k]])
        assert want is not None and np.array_equal(want, data)
        assert np.array_equal(out, data)


def test_push_is_dynamic_bitmap_bookkeeping(torch_cuda):
    # tower/decoder.rs:656-673
    dec = Decoder(Profile.resolve(TOWER, 300, 2, 2))
    assert dec.push(299, bytes([1, 2])) == PushOutcome.advanced(1, 1)
    assert dec.push(299, bytes([3, 4])) == PushOutcome.DEPENDENT
    assert dec.rank() == 1
    assert dec.received() == 2


def test_push_and_finalize_errors(torch_cuda):
    k, m, L = 3, 2, 4
    dec = Decoder(Profile.resolve(TOWER, k, m, L))
    with pytest.raises(DecodeError) as e:
        dec.push(5, bytes(3))
    assert e.value == DecodeError("IndexOutOfRange", index=5, n=5)
    with pytest.raises(DecodeError) as e:
        dec.push(0, bytes(3))
    assert e.value == DecodeError("WrongPayloadLen", expected=4, got=3)
    dec.push(0, bytes(4))
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(12, np.uint8))
    assert e.value == DecodeError("InsufficientRank", rank=1, k=3)
    # InsufficientRank precedes WrongOutputLen
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(11, np.uint8))
    assert e.value.variant == "InsufficientRank"
    for _ in range(4):
        dec.push(0, bytes(4))
    with pytest.raises(DecodeError) as e:
        dec.push(1, bytes(4))
    assert e.value == DecodeError("TooManySymbols", cap=5, received=5)
    dec.reset()
    for i in (0, 1, 4):
        assert dec.push(i, bytes(4)).kind in ("Advanced", "Complete")
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(11, np.uint8))
    assert e.value == DecodeError("WrongOutputLen", expected=12, got=11)
    # A Tower scratch of another geometry (tower/decoder.rs:214-216), after
    # the output length check.
    other = Decoder(Profile.resolve(TOWER, 4, 2, 4)).scratch()
    with pytest.raises(DecodeError) as e:
        dec.finalize_into_with(np.zeros(12, np.uint8), other)
    assert e.value == DecodeError("ScratchMismatch")
    with pytest.raises(DecodeError) as e:
        dec.finalize_into_with(np.zeros(11, np.uint8), other)
    assert e.value.variant == "WrongOutputLen"
    # A recipe cache from a GF(2^8) decoder is the wrong scratch kind (selector.rs:224-240)
    cauchy = Decoder(Profile.resolve(Engine.GOOD_CAUCHY, 3, 2, 4)).scratch()
    with pytest.raises(DecodeError) as e:
        dec.finalize_into_with(np.zeros(11, np.uint8), cauchy)
    assert e.value == DecodeError("ScratchMismatch")


def test_batch_decode_blanket_error_order(torch_cuda):
    # codec.rs:219-240: WrongCount, then push per symbol (IndexOutOfRange ->
    # WrongPayloadLen), Dependent -> DuplicateIndex, then finalize checks.
    k, m, L = 3, 2, 4
    bdec = BatchDecoder(Profile.resolve(TOWER, k, m, L))
    z = bytes(4)
    with pytest.raises(DecodeError) as e:
        bdec.decode([(0, z), (1, z)])
    assert e.value == DecodeError("WrongCount", expected=3, got=2)
    with pytest.raises(DecodeError) as e:
        bdec.decode([(0, z), (9, bytes(2)), (1, z)])
    assert e.value == DecodeError("IndexOutOfRange", index=9, n=5)
    with pytest.raises(DecodeError) as e:
        bdec.decode([(0, z), (1, bytes(2)), (0, z)])
    assert e.value == DecodeError("WrongPayloadLen", expected=4, got=2)
    with pytest.raises(DecodeError) as e:
        bdec.decode([(0, z), (0, z), (1, bytes(2))])
    assert e.value == DecodeError("DuplicateIndex", index=0)
    with pytest.raises(DecodeError) as e:
        bdec.decode_into([(0, z), (1, z), (2, z)], np.zeros(11, np.uint8))
    assert e.value == DecodeError("WrongOutputLen", expected=12, got=11)
    other = BatchDecoder(Profile.resolve(TOWER, 4, 2, 4)).scratch()
    with pytest.raises(DecodeError) as e:
        bdec.decode_into_with([(0, z), (1, z), (2, z)], np.zeros(12, np.uint8), other)
    assert e.value == DecodeError("ScratchMismatch")


def test_streaming_geometry_200_50_8192(torch_cuda):
    """Incremental feed of all k symbols, streaming push of k survivors with
    exactly 50 data symbols erased, finalize_into_with.  Full batch of 256
    codewords on the device (round trip), two codewords against the oracle."""
    torch = torch_cuda
    k, m, L, count = 200, 50, 8192, 256
    p = Profile.resolve(TOWER, k, m, L)
    enc = IncrementalEncoder(p)
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    from scrs_amd.codec import fill_random_dev
    fill_random_dev(data, data.numel(), 0x70E3_0200)
    rep = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(data, rep, count)
    bdec = BatchDecoder(p)
    rng = np.random.default_rng(0x70E3)
    idx = np.zeros((count, k), np.uint16)
    for c in range(count):
        erased = set(rng.choice(k, size=m, replace=False).tolist())
        idx[c] = [i for i in range(k + m) if i not in erased]
    out = torch.empty_like(data)
    bdec.decode_batch_dev(idx, data, rep, out, count)
    torch.cuda.synchronize()
    assert torch.equal(out, data)
    cmat = oracle.tower_cauchy_matrix(k, m)
    dec = Decoder(p)
    scratch = dec.scratch()
    for c in (0, count - 1):
        d = data[c * k * L:(c + 1) * k * L].cpu().numpy()
        r = rep[c * m * L:(c + 1) * m * L].cpu().numpy()
        assert np.array_equal(r, oracle.tower_encode(cmat, k, m, L, d))
        # the incremental path on the host gives the same repairs
        enc.reset()
        for i in rng.permutation(k).tolist():
            enc.feed(i, d[i * L:(i + 1) * L])
        assert enc.repair(7) == r[7 * L:8 * L].tobytes()
        sym = symbols_of(d, r, k, m, L)
        dec.reset()
        order = rng.permutation(idx[c]).tolist()
        for i in order:
            dec.push(int(i), sym[i])
        assert dec.is_complete()
        got = np.zeros(k * L, np.uint8)
        dec.finalize_into_with(got, scratch)
        assert np.array_equal(got, d)


@pytest.mark.parametrize("k,m,L,count", [
    (48, 24, 1024, 3),   # 16-term unrolled chunks, a 16-row and an 8-row group
    (32, 5, 4096, 2),    # 16-term chunks, one 8-row pass
    (40, 19, 2048, 2),   # 8-term chunks, a 16-row and a 4-row group
    (24, 3, 512, 4),     # 8-term chunks, one 4-row pass
    (36, 6, 512, 2),     # neither: the rolled loop
])
def test_batch_round_trip_term_chunks(torch_cuda, k, m, L, count):
    """Device batch encode and decode at term counts that take each chunking
    of the multiply-XOR (multiples of 16, of 8, neither), against the oracle
    for every codeword."""
    torch = torch_cuda
    from scrs_amd.codec import fill_random_dev
    p = Profile.resolve(TOWER, k, m, L)
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), k * 1000 + m)
    rep = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    IncrementalEncoder(p).encode_batch_dev(data, rep, count)
    rng = np.random.default_rng(k + m + L)
    idx = np.zeros((count, k), np.uint16)
    for c in range(count):
        r = int(rng.integers(1, m + 1))
        erased = set(rng.choice(k, size=r, replace=False).tolist())
        idx[c] = rng.permutation([i for i in range(k + m) if i not in erased])[:k]
    out = torch.full_like(data, 0xAB)
    BatchDecoder(p).decode_batch_dev(idx, data, rep, out, count)
    torch.cuda.synchronize()
    cmat = oracle.tower_cauchy_matrix(k, m)
    d_all, r_all = data.cpu().numpy(), rep.cpu().numpy()
    for c in range(count):
        d = d_all[c * k * L:(c + 1) * k * L]
        assert np.array_equal(r_all[c * m * L:(c + 1) * m * L], oracle.tower_encode(cmat, k, m, L, d)), c
    assert torch.equal(out, data)


@pytest.mark.parametrize("k,m,L", [(150, 30, 8192), (64, 9, 4100), (200, 50, 1024)])
@pytest.mark.parametrize("split", [False, True])
def test_lone_codeword_term_split(torch_cuda, k, m, L, split, term_split):
    """One Tower codeword, with and without the term split of the encode and
    the finalize (partial rows XORed into cleared destinations)."""
    torch = torch_cuda
    term_split(split)
    from scrs_amd.codec import fill_random_dev
    p = Profile.resolve(TOWER, k, m, L)
    data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), k * 7 + m)
    rep = torch.full((m * L,), 0xAB, dtype=torch.uint8, device="cuda")
    IncrementalEncoder(p).encode_batch_dev(data, rep, 1)
    torch.cuda.synchronize()
    d = data.cpu().numpy()
    assert np.array_equal(rep.cpu().numpy(), oracle.tower_encode(oracle.tower_cauchy_matrix(k, m), k, m, L, d))
    rng = np.random.default_rng(k)
    erased = set(rng.choice(k, size=m, replace=False).tolist())
    idx = np.array([rng.permutation([i for i in range(k + m) if i not in erased])[:k]], np.uint16)
    out = torch.full_like(data, 0xCD)
    BatchDecoder(p).decode_batch_dev(idx, data, rep, out, 1)
    torch.cuda.synchronize()
    assert torch.equal(out, data)
