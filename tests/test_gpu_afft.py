"""GF(2^16) additive-FFT engine parity: the HIP library against the oracle.

The engine is Reed-Solomon at points = wire index; the oracle evaluates the
same polynomial by Lagrange interpolation.  Shapes follow the reference's
own AFFT tests (afft/encoder.rs:286-416, afft/decoder.rs:846-975,
afft/differential.rs) plus multi-pass transform sizes up to the full
65536-point domain.
"""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from scrs_amd.codec import (BatchDecoder, BatchEncoder, ConfigError, DecodeError, Decoder, EncodeError, Engine,
                            IncrementalEncoder, Profile, PushOutcome, fill_random_dev)

pytestmark = pytest.mark.gpu

AFFT = Engine.AFFT
GOLDEN = Path(__file__).resolve().parent / "golden" / "afft_differential_regressions.json"


def data_block(k, L):
    return np.array([((i * 37) ^ (i << 3) ^ 0x5A) & 0xFF for i in range(k * L)], np.uint8)


def symbols_of(data, rep, k, m, L):
    return [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]


@pytest.mark.parametrize("k,m,L", [
    # fused path: k a power of two, m <= k (afft/encoder.rs:318-345 shapes)
    (2, 1, 2), (4, 2, 64), (16, 4, 40), (256, 64, 48), (512, 128, 40), (4, 4, 10), (64, 64, 6),
    # interpolate-then-evaluate shapes (afft/encoder.rs:290-316), m > k, tiny
    (5, 3, 64), (100, 20, 64), (17, 7, 130), (3, 2, 2), (1, 1, 2), (1, 5, 4), (4, 12, 6), (257, 128, 8),
    # multi-pass general transforms (2^9 .. 2^11 points)
    (300, 200, 18), (1000, 24, 34), (1500, 548, 66), (129, 1, 4098),
])
def test_encode_matches_oracle(torch_cuda, k, m, L):
    rng = np.random.default_rng(k * 7 + m * 3 + L)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    enc = BatchEncoder(Profile.resolve(AFFT, k, m, L))
    got = np.full(m * L, 0xAB, np.uint8)
    enc.encode_into(data, got)
    assert np.array_equal(got, oracle.afft_encode(k, m, L, data))


def test_encode_fused_multi_pass_and_batch_dev(torch_cuda):
    # k = 2^12 (two passes per transform), 3 codewords in one launch chain
    torch = torch_cuda
    k, m, L, count = 4096, 1000, 64, 3
    enc = BatchEncoder(Profile.resolve(AFFT, k, m, L))
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 0xAF77)
    rep = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    enc.encode_batch_dev(data, rep, count)
    torch.cuda.synchronize()
    h = data.cpu().numpy()
    r = rep.cpu().numpy()
    for c in (0, count - 1):
        d = h[c * k * L:(c + 1) * k * L]
        vals = [d[i * L:(i + 1) * L] for i in range(k)]
        outs = [k, k + 1, k + 517, k + m - 1]
        want = oracle.rs16_interp(k, L, range(k), vals, outs)
        got = np.concatenate([r[c * m * L + (x - k) * L:c * m * L + (x - k + 1) * L] for x in outs])
        assert np.array_equal(got, want)


def test_encode_general_multi_pass_matches_oracle(torch_cuda):
    # 2^15-point erasure transform for a non-power-of-two k
    k, m, L = 20000, 10000, 2
    rng = np.random.default_rng(20000)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    enc = BatchEncoder(Profile.resolve(AFFT, k, m, L))
    got = enc.encode(data)
    vals = [data[i * L:(i + 1) * L] for i in range(k)]
    outs = [k, k + 1, k + 4999, k + m - 1]
    want = oracle.rs16_interp(k, L, range(k), vals, outs)
    assert np.array_equal(np.concatenate([got[(x - k) * L:(x - k + 1) * L] for x in outs]), want)


@pytest.mark.parametrize("k,m,L", [(4, 4, 10), (5, 3, 6), (3, 2, 4), (2, 6, 2), (6, 5, 34)])
def test_decode_every_subset(torch_cuda, k, m, L):
    # afft/decoder.rs:846-875 (exhaustive power-of-two and truncated recovery)
    data = data_block(k, L)
    rep = oracle.afft_encode(k, m, L, data)
    sym = symbols_of(data, rep, k, m, L)
    p = Profile.resolve(AFFT, k, m, L)
    bdec = BatchDecoder(p)
    dec = Decoder(p)
    scratch = dec.scratch()
    for sub in itertools.combinations(range(k + m), k):
        out = np.full(k * L, 0xAB, np.uint8)
        bdec.decode_into([(i, sym[i]) for i in sub[::-1]], out)
        assert np.array_equal(out, data), sub
        dec.reset()
        for i in sub:
            dec.push(i, sym[i])
        out2 = np.zeros(k * L, np.uint8)
        dec.finalize_into_with(out2, scratch)
        assert np.array_equal(out2, data), sub


def test_randomized_larger_patterns(torch_cuda):
    # afft/decoder.rs:877-896: k=32, m=16, symbol_len=34, LCG shuffles
    k, m, L = 32, 16, 34
    data = data_block(k, L)
    enc = BatchEncoder(Profile.resolve(AFFT, k, m, L))
    rep = enc.encode(data)
    sym = symbols_of(data, rep, k, m, L)
    dec = Decoder(Profile.resolve(AFFT, k, m, L))
    state = 0x1234_5678_9ABC_DEF0
    mask = (1 << 64) - 1
    for _ in range(20):
        idx = list(range(k + m))
        for pos in range(len(idx) - 1, 0, -1):
            state = (state * 6364136223846793005 + 1) & mask
            j = state % (pos + 1)
            idx[pos], idx[j] = idx[j], idx[pos]
        dec.reset()
        for i in idx[:k]:
            dec.push(i, sym[i])
        assert np.array_equal(dec.finalize(), data)


@pytest.mark.parametrize("k,m,L,drop", [(32, 16, 34, 12), (16, 8, 64, 3)])
def test_rotating_erasure_sets(torch_cuda, k, m, L, drop):
    # afft/decoder.rs:915-975: a fixed-size erasure set rotating over rounds
    data = data_block(k, L)
    rep = oracle.afft_encode(k, m, L, data)
    sym = symbols_of(data, rep, k, m, L)
    dec = Decoder(Profile.resolve(AFFT, k, m, L))
    scratch = dec.scratch()
    out = np.zeros(k * L, np.uint8)
    for rnd in range(6):
        dec.reset()
        received = []
        for off in range(k + m):
            i = (off + rnd) % (k + m)
            if i < drop or len(received) == k:
                continue
            received.append(i)
        for i in range(k + m):
            if len(received) == k:
                break
            if i not in received:
                received.append(i)
        for i in received:
            dec.push(i, sym[i])
        dec.finalize_into_with(out, scratch)
        assert np.array_equal(out, data)


@pytest.mark.parametrize("k,m,L", [(100, 28, 8), (200, 56, 18), (700, 300, 34), (3000, 1000, 10),
                                   (30000, 30000, 4)])
def test_decode_multi_pass_matches_oracle(torch_cuda, k, m, L):
    torch = torch_cuda
    rng = np.random.default_rng(k + m)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    p = Profile.resolve(AFFT, k, m, L)
    rep = BatchEncoder(p).encode(data)
    if k * m <= 300_000:
        assert np.array_equal(rep, oracle.afft_encode(k, m, L, data))
    # erase a random third of the data, fill from repairs
    erased = set(rng.choice(k, size=min(m, k // 3), replace=False).tolist())
    idx = np.array([i for i in range(k + m) if i not in erased][:k], np.uint16)
    d = torch.from_numpy(data).cuda()
    r = torch.from_numpy(rep).cuda()
    out = torch.full((k * L,), 0xAB, dtype=torch.uint8, device="cuda")
    BatchDecoder(p).decode_batch_dev(idx, d, r, out, 1)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), data)


def test_decode_batch_dev_and_ptrs(torch_cuda):
    torch = torch_cuda
    k, m, L, count = 64, 40, 100, 4
    p = Profile.resolve(AFFT, k, m, L)
    data = torch.empty(count * k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 0xD0D0)
    rep = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    BatchEncoder(p).encode_batch_dev(data, rep, count)
    rng = np.random.default_rng(4)
    idx = np.zeros((count, k), np.uint16)
    for c in range(count):
        idx[c] = np.sort(rng.choice(k + m, size=k, replace=False))
    dec = BatchDecoder(p)
    out = torch.zeros_like(data)
    dec.decode_batch_dev(idx, data, rep, out, count)
    torch.cuda.synchronize()
    assert torch.equal(out, data)
    ptrs = np.zeros(count * k, np.uint64)
    for c in range(count):
        for t in range(k):
            i = int(idx[c, t])
            ptrs[c * k + t] = (data.data_ptr() + (c * k + i) * L if i < k
                               else rep.data_ptr() + (c * m + i - k) * L)
    pt = torch.from_numpy(ptrs.view(np.int64)).cuda()
    out2 = torch.zeros_like(data)
    dec.decode_batch_ptrs_dev(idx, pt, out2, count)
    torch.cuda.synchronize()
    assert torch.equal(out2, data)


@pytest.mark.parametrize("shift", [2, 4])
def test_batch_dev_misaligned_buffers(torch_cuda, shift):
    """symbol_len a multiple of 4 with buffers 2 bytes off (byte path) and 4
    bytes off (dword buffer path): the same bytes as the aligned call, and
    the repairs against the oracle."""
    torch = torch_cuda
    k, m, L, count = 24, 12, 64, 3
    p = Profile.resolve(AFFT, k, m, L)
    base = torch.empty(count * k * L + 16, dtype=torch.uint8, device="cuda")
    fill_random_dev(base, base.numel(), 0xA11 + shift)
    data = base[shift:shift + count * k * L]
    rep_buf = torch.full((count * m * L + 16,), 0xAB, dtype=torch.uint8, device="cuda")
    rep = rep_buf[shift:shift + count * m * L]
    BatchEncoder(p).encode_batch_dev(data, rep, count)
    aligned = data.clone()
    rep_al = torch.empty(count * m * L, dtype=torch.uint8, device="cuda")
    BatchEncoder(p).encode_batch_dev(aligned, rep_al, count)
    torch.cuda.synchronize()
    assert torch.equal(rep, rep_al)
    d0 = aligned[:k * L].cpu().numpy()
    want = oracle.afft_encode(k, m, L, d0)
    assert np.array_equal(rep_al[:m * L].cpu().numpy(), want)
    rng = np.random.default_rng(shift)
    idx = np.zeros((count, k), np.uint16)
    for c in range(count):
        idx[c] = np.sort(rng.choice(k + m, size=k, replace=False))
    out_buf = torch.zeros(count * k * L + 16, dtype=torch.uint8, device="cuda")
    out = out_buf[shift:shift + count * k * L]
    BatchDecoder(p).decode_batch_dev(idx, data, rep, out, count)
    torch.cuda.synchronize()
    assert torch.equal(out, aligned)


def test_full_domain_32768_32768_2048(torch_cuda):
    """k = m = 32768, symbol_len = 2048 (128 MiB codeword): fused encode,
    decode from a seeded random half of the 65536 indices, round trip, and
    repair spot checks against the oracle's subgroup Lagrange form."""
    torch = torch_cuda
    k, m, L = 32768, 32768, 2048
    p = Profile.resolve(AFFT, k, m, L)
    data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 0xAF4)
    rep = torch.empty(m * L, dtype=torch.uint8, device="cuda")
    BatchEncoder(p).encode_batch_dev(data, rep, 1)
    rng = np.random.default_rng(0xAF4)
    idx = np.sort(rng.choice(k + m, size=k, replace=False)).astype(np.uint16)
    out = torch.empty_like(data)
    BatchDecoder(p).decode_batch_dev(idx, data, rep, out, 1)
    torch.cuda.synchronize()
    assert torch.equal(out, data)
    # Spot check 3 repairs over the first 64 bytes of every data symbol.
    cols = 64
    h = data.view(k, L)[:, :cols].cpu().numpy()
    vals = [h[i] for i in range(k)]
    outs = [k, k + 12345, k + m - 1]
    want = oracle.rs16_interp(k, cols, range(k), vals, outs).reshape(len(outs), cols)
    got = rep.view(m, L)[[x - k for x in outs], :cols].cpu().numpy()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("j", range(4, 16))
def test_power_of_two_domains_round_trip(torch_cuda, afft_targeted, j):
    """k = 2^j data symbols with m = k and m = k/2 repairs: every way the
    transforms split into passes from 2^5 to 2^16 points, the fused encode,
    and the erasure decode, whose forward over the top bits needs only the
    rows below k when k is half the domain.  Encodes are checked against the
    oracle up to 2^10 symbols, decodes against the data: half of all indices
    erased at random, and (m = k) every data symbol erased."""
    torch = torch_cuda
    afft_targeted(0)  # every decode through the transforms
    k = 1 << j
    for m, L in ((k, 64), (k // 2, 34)):
        p = Profile.resolve(AFFT, k, m, L)
        data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
        fill_random_dev(data, data.numel(), 0x5EED + j * 3 + m)
        rep = torch.empty(m * L, dtype=torch.uint8, device="cuda")
        BatchEncoder(p).encode_batch_dev(data, rep, 1)
        torch.cuda.synchronize()
        if k <= 1024:
            host = data.cpu().numpy()
            assert np.array_equal(rep.cpu().numpy(), oracle.afft_encode(k, m, L, host)), (k, m, L)
        dec = BatchDecoder(p)
        rng = np.random.default_rng(j * 101 + m)
        patterns = [np.sort(rng.choice(k + m, size=k, replace=False))]
        if m == k:
            patterns.append(np.arange(k, k + m))
        for idx in patterns:
            out = torch.full_like(data, 0xA5)
            dec.decode_batch_dev(idx.astype(np.uint16)[None, :], data, rep, out, 1)
            torch.cuda.synchronize()
            assert torch.equal(out, data), (k, m, L, int(idx[0]))


def erase_data(rng, k, m, r):
    """Received indices with exactly r data symbols missing: the first r
    received repairs replace them (every other repair is absent)."""
    missing = set(rng.choice(k, size=r, replace=False).tolist())
    repairs = sorted(rng.choice(m, size=r, replace=False).tolist())
    idx = [i for i in range(k) if i not in missing] + [k + j for j in repairs]
    return np.array(rng.permutation(idx), np.uint16)


@pytest.mark.parametrize("k,m,L", [(32, 16, 34), (256, 64, 48), (1000, 24, 34), (1500, 548, 66),
                                   (4096, 4096, 64), (129, 70, 4098)])
def test_targeted_and_transform_paths_agree(torch_cuda, afft_targeted, k, m, L):
    """Decodes of r missing data symbols through the targeted path (weighted
    closed-form solve + multiply-XOR) and through the transforms give the
    original data, for r up to the targeted path's cap of 64, with the
    layout and the pointer-array forms (afft/decoder.rs:916-975 covers the
    two paths at low and high r)."""
    torch = torch_cuda
    p = Profile.resolve(AFFT, k, m, L)
    data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), k * 31 + m)
    rep = torch.empty(m * L, dtype=torch.uint8, device="cuda")
    BatchEncoder(p).encode_batch_dev(data, rep, 1)
    dec = BatchDecoder(p)
    rng = np.random.default_rng(k + 7 * m)
    for r in (1, 2, 3, 5, 6, 16, 17, 33, 64):
        if r > min(k, m):
            continue
        idx = erase_data(rng, k, m, r)[None, :]
        for limit in (0, 64):
            afft_targeted(limit)
            out = torch.full_like(data, 0xAB)
            dec.decode_batch_dev(idx, data, rep, out, 1)
            torch.cuda.synchronize()
            assert torch.equal(out, data), (r, limit)
        ptrs = np.array([data.data_ptr() + int(i) * L if i < k else rep.data_ptr() + (int(i) - k) * L
                         for i in idx[0]], np.uint64)
        pt = torch.from_numpy(ptrs.view(np.int64)).cuda()
        out = torch.full_like(data, 0xCD)
        dec.decode_batch_ptrs_dev(idx, pt, out, 1)
        torch.cuda.synchronize()
        assert torch.equal(out, data), (r, "ptrs")


def test_targeted_path_full_domain(torch_cuda, afft_targeted):
    """k = m = 32768, symbol_len = 2048: 1, 5 and 64 missing data symbols
    through the targeted path (the term split over 32768 terms), and 5
    through the transforms, all equal to the original data; the streaming
    decoder takes the same path."""
    torch = torch_cuda
    k, m, L = 32768, 32768, 2048
    p = Profile.resolve(AFFT, k, m, L)
    data = torch.empty(k * L, dtype=torch.uint8, device="cuda")
    fill_random_dev(data, data.numel(), 0x7A6)
    rep = torch.empty(m * L, dtype=torch.uint8, device="cuda")
    BatchEncoder(p).encode_batch_dev(data, rep, 1)
    dec = BatchDecoder(p)
    rng = np.random.default_rng(0x7A6)
    afft_targeted(64)
    for r, limit in ((1, 64), (5, 64), (64, 64), (5, 0)):
        afft_targeted(limit)
        idx = erase_data(rng, k, m, r)[None, :]
        out = torch.full_like(data, 0xAB)
        dec.decode_batch_dev(idx, data, rep, out, 1)
        torch.cuda.synchronize()
        assert torch.equal(out, data), (r, limit)


def test_targeted_threshold_setting(torch_cuda, afft_targeted):
    from scrs_amd.codec import afft_set_targeted_max
    afft_targeted(5)
    assert afft_set_targeted_max(1000) == 5
    assert afft_set_targeted_max(-3) == 64
    assert afft_set_targeted_max(5) == 0


def test_engine_selection_and_errors(torch_cuda):
    with pytest.raises(ConfigError) as e:
        IncrementalEncoder(Profile.resolve(AFFT, 4, 2, 8))
    assert e.value == ConfigError("UnsupportedMode", engine=AFFT)
    with pytest.raises(ConfigError) as e:
        Profile.resolve(AFFT, 65535, 2, 2)
    assert e.value == ConfigError("TooManySymbols", cap=65536)
    Profile.resolve(AFFT, 32769, 1, 2)
    enc = BatchEncoder(Profile.resolve(AFFT, 3, 2, 4))
    with pytest.raises(EncodeError) as e:
        enc.encode_into(np.zeros(11, np.uint8), np.zeros(7, np.uint8))
    assert e.value == EncodeError("WrongInputLen", expected=12, got=11)
    with pytest.raises(EncodeError) as e:
        enc.encode_into(np.zeros(12, np.uint8), np.zeros(7, np.uint8))
    assert e.value == EncodeError("WrongOutputLen", expected=8, got=7)
    # an AFFT scratch of another geometry is a resizable workspace: accepted
    other = BatchEncoder(Profile.resolve(AFFT, 9, 9, 8)).scratch()
    d = np.arange(12, dtype=np.uint8)
    out = np.zeros(8, np.uint8)
    enc.encode_into_with(d, out, other)
    assert np.array_equal(out, oracle.afft_encode(3, 2, 4, d))
    gf = BatchEncoder(Profile.resolve(Engine.GOOD_CAUCHY, 3, 2, 4)).scratch()
    with pytest.raises(EncodeError) as e:
        enc.encode_into_with(d, out, gf)
    assert e.value == EncodeError("ScratchMismatch")


def test_decoder_push_and_finalize_errors(torch_cuda):
    # afft/decoder.rs:464-500, 546-568, 233-254
    dec = Decoder(Profile.resolve(AFFT, 300, 2, 2))
    assert dec.push(299, bytes([1, 2])) == PushOutcome.advanced(1, 1)
    assert dec.push(299, bytes([3, 4])) == PushOutcome.DEPENDENT
    dec = Decoder(Profile.resolve(AFFT, 3, 2, 4))
    with pytest.raises(DecodeError) as e:
        dec.push(5, bytes(4))
    assert e.value == DecodeError("IndexOutOfRange", index=5, n=5)
    with pytest.raises(DecodeError) as e:
        dec.push(0, bytes(3))
    assert e.value == DecodeError("WrongPayloadLen", expected=4, got=3)
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(12, np.uint8))
    assert e.value == DecodeError("InsufficientRank", rank=0, k=3)
    for i in (0, 3, 4):
        dec.push(i, bytes(4))
    with pytest.raises(DecodeError) as e:
        dec.finalize_into(np.zeros(10, np.uint8))
    assert e.value == DecodeError("WrongOutputLen", expected=12, got=10)
    other = Decoder(Profile.resolve(AFFT, 4, 2, 4)).scratch()
    with pytest.raises(DecodeError) as e:
        dec.finalize_into_with(np.zeros(12, np.uint8), other)
    assert e.value == DecodeError("ScratchMismatch")
    tower = Decoder(Profile.resolve(Engine.TOWER, 3, 2, 4)).scratch()
    with pytest.raises(DecodeError) as e:
        dec.finalize_into_with(np.zeros(12, np.uint8), tower)
    assert e.value == DecodeError("ScratchMismatch")
    out = np.full(12, 0xAB, np.uint8)
    dec.finalize_into_with(out, dec.scratch())
    assert np.array_equal(out, np.zeros(12, np.uint8))
    bdec = BatchDecoder(Profile.resolve(AFFT, 3, 2, 4))
    with pytest.raises(DecodeError) as e:
        bdec.decode([(0, bytes(4)), (0, bytes(4)), (1, bytes(4))])
    assert e.value == DecodeError("DuplicateIndex", index=0)


@pytest.mark.parametrize("case", [0, 1])
def test_proptest_regressions(torch_cuda, case):
    # proptest-regressions/afft/differential.txt:7-8 through both GF(2^16) engines
    c = json.loads(GOLDEN.read_text())["cases"][case]
    k, m, L = c["k"], c["m"], 2 * c["field_elements"]
    data = np.array(c["data_source"][:k * L], np.uint8)
    idx = sorted(range(k + m), key=lambda i: (c["ordering"][i], i))[:k]
    rep = BatchEncoder(Profile.resolve(AFFT, k, m, L)).encode(data)
    sym = symbols_of(data, rep, k, m, L)
    dec = Decoder(Profile.resolve(AFFT, k, m, L))
    for i in idx:
        dec.push(i, sym[i])
    assert np.array_equal(dec.finalize(), data)
    tenc = IncrementalEncoder(Profile.resolve(Engine.TOWER, k, m, L))
    for i in range(k):
        tenc.feed(i, data[i * L:(i + 1) * L])
    tsym = symbols_of(data, np.frombuffer(b"".join(tenc.repair(j) for j in range(m)), np.uint8), k, m, L)
    tdec = Decoder(Profile.resolve(Engine.TOWER, k, m, L))
    for i in idx:
        tdec.push(i, tsym[i])
    assert np.array_equal(tdec.finalize(), data)
