"""The additive-FFT transform plan on its own (TransformPlan,
afft/transform.rs:24-95) through the C ABI.

Forward is checked against direct evaluation in the novel basis at every
point (the reference's own check, afft/transform.rs:565-615) for sizes up to
4096 and at sampled points above, for every power of two from 1 to 65536;
inverse against forward; the plan against the systematic encoder, whose
repairs are the forward transform of the data's interpolation; and the
device form's columns against the element-slice form.
"""
import numpy as np
import pytest

import oracle
from scrs_amd import Engine, Profile, TransformLengthError, TransformPlan, batch_encoder

pytestmark = pytest.mark.gpu

SIZES = [1 << d for d in range(17)]


def coefficients(n, seed):
    return np.random.default_rng(seed).integers(0, 65536, n, dtype=np.uint16)


@pytest.mark.parametrize("size", SIZES)
def test_forward_matches_direct_novel_basis_evaluation(torch_cuda, size):
    plan = TransformPlan(size)
    assert plan.size == size and plan.log_size == size.bit_length() - 1
    coef = coefficients(size, size)
    got = coef.copy()
    plan.forward(got)
    if size <= 4096:
        points = np.arange(size, dtype=np.uint16)
    else:
        rng = np.random.default_rng(size + 1)
        points = np.unique(np.concatenate([[0, 1, size - 1], rng.integers(0, size, 253)])).astype(np.uint16)
    assert np.array_equal(got[points.astype(np.int64)], oracle.afft_novel_eval(coef, points))


def test_reference_vector_pattern(torch_cuda):
    # afft/transform.rs:579-590: coefficients index * 0x2917 + 0x1357.
    for size in [1, 2, 4, 8, 16, 32]:
        coef = ((np.arange(size, dtype=np.uint32) * 0x2917 + 0x1357) & 0xFFFF).astype(np.uint16)
        got = coef.copy()
        TransformPlan(size).forward(got)
        assert np.array_equal(got, oracle.afft_novel_eval(coef, np.arange(size, dtype=np.uint16)))


@pytest.mark.parametrize("size", SIZES)
def test_forward_inverse_round_trips(torch_cuda, size):
    plan = TransformPlan(size)
    x = coefficients(size, 7 * size + 3)
    y = x.copy()
    plan.forward(y)
    plan.inverse(y)
    assert np.array_equal(y, x)
    plan.inverse(y)
    plan.forward(y)
    assert np.array_equal(y, x)


def test_sizes_and_lengths_are_validated(torch_cuda):
    # afft/transform.rs:568-578
    assert TransformPlan.new(0) is None
    assert TransformPlan.new(3) is None
    assert TransformPlan.new(65536) is not None
    assert TransformPlan.new(131072) is None
    plan = TransformPlan(8)
    with pytest.raises(TransformLengthError) as e:
        plan.forward(np.zeros(4, np.uint16))
    assert e.value.fields == {"expected": 8, "got": 4}
    with pytest.raises(TransformLengthError):
        plan.inverse(np.zeros(9, np.uint16))


@pytest.mark.parametrize("k", [2, 64, 4096, 32768])
def test_plan_reproduces_the_systematic_encoder(torch_cuda, k):
    """Repairs of the fused encode (k a power of two, m = k) are the forward
    transform over 2k points of the inverse transform of the data over k
    points, at the points k..2k-1."""
    data = coefficients(k, k + 11)
    coef = data.copy()
    TransformPlan(k).inverse(coef)
    full = np.concatenate([coef, np.zeros(k, np.uint16)])
    TransformPlan(2 * k).forward(full)
    assert np.array_equal(full[:k], data)
    rep = batch_encoder(Profile.resolve(Engine.AFFT, k, k, 2)).encode(data.view(np.uint8))
    assert np.array_equal(full[k:], rep.view(np.uint16))


@pytest.mark.parametrize("size,symbol_len,count,shift", [(8, 34, 3, 0), (1024, 64, 2, 0), (65536, 6, 1, 0),
                                                         (256, 40, 4, 2), (4096, 12, 5, 0)])
def test_device_form_transforms_every_column(torch_cuda, size, symbol_len, count, shift):
    torch = torch_cuda
    plan = TransformPlan(size)
    rng = np.random.default_rng(size + symbol_len)
    host = rng.integers(0, 256, count * size * symbol_len, dtype=np.uint8)
    buf = torch.zeros(host.size + shift, dtype=torch.uint8, device="cuda")
    buf[shift:] = torch.from_numpy(host).cuda()
    plan.forward_dev(buf[shift:], symbol_len, count)
    torch.cuda.synchronize()
    got = buf[shift:].cpu().numpy().reshape(count, size, symbol_len // 2, 2)
    src = host.reshape(count, size, symbol_len // 2, 2)
    for c in range(count):
        for col in sorted({0, symbol_len // 2 - 1, int(rng.integers(0, symbol_len // 2))}):
            v = np.ascontiguousarray(src[c, :, col, :]).view(np.uint16).reshape(size)
            w = v.copy()
            plan.forward(w)
            assert np.array_equal(np.ascontiguousarray(got[c, :, col, :]).view(np.uint16).reshape(size), w)
    plan.inverse_dev(buf[shift:], symbol_len, count)
    torch.cuda.synchronize()
    assert np.array_equal(buf[shift:].cpu().numpy(), host)
