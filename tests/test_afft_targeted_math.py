"""The additive-FFT targeted decode's coefficients, on the CPU.

The device computes them as the closed-form Cauchy inverse scaled by per-index
weights (w(z) = 1 / P'(z) for data, w(x) = 1 / P(x) for repairs, P the product
of (t + d) over the data points d < k).  Here that closed form is checked
against the reference's own construction (afft/decoder.rs:360-441, 600-621):
generator rows P(x) / ((x + z) P'(z)) from brute-force products, Gauss-Jordan
inversion of the missing-column block, and residual application; and the
resulting coefficients must recover the data of an oracle-encoded codeword.
"""
import itertools

import numpy as np
import pytest

import oracle

mul, inv = oracle.gf16_mul, oracle.gf16_inv


def prod(vals):
    p = 1
    for v in vals:
        p = mul(p, v)
    return p


def div(a, b):
    return mul(a, inv(b))


def gauss_jordan_inverse(a):
    n = len(a)
    aug = [list(row) + [1 if i == j else 0 for j in range(n)] for i, row in enumerate(a)]
    for col in range(n):
        piv = next(r for r in range(col, n) if aug[r][col])
        aug[col], aug[piv] = aug[piv], aug[col]
        s = inv(aug[col][col])
        aug[col] = [mul(s, v) for v in aug[col]]
        for r in range(n):
            if r != col and aug[r][col]:
                f = aug[r][col]
                aug[r] = [v ^ mul(f, p) for v, p in zip(aug[r], aug[col])]
    return [row[n:] for row in aug]


def reference_coefficients(k, received, missing):
    """fused[j][t] by the reference's generator + Gauss-Jordan route."""
    P = lambda x: prod(x ^ d for d in range(k))
    dP = lambda z: prod(z ^ d for d in range(k) if d != z)
    repairs = sorted(i for i in received if i >= k)
    G = {(x, z): div(P(x), mul(x ^ z, dP(z))) for x in repairs for z in range(k)}
    inverse = gauss_jordan_inverse([[G[(x, e)] for e in missing] for x in repairs])
    fused = []
    for j in range(len(missing)):
        row = []
        for t in received:
            if t >= k:
                row.append(inverse[j][repairs.index(t)])
            else:
                v = 0
                for i, x in enumerate(repairs):
                    v ^= mul(inverse[j][i], G[(x, t)])
                row.append(v)
        fused.append(row)
    return fused


def weighted_closed_form(k, received, missing):
    """fused[j][t] as the device computes it: the Cauchy closed form with
    coordinates equal to the indices, times w(t) / w(missing j)."""
    rows = sorted(i for i in received if i >= k)
    cols = list(missing)
    w = lambda i: inv(prod(i ^ d for d in range(k))) if i >= k else inv(prod(i ^ d for d in range(k) if d != i))
    R = {x: div(prod(x ^ e for e in cols), prod(x ^ y for y in rows if y != x)) for x in rows}
    C = {e: div(prod(e ^ x for x in rows), prod(e ^ f for f in cols if f != e)) for e in cols}
    Q = lambda z: div(prod(z ^ e for e in cols), prod(z ^ x for x in rows))
    fused = []
    for e in cols:
        row = []
        for t in received:
            base = mul(R[t], C[e]) if t >= k else mul(C[e], Q(t))
            v = div(base, t ^ e)
            row.append(div(mul(v, w(t)), w(e)))
        fused.append(row)
    return fused


@pytest.mark.parametrize("k,m", [(4, 4), (6, 5), (8, 3), (5, 7)])
def test_weighted_closed_form_equals_reference_construction(k, m):
    rng = np.random.default_rng(k * 10 + m)
    for r in range(1, min(k, m, 5) + 1):
        for _ in range(3):
            missing = sorted(rng.choice(k, size=r, replace=False).tolist())
            reps = sorted((k + rng.choice(m, size=r, replace=False)).tolist())
            received = [int(i) for i in rng.permutation([i for i in range(k) if i not in missing] + reps)]
            assert weighted_closed_form(k, received, missing) == reference_coefficients(k, received, missing)


def test_coefficients_recover_oracle_codeword():
    k, m, L = 6, 5, 8
    data = np.random.default_rng(9).integers(0, 256, k * L, dtype=np.uint8)
    rep = oracle.afft_encode(k, m, L, data)
    sym = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    as16 = lambda b: b.view("<u2").astype(int)
    for missing in itertools.combinations(range(k), 2):
        for reps in ((k, k + 3), (k + 1, k + 4)):
            received = [i for i in range(k) if i not in missing] + list(reps)
            fused = weighted_closed_form(k, received, list(missing))
            for j, e in enumerate(missing):
                out = np.zeros(L // 2, int)
                for t, i in enumerate(received):
                    out ^= np.array([mul(fused[j][t], int(v)) for v in as16(sym[i])])
                assert np.array_equal(out, as16(sym[e])), (missing, reps, e)
