"""Pins of the oracle's additive-FFT codec (Reed-Solomon at points = wire index).

The oracle evaluates the degree < k polynomial through the data points by
Lagrange interpolation.  These tests pin it independently:

* the reference encoder's scalar check (afft/encoder.rs:347-382), recomputed
  here element by element with the KAT-pinned field multiply;
* an expansion in the reference's novel basis X_i(x) = prod over the set bits
  b of i of W_b(x) / W_b(2^b) (afft/transform.rs:229-262, 565-615), solved as
  a linear system at the data points and evaluated at the repair points -
  a different basis and algorithm that must give the same polynomial;
* decode(encode(x)) == x over every k-subset, and the reference's proptest
  regression seeds (proptest-regressions/afft/differential.txt:7-8) through
  both the AFFT and the Tower oracle codecs;
* the power-of-two subgroup shortcut against the general product.
"""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

import oracle


def mul(a, b):
    return oracle.gf16_mul(int(a), int(b))


def inv(a):
    return oracle.gf16_inv(int(a))


def elems(buf):
    return np.frombuffer(np.ascontiguousarray(buf, np.uint8).tobytes(), "<u2").astype(np.int64)


def data_block(k, L):
    # afft/decoder.rs:774-778 (rotate_left(3) of an index below 2^61 is a shift)
    return np.array([((i * 37) ^ (i << 3) ^ 0x5A) & 0xFF for i in range(k * L)], np.uint8)


def test_reference_scalar_formula():
    # afft/encoder.rs:347-382: k=5, m=3, symbol_len=6
    k, m, L = 5, 3, 6
    data = np.array([((i * 29) ^ 0xA5) & 0xFF for i in range(k * L)], np.uint8)
    rep = oracle.afft_encode(k, m, L, data)
    d = elems(data).reshape(k, L // 2)
    for off in range(L // 2):
        for r in range(m):
            x = k + r
            want = 0
            for i in range(k):
                num, den = 1, 1
                for o in range(k):
                    if o != i:
                        num = mul(num, x ^ o)
                        den = mul(den, i ^ o)
                want ^= mul(mul(int(d[i, off]), num), inv(den))
            got = int(elems(rep[r * L:(r + 1) * L])[off])
            assert got == want, (r, off)


def novel_basis(x, count):
    """Normalised subspace-polynomial values W_b(x) / W_b(2^b), b < count, with
    W_0(x) = x and W_{b+1}(x) = W_b(x)^2 + W_b(2^b) W_b(x)."""
    out, w = [], x
    for b in range(count):
        out.append(mul(w, inv(w_at_beta[b])))
        w = mul(w, w) ^ mul(w_at_beta[b], w)
    return out


def _w_at_beta(count):
    vals = []
    for b in range(count):
        # W_b(2^b) via the recursion with the previously computed normalisers
        w = 1 << b
        for c in vals:
            w = mul(w, w) ^ mul(c, w)
        vals.append(w)
    return vals


w_at_beta = _w_at_beta(16)


def basis_row(x, k):
    nb = novel_basis(x, 16)
    row = []
    for i in range(k):
        v = 1
        for b in range(16):
            if (i >> b) & 1:
                v = mul(v, nb[b])
        row.append(v)
    return row


def solve(a, rhs):
    """Gauss-Jordan over GF(2^16): a (k x k) c = rhs (k x w)."""
    k = len(a)
    a = [list(r) + list(s) for r, s in zip(a, rhs)]
    for col in range(k):
        piv = next(r for r in range(col, k) if a[r][col])
        a[col], a[piv] = a[piv], a[col]
        s = inv(a[col][col])
        a[col] = [mul(v, s) for v in a[col]]
        for r in range(k):
            if r != col and a[r][col]:
                f = a[r][col]
                a[r] = [v ^ mul(f, p) for v, p in zip(a[r], a[col])]
    return [row[k:] for row in a]


@pytest.mark.parametrize("k,m,L", [(5, 3, 6), (4, 4, 10), (3, 5, 4), (8, 2, 4), (1, 3, 2)])
def test_novel_basis_expansion_matches_lagrange(k, m, L):
    rng = np.random.default_rng(k * 31 + m)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    d = elems(data).reshape(k, L // 2)
    a = [basis_row(i, k) for i in range(k)]
    coef = solve(a, [list(map(int, d[i])) for i in range(k)])
    rep = elems(oracle.afft_encode(k, m, L, data)).reshape(m, L // 2)
    for r in range(m):
        row = basis_row(k + r, k)
        for off in range(L // 2):
            want = 0
            for j in range(k):
                want ^= mul(row[j], coef[j][off])
            assert int(rep[r, off]) == want


def test_basis_functions_vanish_on_their_subspace():
    # W_b vanishes on span{1, 2, ..., 2^(b-1)} and is 1 at 2^b after normalising.
    for b in range(1, 6):
        for x in range(1 << b):
            assert novel_basis(x, b + 1)[b] == 0
        assert novel_basis(1 << b, b + 1)[b] == 1


@pytest.mark.parametrize("k,m,L", [(4, 4, 10), (5, 3, 6), (3, 2, 4), (2, 6, 2)])
def test_decode_every_subset(k, m, L):
    # afft/decoder.rs:847-875 shapes
    data = data_block(k, L)
    rep = oracle.afft_encode(k, m, L, data)
    sym = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    for sub in itertools.combinations(range(k + m), k):
        got = oracle.afft_decode(k, L, [(i, sym[i]) for i in sub])
        assert np.array_equal(got, data), sub


def test_subgroup_shortcut_matches_general_product():
    k, L = 16, 8
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, k * L, dtype=np.uint8)
    vals = [data[i * L:(i + 1) * L] for i in range(k)]
    outs = list(range(k, k + 9))
    fast = oracle.rs16_interp(k, L, range(k), vals, outs)
    perm = list(range(k))[::-1]  # same points, not in order: the general product
    slow = oracle.rs16_interp(k, L, perm, [vals[i] for i in perm], outs)
    assert np.array_equal(fast, slow)


def test_duplicate_points_rejected():
    with pytest.raises(ValueError):
        oracle.rs16_interp(2, 2, [1, 1], [b"\x00\x00", b"\x00\x00"], [0])


# proptest-regressions/afft/differential.txt:7-8, kept as a fixture
GOLDEN = Path(__file__).resolve().parent / "golden" / "afft_differential_regressions.json"
REGRESSIONS = [(c["k"], c["m"], c["field_elements"], c["data_source"], c["ordering"])
               for c in json.loads(GOLDEN.read_text())["cases"]]


def regression_case(case):
    k, m, fe, src, ordering = case
    assert len(src) == 12 * 16 and len(ordering) == 24
    L = 2 * fe
    data = np.array(src[:k * L], np.uint8)
    # indices sorted by (ordering[index], index), first k kept (differential.rs:39-41)
    idx = sorted(range(k + m), key=lambda i: (ordering[i], i))[:k]
    return k, m, L, data, idx


@pytest.mark.parametrize("case", REGRESSIONS)
def test_proptest_regressions_round_trip(case):
    k, m, L, data, idx = regression_case(case)
    rep = oracle.afft_encode(k, m, L, data)
    sym = [data[i * L:(i + 1) * L] for i in range(k)] + [rep[j * L:(j + 1) * L] for j in range(m)]
    assert np.array_equal(oracle.afft_decode(k, L, [(i, sym[i]) for i in idx]), data)
    c = oracle.tower_cauchy_matrix(k, m)
    trep = oracle.tower_encode(c, k, m, L, data)
    tsym = [data[i * L:(i + 1) * L] for i in range(k)] + [trep[j * L:(j + 1) * L] for j in range(m)]
    assert np.array_equal(oracle.tower_decode_rref(c, k, m, L, [(i, tsym[i]) for i in idx]), data)


def test_direct_novel_evaluation_matches_the_basis_expansion():
    """oracle.afft_novel_eval (the checker of the device transform) against
    the products of normalised subspace polynomials computed here."""
    for n in (1, 2, 4, 8, 32):
        rng = np.random.default_rng(n)
        coef = rng.integers(0, 65536, n, dtype=np.uint16)
        points = np.arange(min(2 * n, 64), dtype=np.uint16)
        got = oracle.afft_novel_eval(coef, points)
        for p, x in enumerate(points):
            nb = novel_basis(int(x), 16)
            want = 0
            for i in range(n):
                xi = 1
                for b in range(n.bit_length() - 1):
                    if (i >> b) & 1:
                        xi = mul(xi, nb[b])
                want ^= mul(int(coef[i]), xi)
            assert want == got[p], (n, int(x))
