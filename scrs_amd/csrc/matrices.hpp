// matrices.hpp - coefficient matrices of the three Cauchy constructions.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "field.hpp"

namespace scrs {

// Engine capacities, k + m (scrs/src/selector.rs:25-32).
constexpr size_t kCapStandard = 256;
constexpr size_t kCapGood = 255;
constexpr size_t kCapTower = 65535;
constexpr size_t kCapAfft = 65536;

// k x m source-major GF(2^8) matrix, C[i*m + j].
//  Standard: X = {0..k-1}, Y = {k..k+m-1} (matrices/cauchy.rs:38,88-99).
//  Good: X = {3^i}, Y = {3^(k+j)} through the diagonal factorisation
//        3^-i * 1/(1 + 3^(k+j-i)) (matrices/good_cauchy.rs:125-151).
std::vector<uint8_t> standard_cauchy_matrix(size_t k, size_t m);
std::vector<uint8_t> good_cauchy_matrix(size_t k, size_t m);

// k x m source-major GF(2^16) Tower matrix with g = 0x0108, built with one
// batch inversion (tower/cauchy.rs:69-137).
std::vector<uint16_t> tower_cauchy_matrix(size_t k, size_t m);

// Packed split-nibble product tables for a T x R coefficient block given
// term-major (coef[t*R + r]).  For term t and row quad q the 32 dwords at
// ((t * quads) + q) * 32 hold lo[n] (n < 16) and hi[n] (16 + n); byte s of
// each dword is the product for row 4q + s: coef * n and coef * (n << 4).
// Rows at or beyond R are zero.  `quads` >= ceil(R / 4).
void build_nibble_tables(const uint8_t *coef, size_t T, size_t R, size_t quads,
                         uint32_t *out);

}  // namespace scrs
