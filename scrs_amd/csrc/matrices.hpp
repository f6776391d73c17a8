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

// Split-nibble product tables for a T x R coefficient block given term-major
// (coef[t*R + r]).  For term t and 16-row group g, the 512 bytes at dword
// (t * groups + g) * 128 hold 32 entries of 16 bytes: entry n (n < 16) is
// the row vector coef[t][16g + s] * n, entry 16 + n is coef * (n << 4); byte
// s of an entry belongs to row 16g + s.  Rows at or beyond R are zero.
// `groups` >= ceil(R / 16).
void build_nibble_tables(const uint8_t *coef, size_t T, size_t R, size_t groups,
                         uint32_t *out);

// GF(2^16) product tables for a T x R block (coef[t*R + r]).  For term t and
// 16-row group g, the 2048 bytes at dword (t * groups + g) * 512 hold 64
// entries of 32 bytes, entry (q * 16 + n) being the row products
// coef[t][16g + s] * (n << 4q), little-endian 2 bytes per row s.
void build_nibble_tables16(const uint16_t *coef, size_t T, size_t R, size_t groups,
                           uint32_t *out);

}  // namespace scrs
