// Layout of the MFMA operand table of the folded swap blocks.
//
// For degree n the folded blocks (operator_data.cpp:47-147) are cut into
// row blocks b of 16 orders and K-chunks c of 4 orders; fragment (b, c) is
// the 16x4 A operand of v_mfma_f64_16x16x4f64 in lane order: lane
// r + 16k holds block entry [m = 16b + r][l = 4c + k] (zero outside
// m, l <= n).  Variant v = 0 holds the blocks, v = 1 their transposes; part
// p = 0 the real block, p = 1 the imaginary one.
#pragma once

#include <cstddef>

namespace sfmm {
namespace mfma {

// Degrees 0..kMaxOrder-1 are tabulated.
inline constexpr int kMaxOrder = 32;

constexpr int row_blocks(int n) { return (n + 16) / 16; }
constexpr int k_chunks(int n) { return (n + 4) / 4; }

constexpr int degree_base(int n)
{
    int s = 0;
    for (int j = 0; j < n; ++j) s += row_blocks(j) * k_chunks(j);
    return s;
}

constexpr int fragments_per_part() { return degree_base(kMaxOrder); }

// Index (in fragments of 64 doubles) of fragment (v, p, n, b, c).
constexpr int frag_index(int v, int p, int n, int b, int c)
{
    return (v * 2 + p) * fragments_per_part() + degree_base(n) + b * k_chunks(n) + c;
}

constexpr size_t table_doubles() { return size_t(4) * fragments_per_part() * 64; }

// Fill `out` (table_doubles() reals) on the host.
void build_table(double *out);

// Device copy of the table on the current device, built on first use and
// kept until the process ends.  Returns nullptr on failure.
const double *device_table();

}  // namespace mfma
}  // namespace sfmm
