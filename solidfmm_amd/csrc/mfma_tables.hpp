// Dense copies of the folded swap blocks for the MFMA kernels.
//
// Degree n holds the real block SR_n and then the imaginary block SI_n
// (operator_data.cpp:47-147, row = output order m, column = input order l),
// each (n+1) rows at an odd row stride, so the 16 rows one MFMA operand
// reads fall in distinct LDS banks.  The transposed blocks the L-type
// pipeline needs are not stored: TR = D^-1 SR^T D and TI = D^-1 SI^T D with
// D = diag(1, 2, ..., 2), so a transposed swap reads SR^T / SI^T from the
// same table, doubles its inputs of order l >= 1 and halves its outputs of
// order m >= 1 (exact in binary floating point).
#pragma once

#include <cstddef>

namespace sfmm {
namespace mfma {

// Degrees 0..kMaxOrder-1 are tabulated (the FP64 limit of the reference,
// operator_data.cpp:35-45, is order 86).
inline constexpr int kMaxOrder = 86;

constexpr int dense_stride(int n) { return (n + 1) | 1; }
constexpr int dense_part(int n) { return (n + 1) * dense_stride(n); }

// Offset (in doubles) of degree n's SR block; SI follows at + dense_part(n).
constexpr int dense_base(int n)
{
    int s = 0;
    for (int j = 0; j < n; ++j) s += 2 * dense_part(j);
    return s;
}

// Doubles needed for degrees 0..P-1.
constexpr int dense_size(int P) { return dense_base(P); }

constexpr size_t table_doubles() { return size_t(dense_size(kMaxOrder)); }

// Fill `out` (table_doubles() reals) on the host (swap_host.cpp).
void build_table(double *out);

// Device copy of the table on the current device, built on first use and
// kept until the process ends.  Returns nullptr on failure.
const double *device_table();

// Device factorials n! for n < 2 kMaxOrder, followed by 1/n! for
// n < 2 kMaxOrder (operator_data.cpp:195-214 arithmetic).
const double *device_factorials();

}  // namespace mfma
}  // namespace sfmm
