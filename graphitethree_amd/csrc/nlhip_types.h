// Plain structs shared by host code and HIP kernels.
#pragma once
#include <cstdint>

namespace nlhip {

// Column indices of a CRS matrix as the SpMV engines read them.  `ci` holds
// the plain 32-bit indices.  When the matrix is banded enough, `ci16` holds
// the same indices as 16-bit offsets from a per-segment base: entry j of the
// nonzero stream has column segb[j >> kSegShift] + ci16[j], and segb holds one
// base per 1024-entry segment plus one padding entry (ci16 and segb are null
// otherwise).
constexpr int kSegShift = 10;
constexpr uint32_t kSegEntries = 1u << kSegShift;

struct Cols {
    const uint32_t* ci;
    const uint16_t* ci16;
    const uint32_t* segb;
};

// Device-resident state of one PCG solve.  Each field has exactly one writer
// per kernel launch (block 0 of the kernel named in the comment); every other
// block only reads fields written by an earlier launch.
template <typename T>
struct CGStatus {
    T rho;           // r.h of the current iteration      (written by spmv step)
    T rho_prev;      // r.h of the previous iteration     (written by update step)
    T thresh;        // eps^2 * b.b                        (init)
    T bb;            // b.b                                (init)
    T err;           // r.r when the loop stopped          (spmv step)
    T alpha, beta;   // last step sizes, for diagnostics
    int32_t iter_a;  // completed iterations, read by the direction step
    int32_t iter_b;  // completed iterations, read by the update step
    int32_t done;    // 1 once the loop has stopped        (spmv step)
    int32_t iters;   // iterations performed when done
    int32_t max_iter;
    int32_t stop;    // 1 once the stopping test held      (direction step)
};

// ---- 2x2 block values ------------------------------------------------------
// The four values of block k: its upper pair (a00, a01: row 2I of its block
// row, columns 2J and 2J + 1) and its lower pair (a10, a11: row 2I + 1).
// FP64: in groups of 64 blocks, the 64 upper pairs of the group, then its
// 64 lower pairs, so that a wavefront loading the upper (lower) pairs of 64
// consecutive blocks reads one or two contiguous runs of 16-byte words
// rather than every other word of 2 KiB.  FP32: block by block (a00 a01 a10
// a11), so that one 16-byte load per lane reads a whole block and a
// wavefront 1 KiB of consecutive words.  Offsets in values of T, and the
// array length for nblk blocks (whole groups of 64).
template <typename T>
constexpr uint64_t block_upper(uint64_t k) {
    return sizeof(T) == 8 ? 256 * (k >> 6) + 2 * (k & 63) : 4 * k;
}
template <typename T>
constexpr uint64_t block_lower(uint64_t k) {
    return block_upper<T>(k) + (sizeof(T) == 8 ? 128 : 2);
}
constexpr uint64_t block_vals_size(uint64_t nblk) { return 4 * ((nblk + 63) & ~uint64_t(63)); }

}  // namespace nlhip
