// Plain structs shared by host code and HIP kernels.
#pragma once
#include <cstdint>

namespace nlhip {

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
    int32_t iter_a;  // completed iterations, read by the spmv step
    int32_t iter_b;  // completed iterations, read by the update step
    int32_t done;    // 1 once the stopping test held      (spmv step)
    int32_t iters;   // iterations performed when done
    int32_t max_iter;
    int32_t pad;
};

}  // namespace nlhip
