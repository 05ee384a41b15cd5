// Host-callable launchers for the HIP kernels (defined in k_*.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "nlhip_types.h"

namespace nlhip {

// Rows-per-wavefront dispatch over the compiled SpMV engines (codes of
// dev_common.h's EngineOf): 2, 4, 8, 16 -> 32, 16, 8, 4 lanes per row;
// 65..72 -> stream engines (one row per lane); 73..79 -> 2x2 block stream
// engines (their callers pass the block arrays).  Launchers pass the code
// through engine_rpw first, which resolves 64.
#define NLHIP_RPW_SWITCH(rpw, CALL)                       \
    switch (rpw) {                                        \
        case 2: { constexpr int RPW = 2; CALL; } break;   \
        case 4: { constexpr int RPW = 4; CALL; } break;   \
        case 8: { constexpr int RPW = 8; CALL; } break;   \
        case 16: { constexpr int RPW = 16; CALL; } break; \
        case 65: { constexpr int RPW = 65; CALL; } break; \
        case 66: { constexpr int RPW = 66; CALL; } break; \
        case 67: { constexpr int RPW = 67; CALL; } break; \
        case 68: { constexpr int RPW = 68; CALL; } break; \
        case 69: { constexpr int RPW = 69; CALL; } break; \
        case 70: { constexpr int RPW = 70; CALL; } break; \
        case 71: { constexpr int RPW = 71; CALL; } break; \
        case 72: { constexpr int RPW = 72; CALL; } break; \
        case 73: { constexpr int RPW = 73; CALL; } break; \
        case 74: { constexpr int RPW = 74; CALL; } break; \
        case 75: { constexpr int RPW = 75; CALL; } break; \
        case 76: { constexpr int RPW = 76; CALL; } break; \
        case 77: { constexpr int RPW = 77; CALL; } break; \
        case 78: { constexpr int RPW = 78; CALL; } break; \
        case 79: { constexpr int RPW = 79; CALL; } break; \
        default: return hipErrorInvalidValue;             \
    }

// Launch shapes a caller may set: 2, 4, 8, 16, and the stream engines 64
// (the default stream engine) and 65..79 (variants kept for sweeps; they
// compute the same sums as 64; 73..79 run as 64 on matrices without 2x2
// block form, and as their twin for the matrix's block-column format).
inline bool valid_rpw(int rpw) {
    return rpw == 2 || rpw == 4 || rpw == 8 || rpw == 16 || (rpw >= 64 && rpw <= 79);
}

// Smallest system (rows) on which code 64 runs the 2x2 block engine: below
// it the scalar engine's shorter chunk chain per wavefront is faster.
constexpr uint32_t kBlockMinRows = 131072;

// Default stream engines: with and without 16-bit column offsets (each wave
// streams one contiguous run of tiles; the XCD-interleaved runs of 69/70
// gather less but are no faster on large systems and slower on small ones).
constexpr int kStreamAuto = 68;
constexpr int kStreamAutoPlain = 71;

// Block engine for a requested block code and the matrix's block-column
// format: 73..77 read 16-bit offsets, 78 / 79 (4 / 8 blocks per lane) the
// 32-bit block columns; each maps to its twin for the other format.
inline int block_engine(int code, bool d16) {
    if (d16) return code == 78 ? 73 : (code == 79 ? 76 : code);
    return code == 76 ? 79 : (code == 78 || code == 79 ? code : 78);
}

// Engine a launcher runs for a requested code: 64 becomes the default stream
// engine for the matrix's column format; a variant that reads 16-bit offsets
// falls back to its 32-bit twin when the matrix has none.
inline int engine_rpw(int rpw, const Cols& cols) {
    if (rpw == 64) return cols.ci16 ? kStreamAuto : kStreamAutoPlain;
    if (!cols.ci16 && rpw == 68) return 71;
    if (!cols.ci16 && rpw == 70) return 69;
    if (!cols.ci16 && rpw == 72) return 67;
    return rpw;
}

// SpMV work per workgroup for `grid` workgroups: rows (vector engines,
// a multiple of the rows one pass covers) or 64-row tiles (stream engine).
uint32_t spmv_work_per_block(uint32_t m, uint32_t grid, int rpw);

template <typename T>
hipError_t spmv(uint32_t m, const uint32_t* rp, const Cols& cols, const T* val, const T* x,
                T* y, int rpw, uint32_t grid, hipStream_t s);

// r = b - A x (r may be null), partial sums of r.r and b.b per workgroup.
hipError_t residual64(uint32_t m, const uint32_t* rp, const Cols& cols, const double* val,
                      const double* x, const double* b, double* r, double* part_rr,
                      double* part_bb, int rpw, uint32_t grid, hipStream_t s);

template <typename T>
hipError_t jacobi_dinv(uint32_t m, const uint32_t* rp, const uint32_t* ci, const double* val,
                       T* dinv, hipStream_t s);

hipError_t cast_f64_f32(uint64_t n, const double* a, float* b, hipStream_t s);

// 16-bit column offsets (Cols): segb[s] = smallest column of segment s and
// *span = max over segments of (largest - smallest column), atomically
// max-combined into *span (set it to 0 first); then ci16[j] = ci[j] -
// segb[j >> kSegShift] once the span is known to fit in 16 bits.
hipError_t seg_bases(uint64_t nnz, const uint32_t* ci, uint32_t* segb, uint32_t* span,
                     hipStream_t s);
// 2x2 block form over nb = m/2 row pairs: check (sets *bad, cleared first by
// the caller), then pack brp (nb + 1), bcol (nnz/4, may be null) and bval
// (nnz values, 4 per block).
hipError_t block_check(uint32_t nb, const uint32_t* rp, const uint32_t* ci, uint32_t* bad,
                       hipStream_t s);
template <typename T>
hipError_t block_pack(uint32_t nb, const uint32_t* rp, const uint32_t* ci, const double* val,
                      uint32_t* brp, uint32_t* bcol, T* bval, hipStream_t s);
// FP32 block values from the FP64 ones of the same nblk blocks (each in its
// type's layout, block_upper / block_lower).
hipError_t block_vals_f32(uint64_t nblk, const double* src, float* dst, hipStream_t s);
hipError_t seg_encode(uint64_t nnz, const uint32_t* ci, const uint32_t* segb, uint16_t* ci16,
                      hipStream_t s);

// BLAS-1
hipError_t blas_dscal(uint64_t n, double a, double* x, hipStream_t s);
hipError_t blas_daxpy(uint64_t n, double a, const double* x, double* y, hipStream_t s);
hipError_t blas_dmul(uint64_t n, const double* a, const double* x, double* y, hipStream_t s);
// dst[0..16 n16) = src[0..16 n16) with `grid` workgroups.
hipError_t copy16(uint64_t n16, const void* src, void* dst, uint32_t grid, hipStream_t s);
// Reads n16 16-byte words; out receives grid x 256 words.
hipError_t read16(uint64_t n16, const void* src, void* out, uint32_t grid, bool nt,
                  hipStream_t s);
// Each wavefront reads one contiguous run of the buffer (non-temporal).
hipError_t read16_runs(uint64_t n16, const void* src, void* out, uint32_t grid, hipStream_t s);
// The SpMV's traffic without its arithmetic (k_blas.hip): nblk 32-byte
// blocks and 2-byte offsets streamed, a 16-byte gather per block from an
// L2-resident vector of nblk / bpr words (xg), and per row pair (bpr
// blocks) a 16-byte read from rin and, when write, a 16-byte store to rout
// (both nblk / bpr words); out: 16 bytes.
hipError_t stream_mix(uint64_t nblk, uint32_t bpr, const double* vals, const uint16_t* off,
                      const double* xg, const double* rin, double* rout, double* out,
                      bool write, uint32_t grid, hipStream_t s);
template <typename T>
hipError_t fill(uint64_t n, T v, T* x, hipStream_t s);
// out[0] = x.y, out[1] = sqrt(x.y); part holds `grid` doubles.
hipError_t blas_ddot(uint64_t n, const double* x, const double* y, double* part, uint32_t grid,
                     double* out, hipStream_t s);
// out[0] = sum(part[0..P)), out[1] = sqrt(out[0]).
hipError_t reduce_final(const double* part, uint32_t P, double* out, hipStream_t s);

// The SpMV step of the fused PCG runs on up to this many times the
// workgroups of the other launches (more waves in flight for its gathers).
constexpr uint32_t kSpmvGridMul = 2;

// Fused PCG.  All arrays are device memory.
template <typename T>
struct CGArgs {
    uint32_t m;      // rows
    uint32_t chunk;  // rows per workgroup of the element-wise launches
    uint32_t sp;     // SpMV work per workgroup of the init launch (P workgroups)
    uint32_t P;      // workgroups per launch = number of partial sums
    uint32_t Ps;     // workgroups of the SpMV step (d.q partials), Ps <= kSpmvGridMul P
    uint32_t sps;    // SpMV work per workgroup of the SpMV step
    const uint32_t* rp;
    Cols cols;
    const T* val;
    const T* dinv;   // null: no preconditioner
    T* x;
    const T* b;
    T* r;
    T* h;
    T* d;
    T* q;
    T* part_dq;
    T* part_rh;
    T* part_rr;
    T* part_bb;
    CGStatus<T>* st;
};

template <typename T>
hipError_t cg_start(const CGArgs<T>& a, int rpw, bool precond, T eps, int max_iter,
                    hipStream_t s);
// One iteration: direction step, SpMV step, update step.
template <typename T>
hipError_t cg_step(const CGArgs<T>& a, int rpw, bool precond, hipStream_t s);
// One launch of an iteration: part 0 = direction, 1 = SpMV, 2 = update.
template <typename T>
hipError_t cg_step_part(const CGArgs<T>& a, int rpw, bool precond, int part, hipStream_t s);

// ---- batched systems (k_batch.hip) ----------------------------------------

// A batch tile is 4 passes of 64 rows, or of 64 block rows (row pairs) for
// a system in 2x2 block form; its partial sums are formed pass by pass.
constexpr uint32_t kTilePassRows = 64;
constexpr uint32_t kTileRows = 4 * kTilePassRows;       // scalar form
constexpr uint32_t kBlockTileRows = 2 * kTileRows;      // 2x2 block form
constexpr uint32_t kNoBlocks = 0xffffffffu;

template <typename T>
struct BatchArgs {
    uint32_t nsys;
    uint32_t ntiles;
    const uint32_t* row_off;    // nsys + 1
    const uint32_t* tile_off;   // nsys + 1: tiles of system s
    const uint32_t* tile_sys;   // ntiles
    const uint32_t* tile_row0;  // ntiles: first (global) row of the tile
    const uint32_t* rp;         // global rowptr (n + 1)
    const uint32_t* ci;         // global column indices
    const uint16_t* ci16;       // their 16-bit segment offsets, or null
    const uint32_t* segb;       // segment bases of ci16
    // 2x2 block form of the systems that have it (BlockStreamEngine):
    // sys_pair0[s] = first block row of system s (kNoBlocks: scalar form;
    // the array is null when no system has the form); the block rows of s
    // are its row pairs (row_off[s] + 2p, +1), i.e. block row B holds rows
    // 2B + delta_s with delta_s = row_off[s] - 2 sys_pair0[s]; brp over all
    // block rows; block columns as 16-bit offsets (bci16, bsegb) or, when
    // they do not fit, 32-bit (bcol)
    const uint32_t* sys_pair0;
    uint32_t all_blocks;        // every system has the block form
    uint32_t spmv_r2_tiles;     // active tiles from which SpMV wavefronts
    uint32_t spmv_r4_tiles;     //   stream 2 (4) passes of a tile
    const uint32_t* brp;
    const uint32_t* bcol;
    const uint16_t* bci16;
    const uint32_t* bsegb;
    const T* bval;
    const T* val;
    const T* dinv;              // null: no preconditioner
    T* x;
    const T* b;
    T* r;
    T* h;
    T* d;
    T* q;
    T* part_dq;                 // per tile
    T* part_rh;
    T* part_rr;
    T* part_bb;
    T* rho;                     // per system: r.h of the current iteration
    T* rho_prev;                //   r.h of the previous one
    T* alpha;
    T* beta;
    T* err;
    T* thresh;
    T* bb;
    int32_t* iters;
    int32_t* done;
    int32_t* stop;              // stopping test held (turned into done next launch)
    const int32_t* skip;        // per system, may be null: start as done
    const int32_t* cap;         // per system, may be null: iteration cap (the
                                //   launch's max_iter applies when smaller)
    uint32_t* act_tiles;
    uint32_t* n_act_tiles;
    uint32_t* act_sys;
    uint32_t* n_act_sys;
};

template <typename T>
hipError_t batch_start(const BatchArgs<T>& a, bool precond, T eps, int max_iter, uint32_t grid,
                       hipStream_t s);
template <typename T>
hipError_t batch_iteration(const BatchArgs<T>& a, bool precond, int max_iter, uint32_t grid,
                           hipStream_t s);
template <typename T>
hipError_t batch_compact(const BatchArgs<T>& a, uint32_t grid, hipStream_t s);
// n_iter iterations and then the list rebuild in one cooperative launch of
// `grid` workgroups (at most batch_persistent_grid, all resident at once).
template <typename T>
hipError_t batch_persistent_grid(const BatchArgs<T>& a, bool precond, uint32_t* grid);
template <typename T>
hipError_t batch_persistent(const BatchArgs<T>& a, bool precond, int max_iter, uint32_t n_iter,
                            uint32_t grid, hipStream_t s);
// r = b - A x (FP64) and per-system |r|^2 -> err[], |b|^2 -> bb[].
hipError_t batch_residual64(const BatchArgs<double>& a, uint32_t grid, hipStream_t s);
// b32 = float(r64 * scale[s]), x32 = 0.
hipError_t batch_scale_cast(const BatchArgs<double>& a64, const BatchArgs<float>& a32,
                            const double* scale, uint32_t grid, hipStream_t s);
// x64 += unscale[s] * x32 where a32.skip[s] == 0.
hipError_t batch_ir_update(const BatchArgs<double>& a64, const BatchArgs<float>& a32,
                           const double* unscale, uint32_t grid, hipStream_t s);
// dst = src on the rows of the systems with flag[s] (FP64 batch vectors).
hipError_t batch_copy_systems(const BatchArgs<double>& a, double* dst, const double* src,
                              const int32_t* flag, uint32_t grid, hipStream_t s);

// ---- multiple right-hand sides (k_spmm.hip) ---------------------------------

template <typename T>
struct MultiArgs {
    uint32_t m;               // rows
    uint32_t k;               // right-hand sides (<= 32)
    uint32_t G;               // lanes per row: k rounded up to a power of two
    uint32_t P;               // workgroups of the row-shaped launches
    uint32_t rows_per_block;
    uint32_t one_chunk;       // every SpMM pass fits one LDS chunk (spmm_window_max)
    const uint32_t* rp;
    const uint32_t* ci;
    const T* val;
    const T* dinv;            // per row; null: no preconditioner
    T* x;                     // n x k interleaved
    const T* b;
    T* r;
    T* h;
    T* d;
    T* q;
    T* part_dq;               // P x k
    T* part_rh;
    T* part_rr;
    T* part_bb;
    T* rho;                   // per column
    T* alpha;
    T* beta;
    T* err;
    T* thresh;
    T* bb;
    int32_t* iters;
    int32_t* done;
    const int32_t* skip;      // may be null
    // null, or a flag that reads 0 when X is all zeros: the start then takes
    // every row sum of A X as +0 without reading the matrix (the sum of
    // products with zeros, from +0, is +0 whatever their signs)
    const int32_t* x_nonzero;
    // SpMM step: with sp > 0 (k <= kMultiStreamMaxK) the stream engine, Pdq
    // workgroups of sp 64-row tiles each, reading the column indices as
    // `cols` (16-bit offsets when cols.ci16 is set); with sp = 0 the
    // row-group engine on P workgroups.  Pdq partials of d.q per column.
    Cols cols;
    uint32_t sp;
    uint32_t Pdq;
};

// Right-hand sides up to which the multi-RHS SpMM step streams the matrix.
constexpr uint32_t kMultiStreamMaxK = 4;
// Workgroups of the stream SpMM step (k right-hand sides, 16-bit column
// offsets or not) that one CU holds at once.
template <typename T>
uint32_t multi_stream_per_cu(uint32_t k, bool d16);

template <typename T>
hipError_t multi_start(const MultiArgs<T>& a, bool precond, T eps, int max_iter, hipStream_t s);
template <typename T>
hipError_t multi_iteration(const MultiArgs<T>& a, bool precond, int max_iter, hipStream_t s);
// nonzero (to_interleaved only, may be null): set to 1 when a value moved
// is not zero.
template <typename T>
hipError_t multi_layout(uint32_t n, uint32_t k, const T* src, T* dst, bool to_interleaved,
                        hipStream_t s, int32_t* nonzero = nullptr);
// Re-opens the FP64 columns flagged in again[k] (done -> 0 with the beta /
// rho update their stopping launch skipped), then recounts the active ones.
hipError_t multi_reopen(const MultiArgs<double>& a, const int32_t* again, hipStream_t s);
hipError_t multi_residual64(const MultiArgs<double>& a, hipStream_t s);
hipError_t multi_scale_cast(const MultiArgs<double>& a64, const MultiArgs<float>& a32,
                            const double* scale, hipStream_t s);
hipError_t multi_ir_update(const MultiArgs<double>& a64, const MultiArgs<float>& a32,
                           const double* unscale, hipStream_t s);
// dst[:, j] = src[:, j] (n x k interleaved) for the columns with flag[j].
hipError_t multi_copy_cols(uint32_t n, uint32_t k, double* dst, const double* src,
                           const int32_t* flag, hipStream_t s);

// Mixed-precision refinement helpers.
hipError_t ir_scale_cast(uint64_t n, const double* r, double s, float* out, hipStream_t st);
hipError_t ir_update(uint64_t n, double* x, const float* c, double s, hipStream_t st);

}  // namespace nlhip
