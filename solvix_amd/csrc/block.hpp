// Multi-RHS Jacobi-PCG on one shared CSR matrix (R <= 64 load cases).
#pragma once
#include "svx_internal.hpp"

namespace svx {

constexpr int kMaxRhs = 64;
constexpr int kGroupDefault = 4;   // nodes per group of the block SpMV (a 2x2x1 patch of a solid mesh)
constexpr int kReduceGroups = 32;  // first-level workgroups of a reduction

struct alignas(16) BlockScal {
    double bnorm[kMaxRhs];
    double rz[kMaxRhs];
    double alpha[kMaxRhs];
    double beta[kMaxRhs];
    double rr[kMaxRhs];
    double res[kMaxRhs];
    int64_t it[kMaxRhs];
    int32_t status[kMaxRhs];
    double best[kMaxRhs];    // FP32-IR: best true residual of the FP32 phase so far
    double rz_old[kMaxRhs];  // FP32-IR: r.z before the last beta (beta after a refresh)
    double tol;
    int64_t max_iter;
    int32_t any_running;
    int32_t pad_;
};

struct BlockArgs {
    int32_t n;
    int32_t R;
    int32_t rows_per_tile;  // gathering SpMV tiles (no group plan): a multiple of bs
    int32_t n_tiles;        // SpMV tiles: four groups each, or gathering tiles
    int32_t vrows_per_tile; // tiles of the vector kernels (initialisation, update)
    int32_t n_vtiles;
    int32_t bs;  // node block size (2, 3) for the node-row SpMV; 1 = row form
    int32_t xcd; // SpMV tiles in XCD-contiguous order (1) or tile = workgroup (0)
    int32_t group;     // nodes per group of the group SpMV (0: gathering SpMV)
    int32_t n_groups;
    const int32_t* row_ptr;
    const int32_t* brow_ptr;  // [n/bs + 1] (bs > 1)
    const int32_t* bcol;      // node-block columns (bs > 1)
    const int32_t* col;
    const double* val;
    const double* dinv;
    const int32_t* grp;     // group SpMV: [n_groups][4] first union column, end, first packed block, end
    const int32_t* gnodes;  // [n_groups * group] members
    const int32_t* ucol;    // [n_cols][2] union column node, member mask
    const double* pval;     // packed block values [nblk][bs][bs], group by group
    const double* F;  // [n][R]
    double* U;
    double* Rv;
    double* P;   // direction, even iterations read it and the odd ones write it
    double* P2;  // the other direction buffer
    double* Q;
    // FP32-IR inner iterations (group SpMV only): FP32 packed values,
    // correction since the last refresh, residual, direction (two buffers), q
    const float* pval32;
    float* X32;
    float* R32;
    float* P32;
    float* P32b;
    float* Q32;
    double* partials;  // slot-major per-tile, per-load-case sums: [1][n_tiles][64] (p.q), [2][n_vtiles][64] (r.r, r.z)
    double* l1;        // first-level sums [2][kReduceGroups][64]
    BlockScal* S;
};

void blk_init(const BlockArgs& a, hipStream_t s);
// Iteration `it` of a solve (its parity picks the direction buffers).
void blk_iteration(const BlockArgs& a, int it, hipStream_t s);
void blk_count_running(const BlockArgs& a, int32_t* out, hipStream_t s);
// One SpMV of iteration 0 (with a group plan: the direction step, then the
// product).
void blk_spmv_only(const BlockArgs& a, hipStream_t s);
// The SpMV kernel alone (with a group plan, the product of the p' the last
// step wrote; the gathering SpMV evaluates the step itself).
void blk_product_only(const BlockArgs& a, hipStream_t s);
// FP32-IR block solve (a mesh with a group plan): the FP32 phase starts from
// u = 0 (blk32_init), runs FP32 iterations (blk32_iteration) and refreshes
// (blk32_refresh: u += x32, r = f - K u in FP64, hand-over test per load
// case); blk_warm_start then sets up the FP64 block iterations from every
// handed-over case's u and r.
void blk32_init(const BlockArgs& a, hipStream_t s);
void blk32_iteration(const BlockArgs& a, int it, hipStream_t s);
void blk32_refresh(const BlockArgs& a, hipStream_t s);
void blk_warm_start(const BlockArgs& a, hipStream_t s);
void blk_pack(const double* in, double* out, int32_t n, int32_t R, const uint8_t* mask, hipStream_t s);
void blk_unpack(const double* in, double* out, int32_t n, int32_t R, hipStream_t s);
// Packs the CSR values block by block in group order (BlockPlan's pk_src /
// pk_stride) into pval [nblk][bs][bs].
void blk_pack_values(const double* val, const int32_t* pk_src, const int32_t* pk_stride, int64_t nblk, int bs,
                     double* pval, hipStream_t s);
// The same, rounded to FP32.
void blk_pack_values32(const double* val, const int32_t* pk_src, const int32_t* pk_stride, int64_t nblk, int bs,
                       float* pval, hipStream_t s);

}  // namespace svx
