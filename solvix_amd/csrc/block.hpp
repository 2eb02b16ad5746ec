// Multi-RHS Jacobi-PCG on one shared CSR matrix (R <= 64 load cases).
#pragma once
#include "svx_internal.hpp"

namespace svx {

constexpr int kMaxRhs = 64;

struct alignas(16) BlockScal {
    double bnorm[kMaxRhs];
    double rz[kMaxRhs];
    double alpha[kMaxRhs];
    double beta[kMaxRhs];
    double rr[kMaxRhs];
    double res[kMaxRhs];
    int64_t it[kMaxRhs];
    int32_t status[kMaxRhs];
    double tol;
    int64_t max_iter;
    int32_t any_running;
    int32_t pad_;
};

struct BlockArgs {
    int32_t n;
    int32_t R;
    int32_t rows_per_tile;  // SpMV tiles: a multiple of bs
    int32_t n_tiles;
    int32_t vrows_per_tile;  // tiles of the vector kernels (initialisation, updates)
    int32_t n_vtiles;
    int32_t bs;  // node block size (2, 3) for the node-row SpMV; 1 = row form
    int32_t xcd; // SpMV tiles in XCD-contiguous order (1) or tile = workgroup (0)
    const int32_t* row_ptr;
    const int32_t* brow_ptr;  // [n/bs + 1] (bs > 1)
    const int32_t* bcol;      // node-block columns (bs > 1)
    const int32_t* col;
    const double* val;
    const double* dinv;
    const double* F;  // [n][R]
    double* U;
    double* Rv;
    double* P;
    double* Z;
    double* Q;
    double* partials;  // slot-major per-tile, per-load-case sums: [1][n_tiles][64] (p.q), [2][n_vtiles][64] (r.r, r.z)
    BlockScal* S;
};

void blk_init(const BlockArgs& a, hipStream_t s);
void blk_iteration(const BlockArgs& a, hipStream_t s);
void blk_count_running(const BlockArgs& a, int32_t* out, hipStream_t s);
void blk_spmv_only(const BlockArgs& a, hipStream_t s);
void blk_pack(const double* in, double* out, int32_t n, int32_t R, const uint8_t* mask, hipStream_t s);
void blk_unpack(const double* in, double* out, int32_t n, int32_t R, hipStream_t s);

}  // namespace svx
