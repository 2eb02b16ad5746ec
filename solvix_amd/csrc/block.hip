// Multi-RHS Jacobi-PCG: R <= 64 load cases on one shared matrix.
//
// Vectors are stored row-major [n][R]: one 64-wide wavefront owns one matrix
// row and its lane j owns load case j. The SpMV reads each row's values and
// column indices once (wave-uniform scalar loads) and multiplies them with R
// right-hand sides at once, so the matrix streams from HBM once per iteration
// for all load cases (block SpMV). Every load case runs the recurrence of
// solver.rs:82-149 with its own alpha, beta, residual and convergence mask;
// dot products are reduced per load case in a fixed order (lane-local sums,
// waves of the block in order, tiles in order).
#include "block.hpp"

namespace svx {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;  // rows per block pass

// Per-load-case partials of this block: lanes sum their waves in order.
__device__ __forceinline__ void block_rhs_sum2(double& a, double& b) {
    __shared__ double sa[kWavesPerBlock][kWave], sb[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    sa[w][lane] = a;
    sb[w][lane] = b;
    __syncthreads();
    if (w == 0) {
        a = sa[0][lane], b = sb[0][lane];
#pragma unroll
        for (int i = 1; i < kWavesPerBlock; ++i) a += sa[i][lane], b += sb[i][lane];
    }
    __syncthreads();
}

// Publish wave 0's per-lane partials for this tile; true in the last tile.
__device__ __forceinline__ bool publish_rhs(double a, double b, double* partials, int tile, uint32_t* counter,
                                            int n_tiles) {
    __shared__ int s_last;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (w == 0) {
        partials[(size_t)tile * 2 * kWave + lane] = a;
        partials[(size_t)tile * 2 * kWave + kWave + lane] = b;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(counter, 1u) == (uint32_t)(n_tiles - 1);
    __syncthreads();
    bool last = s_last != 0;
    if (last) __threadfence();
    return last;
}

// Sum over tiles (in order) for every load case, computed by wave 0.
__device__ __forceinline__ void reduce_rhs(const double* partials, int n_tiles, double& a, double& b) {
    const int lane = threadIdx.x % kWave;
    a = 0.0, b = 0.0;
    for (int t = 0; t < n_tiles; ++t) {
        a += partials[(size_t)t * 2 * kWave + lane];
        b += partials[(size_t)t * 2 * kWave + kWave + lane];
    }
}

__global__ __launch_bounds__(kBlock) void blk_init_kernel(BlockArgs A) {
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const bool on = lane < A.R;
    double ff = 0.0, rz = 0.0;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        if (!on) continue;
        size_t k = (size_t)i * A.R + lane;
        double fi = A.F[k];
        double zi = fi * A.dinv[i];
        A.U[k] = 0.0;
        A.Rv[k] = fi;
        A.Z[k] = zi;
        A.P[k] = zi;
        ff += fi * fi;
        rz += fi * zi;
    }
    block_rhs_sum2(ff, rz);
    if (!publish_rhs(ff, rz, A.partials, tile, &A.S->cnt_a, A.n_tiles)) return;
    if (w == 0) {
        reduce_rhs(A.partials, A.n_tiles, ff, rz);
        if (on) {
            A.S->bnorm[lane] = sqrt(ff);
            A.S->rr[lane] = ff;
            A.S->rz[lane] = rz;
            A.S->it[lane] = 0;
            A.S->status[lane] = (ff == 0.0) ? ST_CONVERGED : ST_RUNNING;
            A.S->res[lane] = 0.0;
        }
        if (lane == 0) A.S->cnt_a = 0;
    }
}

__device__ __forceinline__ bool any_running(const BlockScal* S, int R) {
    for (int j = 0; j < R; ++j)
        if (S->status[j] == ST_RUNNING) return true;
    return false;
}

__global__ __launch_bounds__(kBlock) void blk_spmv_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const bool on = lane < A.R;
    double pq = 0.0;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        double acc = 0.0;
        const int32_t kb = A.row_ptr[i], ke = A.row_ptr[i + 1];
        if (on)
            for (int32_t k = kb; k < ke; ++k) acc += A.val[k] * A.P[(size_t)A.col[k] * A.R + lane];
        if (on) {
            size_t o = (size_t)i * A.R + lane;
            A.Q[o] = acc;
            pq += A.P[o] * acc;
        }
    }
    double dummy = 0.0;
    block_rhs_sum2(pq, dummy);
    if (!publish_rhs(pq, 0.0, A.partials, tile, &A.S->cnt_a, A.n_tiles)) return;
    if (w == 0) {
        reduce_rhs(A.partials, A.n_tiles, pq, dummy);
        if (on && A.S->status[lane] == ST_RUNNING) {
            if (pq <= 0.0 || pq != pq) {
                A.S->it[lane] += 1;
                A.S->res[lane] = sqrt(A.S->rr[lane]) / A.S->bnorm[lane];
                A.S->status[lane] = ST_BREAKDOWN;
            } else {
                A.S->alpha[lane] = A.S->rz[lane] / pq;
            }
        }
        if (lane == 0) A.S->cnt_a = 0;
    }
}

__global__ __launch_bounds__(kBlock) void blk_update_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const bool on = lane < A.R && A.S->status[lane] == ST_RUNNING;
    const double alpha = on ? A.S->alpha[lane] : 0.0;
    double rr = 0.0, rz = 0.0;
    if (on)
        for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
            size_t k = (size_t)i * A.R + lane;
            A.U[k] += alpha * A.P[k];
            double ri = A.Rv[k] - alpha * A.Q[k];
            A.Rv[k] = ri;
            double zi = ri * A.dinv[i];
            A.Z[k] = zi;
            rr += ri * ri;
            rz += ri * zi;
        }
    block_rhs_sum2(rr, rz);
    if (!publish_rhs(rr, rz, A.partials, tile, &A.S->cnt_b, A.n_tiles)) return;
    if (w == 0) {
        reduce_rhs(A.partials, A.n_tiles, rr, rz);
        if (lane < A.R && A.S->status[lane] == ST_RUNNING) {
            A.S->rr[lane] = rr;
            A.S->it[lane] += 1;
            double res = sqrt(rr) / A.S->bnorm[lane];
            A.S->res[lane] = res;
            if (res <= A.S->tol) {
                A.S->status[lane] = ST_CONVERGED;
            } else {
                A.S->beta[lane] = rz / A.S->rz[lane];
                A.S->rz[lane] = rz;
                if (A.S->it[lane] >= A.S->max_iter) A.S->status[lane] = ST_MAXITER;
            }
        }
        if (lane == 0) A.S->cnt_b = 0;
    }
}

__global__ __launch_bounds__(kBlock) void blk_pupdate_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const bool on = lane < A.R && A.S->status[lane] == ST_RUNNING;
    if (!on) return;
    const double beta = A.S->beta[lane];
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        size_t k = (size_t)i * A.R + lane;
        A.P[k] = A.Z[k] + beta * A.P[k];
    }
}

// One lane per load case: recompute the running flag read by every kernel.
__global__ void blk_flag_kernel(BlockArgs A, int32_t* running_out) {
    const int lane = threadIdx.x;
    int run = lane < A.R && A.S->status[lane] == ST_RUNNING;
    unsigned long long m = __ballot(run);
    if (lane == 0) {
        A.S->any_running = m != 0ull;
        if (running_out) *running_out = __popcll(m);
    }
}

// [R][n] <-> [n][R] (lane-major), zeroing the constrained rows on the way in.
__global__ void blk_pack_kernel(const double* __restrict__ in, double* __restrict__ out, int32_t n, int32_t R,
                                const uint8_t* __restrict__ mask) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * R) return;
    int32_t i = (int32_t)(t / R), j = (int32_t)(t % R);
    out[t] = (mask && mask[i]) ? 0.0 : in[(size_t)j * n + i];
}

__global__ void blk_unpack_kernel(const double* __restrict__ in, double* __restrict__ out, int32_t n, int32_t R) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * R) return;
    int32_t j = (int32_t)(t / n), i = (int32_t)(t % n);
    out[t] = in[(size_t)i * R + j];
}

}  // namespace

void blk_pack(const double* in, double* out, int32_t n, int32_t R, const uint8_t* mask, hipStream_t s) {
    int64_t tot = (int64_t)n * R;
    blk_pack_kernel<<<(unsigned)((tot + kBlock - 1) / kBlock), kBlock, 0, s>>>(in, out, n, R, mask);
    SVX_HIP(hipGetLastError());
}

void blk_unpack(const double* in, double* out, int32_t n, int32_t R, hipStream_t s) {
    int64_t tot = (int64_t)n * R;
    blk_unpack_kernel<<<(unsigned)((tot + kBlock - 1) / kBlock), kBlock, 0, s>>>(in, out, n, R);
    SVX_HIP(hipGetLastError());
}

void blk_init(const BlockArgs& a, hipStream_t s) {
    blk_init_kernel<<<a.n_tiles, kBlock, 0, s>>>(a);
    blk_flag_kernel<<<1, kWave, 0, s>>>(a, nullptr);
    SVX_HIP(hipGetLastError());
}

void blk_iteration(const BlockArgs& a, hipStream_t s) {
    blk_spmv_kernel<<<a.n_tiles, kBlock, 0, s>>>(a);
    blk_update_kernel<<<a.n_tiles, kBlock, 0, s>>>(a);
    blk_pupdate_kernel<<<a.n_tiles, kBlock, 0, s>>>(a);
    blk_flag_kernel<<<1, kWave, 0, s>>>(a, nullptr);
    SVX_HIP(hipGetLastError());
}

void blk_count_running(const BlockArgs& a, int32_t* out, hipStream_t s) {
    blk_flag_kernel<<<1, kWave, 0, s>>>(a, out);
    SVX_HIP(hipGetLastError());
}

void blk_spmv_only(const BlockArgs& a, hipStream_t s) {
    blk_spmv_kernel<<<a.n_tiles, kBlock, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
