// SpMV and Jacobi-PCG iteration kernels on gfx950.
//
// All Krylov kernels work on a *set* of instances (one system, or hundreds
// of uneven ones): the launch grid is a tile table, each workgroup owns one
// contiguous row range of one instance. A finished instance costs one scalar
// load per tile. Dot products are reduced deterministically: a fixed
// in-block tree, then the last workgroup of the instance (last-block-done
// counter) sums the per-tile partials in tile order. Results therefore do
// not depend on scheduling, nor on which other instances share the launch.
//
// FP64 arm: the recurrence of solver.rs:82-149 (u, r, z, p, alpha, beta,
// breakdown and budget tests), CSR SpMV with `lanes` lanes per row.
// FP32-IR arm: the same recurrence on an FP32 node-block (BSR) copy of K and
// FP32 vectors with FP64-accumulated dots; every `period` iterations the
// correction is folded into the FP64 solution and the residual is replaced by
// the true FP64 residual f - K u (reliable-update refinement). Convergence is
// declared only on that FP64 residual.
#include "svx_internal.hpp"

namespace svx {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Deterministic block sum of three values; results valid in thread 0.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c) {
    __shared__ double sa[kBlock / kWave], sb[kBlock / kWave], sc[kBlock / kWave];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) sa[w] = a, sb[w] = b, sc[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        a = sa[0], b = sb[0], c = sc[0];
#pragma unroll
        for (int i = 1; i < kBlock / kWave; ++i) a += sa[i], b += sb[i], c += sc[i];
    }
    __syncthreads();
}

__device__ __forceinline__ void block_sum2(double& a, double& b) {
    double c = 0.0;
    block_sum3(a, b, c);
}

// Publish this tile's partials (4 slots per tile); returns true in every
// thread of the last tile of the instance to arrive, which then sees all.
__device__ __forceinline__ bool publish_partials(double a, double b, double c, double* partials, int tile,
                                                 uint32_t* counter, int tile_count) {
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        partials[4 * tile] = a;
        partials[4 * tile + 1] = b;
        partials[4 * tile + 2] = c;
        __threadfence();
        uint32_t prev = atomicAdd(counter, 1u);
        s_last = (prev == (uint32_t)(tile_count - 1));
    }
    __syncthreads();
    bool last = s_last != 0;
    if (last) __threadfence();
    return last;
}
__device__ __forceinline__ bool publish_partials(double a, double b, double* partials, int tile, uint32_t* counter,
                                                 int tile_count) {
    return publish_partials(a, b, 0.0, partials, tile, counter, tile_count);
}

// Sum of the instance's partials in tile order (strided, then block tree).
__device__ __forceinline__ void reduce_partials(const double* partials, int tile_begin, int tile_count, double& a,
                                                double& b, double& c) {
    a = 0.0, b = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < tile_count; k += kBlock) {
        a += partials[4 * (tile_begin + k)];
        b += partials[4 * (tile_begin + k) + 1];
        c += partials[4 * (tile_begin + k) + 2];
    }
    block_sum3(a, b, c);
}
__device__ __forceinline__ void reduce_partials(const double* partials, int tile_begin, int tile_count, double& a,
                                                double& b) {
    double c;
    reduce_partials(partials, tile_begin, tile_count, a, b, c);
}

template <int LANES>
__device__ __forceinline__ double csr_row(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                          const double* __restrict__ val, const double* __restrict__ x, int32_t row,
                                          int lane) {
    double acc = 0.0;
    for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += LANES) acc += val[k] * x[col[k]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    return acc;
}

// One block row of the FP32 node-block matrix (values stored component-major:
// bval[(a*BS+b)*nblk + k]) times x; FP64 accumulation.
template <int BS, int LANES>
__device__ __forceinline__ void bsr_row(const int32_t* __restrict__ brp, const int32_t* __restrict__ bcol,
                                        const float* __restrict__ bval, int32_t nblk, const float* __restrict__ x,
                                        int32_t br, int lane, double acc[BS]) {
#pragma unroll
    for (int a = 0; a < BS; ++a) acc[a] = 0.0;
    for (int32_t k = brp[br] + lane; k < brp[br + 1]; k += LANES) {
        int32_t j = bcol[k];
        double xj[BS];
#pragma unroll
        for (int b = 0; b < BS; ++b) xj[b] = (double)x[j * BS + b];
#pragma unroll
        for (int a = 0; a < BS; ++a)
#pragma unroll
            for (int b = 0; b < BS; ++b) acc[a] += (double)bval[(int64_t)(a * BS + b) * nblk + k] * xj[b];
    }
#pragma unroll
    for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1) acc[a] += __shfl_xor(acc[a], off, LANES);
}

// ------------------------------------------------------------ plain SpMV
__global__ void spmv_ordered_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                    const double* __restrict__ val, const double* __restrict__ x,
                                    double* __restrict__ y) {
#pragma clang fp contract(off)
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double o = 0.0;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) o += val[k] * x[col[k]];  // serial, unfused
    y[r] = o;
}

template <int LANES>
__global__ __launch_bounds__(kBlock) void spmv_csr64_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ col,
                                                            const double* __restrict__ val,
                                                            const double* __restrict__ x, double* __restrict__ y) {
    const int lane = threadIdx.x % LANES;
    int64_t row = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / LANES;
    if (row >= n) return;
    double acc = csr_row<LANES>(rp, col, val, x, (int32_t)row, lane);
    if (lane == 0) y[row] = acc;
}

template <int BS, int LANES>
__global__ __launch_bounds__(kBlock) void spmv_bsr32_kernel(int64_t nb, const int32_t* __restrict__ brp,
                                                            const int32_t* __restrict__ bcol,
                                                            const float* __restrict__ bval,
                                                            const float* __restrict__ x, float* __restrict__ y) {
    const int lane = threadIdx.x % LANES;
    int64_t br = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / LANES;
    if (br >= nb) return;
    double acc[BS];
    bsr_row<BS, LANES>(brp, bcol, bval, brp[nb], x, (int32_t)br, lane, acc);
    if (lane < BS) {
        double v = acc[0];
#pragma unroll
        for (int a = 1; a < BS; ++a)
            if (lane == a) v = acc[a];
        y[br * BS + lane] = (float)v;
    }
}

// ------------------------------------------------------------ FP64 PCG
__global__ __launch_bounds__(kBlock) void cg64_init_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || !(S->flags & FL_FRESH) || (S->flags & FL_WARM)) return;
    double ff = 0.0, rz = 0.0;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        double fi = D.f[i];
        double zi = fi * D.dinv[i];
        D.u[i] = 0.0;
        D.r[i] = fi;
        D.z[i] = zi;
        D.p[i] = zi;
        ff += fi * fi;
        rz += fi * zi;
    }
    block_sum2(ff, rz);
    if (!publish_partials(ff, rz, partials, blockIdx.x, &S->cnt_a, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, ff, rz);
    if (threadIdx.x == 0) {
        S->cnt_a = 0;
        S->bnorm = sqrt(ff);
        S->rr = ff;
        S->rz = rz;
        S->it = 0;
        S->flags = 0;
        if (S->bnorm == 0.0) {  // solver.rs:93-95
            S->res = 0.0;
            S->status = ST_CONVERGED;
        }
    }
}

// FP64 recurrence started from the current u: r = f - K u, z = r / diag,
// p = z; the iteration count carries on from the FP32 phase.
template <int LANES>
__global__ __launch_bounds__(kBlock) void cg64_init_warm_kernel(const Tile* __restrict__ tiles,
                                                                const InstDesc* __restrict__ descs,
                                                                double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & (FL_FRESH | FL_WARM)) != (FL_FRESH | FL_WARM)) return;
    const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
    constexpr int G = kBlock / LANES;
    double rr = 0.0, rz = 0.0;
    for (int32_t row = t.row_begin + grp; row < t.row_end; row += G) {
        double ku = csr_row<LANES>(D.row_ptr, D.col, D.val, D.u, row, lane);
        if (lane == 0) {
            double ri = D.f[row] - ku;
            double zi = ri * D.dinv[row];
            D.r[row] = ri;
            D.z[row] = zi;
            D.p[row] = zi;
            rr += ri * ri;
            rz += ri * zi;
        }
    }
    block_sum2(rr, rz);
    if (!publish_partials(rr, rz, partials, blockIdx.x, &S->cnt_a, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, rr, rz);
    if (threadIdx.x == 0) {
        S->cnt_a = 0;
        S->rr = rr;
        S->rz = rz;
        S->flags = 0;
        S->res = sqrt(rr) / S->bnorm;
        if (S->res <= S->tol) S->status = ST_CONVERGED;
    }
}

template <int LANES>
__global__ __launch_bounds__(kBlock) void cg64_spmv_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & FL_FRESH)) return;
    const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
    constexpr int G = kBlock / LANES;
    double pq = 0.0;
    for (int32_t row = t.row_begin + grp; row < t.row_end; row += G) {
        double acc = csr_row<LANES>(D.row_ptr, D.col, D.val, D.p, row, lane);
        if (lane == 0) {
            D.q[row] = acc;
            pq += D.p[row] * acc;
        }
    }
    double dummy = 0.0;
    block_sum2(pq, dummy);
    if (!publish_partials(pq, 0.0, partials, blockIdx.x, &S->cnt_a, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, pq, dummy);
    if (threadIdx.x == 0) {
        S->cnt_a = 0;
        S->pap = pq;
        if (pq <= 0.0 || pq != pq) {  // solver.rs:118-123
            S->it += 1;
            S->res = sqrt(S->rr) / S->bnorm;
            S->status = ST_BREAKDOWN;
        } else {
            S->alpha = S->rz / pq;
        }
    }
}

__global__ __launch_bounds__(kBlock) void cg64_update_kernel(const Tile* __restrict__ tiles,
                                                             const InstDesc* __restrict__ descs,
                                                             double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & FL_FRESH)) return;
    const double alpha = S->alpha;
    double rr = 0.0, rz = 0.0;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        double pi = D.p[i];
        D.u[i] += alpha * pi;
        double ri = D.r[i] - alpha * D.q[i];
        D.r[i] = ri;
        double zi = ri * D.dinv[i];
        D.z[i] = zi;
        rr += ri * ri;
        rz += ri * zi;
    }
    block_sum2(rr, rz);
    if (!publish_partials(rr, rz, partials, blockIdx.x, &S->cnt_b, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, rr, rz);
    if (threadIdx.x == 0) {
        S->cnt_b = 0;
        S->rr = rr;
        S->it += 1;
        double res = sqrt(rr) / S->bnorm;
        S->res = res;
        if (res <= S->tol) {
            S->status = ST_CONVERGED;
        } else {
            S->beta = rz / S->rz;
            S->rz = rz;
            if (S->it >= S->max_iter) S->status = ST_MAXITER;
        }
    }
}

__global__ __launch_bounds__(kBlock) void cg64_pupdate_kernel(const Tile* __restrict__ tiles,
                                                              const InstDesc* __restrict__ descs) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & FL_FRESH)) return;
    const double beta = S->beta;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) D.p[i] = D.z[i] + beta * D.p[i];
}

// ------------------------------------------------------------ FP32-IR
// Reliable-update refinement (residual replacement): the FP32 loop runs the
// PCG recurrence; every `period` iterations, an instance whose recursive
// residual has fallen below ru_delta times its largest value since the last
// replacement (or below tol) folds its FP32 correction into the FP64 u,
// recomputes r = f - K u in FP64, and continues with p = z + beta p where z
// and beta come from the replaced residual. The solve ends when the FP64
// residual satisfies ||f - K u|| <= tol ||f|| + floor_scale || |K||u| + |f| ||,
// i.e. reaches tol or the rounding level at which FP64 can still resolve
// f - K u (the reference's own solutions sit at that level on ill-conditioned
// meshes).
enum : int32_t { FL_HOLD = 2 };
// A refresh at least this often, so an instance whose FP32 estimate never
// drops still gets its FP64 residual checked (and can fall back).
constexpr int64_t kForcedRefresh = 256;

__global__ __launch_bounds__(kBlock) void ir32_init_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || !(S->flags & FL_FRESH)) return;
    double ff = 0.0, rz = 0.0;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        double fi = D.f[i];
        float ri = (float)fi;
        float zi = ri * D.dinv32[i];
        D.u[i] = 0.0;
        D.x32[i] = 0.0f;
        D.r32[i] = ri;
        D.z32[i] = zi;
        D.p32[i] = zi;
        ff += fi * fi;
        rz += (double)ri * (double)zi;
    }
    block_sum2(ff, rz);
    if (!publish_partials(ff, rz, partials, blockIdx.x, &S->cnt_a, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, ff, rz);
    if (threadIdx.x == 0) {
        S->cnt_a = 0;
        S->bnorm = sqrt(ff);
        S->rr = ff;
        S->rz = rz;
        S->rz_prev = rz;
        S->rmax = S->bnorm;
        S->it = 0;
        S->refreshes = 0;
        S->flags = 0;
        S->floor = 0.0;
        S->res_best = 1.0;
        S->stall = 0;
        S->fp64_it = 0;
        S->it_refresh = 0;
        if (S->bnorm == 0.0) {
            S->res = 0.0;
            S->status = ST_CONVERGED;
        }
    }
}

template <int BS, int LANES>
__global__ __launch_bounds__(kBlock) void ir32_spmv_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & (FL_HOLD | FL_FRESH))) return;
    const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
    constexpr int G = kBlock / LANES;
    const int32_t nblk = D.brow_ptr[D.n / BS];
    double pq = 0.0;
    for (int32_t br = t.row_begin / BS + grp; br < t.row_end / BS; br += G) {
        double acc[BS];
        bsr_row<BS, LANES>(D.brow_ptr, D.bcol, D.bval, nblk, D.p32, br, lane, acc);
        if (lane < BS) {
            double v = acc[0];
#pragma unroll
            for (int a = 1; a < BS; ++a)
                if (lane == a) v = acc[a];
            float qv = (float)v;
            D.q32[br * BS + lane] = qv;
            pq += (double)D.p32[br * BS + lane] * (double)qv;
        }
    }
    double dummy = 0.0;
    block_sum2(pq, dummy);
    if (!publish_partials(pq, 0.0, partials, blockIdx.x, &S->cnt_a, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, pq, dummy);
    if (threadIdx.x == 0) {
        S->cnt_a = 0;
        S->pap = pq;
        if (pq <= 0.0 || pq != pq) {
            S->it += 1;
            S->res = sqrt(S->rr) / S->bnorm;
            S->status = ST_BREAKDOWN;
        } else {
            S->alpha = S->rz / pq;
        }
    }
}

__global__ __launch_bounds__(kBlock) void ir32_update_kernel(const Tile* __restrict__ tiles,
                                                             const InstDesc* __restrict__ descs,
                                                             double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & (FL_HOLD | FL_FRESH))) return;
    const float alpha = (float)S->alpha;
    double rr = 0.0, rz = 0.0;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        float pi = D.p32[i];
        D.x32[i] += alpha * pi;
        float ri = D.r32[i] - alpha * D.q32[i];
        D.r32[i] = ri;
        float zi = ri * D.dinv32[i];
        D.z32[i] = zi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    block_sum2(rr, rz);
    if (!publish_partials(rr, rz, partials, blockIdx.x, &S->cnt_b, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, rr, rz);
    if (threadIdx.x == 0) {
        S->cnt_b = 0;
        S->rr = rr;
        S->it += 1;
        double est = sqrt(rr);
        S->rmax = fmax(S->rmax, est);
        S->beta = rz / S->rz;
        S->rz_prev = S->rz;
        S->rz = rz;
        if (S->it >= S->max_iter) {
            S->flags |= FL_REFRESH | FL_HOLD;
        } else if (est / S->bnorm <= S->tol || est <= S->ru_delta * S->rmax ||
                   S->it - S->it_refresh >= kForcedRefresh) {
            S->flags |= FL_REFRESH;
        }
    }
}

// p = z + beta p. At the last iteration of a period an instance flagged for
// refresh skips this: its direction is rebuilt from the replaced residual.
__global__ __launch_bounds__(kBlock) void ir32_pupdate_kernel(const Tile* __restrict__ tiles,
                                                              const InstDesc* __restrict__ descs, int period_end) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || (S->flags & (FL_HOLD | FL_FRESH))) return;
    if (period_end && (S->flags & FL_REFRESH)) return;
    const float beta = (float)S->beta;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) D.p32[i] = D.z32[i] + beta * D.p32[i];
}

// Refresh step 1: fold the FP32 correction into the FP64 solution.
__global__ __launch_bounds__(kBlock) void ir32_fold_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || !(S->flags & FL_REFRESH)) return;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        D.u[i] += (double)D.x32[i];
        D.x32[i] = 0.0f;
    }
}

// Refresh step 2: r = f - K u in FP64 with its evaluation floor, replace
// r32 / z32, test convergence, set beta from the replaced residual.
template <int LANES>
__global__ __launch_bounds__(kBlock) void ir32_residual_kernel(const Tile* __restrict__ tiles,
                                                               const InstDesc* __restrict__ descs,
                                                               double* __restrict__ partials) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || !(S->flags & FL_REFRESH)) return;
    const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
    constexpr int G = kBlock / LANES;
    double rr = 0.0, rz = 0.0, fl = 0.0;
    for (int32_t row = t.row_begin + grp; row < t.row_end; row += G) {
        double ku = 0.0, aku = 0.0;
        for (int32_t k = D.row_ptr[row] + lane; k < D.row_ptr[row + 1]; k += LANES) {
            double v = D.val[k] * D.u[D.col[k]];
            ku += v;
            aku += fabs(v);
        }
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1) {
            ku += __shfl_xor(ku, off, LANES);
            aku += __shfl_xor(aku, off, LANES);
        }
        if (lane == 0) {
            double fi = D.f[row];
            double ri = fi - ku;
            float r32 = (float)ri;
            float z32 = r32 * D.dinv32[row];
            D.r32[row] = r32;
            D.z32[row] = z32;
            rr += ri * ri;
            rz += (double)r32 * (double)z32;
            double a = aku + fabs(fi);
            fl += a * a;
        }
    }
    block_sum3(rr, rz, fl);
    if (!publish_partials(rr, rz, fl, partials, blockIdx.x, &S->cnt_a, D.tile_count)) return;
    reduce_partials(partials, D.tile_begin, D.tile_count, rr, rz, fl);
    if (threadIdx.x == 0) {
        S->cnt_a = 0;
        double res = sqrt(rr) / S->bnorm;
        double floor = S->floor_scale * sqrt(fl) / S->bnorm;
        S->res = res;
        S->floor = floor;
        S->rr = rr;
        S->refreshes += 1;
        S->it_refresh = S->it;
        if (res < 0.5 * S->res_best) {
            S->res_best = res;
            S->stall = 0;
        } else {
            S->stall += 1;
        }
        if (res <= S->tol + floor) {
            S->status = ST_CONVERGED;
        } else if (S->flags & FL_HOLD) {
            S->status = ST_MAXITER;
        } else if (S->stall >= 4) {
            S->status = ST_FALLBACK;  // FP32 cannot resolve this system further
        } else {
            S->rmax = sqrt(rr);
            S->beta = rz / S->rz_prev;
            S->rz = rz;
        }
    }
}

// Refresh step 3: p = z + beta p for the refreshed instances; the last tile
// of each clears its refresh flag.
__global__ __launch_bounds__(kBlock) void ir32_repoint_kernel(const Tile* __restrict__ tiles,
                                                              const InstDesc* __restrict__ descs) {
    const Tile t = tiles[blockIdx.x];
    const InstDesc& D = descs[t.inst];
    InstScal* S = D.s;
    if (S->status != ST_RUNNING || !(S->flags & FL_REFRESH)) return;
    const float beta = (float)S->beta;
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) D.p32[i] = D.z32[i] + beta * D.p32[i];
    __syncthreads();
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&S->cnt_c, 1u) == (uint32_t)(D.tile_count - 1);
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        S->cnt_c = 0;
        S->flags = 0;
    }
}

__global__ void count_running_kernel(const InstDesc* __restrict__ descs, int32_t n_inst, int32_t* out) {
    __shared__ int32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int32_t local = 0;
    for (int32_t i = threadIdx.x; i < n_inst; i += blockDim.x) local += descs[i].s->status == ST_RUNNING;
    atomicAdd(&cnt, local);
    __syncthreads();
    if (threadIdx.x == 0) *out = cnt;
}

__global__ void gather_scalars_kernel(const InstDesc* __restrict__ descs, int32_t n_inst, InstScal* out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) out[i] = *descs[i].s;
}

__global__ void promote_fallback_kernel(const InstDesc* __restrict__ descs, int32_t n_inst) {
    for (int32_t i = threadIdx.x; i < n_inst; i += blockDim.x) {
        InstScal* S = descs[i].s;
        if (S->status == ST_FALLBACK) {
            S->status = ST_RUNNING;
            S->flags = FL_FRESH | FL_WARM;
            S->fp64_it = -S->it;  // completed when the FP64 phase ends
        }
    }
}

__global__ void set_running_kernel(const InstDesc* __restrict__ descs, int32_t n_inst) {
    for (int32_t i = threadIdx.x; i < n_inst; i += blockDim.x) {
        descs[i].s->status = ST_RUNNING;
        descs[i].s->flags = 0;
    }
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

#define SVX_LANES_DISPATCH(lanes, MACRO) \
    switch (lanes) {                     \
        case 2: MACRO(2); break;         \
        case 4: MACRO(4); break;         \
        case 8: MACRO(8); break;         \
        case 16: MACRO(16); break;       \
        case 32: MACRO(32); break;       \
        default: MACRO(64); break;       \
    }

}  // namespace

void launch_spmv_ordered(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* x,
                         double* y, hipStream_t s) {
    if (n <= 0) return;
    spmv_ordered_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, row_ptr, col, val, x, y);
    SVX_HIP(hipGetLastError());
}

void launch_spmv_csr64(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* x,
                       double* y, int lanes, hipStream_t s) {
    if (n <= 0) return;
#define L(LN) spmv_csr64_kernel<LN><<<grid_for(n * LN, kBlock), kBlock, 0, s>>>(n, row_ptr, col, val, x, y)
    SVX_LANES_DISPATCH(lanes, L)
#undef L
    SVX_HIP(hipGetLastError());
}

void launch_spmv_bsr32(int64_t nb, int bs, const int32_t* brow_ptr, const int32_t* bcol, const float* bval,
                       const float* x, float* y, int lanes, hipStream_t s) {
    if (nb <= 0) return;
    if (bs == 3) {
#define L(LN) spmv_bsr32_kernel<3, LN><<<grid_for(nb * LN, kBlock), kBlock, 0, s>>>(nb, brow_ptr, bcol, bval, x, y)
        SVX_LANES_DISPATCH(lanes, L)
#undef L
    } else {
#define L(LN) spmv_bsr32_kernel<2, LN><<<grid_for(nb * LN, kBlock), kBlock, 0, s>>>(nb, brow_ptr, bcol, bval, x, y)
        SVX_LANES_DISPATCH(lanes, L)
#undef L
    }
    SVX_HIP(hipGetLastError());
}

void cg64_init(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    cg64_init_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials);
#define L(LN) cg64_init_warm_kernel<LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
    SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    SVX_HIP(hipGetLastError());
}

void cg64_iteration(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
#define L(LN) cg64_spmv_kernel<LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
    SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    cg64_update_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials);
    cg64_pupdate_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs);
    SVX_HIP(hipGetLastError());
}

void ir32_init(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    ir32_init_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials);
    SVX_HIP(hipGetLastError());
}

void ir32_iteration(const KrylovSet& k, hipStream_t s, bool period_end) {
    if (k.n_tiles <= 0) return;
    if (k.bs == 3) {
#define L(LN) ir32_spmv_kernel<3, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.blanes, L)
#undef L
    } else if (k.bs == 2) {
#define L(LN) ir32_spmv_kernel<2, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.blanes, L)
#undef L
    } else {
#define L(LN) ir32_spmv_kernel<1, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.blanes, L)
#undef L
    }
    ir32_update_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials);
    ir32_pupdate_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, period_end ? 1 : 0);
    SVX_HIP(hipGetLastError());
}

void ir32_refresh(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    ir32_fold_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs);
#define L(LN) ir32_residual_kernel<LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
    SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    ir32_repoint_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs);
    SVX_HIP(hipGetLastError());
}

void launch_gather_scalars(const InstDesc* descs, int32_t n_inst, InstScal* out, hipStream_t s) {
    if (n_inst <= 0) return;
    gather_scalars_kernel<<<grid_for(n_inst, kBlock), kBlock, 0, s>>>(descs, n_inst, out);
    SVX_HIP(hipGetLastError());
}

void launch_promote_fallback(const InstDesc* descs, int32_t n_inst, hipStream_t s) {
    promote_fallback_kernel<<<1, kBlock, 0, s>>>(descs, n_inst);
    SVX_HIP(hipGetLastError());
}

void launch_set_running(const InstDesc* descs, int32_t n_inst, hipStream_t s) {
    set_running_kernel<<<1, kBlock, 0, s>>>(descs, n_inst);
    SVX_HIP(hipGetLastError());
}

void spmv_phase(const KrylovSet& k, int precision, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    if (precision == SVX_FP64) {
#define L(LN) cg64_spmv_kernel<LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    } else if (k.bs == 3) {
#define L(LN) ir32_spmv_kernel<3, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.blanes, L)
#undef L
    } else if (k.bs == 2) {
#define L(LN) ir32_spmv_kernel<2, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.blanes, L)
#undef L
    } else {
#define L(LN) ir32_spmv_kernel<1, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials)
        SVX_LANES_DISPATCH(k.blanes, L)
#undef L
    }
    SVX_HIP(hipGetLastError());
}

void launch_count_running(const InstDesc* descs, int32_t n_inst, int32_t* out, hipStream_t s) {
    count_running_kernel<<<1, kBlock, 0, s>>>(descs, n_inst, out);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
