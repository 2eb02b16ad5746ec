// Multi-RHS Jacobi-PCG: R <= 64 load cases on one shared matrix.
//
// Vectors are stored row-major [n][R]: one 64-wide wavefront owns one matrix
// row and its lane j owns load case j. The SpMV reads each row's values and
// column indices once (wave-uniform loads) and multiplies them with R
// right-hand sides at once, so the matrix streams from HBM once per iteration
// for all load cases (block SpMV). Every load case runs the recurrence of
// solver.rs:82-149 with its own alpha, beta, residual and convergence mask.
//
// Plane and solid meshes use a node-row SpMV (one wavefront per node, each
// gathered row of P serving the node's ndof rows).
// One iteration is five launches: SpMV (partial p.q per tile and load case),
// a one-workgroup reduction (alpha, breakdown), the update of u, r, z
// (partials r.r, r.z), a second reduction (residual, convergence, beta) and
// the p update. Reductions are per load case in a fixed order (lane-local
// sums, tiles strided over the reduction's waves, waves in order), with no
// atomics or fences; the kernel boundaries order the data.
#include "block.hpp"

namespace svx {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;  // rows per block pass
constexpr int kReduceThreads = 1024;
constexpr int kReduceWaves = kReduceThreads / kWave;

// Tile of an SpMV workgroup: with xcd set, each of the eight XCDs (which the
// hardware deals workgroups to round robin) takes a contiguous run of tiles,
// so neighbouring nodes' rows of P are gathered through one L2. Tiles and
// their partial-sum columns are unchanged.
__device__ __forceinline__ int spmv_tile(int xcd) {
    const int n = (int)gridDim.x, b = (int)blockIdx.x;
    if (!xcd) return b;
    const int q = n / 8, r = n % 8, x = b % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

// Per-load-case partials of this block (waves summed in order) into tile's slot.
__device__ __forceinline__ void write_rhs_partials(double a, double b, double* partials, int tile, int n_tiles,
                                                   int nv) {
    __shared__ double sa[kWavesPerBlock][kWave], sb[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    sa[w][lane] = a;
    sb[w][lane] = b;
    __syncthreads();
    if (w == 0) {
        a = sa[0][lane], b = sb[0][lane];
#pragma unroll
        for (int i = 1; i < kWavesPerBlock; ++i) a += sa[i][lane], b += sb[i][lane];
        partials[(size_t)tile * kWave + lane] = a;
        if (nv > 1) partials[((size_t)n_tiles + tile) * kWave + lane] = b;
    }
}

// Sums over tiles of slot s (0 or 1) for every load case: wave w adds tiles
// w, w + 16, ... in turn; the 16 wave sums are then added in order.
__device__ __forceinline__ double reduce_slot(const double* __restrict__ partials, int n_tiles, int s) {
    __shared__ double sh[kReduceWaves][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const double* base = partials + (size_t)s * n_tiles * kWave;
    double acc = 0.0;
    int t = w;
    for (; t + 3 * kReduceWaves < n_tiles; t += 4 * kReduceWaves) {  // four loads in flight
        const double v0 = base[(size_t)t * kWave + lane];
        const double v1 = base[(size_t)(t + kReduceWaves) * kWave + lane];
        const double v2 = base[(size_t)(t + 2 * kReduceWaves) * kWave + lane];
        const double v3 = base[(size_t)(t + 3 * kReduceWaves) * kWave + lane];
        acc += v0;
        acc += v1;
        acc += v2;
        acc += v3;
    }
    for (; t < n_tiles; t += kReduceWaves) acc += base[(size_t)t * kWave + lane];
    sh[w][lane] = acc;
    __syncthreads();
    double sum = sh[0][lane];
#pragma unroll
    for (int i = 1; i < kReduceWaves; ++i) sum += sh[i][lane];
    __syncthreads();
    return sum;
}

__global__ __launch_bounds__(kBlock) void blk_init_kernel(BlockArgs A) {
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
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
    write_rhs_partials(ff, rz, A.partials, tile, A.n_vtiles, 2);
}

// One workgroup: initial scalars per load case.
__global__ __launch_bounds__(kReduceThreads) void blk_init_finish_kernel(BlockArgs A) {
    const double ff = reduce_slot(A.partials, A.n_vtiles, 0);
    const double rz = reduce_slot(A.partials, A.n_vtiles, 1);
    const int lane = threadIdx.x;
    if (lane >= kWave) return;
    const bool on = lane < A.R;
    if (on) {
        A.S->bnorm[lane] = sqrt(ff);
        A.S->rr[lane] = ff;
        A.S->rz[lane] = rz;
        A.S->it[lane] = 0;
        A.S->status[lane] = (ff == 0.0) ? ST_CONVERGED : ST_RUNNING;  // solver.rs:93-95
        A.S->res[lane] = 0.0;
    }
    unsigned long long m = __ballot(on && ff != 0.0);
    if (lane == 0) A.S->any_running = m != 0ull;
}

__global__ __launch_bounds__(kBlock) void blk_spmv_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = spmv_tile(A.xcd);
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const bool on = lane < A.R;
    const int R = A.R;
    double pq = 0.0;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        double acc = 0.0;
        const int32_t kb = A.row_ptr[i], ke = A.row_ptr[i + 1];
        if (on) {
            int32_t k = kb;
            for (; k + 3 < ke; k += 4) {  // four gathers in flight
                const int32_t c0 = A.col[k], c1 = A.col[k + 1], c2 = A.col[k + 2], c3 = A.col[k + 3];
                const double v0 = A.val[k], v1 = A.val[k + 1], v2 = A.val[k + 2], v3 = A.val[k + 3];
                const double p0 = A.P[(size_t)c0 * R + lane], p1 = A.P[(size_t)c1 * R + lane];
                const double p2 = A.P[(size_t)c2 * R + lane], p3 = A.P[(size_t)c3 * R + lane];
                acc += v0 * p0;
                acc += v1 * p1;
                acc += v2 * p2;
                acc += v3 * p3;
            }
            for (; k < ke; ++k) acc += A.val[k] * A.P[(size_t)A.col[k] * R + lane];
            size_t o = (size_t)i * R + lane;
            A.Q[o] = acc;
            pq += A.P[o] * acc;
        }
    }
    write_rhs_partials(pq, 0.0, A.partials, tile, A.n_tiles, 1);
}

// Node-row form of the block SpMV for plane / solid meshes: one wavefront
// owns the BS rows of one node, which share their columns, so each gathered
// row of P (one load case per lane) serves all BS rows; the values are read
// as node blocks of the CSR storage (row a of the node holds block k's
// entries at val[row_ptr[node*BS] + a*BS*nb + BS*k + b]). Each row still sums
// its entries in column order, so the results equal the row form's.
// The node is uniform over the wavefront and the arrays reach the loop as
// restrict-qualified parameters (the matrix is not written in this kernel),
// so the block columns and matrix values (the same for all 64 load cases)
// are scalar loads and the vector memory pipeline carries only the gathers
// of P and the stores of Q.
template <int BS>
__device__ __forceinline__ double blk_node_rows(int32_t n0, int32_t n1, int w, int lane, int R, bool on,
                                                const int32_t* __restrict__ row_ptr,
                                                const int32_t* __restrict__ brow_ptr,
                                                const int32_t* __restrict__ bcol, const double* __restrict__ val,
                                                const double* __restrict__ P, double* __restrict__ Q) {
    double pq = 0.0;
    for (int32_t node = n0 + w; node < n1; node += kWavesPerBlock) {
        double acc[BS];
#pragma unroll
        for (int a = 0; a < BS; ++a) acc[a] = 0.0;
        const int32_t kb = brow_ptr[node], nb = brow_ptr[node + 1] - kb;
        const int32_t rs = row_ptr[node * BS];
        if (on) {
            if (row_ptr[node * BS + BS] - rs != BS * BS * nb) {  // isolated node: diagonal entries only
#pragma unroll
                for (int a = 0; a < BS; ++a)
                    acc[a] = val[row_ptr[node * BS + a]] * P[(size_t)(node * BS + a) * R + lane];
            } else {
                const double* v0 = val + rs;
#pragma unroll 2
                for (int32_t k = 0; k < nb; ++k) {
                    const int32_t j = bcol[kb + k];
                    double p[BS];
#pragma unroll
                    for (int b = 0; b < BS; ++b) p[b] = P[(size_t)(j * BS + b) * R + lane];
#pragma unroll
                    for (int a = 0; a < BS; ++a)
#pragma unroll
                        for (int b = 0; b < BS; ++b) acc[a] += v0[a * BS * nb + BS * k + b] * p[b];
                }
            }
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const size_t o = (size_t)(node * BS + a) * R + lane;
                Q[o] = acc[a];
                pq += P[o] * acc[a];
            }
        }
    }
    return pq;
}

template <int BS>
__global__ __launch_bounds__(kBlock) void blk_spmv_node_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
    const int tile = spmv_tile(A.xcd);
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const double pq = blk_node_rows<BS>(r0 / BS, r1 / BS, w, lane, A.R, lane < A.R, A.row_ptr, A.brow_ptr, A.bcol,
                                        A.val, A.P, A.Q);
    write_rhs_partials(pq, 0.0, A.partials, tile, A.n_tiles, 1);
}

// One workgroup: alpha per load case, breakdown test (solver.rs:118-123).
__global__ __launch_bounds__(kReduceThreads) void blk_alpha_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const double pq = reduce_slot(A.partials, A.n_tiles, 0);
    const int lane = threadIdx.x;
    if (lane >= kWave || lane >= A.R || A.S->status[lane] != ST_RUNNING) return;
    if (pq <= 0.0 || pq != pq) {
        A.S->it[lane] += 1;
        A.S->res[lane] = sqrt(A.S->rr[lane]) / A.S->bnorm[lane];
        A.S->status[lane] = ST_BREAKDOWN;
        A.S->alpha[lane] = 0.0;
    } else {
        A.S->alpha[lane] = A.S->rz[lane] / pq;
    }
}

__global__ __launch_bounds__(kBlock) void blk_update_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
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
    write_rhs_partials(rr, rz, A.partials, tile, A.n_vtiles, 2);
}

// One workgroup: residual, convergence, budget and beta per load case; the
// running flag every kernel reads.
__global__ __launch_bounds__(kReduceThreads) void blk_beta_kernel(BlockArgs A, int32_t* running_out) {
    if (!A.S->any_running) {
        if (threadIdx.x == 0 && running_out) *running_out = 0;
        return;
    }
    const double rr = reduce_slot(A.partials, A.n_vtiles, 0);
    const double rz = reduce_slot(A.partials, A.n_vtiles, 1);
    const int lane = threadIdx.x;
    if (lane >= kWave) return;
    bool run = false;
    if (lane < A.R && A.S->status[lane] == ST_RUNNING) {
        A.S->rr[lane] = rr;
        A.S->it[lane] += 1;
        double res = sqrt(rr) / A.S->bnorm[lane];
        A.S->res[lane] = res;
        if (res <= A.S->tol) {  // solver.rs:131-134
            A.S->status[lane] = ST_CONVERGED;
        } else {
            A.S->beta[lane] = rz / A.S->rz[lane];
            A.S->rz[lane] = rz;
            if (A.S->it[lane] >= A.S->max_iter)
                A.S->status[lane] = ST_MAXITER;
            else
                run = true;
        }
    }
    unsigned long long m = __ballot(run);
    if (lane == 0) {
        A.S->any_running = m != 0ull;
        if (running_out) *running_out = __popcll(m);
    }
}

__global__ __launch_bounds__(kBlock) void blk_pupdate_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    const bool on = lane < A.R && A.S->status[lane] == ST_RUNNING;
    if (!on) return;
    const double beta = A.S->beta[lane];
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        size_t k = (size_t)i * A.R + lane;
        A.P[k] = A.Z[k] + beta * A.P[k];
    }
}

// One lane per load case: the number of running load cases.
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
    blk_init_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    blk_init_finish_kernel<<<1, kReduceThreads, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void launch_blk_spmv(const BlockArgs& a, hipStream_t s) {
    if (a.bs == 3)
        blk_spmv_node_kernel<3><<<a.n_tiles, kBlock, 0, s>>>(a);
    else if (a.bs == 2)
        blk_spmv_node_kernel<2><<<a.n_tiles, kBlock, 0, s>>>(a);
    else
        blk_spmv_kernel<<<a.n_tiles, kBlock, 0, s>>>(a);
}

void blk_iteration(const BlockArgs& a, hipStream_t s) {
    launch_blk_spmv(a, s);
    blk_alpha_kernel<<<1, kReduceThreads, 0, s>>>(a);
    blk_update_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    blk_beta_kernel<<<1, kReduceThreads, 0, s>>>(a, nullptr);
    blk_pupdate_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk_count_running(const BlockArgs& a, int32_t* out, hipStream_t s) {
    blk_flag_kernel<<<1, kWave, 0, s>>>(a, out);
    SVX_HIP(hipGetLastError());
}

void blk_spmv_only(const BlockArgs& a, hipStream_t s) {
    launch_blk_spmv(a, s);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
