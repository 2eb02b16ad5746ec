// Multi-RHS Jacobi-PCG: R <= 64 load cases on one shared matrix.
//
// Vectors are stored row-major [n][R]: lane j of a 64-wide wavefront owns load
// case j of whichever row the wavefront is on. The SpMV reads each row's
// values and column indices once (wave-uniform loads) and multiplies them
// with all R right-hand sides at once, so the matrix streams from HBM once per
// iteration for every load case. Each load case runs the recurrence of
// solver.rs:82-149 with its own alpha, beta, residual and convergence mask.
//
// One iteration is seven launches (six for the gathering SpMV):
//  1. the pending direction step p' = z + beta p with z = r / diag (p is
//     double-buffered: iteration parity `par` reads P[par] and writes p'
//     into P[1 - par]); the gathering SpMV evaluates it on the fly instead,
//     wherever a row of p' is needed, and the owner of a row stores it;
//  2. SpMV q = K p', partial p'.q per tile and load case;
//  3-4. a first-level reduction over groups of tiles, then one wavefront:
//     alpha, breakdown (solver.rs:118-123);
//  5. the update of u and r (z is not stored), partials r.r and r.z;
//  6-7. a first-level reduction, then one wavefront: residual, convergence,
//     budget and beta (solver.rs:126-148).
// Reductions are per load case in a fixed order (first-level groups of
// consecutive tiles with the group's waves taking tiles in turn, then the
// groups in order), with no atomics or fences; kernel boundaries order the
// data.
//
// SpMV forms: meshes with a group plan take the group SpMV
// (blk_spmv_group_kernel, one wavefront per group of nodes); otherwise the
// node-row gathering SpMV (one wavefront per node) or, for scalar systems,
// the row form.
#include "block.hpp"

namespace svx {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;  // rows per block pass

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

// Direction step of one row for this lane's load case: p' = z + beta p with
// z = r / diag, the values pupdate-style PCG stores (solver.rs:139-143).
__device__ __forceinline__ double dir_step(double beta, double p, double r, double d) { return fma(beta, p, r * d); }

// Per-load-case partials of this 256-thread block (waves summed in order) into tile's slot.
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

// Sum of the first-level sums of slot s for this lane's load case, groups in order.
__device__ __forceinline__ double l1_sum(const double* __restrict__ l1, int s, int lane) {
    const double* b = l1 + (size_t)s * kReduceGroups * kWave + lane;
    double v[kReduceGroups];
#pragma unroll
    for (int g = 0; g < kReduceGroups; ++g) v[g] = b[g * kWave];
    double acc = v[0];
#pragma unroll
    for (int g = 1; g < kReduceGroups; ++g) acc += v[g];
    return acc;
}

// First-level sums: group g (a workgroup) adds the tiles [g T / G, (g + 1) T / G)
// of each slot for every load case, its waves taking tiles in turn and then
// added in order. `gate` (when given) skips the work once no case runs.
__global__ __launch_bounds__(kBlock) void blk_reduce1_kernel(const double* __restrict__ partials, int n_tiles,
                                                              int nslots, double* __restrict__ l1,
                                                              const BlockScal* gate) {
    if (gate && !gate->any_running) return;
    __shared__ double sh[kWavesPerBlock][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave, g = blockIdx.x;
    const int t0 = (int)((int64_t)g * n_tiles / kReduceGroups), t1 = (int)((int64_t)(g + 1) * n_tiles / kReduceGroups);
    for (int s = 0; s < nslots; ++s) {
        const double* base = partials + (size_t)s * n_tiles * kWave + lane;
        double acc = 0.0;
        int t = t0 + w;
        for (; t + kWavesPerBlock < t1; t += 2 * kWavesPerBlock) {  // two loads in flight
            const double v0 = base[(size_t)t * kWave], v1 = base[(size_t)(t + kWavesPerBlock) * kWave];
            acc += v0;
            acc += v1;
        }
        if (t < t1) acc += base[(size_t)t * kWave];
        sh[w][lane] = acc;
        __syncthreads();
        if (w == 0) {
            double sum = sh[0][lane];
#pragma unroll
            for (int i = 1; i < kWavesPerBlock; ++i) sum += sh[i][lane];
            l1[((size_t)s * kReduceGroups + g) * kWave + lane] = sum;
        }
        __syncthreads();
    }
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
        A.P[k] = 0.0;  // with beta = 0 the first direction step gives p' = z (solver.rs:90)
        ff += fi * fi;
        rz += fi * zi;
    }
    write_rhs_partials(ff, rz, A.partials, tile, A.n_vtiles, 2);
}

// One wavefront: initial scalars per load case.
__global__ void blk_init_finish_kernel(BlockArgs A) {
    const int lane = threadIdx.x;
    const double ff = l1_sum(A.l1, 0, lane);
    const double rz = l1_sum(A.l1, 1, lane);
    const bool on = lane < A.R;
    if (on) {
        A.S->bnorm[lane] = sqrt(ff);
        A.S->rr[lane] = ff;
        A.S->rz[lane] = rz;
        A.S->beta[lane] = 0.0;
        A.S->it[lane] = 0;
        A.S->best[lane] = 1.0;  // ||f - K 0|| / ||f||
        A.S->status[lane] = (ff == 0.0) ? ST_CONVERGED : ST_RUNNING;  // solver.rs:93-95
        A.S->res[lane] = 0.0;
    }
    unsigned long long m = __ballot(on && ff != 0.0);
    if (lane == 0) A.S->any_running = m != 0ull;
}

// ---------------------------------------------------------------- group SpMV
// With a group plan an iteration takes the direction step in its own pass
// (blk_step_kernel: p' = z + beta p over the vector tiles, p double-buffered)
// and the group SpMV gathers p' alone: one row of p' per union column and
// dof, where evaluating the step on the fly would gather r and the previous
// direction (twice the vector loads, and vector-load return is what bounds
// the product).
__global__ __launch_bounds__(kBlock) void blk_step_kernel(BlockArgs A, int par) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    if (lane >= A.R) return;
    const double beta = A.S->status[lane] == ST_RUNNING ? A.S->beta[lane] : 0.0;
    const double* __restrict__ Pold = par ? A.P2 : A.P;
    double* __restrict__ Pnew = par ? A.P : A.P2;
    const double* __restrict__ Rv = A.Rv;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        const size_t k = (size_t)i * A.R + lane;
        Pnew[k] = dir_step(beta, Pold[k], Rv[k], A.dinv[i]);
    }
}

// Products of one union column's members OFF + i for the set bits i of the
// 4-bit mask M (their blocks consecutive at v, in ascending position): every
// block load is issued before the first product, so a column costs one
// scalar-load round trip per nibble of its member mask.
template <int BS, int G, int OFF, uint32_t M, typename T>
__device__ __forceinline__ void col_products(const T* __restrict__ v, const T (&p)[BS], T (&acc)[G][BS]) {
    constexpr int NM = __builtin_popcount(M);
    T kv[NM][BS * BS];
#pragma unroll
    for (int k = 0; k < NM; ++k)
#pragma unroll
        for (int e = 0; e < BS * BS; ++e) kv[k][e] = v[k * BS * BS + e];
    int k = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if ((M >> i) & 1u) {
#pragma unroll
            for (int a = 0; a < BS; ++a)
#pragma unroll
                for (int b = 0; b < BS; ++b) acc[OFF + i][a] += kv[k][a * BS + b] * p[b];
            ++k;
        }
    }
}

// Members OFF..OFF + 3 of one column (mask nibble m); then the block pointer
// moves past their blocks.
template <int BS, int G, int OFF, typename T>
__device__ __forceinline__ void col_nibble(const T* __restrict__& v, uint32_t m, const T (&p)[BS], T (&acc)[G][BS]) {
    switch (m) {
#define SVX_COL(M)                                 \
    case M:                                        \
        col_products<BS, G, OFF, M, T>(v, p, acc); \
        break;
        SVX_COL(1) SVX_COL(2) SVX_COL(3) SVX_COL(4) SVX_COL(5) SVX_COL(6) SVX_COL(7) SVX_COL(8)
        SVX_COL(9) SVX_COL(10) SVX_COL(11) SVX_COL(12) SVX_COL(13) SVX_COL(14) SVX_COL(15)
#undef SVX_COL
        default:
            break;
    }
    v += __builtin_popcount(m) * BS * BS;
}

// One union column with member mask m (groups of 4 or 8 nodes).
template <int BS, int G, typename T>
__device__ __forceinline__ void col_step(const T* __restrict__& v, uint32_t m, const T (&p)[BS], T (&acc)[G][BS]) {
    static_assert(G == 4 || G == 8, "groups of 4 or 8 nodes");
    col_nibble<BS, G, 0, T>(v, m & 15u, p, acc);
    if constexpr (G == 8) col_nibble<BS, G, 4, T>(v, m >> 4, p, acc);
}

template <int BS, typename T>
__device__ __forceinline__ void gather_rows(const T* __restrict__ Pd, int32_t node, int R, int ln, T (&p)[BS]) {
#pragma unroll
    for (int b = 0; b < BS; ++b) p[b] = Pd[(size_t)(node * BS + b) * R + ln];
}

// One wavefront per node group (consecutive nodes in Morton order, see
// build_blk_plan): it walks the union of the members' block columns in
// ascending order, gathers each column's rows of p' once (one load case per
// lane) and applies them to every member block in that column. Per block row
// the blocks are added in ascending column order, the order of the CSR row,
// so each row's sum is the gathering SpMV's. The union column, its member
// mask and the packed values are wave-uniform (scalar loads, one contiguous
// run of blocks per group); the gathers are branch-free (lanes past the last
// load case read that case's entries and discard them).
// T is the type of the values, the direction and q (double, or float for
// FP32-IR's inner iterations); p'.q is accumulated in FP64.
template <int BS, int G, typename T>
__device__ __forceinline__ double group_rows(int gi, int lane, int R, bool on, const int32_t* __restrict__ grp,
                                             const int32_t* __restrict__ gnodes, const int2* __restrict__ ucol,
                                             const T* __restrict__ pval, const T* __restrict__ Pd,
                                             T* __restrict__ Q) {
    const int32_t c0 = grp[4 * gi], c1 = grp[4 * gi + 1], t0 = grp[4 * gi + 2];
    const int ln = on ? lane : R - 1;
    T acc[G][BS];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int a = 0; a < BS; ++a) acc[g][a] = T(0);
    const T* __restrict__ v = pval + (size_t)t0 * BS * BS;
    // Software pipeline over the union columns, two per trip with two row
    // buffers: column c's products run while column c + 1's rows of p' and
    // column c + 2's entry are in flight. Past the group's end the last
    // column is loaded again, unused.
    int2 ua = ucol[c0], ub = ucol[min(c0 + 1, c1 - 1)];
    T pa[BS], pb[BS];
    gather_rows<BS, T>(Pd, ua.x, R, ln, pa);
    for (int32_t c = c0;; c += 2) {
        gather_rows<BS, T>(Pd, ub.x, R, ln, pb);
        const int2 ua2 = ucol[min(c + 2, c1 - 1)];
        col_step<BS, G, T>(v, (uint32_t)ua.y, pa, acc);
        if (c + 1 >= c1) break;
        gather_rows<BS, T>(Pd, ua2.x, R, ln, pa);
        const int2 ub2 = ucol[min(c + 3, c1 - 1)];
        col_step<BS, G, T>(v, (uint32_t)ub.y, pb, acc);
        if (c + 2 >= c1) break;
        ua = ua2;
        ub = ub2;
    }
    double pq = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int32_t node = gnodes[gi * G + g];
        if (node >= 0) {
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                const size_t o = (size_t)(node * BS + a) * R + ln;
                const T own = Pd[o];
                if (on) {
                    Q[o] = acc[g][a];
                    pq += (double)own * (double)acc[g][a];
                }
            }
        }
    }
    return pq;
}

// q = K p' over every group (values pval, direction Pd), partial p'.q per
// tile and load case.
template <int BS, int G, typename T>
__global__ __launch_bounds__(kBlock) void blk_spmv_group_kernel(BlockArgs A, const T* __restrict__ Pd,
                                                                  T* __restrict__ Q, const T* __restrict__ pval) {
    if (!A.S->any_running) return;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
    const int tile = spmv_tile(A.xcd);
    const int gi = tile * kWavesPerBlock + w;
    const bool on = lane < A.R;
    double pq = 0.0;
    if (gi < A.n_groups)
        pq = group_rows<BS, G, T>(gi, lane, A.R, on, A.grp, A.gnodes, (const int2*)A.ucol, pval, Pd, Q);
    write_rhs_partials(pq, 0.0, A.partials, tile, A.n_tiles, 1);
}

// Block values in group order: entry (a, b) of packed block t is
// val[src + a * stride + b], or, for a diagonal-only row (stride 0), its
// diagonal val[src + a] when a == b and 0 otherwise.
template <int BS, typename T>
__global__ void blk_pack_values_kernel(const double* __restrict__ val, const int32_t* __restrict__ src,
                                       const int32_t* __restrict__ stride, int64_t nblk, T* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk * BS * BS) return;
    const int64_t blk = t / (BS * BS);
    const int e = (int)(t % (BS * BS)), a = e / BS, b = e % BS;
    const int32_t st = stride[blk], s0 = src[blk];
    out[t] = (T)(st == 0 ? (a == b ? val[s0 + a] : 0.0) : val[(int64_t)s0 + (int64_t)a * st + b]);
}

// ------------------------------------------------------------ gathering SpMV
// Without a group plan: each gathered column's p' is evaluated from r and the
// previous direction on the fly (the owner of the row stores the same value).
__global__ __launch_bounds__(kBlock) void blk_spmv_kernel(BlockArgs A, int par) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = spmv_tile(A.xcd);
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const bool on = lane < A.R;
    const int R = A.R;
    const double beta = (on && A.S->status[lane] == ST_RUNNING) ? A.S->beta[lane] : 0.0;
    const double* __restrict__ Pold = par ? A.P2 : A.P;
    double* __restrict__ Pnew = par ? A.P : A.P2;
    auto pstep = [&](int32_t c) {
        const size_t o = (size_t)c * R + lane;
        return dir_step(beta, Pold[o], A.Rv[o], A.dinv[c]);
    };
    double pq = 0.0;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        double acc = 0.0;
        const int32_t kb = A.row_ptr[i], ke = A.row_ptr[i + 1];
        if (on) {
            int32_t k = kb;
            for (; k + 1 < ke; k += 2) {
                const int32_t c0 = A.col[k], c1 = A.col[k + 1];
                const double v0 = A.val[k], v1 = A.val[k + 1];
                const double p0 = pstep(c0), p1 = pstep(c1);
                acc += v0 * p0;
                acc += v1 * p1;
            }
            for (; k < ke; ++k) acc += A.val[k] * pstep(A.col[k]);
            const size_t o = (size_t)i * R + lane;
            const double own = pstep(i);
            Pnew[o] = own;
            A.Q[o] = acc;
            pq += own * acc;
        }
    }
    write_rhs_partials(pq, 0.0, A.partials, tile, A.n_tiles, 1);
}

template <int BS>
__global__ __launch_bounds__(kBlock) void blk_spmv_node_kernel(BlockArgs A, int par) {
    if (!A.S->any_running) return;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
    const int tile = spmv_tile(A.xcd);
    const int32_t r0 = tile * A.rows_per_tile, r1 = min(A.n, r0 + A.rows_per_tile);
    const int R = A.R;
    const bool on = lane < R;
    const double beta = (on && A.S->status[lane] == ST_RUNNING) ? A.S->beta[lane] : 0.0;
    const double* __restrict__ Pold = par ? A.P2 : A.P;
    double* __restrict__ Pnew = par ? A.P : A.P2;
    const double* __restrict__ Rv = A.Rv;
    const double* __restrict__ dinv = A.dinv;
    const double* __restrict__ val = A.val;
    const int32_t* __restrict__ row_ptr = A.row_ptr;
    const int32_t* __restrict__ brow_ptr = A.brow_ptr;
    const int32_t* __restrict__ bcol = A.bcol;
    auto pstep = [&](int32_t c) {
        const size_t o = (size_t)c * R + lane;
        return dir_step(beta, Pold[o], Rv[o], dinv[c]);
    };
    double pq = 0.0;
    for (int32_t node = r0 / BS + w; node < r1 / BS; node += kWavesPerBlock) {
        double acc[BS];
#pragma unroll
        for (int a = 0; a < BS; ++a) acc[a] = 0.0;
        const int32_t kb = brow_ptr[node], nb = brow_ptr[node + 1] - kb;
        const int32_t rs = row_ptr[node * BS];
        if (!on) continue;
        if (row_ptr[node * BS + BS] - rs != BS * BS * nb) {  // a node of no element: diagonal only
#pragma unroll
            for (int a = 0; a < BS; ++a) acc[a] = val[row_ptr[node * BS + a]] * pstep(node * BS + a);
        } else {
            const double* v0 = val + rs;
#pragma unroll 2
            for (int32_t k = 0; k < nb; ++k) {
                const int32_t j = bcol[kb + k];
                double p[BS];
#pragma unroll
                for (int b = 0; b < BS; ++b) p[b] = pstep(j * BS + b);
#pragma unroll
                for (int a = 0; a < BS; ++a)
#pragma unroll
                    for (int b = 0; b < BS; ++b) acc[a] += v0[a * BS * nb + BS * k + b] * p[b];
            }
        }
#pragma unroll
        for (int a = 0; a < BS; ++a) {
            const size_t o = (size_t)(node * BS + a) * R + lane;
            const double own = pstep(node * BS + a);
            Pnew[o] = own;
            A.Q[o] = acc[a];
            pq += own * acc[a];
        }
    }
    write_rhs_partials(pq, 0.0, A.partials, tile, A.n_tiles, 1);
}

// ------------------------------------------------------------- scalars, update
// One wavefront: alpha per load case, breakdown test (solver.rs:118-123).
__global__ void blk_alpha_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int lane = threadIdx.x;
    const double pq = l1_sum(A.l1, 0, lane);
    if (lane >= A.R || A.S->status[lane] != ST_RUNNING) return;
    if (pq <= 0.0 || pq != pq) {
        A.S->it[lane] += 1;
        A.S->res[lane] = sqrt(A.S->rr[lane]) / A.S->bnorm[lane];
        A.S->status[lane] = ST_BREAKDOWN;
        A.S->alpha[lane] = 0.0;
    } else {
        A.S->alpha[lane] = A.S->rz[lane] / pq;
    }
}

__global__ __launch_bounds__(kBlock) void blk_update_kernel(BlockArgs A, const double* __restrict__ Pn) {
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
            A.U[k] += alpha * Pn[k];
            double ri = A.Rv[k] - alpha * A.Q[k];
            A.Rv[k] = ri;
            double zi = ri * A.dinv[i];
            rr += ri * ri;
            rz += ri * zi;
        }
    write_rhs_partials(rr, rz, A.partials, tile, A.n_vtiles, 2);
}

// One wavefront: residual, convergence, budget and beta per load case; the
// running flag every kernel reads.
__global__ void blk_beta_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int lane = threadIdx.x;
    const double rr = l1_sum(A.l1, 0, lane);
    const double rz = l1_sum(A.l1, 1, lane);
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
    if (lane == 0) A.S->any_running = m != 0ull;
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

// ------------------------------------------- FP32-IR: FP32 phase, hand-over
// The FP32 phase runs the block iteration above on the FP32 packed values
// and FP32 vectors (x32 the correction since the last refresh, r32, p32 in
// two buffers, q32), with dots accumulated in FP64 (z = r / diag in FP32).
// A refresh folds x32 into the FP64 u, recomputes r = f - K u in FP64 and
// hands a load case over to FP64 (status ST_FALLBACK) once its true residual
// reaches tol or fails to halve its best so far; otherwise CG continues with
// the replaced residual (beta from its r.z). A breakdown also hands over
// (from the u and r of the last refresh). The FP32 phase never converges a
// case: the FP64 block iteration finishes every case (blk_warm_start).

__global__ __launch_bounds__(kBlock) void blk32_init_kernel(BlockArgs A) {
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    const bool on = lane < A.R;
    double ff = 0.0, rz = 0.0;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        if (!on) continue;
        const size_t k = (size_t)i * A.R + lane;
        const double fi = A.F[k];
        const float r = (float)fi;
        const float z = r * (float)A.dinv[i];
        A.U[k] = 0.0;
        A.Rv[k] = fi;
        A.X32[k] = 0.0f;
        A.R32[k] = r;
        A.P32[k] = 0.0f;  // with beta = 0 the first step gives p' = z
        ff += fi * fi;
        rz += (double)r * (double)z;
    }
    write_rhs_partials(ff, rz, A.partials, tile, A.n_vtiles, 2);
}

__global__ __launch_bounds__(kBlock) void blk32_step_kernel(BlockArgs A, int par) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    if (lane >= A.R) return;
    const float beta = A.S->status[lane] == ST_RUNNING ? (float)A.S->beta[lane] : 0.0f;
    const float* __restrict__ Pold = par ? A.P32b : A.P32;
    float* __restrict__ Pnew = par ? A.P32 : A.P32b;
    const float* __restrict__ R32 = A.R32;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        const size_t k = (size_t)i * A.R + lane;
        Pnew[k] = fmaf(beta, Pold[k], R32[k] * (float)A.dinv[i]);
    }
}

// One wavefront: alpha; a breakdown hands the case over to FP64.
__global__ void blk32_alpha_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int lane = threadIdx.x;
    const double pq = l1_sum(A.l1, 0, lane);
    if (lane >= A.R || A.S->status[lane] != ST_RUNNING) return;
    if (pq <= 0.0 || pq != pq) {
        A.S->status[lane] = ST_FALLBACK;
        A.S->alpha[lane] = 0.0;
    } else {
        A.S->alpha[lane] = A.S->rz[lane] / pq;
    }
}

__global__ __launch_bounds__(kBlock) void blk32_update_kernel(BlockArgs A, const float* __restrict__ Pn) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    const bool on = lane < A.R && A.S->status[lane] == ST_RUNNING;
    const float alpha = on ? (float)A.S->alpha[lane] : 0.0f;
    double rr = 0.0, rz = 0.0;
    if (on)
        for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
            const size_t k = (size_t)i * A.R + lane;
            A.X32[k] = fmaf(alpha, Pn[k], A.X32[k]);
            const float r = fmaf(-alpha, A.Q32[k], A.R32[k]);
            A.R32[k] = r;
            const float z = r * (float)A.dinv[i];
            rr += (double)r * (double)r;
            rz += (double)r * (double)z;
        }
    write_rhs_partials(rr, rz, A.partials, tile, A.n_vtiles, 2);
}

// One wavefront: iteration count, budget and beta (no convergence test on
// the FP32 recursive residual).
__global__ void blk32_beta_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int lane = threadIdx.x;
    const double rr = l1_sum(A.l1, 0, lane);
    const double rz = l1_sum(A.l1, 1, lane);
    bool run = false;
    if (lane < A.R && A.S->status[lane] == ST_RUNNING) {
        A.S->rr[lane] = rr;
        A.S->it[lane] += 1;
        if (A.S->it[lane] >= A.S->max_iter) {
            A.S->res[lane] = sqrt(rr) / A.S->bnorm[lane];
            A.S->status[lane] = ST_MAXITER;
        } else {
            A.S->beta[lane] = rz / A.S->rz[lane];
            A.S->rz_old[lane] = A.S->rz[lane];
            A.S->rz[lane] = rz;
            run = true;
        }
    }
    unsigned long long m = __ballot(run);
    if (lane == 0) A.S->any_running = m != 0ull;
}

// Refresh, first pass: u += x32, x32 = 0 for the running cases.
__global__ __launch_bounds__(kBlock) void blk32_fold_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    if (lane >= A.R || A.S->status[lane] != ST_RUNNING) return;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        const size_t k = (size_t)i * A.R + lane;
        A.U[k] += (double)A.X32[k];
        A.X32[k] = 0.0f;
    }
}

// Refresh, after q = K u in FP64: r = f - q (FP64 and FP32 copies) for the
// running cases, partials r.r (FP64) and r32.z32.
__global__ __launch_bounds__(kBlock) void blk32_residual_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    const bool on = lane < A.R && A.S->status[lane] == ST_RUNNING;
    double rr = 0.0, rz = 0.0;
    if (on)
        for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
            const size_t k = (size_t)i * A.R + lane;
            const double r = A.F[k] - A.Q[k];
            const float r32 = (float)r;
            const float z = r32 * (float)A.dinv[i];
            A.Rv[k] = r;
            A.R32[k] = r32;
            rr += r * r;
            rz += (double)r32 * (double)z;
        }
    write_rhs_partials(rr, rz, A.partials, tile, A.n_vtiles, 2);
}

// Refresh, one wavefront: true residual and the hand-over test per case.
__global__ void blk32_decide_kernel(BlockArgs A) {
    if (!A.S->any_running) return;
    const int lane = threadIdx.x;
    const double rr = l1_sum(A.l1, 0, lane);
    const double rz = l1_sum(A.l1, 1, lane);
    bool run = false;
    if (lane < A.R && A.S->status[lane] == ST_RUNNING) {
        const double res = sqrt(rr) / A.S->bnorm[lane];
        A.S->res[lane] = res;
        A.S->rr[lane] = rr;
        if (res <= A.S->tol || !(res <= 0.5 * A.S->best[lane])) {
            A.S->status[lane] = ST_FALLBACK;
        } else {
            A.S->best[lane] = res;
            A.S->beta[lane] = rz / A.S->rz_old[lane];
            A.S->rz[lane] = rz;
            run = true;
        }
    }
    unsigned long long m = __ballot(run);
    if (lane == 0) A.S->any_running = m != 0ull;
}

// FP64 set-up from the handed-over cases' u and r = f - K u: p = 0 (the
// first FP64 step gives p' = z), partials r.r and r.z.
__global__ __launch_bounds__(kBlock) void blk_warm_kernel(BlockArgs A) {
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tile = blockIdx.x;
    const int32_t r0 = tile * A.vrows_per_tile, r1 = min(A.n, r0 + A.vrows_per_tile);
    const bool on = lane < A.R;
    double rr = 0.0, rz = 0.0;
    for (int32_t i = r0 + w; i < r1; i += kWavesPerBlock) {
        if (!on) continue;
        const size_t k = (size_t)i * A.R + lane;
        const double ri = A.Rv[k];
        const double zi = ri * A.dinv[i];
        A.P[k] = 0.0;
        rr += ri * ri;
        rz += ri * zi;
    }
    write_rhs_partials(rr, rz, A.partials, tile, A.n_vtiles, 2);
}

// One wavefront: handed-over cases resume as FP64 iterations, or converge
// at once when their true residual is already at tol.
__global__ void blk_warm_finish_kernel(BlockArgs A) {
    const int lane = threadIdx.x;
    const double rr = l1_sum(A.l1, 0, lane);
    const double rz = l1_sum(A.l1, 1, lane);
    if (lane < A.R && A.S->status[lane] == ST_FALLBACK) {
        const double res = sqrt(rr) / A.S->bnorm[lane];
        A.S->rr[lane] = rr;
        A.S->rz[lane] = rz;
        A.S->beta[lane] = 0.0;
        A.S->res[lane] = res;
        A.S->status[lane] = res <= A.S->tol ? ST_CONVERGED : ST_RUNNING;
    }
    const int run = lane < A.R && A.S->status[lane] == ST_RUNNING;
    unsigned long long m = __ballot(run);
    if (lane == 0) A.S->any_running = m != 0ull;
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

// The SpMV kernel of iteration parity `par` (with a group plan, the product
// of the p' that blk_step_kernel wrote).
template <typename T>
void launch_group_product(const BlockArgs& a, const T* Pd, T* Q, const T* pval, hipStream_t s) {
    const int bs = a.bs, g = a.group;
    if (bs == 3 && g == 8)
        blk_spmv_group_kernel<3, 8, T><<<a.n_tiles, kBlock, 0, s>>>(a, Pd, Q, pval);
    else if (bs == 3)
        blk_spmv_group_kernel<3, 4, T><<<a.n_tiles, kBlock, 0, s>>>(a, Pd, Q, pval);
    else if (bs == 2 && g == 8)
        blk_spmv_group_kernel<2, 8, T><<<a.n_tiles, kBlock, 0, s>>>(a, Pd, Q, pval);
    else if (bs == 2)
        blk_spmv_group_kernel<2, 4, T><<<a.n_tiles, kBlock, 0, s>>>(a, Pd, Q, pval);
    else if (g == 8)
        blk_spmv_group_kernel<1, 8, T><<<a.n_tiles, kBlock, 0, s>>>(a, Pd, Q, pval);
    else
        blk_spmv_group_kernel<1, 4, T><<<a.n_tiles, kBlock, 0, s>>>(a, Pd, Q, pval);
}

void launch_blk_product(const BlockArgs& a, int par, hipStream_t s) {
    if (a.group > 0) {
        launch_group_product<double>(a, par ? a.P : a.P2, a.Q, a.pval, s);
    } else if (a.bs == 3) {
        blk_spmv_node_kernel<3><<<a.n_tiles, kBlock, 0, s>>>(a, par);
    } else if (a.bs == 2) {
        blk_spmv_node_kernel<2><<<a.n_tiles, kBlock, 0, s>>>(a, par);
    } else {
        blk_spmv_kernel<<<a.n_tiles, kBlock, 0, s>>>(a, par);
    }
}

void launch_blk_spmv(const BlockArgs& a, int par, hipStream_t s) {
    if (a.group > 0) blk_step_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a, par);
    launch_blk_product(a, par, s);
}

}  // namespace

namespace {
template <typename T>
void pack_values(const double* val, const int32_t* pk_src, const int32_t* pk_stride, int64_t nblk, int bs, T* pval,
                 hipStream_t s) {
    const int64_t tot = nblk * bs * bs;
    if (tot == 0) return;
    const unsigned grid = (unsigned)((tot + kBlock - 1) / kBlock);
    if (bs == 3)
        blk_pack_values_kernel<3, T><<<grid, kBlock, 0, s>>>(val, pk_src, pk_stride, nblk, pval);
    else if (bs == 2)
        blk_pack_values_kernel<2, T><<<grid, kBlock, 0, s>>>(val, pk_src, pk_stride, nblk, pval);
    else
        blk_pack_values_kernel<1, T><<<grid, kBlock, 0, s>>>(val, pk_src, pk_stride, nblk, pval);
    SVX_HIP(hipGetLastError());
}
}  // namespace

void blk_pack_values(const double* val, const int32_t* pk_src, const int32_t* pk_stride, int64_t nblk, int bs,
                     double* pval, hipStream_t s) {
    pack_values<double>(val, pk_src, pk_stride, nblk, bs, pval, s);
}

void blk_pack_values32(const double* val, const int32_t* pk_src, const int32_t* pk_stride, int64_t nblk, int bs,
                       float* pval, hipStream_t s) {
    pack_values<float>(val, pk_src, pk_stride, nblk, bs, pval, s);
}

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
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_vtiles, 2, a.l1, nullptr);
    blk_init_finish_kernel<<<1, kWave, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk_iteration(const BlockArgs& a, int it, hipStream_t s) {
    const int par = it & 1;
    launch_blk_spmv(a, par, s);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_tiles, 1, a.l1, a.S);
    blk_alpha_kernel<<<1, kWave, 0, s>>>(a);
    blk_update_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a, par ? a.P : a.P2);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_vtiles, 2, a.l1, a.S);
    blk_beta_kernel<<<1, kWave, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk32_init(const BlockArgs& a, hipStream_t s) {
    blk32_init_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_vtiles, 2, a.l1, nullptr);
    blk_init_finish_kernel<<<1, kWave, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk32_iteration(const BlockArgs& a, int it, hipStream_t s) {
    const int par = it & 1;
    blk32_step_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a, par);
    launch_group_product<float>(a, par ? a.P32 : a.P32b, a.Q32, a.pval32, s);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_tiles, 1, a.l1, a.S);
    blk32_alpha_kernel<<<1, kWave, 0, s>>>(a);
    blk32_update_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a, par ? a.P32 : a.P32b);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_vtiles, 2, a.l1, a.S);
    blk32_beta_kernel<<<1, kWave, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk32_refresh(const BlockArgs& a, hipStream_t s) {
    blk32_fold_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    launch_group_product<double>(a, a.U, a.Q, a.pval, s);
    blk32_residual_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_vtiles, 2, a.l1, a.S);
    blk32_decide_kernel<<<1, kWave, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk_warm_start(const BlockArgs& a, hipStream_t s) {
    blk_warm_kernel<<<a.n_vtiles, kBlock, 0, s>>>(a);
    blk_reduce1_kernel<<<kReduceGroups, kBlock, 0, s>>>(a.partials, a.n_vtiles, 2, a.l1, nullptr);
    blk_warm_finish_kernel<<<1, kWave, 0, s>>>(a);
    SVX_HIP(hipGetLastError());
}

void blk_count_running(const BlockArgs& a, int32_t* out, hipStream_t s) {
    blk_flag_kernel<<<1, kWave, 0, s>>>(a, out);
    SVX_HIP(hipGetLastError());
}

void blk_spmv_only(const BlockArgs& a, hipStream_t s) {
    launch_blk_spmv(a, 0, s);
    SVX_HIP(hipGetLastError());
}

void blk_product_only(const BlockArgs& a, hipStream_t s) {
    launch_blk_product(a, 0, s);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
