// SpMV and Jacobi-PCG iteration kernels on gfx950.
//
// All Krylov kernels work on a *set* of instances (one system, or hundreds
// of uneven ones): the launch grid is a tile table, each workgroup owns one
// contiguous row range of one instance. A finished instance costs one scalar
// load per tile.
//
// Reductions use no atomics and no device-scope fences (on gfx950 an
// agent-scope release fence writes the XCD's L2 back; one per workgroup per
// phase costs more than the phase). A phase writes one partial sum per tile;
// every tile of the *next* phase re-reduces its instance's partials in tile
// order (fixed strided order, then a fixed in-block tree), so all tiles
// derive bit-identical scalars, and the kernel boundary orders the data. The
// instance's first tile (the leader) alone writes the live scalar state; the
// other tiles read only fields the leader does not write in that phase
// (InstScal::snap_*, configuration). Results therefore depend neither on
// scheduling nor on which other instances share the launch.
//
// FP64 arm: the recurrence of solver.rs:82-149 (u, r, z, p, alpha, beta,
// breakdown and budget tests). Plane and solid meshes multiply node rows (the
// ndof rows of one node share their column structure) as node blocks, with
// `lanes` lanes per node row; raw CSR input multiplies row by row.
// FP32-IR arm: the same recurrence on an FP32 node-block (BSR) copy of K and
// FP32 vectors with FP64-accumulated dots, with reliable-update refinement
// (residual replacement): every `period` iterations an instance whose
// recursive residual has fallen below ru_delta times its largest value since
// the last replacement (or below tol, or that has not been checked for
// kForcedRefresh iterations) folds its FP32 correction into the FP64 u,
// recomputes r = f - K u in FP64, and continues with p = z + beta p built from
// the replaced residual. FP64 cannot evaluate f - K u below the floor
// floor_scale || |K||u| + |f| || / ||f||, and on these ill-conditioned
// systems the norm of the true residual does not bound the error the way the
// reference's test does: FP64 PCG stops on its recursive residual <= tol
// (solver.rs:129-131) with its true residual at the floor and its smooth
// error components well below it. So the FP32 phase ends when its FP64 true
// residual is within the floor of tol (within kFloorSlack floors once a
// replacement failed to halve it), or below tol, and FP64 PCG
// finishes from the current u with the reference's recursive test, to
// min(tol, residual at the hand-over / 10): at least a decade of FP64 CG
// recursion, which reduces the error as the reference's last iterations do.
// An instance whose residual stops improving above that level (a
// replacement that does not halve the best residual so far) hands over the
// same way, to tol: past that point its FP32 iterations no longer shorten
// the FP64 finish. Measured on the 512-instance plane batch (FP64 phase on
// the matrix-free operator), with the same error against the oracle:
// handing over at the first such replacement instead of the fourth took the
// batched solve from 3.93 s to 3.19 s, and checking at least every 64
// iterations instead of 256 (so a stall is seen sooner) to 2.88 s; every 32
// iterations saves another 1-3% (batch, single C3D8 block and block solve).
#include <algorithm>

#include "mf_element.hpp"
#include "svx_internal.hpp"

namespace svx {
namespace {

constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;
constexpr int32_t FL_HOLD = 2;          // FP32-IR: budget reached, awaiting the final FP64 check
constexpr int64_t kForcedRefresh = 32;  // FP32-IR: check the FP64 residual at least this often
constexpr double kFloorSlack = 4.0;     // FP32-IR: floors accepted once the residual stagnates
constexpr int kStallLimit = 1;          // FP32-IR: refreshes without halving the best residual before FP64 takes over

// Tile of this workgroup in the gathering kernels (SpMV, FP64 residual, warm
// start, matrix-free product). The hardware deals workgroups out to the
// eight XCDs round robin (workgroup b runs on XCD b % 8), and each XCD has
// its own L2. Mapping workgroup b to tile (b % 8) * (n / 8) + b / 8 (the
// first n % 8 XCDs take one tile more) gives each XCD a contiguous run of
// tiles, so tiles whose rows gather each other's z and p (neighbouring
// chunks of one mesh, neighbouring node patches) share an L2. Tiles, their
// partial-sum columns and hence every sum are unchanged: only which XCD runs
// a tile moves.
// xcd = 0 keeps tile = workgroup (the launch shape's choice).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int xcd) {
    const int n = (int)gridDim.x, b = (int)blockIdx.x;
    if (!xcd) return b;
    const int q = n / kXcds, r = n % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// Instance arrays typed as global memory. The pointers come out of the
// descriptor table, so the compiler cannot infer their address space: as
// plain pointers they compile to flat loads, whose waits also cover LDS
// traffic and cannot retire one load ahead of the others (every column index
// would wait for the value stream issued beside it). Typed global, they
// compile to global loads with partial waits.
#define SVX_GLOBAL __attribute__((address_space(1)))
template <typename T>
using gptr = SVX_GLOBAL T*;
template <typename T>
__device__ __forceinline__ gptr<T> global(T* p) {
    return (gptr<T>)p;
}
// A load with the nontemporal hint (NT) or an ordinary one. Measured on the
// 265,923-DOF C3D8 block: a matrix stream with the hint no longer evicts the
// gathered vectors when it comes from HBM (36.2 us against 40.9 us per
// product), but it is not kept in the last-level cache either, so a matrix
// that fits there is read faster without it (29.0 us against 35.2 us).
template <bool NT, typename T>
__device__ __forceinline__ T load_hint(gptr<const T> p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
// The store counterpart. The update kernels stream their vectors with both
// hints once a launch's vectors exceed the last-level cache: on 16 M rows of
// five FP64 reads and three writes, 6.28 TB/s against 5.50 TB/s without
// hints and 5.6 TB/s with either hint alone.
template <bool NT, typename T>
__device__ __forceinline__ void store_hint(gptr<T> p, T v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

struct Dev {
    int32_t n;
    int32_t stepk;   // InstDesc::step_kernel
    int32_t alphak;  // InstDesc::alpha_kernel
    gptr<const int32_t> row_ptr, col, brow_ptr, bcol;
    gptr<const int32_t> ell_off, ell_nb, ell_col;
    gptr<const double> val, ell_val, dinv, f;
    gptr<const float> bval, ell_val32, dinv32;
    gptr<double> u, r, p, z, q, p_alt;
    gptr<float> x32, r32, p32, z32, q32, p32_alt;
};

__device__ __forceinline__ Dev view(const InstDesc& d) {
    Dev v;
    v.n = d.n;
    v.stepk = d.step_kernel;
    v.alphak = d.alpha_kernel;
    v.row_ptr = global(d.row_ptr);
    v.col = global(d.col);
    v.brow_ptr = global(d.brow_ptr);
    v.bcol = global(d.bcol);
    v.ell_off = global(d.ell_off);
    v.ell_nb = global(d.ell_nb);
    v.ell_col = global(d.ell_col);
    v.ell_val = global(d.ell_val);
    v.ell_val32 = global(d.ell_val32);
    v.val = global(d.val);
    v.dinv = global(d.dinv);
    v.f = global(d.f);
    v.bval = global(d.bval);
    v.dinv32 = global(d.dinv32);
    v.u = global(d.u);
    v.r = global(d.r);
    v.p = global(d.p);
    v.z = global(d.z);
    v.q = global(d.q);
    v.x32 = global(d.x32);
    v.r32 = global(d.r32);
    v.p32 = global(d.p32);
    v.z32 = global(d.z32);
    v.q32 = global(d.q32);
    v.p_alt = global(d.p_alt);
    v.p32_alt = global(d.p32_alt);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Deterministic block all-reduce of NV values: every thread receives the sums.
template <int NV>
__device__ __forceinline__ void block_allreduce(double (&v)[NV]) {
    __shared__ double sh[NV][kWaves];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = wave_sum(v[i]);
        if (lane == 0) sh[i][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = sh[i][0];
#pragma unroll
        for (int k = 1; k < kWaves; ++k) s += sh[i][k];
        v[i] = s;
    }
    __syncthreads();
}

// This tile's partials (block-reduced) into its slots. Partials are stored
// slot-major, partials[slot * stride + tile], so that one slot of an
// instance is a contiguous run of its tiles.
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], const int (&slot)[NV], double* partials, int stride,
                                               int tile) {
    block_allreduce<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[(size_t)slot[i] * stride + tile] = v[i];
    }
}

// The instance's sums of the given slots over its tiles, in tile order
// (thread k adds tiles k, k + kBlock, ... in turn; the first four loads of a
// thread are issued together).
template <int NV>
__device__ __forceinline__ void reduce_tiles(const double* __restrict__ partials, int stride, int tile_begin,
                                             int tile_count, const int (&slot)[NV], double (&out)[NV]) {
    constexpr int kAhead = 4;
    double buf[kAhead][NV];
#pragma unroll
    for (int j = 0; j < kAhead; ++j) {
        const int k = threadIdx.x + j * kBlock;
#pragma unroll
        for (int i = 0; i < NV; ++i) buf[j][i] = k < tile_count ? partials[(size_t)slot[i] * stride + tile_begin + k] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        out[i] = 0.0;
#pragma unroll
        for (int j = 0; j < kAhead; ++j) out[i] += buf[j][i];
    }
    for (int k = threadIdx.x + kAhead * kBlock; k < tile_count; k += kBlock) {
#pragma unroll
        for (int i = 0; i < NV; ++i) out[i] += partials[(size_t)slot[i] * stride + tile_begin + k];
    }
    block_allreduce<NV>(out);
}

// x_j as an SpMV reads it: a stored vector, or a direction step p' = z + beta p
// evaluated on the fly with one fused multiply-add in the vector precision
// (the owner of row j stores exactly this value).
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename V>
struct Plain {
    gptr<const V> x;
    __device__ __forceinline__ V operator()(int32_t i) const { return x[i]; }
};

template <typename V>
struct Step {
    gptr<const V> z, p;
    V beta;
    __device__ __forceinline__ V operator()(int32_t i) const { return fmadd(beta, p[i], z[i]); }
};

template <int LANES, typename X>
__device__ __forceinline__ double csr_row(gptr<const int32_t> rp, gptr<const int32_t> col, gptr<const double> val,
                                          const X& x, int32_t row, int lane) {
    double acc = 0.0;
    const int32_t kb = rp[row], ke = rp[row + 1];
#pragma unroll 4
    for (int32_t k = kb + lane; k < ke; k += LANES) acc += val[k] * (double)x(col[k]);
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    return acc;
}

// One row of the FP32 copy of a raw CSR matrix times x; FP64 accumulation.
template <int LANES, typename X>
__device__ __forceinline__ double csr32_row(gptr<const int32_t> rp, gptr<const int32_t> col, gptr<const float> val,
                                            const X& x, int32_t row, int lane) {
    double acc = 0.0;
    const int32_t kb = rp[row], ke = rp[row + 1];
#pragma unroll 4
    for (int32_t k = kb + lane; k < ke; k += LANES) acc += (double)val[k] * (double)x(col[k]);
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LANES);
    return acc;
}

// One node row of the node-row sliced ELL layout times x over the blocks
// k = k0, k0 + ks, k0 + 2 ks, ... < nb: cl / vl point at the row's first
// block column / value (chunk offset plus lane), nb is the chunk's padded
// block count. The loads of UNR blocks are issued together; the sums run in
// block order whatever UNR, so UNR changes timing only. The block columns of
// the next UNR blocks are loaded one group ahead, so a group's x gathers go
// out together with its value loads (one memory round trip per group, not
// two).
// ABS: also accumulate |K_ij x_j| (the FP64 evaluation floor of f - K x).
// NT: the matrix values are loaded nontemporal (read once per launch; they do
// not displace the gathered vectors from the caches).
template <int BS, int UNR, typename V, bool ABS, bool NT, typename X>
__device__ __forceinline__ void ell_row(int32_t nb, int32_t k0, int32_t ks, gptr<const int32_t> cl,
                                        gptr<const V> vl, const X& x, double (&acc)[BS], double (&aacc)[BS]) {
    constexpr int BB = BS * BS;
#pragma unroll
    for (int a = 0; a < BS; ++a) acc[a] = 0.0, aacc[a] = 0.0;
    int32_t k = k0;
    if (k + (UNR - 1) * ks < nb) {
        int32_t j[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) j[u] = cl[(k + u * ks) * kEllRows];
        for (;;) {
            V w[UNR][BB];
            double xj[UNR][BS];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
#pragma unroll
                for (int m = 0; m < BB; ++m) w[u][m] = load_hint<NT>(vl + ((k + u * ks) * BB + m) * kEllRows);
#pragma unroll
                for (int b = 0; b < BS; ++b) xj[u][b] = (double)x(j[u] * BS + b);
            }
            const int32_t kn = k + UNR * ks;
            const bool more = kn + (UNR - 1) * ks < nb;  // uniform: k, ks and nb are per wavefront
            if (more) {
#pragma unroll
                for (int u = 0; u < UNR; ++u) j[u] = cl[(kn + u * ks) * kEllRows];
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int a = 0; a < BS; ++a)
#pragma unroll
                    for (int b = 0; b < BS; ++b) {
                        const double t = (double)w[u][a * BS + b] * xj[u][b];
                        acc[a] += t;
                        if constexpr (ABS) aacc[a] += fabs(t);
                    }
            k = kn;
            if (!more) break;
        }
    }
    for (; k < nb; k += ks) {
        const int32_t jj = cl[k * kEllRows];
        V w[BB];
#pragma unroll
        for (int m = 0; m < BB; ++m) w[m] = load_hint<NT>(vl + (k * BB + m) * kEllRows);
        double xj[BS];
#pragma unroll
        for (int b = 0; b < BS; ++b) xj[b] = (double)x(jj * BS + b);
#pragma unroll
        for (int a = 0; a < BS; ++a)
#pragma unroll
            for (int b = 0; b < BS; ++b) {
                const double t = (double)w[a * BS + b] * xj[b];
                acc[a] += t;
                if constexpr (ABS) aacc[a] += fabs(t);
            }
    }
}

// (K x)[r] (and (|K||x|)[r] with ABS) for every node row r of the ELL chunks
// [c0, c1) of an instance with n_nodes node rows, handed to epi(r, acc, aacc)
// (lanes past the last node row are skipped). Lane l of a wavefront owns node
// row 64c + l; every load of a wavefront covers 64 consecutive words (values,
// columns) and the column gathers hit the caches. Four or more chunks: one
// chunk per wavefront at a time. One or two chunks (the tiles of a small
// instance, and the two-chunk tiles of a large assembled 3x3-block one):
// the S = 4 or 2 wavefronts of a chunk take every S-th block each
// and the owner adds their sums in LDS in wavefront order, so a small
// instance still spreads over all four wavefronts. The split is a function of
// the tile, hence of the instance alone.
template <int BS, int UNR, typename V, bool ABS, bool NT, typename X, typename E>
__device__ __forceinline__ void ell_tile(int32_t c0, int32_t c1, int32_t n_nodes, gptr<const int32_t> off,
                                         gptr<const int32_t> nbs, gptr<const int32_t> col, gptr<const V> val,
                                         const X& x, E&& epi) {
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int32_t nc = c1 - c0;
    double acc[BS], aacc[BS];
    if (nc > 2) {
        for (int32_t c = c0 + w; c < c1; c += kWaves) {
            const int64_t o = off[c];
            ell_row<BS, UNR, V, ABS, NT>(nbs[c], 0, 1, col + o + lane, val + o * (BS * BS) + lane, x, acc, aacc);
            const int32_t r = c * kEllRows + lane;
            if (r < n_nodes) epi(r, acc, aacc);
        }
        return;
    }
    constexpr int NS = ABS ? 2 * BS : BS;
    __shared__ double red[kWaves][NS][kWave];
    const int S = nc == 1 ? kWaves : kWaves / 2;
    const int32_t c = c0 + w / S;
    const int32_t sw = w % S;
    const int64_t o = off[c];
    ell_row<BS, UNR, V, ABS, NT>(nbs[c], sw, S, col + o + lane, val + o * (BS * BS) + lane, x, acc, aacc);
#pragma unroll
    for (int a = 0; a < BS; ++a) {
        red[w][a][lane] = acc[a];
        if constexpr (ABS) red[w][BS + a][lane] = aacc[a];
    }
    __syncthreads();
    if (sw == 0) {
        for (int t = 1; t < S; ++t)
#pragma unroll
            for (int a = 0; a < BS; ++a) {
                acc[a] += red[w + t][a][lane];
                if constexpr (ABS) aacc[a] += red[w + t][BS + a][lane];
            }
        const int32_t r = c * kEllRows + lane;
        if (r < n_nodes) epi(r, acc, aacc);
    }
}

// q = K p' over the ELL chunks [c0, c1), where p' is the direction x(.)
// evaluates (the stored p, or z + beta p); p' of the own rows is stored into
// nxt. Returns this thread's share of p'.q.
template <int BS, int UNR, typename V, bool NT, typename X>
__device__ __forceinline__ double ell_rows(int32_t c0, int32_t c1, int32_t n_nodes, gptr<const int32_t> off,
                                           gptr<const int32_t> nbs, gptr<const int32_t> col, gptr<const V> val,
                                           const X& x, gptr<V> nxt, gptr<V> q) {
    double dot = 0.0;
    ell_tile<BS, UNR, V, false, NT>(c0, c1, n_nodes, off, nbs, col, val, x,
                                [&](int32_t r, const double (&acc)[BS], const double (&)[BS]) {
#pragma unroll
                                    for (int a = 0; a < BS; ++a) {
                                        const V pa = x(r * BS + a);
                                        const V qa = (V)acc[a];
                                        nxt[r * BS + a] = pa;
                                        q[r * BS + a] = qa;
                                        dot += (double)pa * (double)qa;
                                    }
                                });
    return dot;
}

// The ELL chunks of a tile (tiles of node-block instances are whole chunks).
template <int BS>
__device__ __forceinline__ void tile_chunks(const Tile& t, int32_t& c0, int32_t& c1) {
    constexpr int32_t cr = kEllRows * BS;
    c0 = t.row_begin / cr;
    c1 = (t.row_end + cr - 1) / cr;
}

__device__ __forceinline__ bool is_leader(const Tile& t) { return t.row_begin == 0 && threadIdx.x == 0; }

// An empty asm statement taking uniform values as operands: their loads
// complete before it, so loads written ahead of an instance's status test
// are issued together with the test's own loads instead of being sunk past
// the early return.
template <typename T>
__device__ __forceinline__ int issue_one(const T& v) {
    asm volatile("" ::"s"(v));
    return 0;
}
template <typename... T>
__device__ __forceinline__ void issue_before_test(const T&... v) {
    const int done[] = {issue_one(v)...};
    (void)done;
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
    double acc = csr_row<LANES>(global(rp), global(col), global(val), Plain<double>{global(x)}, (int32_t)row, lane);
    if (lane == 0) y[row] = acc;
}

// ------------------------------------------------------------ initialisation
// Cold start (u = 0), FP64 vectors.
__global__ __launch_bounds__(kBlock) void cg64_init_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials, int pstride) {
    const Tile t = tiles[blockIdx.x];
    const Dev D = view(descs[t.inst]);
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || (S->flags & (FL_FRESH | FL_WARM)) != FL_FRESH) return;
    double v[2] = {0.0, 0.0};
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        double fi = D.f[i];
        double zi = fi * D.dinv[i];
        D.u[i] = 0.0;
        D.r[i] = fi;
        D.z[i] = zi;
        D.p[i] = zi;
        v[0] += fi * fi;
        v[1] += fi * zi;
    }
    const int slot[2] = {P_IFF, P_IRZ};
    write_partials<2>(v, slot, partials, pstride, blockIdx.x);
}

// Warm start: FP64 recurrence from the current u, r = f - K u. BS = 1: CSR
// rows with W lanes each; BS = 2, 3: ELL node rows, W blocks in flight.
template <int BS, int W>
__global__ __launch_bounds__(kBlock) void cg64_init_warm_kernel(const Tile* __restrict__ tiles,
                                                                const InstDesc* __restrict__ descs,
                                                                double* __restrict__ partials, int pstride, int xcd) {
    const Tile t = tiles[xcd_tile(xcd)];
    const Dev D = view(descs[t.inst]);
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || (S->flags & (FL_FRESH | FL_WARM)) != (FL_FRESH | FL_WARM)) return;
    double v[2] = {0.0, 0.0};
    if constexpr (BS == 1) {
        const int lane = threadIdx.x % W, grp = threadIdx.x / W;
        constexpr int G = kBlock / W;
        for (int32_t row = t.row_begin + grp; row < t.row_end; row += G) {
            double ku = csr_row<W>(D.row_ptr, D.col, D.val, Plain<double>{D.u}, row, lane);
            if (lane == 0) {
                double ri = D.f[row] - ku;
                double zi = ri * D.dinv[row];
                D.r[row] = ri;
                D.z[row] = zi;
                D.p[row] = zi;
                v[0] += ri * ri;
                v[1] += ri * zi;
            }
        }
    } else {
        int32_t c0, c1;
        tile_chunks<BS>(t, c0, c1);
        ell_tile<BS, W, double, false, false>(c0, c1, D.n / BS, D.ell_off, D.ell_nb, D.ell_col, D.ell_val,
                                       Plain<double>{D.u},
                                       [&](int32_t r, const double (&ku)[BS], const double (&)[BS]) {
#pragma unroll
                                           for (int a = 0; a < BS; ++a) {
                                               const int32_t i = r * BS + a;
                                               const double ri = D.f[i] - ku[a];
                                               const double zi = ri * D.dinv[i];
                                               D.r[i] = ri;
                                               D.z[i] = zi;
                                               D.p[i] = zi;
                                               v[0] += ri * ri;
                                               v[1] += ri * zi;
                                           }
                                       });
    }
    const int slot[2] = {P_IFF, P_IRZ};
    write_partials<2>(v, slot, partials, pstride, xcd_tile(xcd));
}

// Cold start (u = 0), FP32-IR vectors.
__global__ __launch_bounds__(kBlock) void ir32_init_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials, int pstride) {
    const Tile t = tiles[blockIdx.x];
    const Dev D = view(descs[t.inst]);
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || (S->flags & (FL_FRESH | FL_WARM)) != FL_FRESH) return;
    double v[2] = {0.0, 0.0};
    for (int32_t i = t.row_begin + threadIdx.x; i < t.row_end; i += kBlock) {
        double fi = D.f[i];
        float ri = (float)fi;
        float zi = ri * D.dinv32[i];
        D.u[i] = 0.0;
        D.x32[i] = 0.0f;
        D.r32[i] = ri;
        D.z32[i] = zi;
        D.p32[i] = zi;
        v[0] += fi * fi;
        v[1] += (double)ri * (double)zi;
    }
    const int slot[2] = {P_IFF, P_IRZ};
    write_partials<2>(v, slot, partials, pstride, blockIdx.x);
}

// Leader tile only: scalars of the freshly initialised instances.
__global__ __launch_bounds__(kBlock) void init_finish_kernel(const Tile* __restrict__ tiles,
                                                             const InstDesc* __restrict__ descs,
                                                             const double* __restrict__ partials, int pstride) {
    const Tile t = tiles[blockIdx.x];
    if (t.row_begin != 0) return;
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || !(S->flags & FL_FRESH)) return;
    const int slot[2] = {P_IFF, P_IRZ};
    double v[2];
    reduce_tiles<2>(partials, pstride, t.tile_begin, t.tile_count, slot, v);
    if (threadIdx.x != 0) return;
    if (S->flags & FL_WARM) {  // v[0] = r . r of the warm residual
        S->rr = v[0];
        S->rz = v[1];
        S->res = sqrt(v[0]) / S->bnorm;
        S->flags = 0;
        S->dir = 1;  // p = z is a complete direction
        if (S->res <= S->tol) S->status = ST_CONVERGED, S->stop_rule = SVX_STOP_TRUE;  // r = f - K u in FP64
        return;
    }
    S->bnorm = sqrt(v[0]);
    S->rr = v[0];
    S->rz = v[1];
    S->it = 0;
    S->rmax = S->bnorm;
    S->res_best = HUGE_VAL;  // the first refresh sets the reference (CG residuals may first grow above 1)
    S->stall = 0;
    S->refreshes = 0;
    S->it_refresh = 0;
    S->fp64_it = 0;
    S->evals = 0;
    S->stop_rule = 0;
    S->fb_reason = 0;
    S->floor = 0.0;
    S->flags = 0;
    S->dir = 1;  // p = z is a complete direction
    if (S->bnorm == 0.0) {  // solver.rs:93-95
        S->res = 0.0;
        S->status = ST_CONVERGED;
    }
}

// ------------------------------------------------------------ FP64 PCG
// The instance arrays reach the row loops as restrict-qualified parameters:
// q, r, u, z and p are distinct arrays, and saying so lets the compiler issue
// every load of a row before its stores and keep stores out of the waits.

// q = K p' over CSR rows [b0, b1) of a raw CSR instance, p' as x(.) evaluates
// it, stored into nxt for the own rows; this thread's share of p'.q.
template <int LANES, typename V, typename X>
__device__ __forceinline__ double csr_rows(int32_t b0, int32_t b1, gptr<const int32_t> rp, gptr<const int32_t> col,
                                           gptr<const V> val, const X& x, gptr<V> nxt, gptr<V> q) {
    const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
    constexpr int G = kBlock / LANES;
    double dot = 0.0;
    for (int32_t row = b0 + grp; row < b1; row += G) {
        const V pr = x(row);  // dot operand, issued with the row's loads
        V qv;
        if constexpr (sizeof(V) == 8)
            qv = (V)csr_row<LANES>(rp, col, (gptr<const double>)val, x, row, lane);
        else
            qv = (V)csr32_row<LANES>(rp, col, (gptr<const float>)val, x, row, lane);
        if (lane == 0) {
            nxt[row] = pr;
            q[row] = qv;
            dot += (double)pr * (double)qv;
        }
    }
    return dot;
}

// q = K p' over a tile's rows with the direction x(.): CSR rows (BS = 1, W
// lanes each) or ELL node rows (BS = 2, 3, W blocks in flight).
template <int BS, int W, typename V, bool NT, typename X>
__device__ __forceinline__ double tile_spmv(const Tile& t, const Dev& D, gptr<const V> ell_val, gptr<const V> csr_val,
                                           const X& x, gptr<V> nxt, gptr<V> q) {
    if constexpr (BS == 1) {
        return csr_rows<W, V>(t.row_begin, t.row_end, D.row_ptr, D.col, csr_val, x, nxt, q);
    } else {
        int32_t c0, c1;
        tile_chunks<BS>(t, c0, c1);
        return ell_rows<BS, W, V, NT>(c0, c1, D.n / BS, D.ell_off, D.ell_nb, D.ell_col, ell_val, x, nxt, q);
    }
}

// Rows b + threadIdx.x + j * kBlock (j < kPre) of a tile are loaded before
// the tile reduces its instance's partials, so that the vector loads and the
// partial loads are in flight together; rows past b + kPre * kBlock (tiles
// longer than the default shapes) go through the row loops afterwards.
constexpr int kPre = 3;

template <typename T, bool NT = false>
__device__ __forceinline__ void preload(gptr<const T> a, int32_t b, int32_t e, T (&v)[kPre]) {
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
        const int32_t i = b + (int32_t)threadIdx.x + j * kBlock;
        v[j] = i < e ? load_hint<NT>(a + i) : T(0);
    }
}

// Direction step and q = K p of one FP64 iteration (solver.rs:100-148 in the
// order direction, product). The current direction is in P[par] (p, or
// p_alt for par = 1); the new one goes to P[1 - par], so the gathers of every
// tile read an array no tile writes in this phase.
// Pending step: the tile re-reduces r.r and r.z of the last update, tests
// convergence and the budget, and takes p' = z + beta p, evaluated with the
// same fused multiply-add at its own rows and at every gathered column.
// dir set (after an initialisation): the direction is complete, p' = p.
// force_dir (timing only): beta = 1, no reduction, no scalar writes.
// NT: nontemporal loads of the ELL values (KrylovSet::nt).
template <int BS, int W, bool NT>
__global__ __launch_bounds__(kBlock) void cg64_spmv_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials, int pstride, int par,
                                                           int force_dir, int xcd) {
    const Tile t = tiles[xcd_tile(xcd)];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags, dir = S->dir;
    const double sbeta = S->beta;
    const Dev D = view(descs[t.inst]);  // read with the state, so the loads overlap
    issue_before_test(D.p, D.p_alt, D.z, D.q, D.ell_off, D.ell_nb, D.ell_col, D.ell_val, par, force_dir, D.stepk);
    if (status != ST_RUNNING || (flags & FL_FRESH)) return;
    const gptr<const double> cur = par ? D.p_alt : D.p;
    const gptr<double> nxt = par ? D.p : D.p_alt;
    double v[1];
    if (!force_dir && dir) {
        v[0] = tile_spmv<BS, W, double, NT>(t, D, D.ell_val, D.val, Plain<double>{cur}, nxt, D.q);
    } else {
        double beta = 1.0;
        if (!force_dir && D.stepk) {
            beta = sbeta;  // cg64_mf_step_kernel took the pending step (convergence, budget, beta)
        } else if (!force_dir) {
            const double snap_rz = S->snap_rz;
            const int64_t snap_it = S->snap_it;
            const int sl[2] = {P_RR, P_RZ};
            double rr[2];
            reduce_tiles<2>(partials, pstride, t.tile_begin, t.tile_count, sl, rr);
            const double res = sqrt(rr[0]) / S->bnorm;
            const bool conv = res <= S->tol;
            const int64_t it = snap_it + 1;
            const bool budget = it >= S->max_iter;
            beta = rr[1] / snap_rz;
            if (is_leader(t)) {
                S->it = it;
                S->rr = rr[0];
                S->res = res;
                if (conv) {
                    S->status = ST_CONVERGED;
                } else {
                    S->beta = beta;
                    S->rz = rr[1];
                    if (budget) S->status = ST_MAXITER;
                }
            }
            if (conv || budget) return;
        }
        v[0] = tile_spmv<BS, W, double, NT>(t, D, D.ell_val, D.val, Step<double>{D.z, cur, beta}, nxt, D.q);
    }
    const int slot[1] = {P_PQ};
    write_partials<1>(v, slot, partials, pstride, xcd_tile(xcd));
}

// ------------------------------------------------ matrix-free plane stress
// q = K p' for CPS4 / CPS3 meshes of one material without the assembled
// matrix: K p' = sum_e K_e p'_e with K_e p'_e evaluated from the element
// formulas of element.rs:91-211 (the B matrices of cps3_b_matrix and quad_b,
// D of d_matrix, 2x2 Gauss points in the order of cps4_stiffness), i.e. the
// product the assembled values give, up to rounding. One workgroup per node
// patch (up to kBlock owned nodes, a run of the mesh's Morton order):
//  1. the patch's nodes (owned, then halo): coordinates and p' = z + beta p
//     into LDS (the owned nodes' p' is the value the tile stores);
//  2. one thread per incident element: its nodal forces f_e, kept in
//     registers until every thread has read the node arrays, then stored
//     over them in LDS;
//  3. one thread per owned node: q = the sum of its elements' f_e in element
//     order (a property of the mesh, so independent of the tiling and of the
//     other instances in the launch); constrained dofs take q = p', the
//     identity rows of the constrained matrix (solver.rs:43-62).
// Element work per patch is one (CPS4) or two (CPS3) element slots per
// thread.
struct MfDev {
    gptr<const MfPatch> patch;
    gptr<const int32_t> nodes;
    gptr<const uint64_t> econn;  // four 16-bit local node positions per element slot
    gptr<const uint16_t> adj;
    gptr<const double> xy;
    gptr<const uint8_t> mask;
    MfMaterial mat;
};

__device__ __forceinline__ MfDev mf_view(const InstDesc& d) {
    MfDev m;
    m.patch = global(d.mf_patch);
    m.nodes = global(d.mf_nodes);
    m.econn = global((const uint64_t*)d.mf_econn);
    m.adj = global(d.mf_adj);
    m.xy = global(d.mf_xy);
    m.mask = global(d.mf_mask);
    m.mat = MfMaterial{d.mf_c11, d.mf_c12, d.mf_c33, d.mf_t};
    return m;
}

// LDS of one patch: coordinates and p' of cap_nodes nodes; once every
// element has read them, the same bytes hold the nodal forces of cap_elem
// element slots plus a zero slot, component-major (the smaller footprint
// lets six CPS3 patches share a compute unit instead of four).
inline size_t mf_lds_bytes(int nen, int cap_nodes, int cap_elem) {
    return sizeof(double) * std::max(4 * (size_t)cap_nodes, 2 * (size_t)nen * (cap_elem + 1));
}

// One patch of q = K p' (the pending direction step evaluated by x); stores
// p' and q of the owned nodes; returns this thread's p'.q. NT: the patch's
// own lists (incidence codes, element connectivity, node list), read once
// per product, are loaded with the nontemporal hint.
template <int NEN, bool NT, typename X>
__device__ __forceinline__ double mf_patch_spmv(const MfDev& M, const MfPatch& P, const X& x, gptr<double> nxt,
                                                gptr<double> q, int cap_nodes, int cap_elem) {
    constexpr int MAXA = NEN == 4 ? 4 : 8;
    constexpr int ROUNDS = NEN == 4 ? 1 : 2;
    extern __shared__ double mf_lds[];
    double* sx = mf_lds;
    double* sy = sx + cap_nodes;
    double* v0 = sy + cap_nodes;
    double* v1 = v0 + cap_nodes;
    double* fe = mf_lds;  // written once the node arrays are no longer read
    const int fs = cap_elem + 1;
    const int tid = threadIdx.x;
    const bool owner = tid < P.n_owned;
    // the owned node's incidence codes, two 16-bit codes per register
    uint32_t code2[MAXA / 2];
#pragma unroll
    for (int k = 0; k < MAXA; k += 2) {
        const uint32_t lo = (k < P.n_adj && owner) ? (uint32_t)load_hint<NT>(M.adj + P.adj_off + k * kBlock + tid) : (uint32_t)(P.n_elem * 4);
        const uint32_t hi =
            (k + 1 < P.n_adj && owner) ? (uint32_t)load_hint<NT>(M.adj + P.adj_off + (k + 1) * kBlock + tid) : (uint32_t)(P.n_elem * 4);
        code2[k / 2] = lo | (hi << 16);
    }
    uint64_t ec[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int sl = tid + r * kBlock;
        ec[r] = sl < P.n_elem ? load_hint<NT>(M.econn + P.elem_off + sl) : 0ull;
    }
    // every global load of the staging phase is issued before its LDS
    // stores: the node list (clamped to the patch, so no branch), then the
    // coordinates and p' of both rounds
    constexpr int NR = kMfMaxNodes / kBlock;
    int32_t g[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) g[r] = load_hint<NT>(M.nodes + P.node_off + min(tid + r * kBlock, P.n_nodes - 1));
    double cx[NR], cy[NR], va[NR], vb[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        cx[r] = M.xy[2 * g[r]], cy[r] = M.xy[2 * g[r] + 1];
        va[r] = x(2 * g[r]), vb[r] = x(2 * g[r] + 1);
    }
    const uint32_t mk = owner ? M.mask[g[0]] : 0u;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int j = tid + r * kBlock;
        if (j < P.n_nodes) sx[j] = cx[r], sy[j] = cy[r], v0[j] = va[r], v1[j] = vb[r];
    }
    const double own0 = va[0], own1 = vb[0];
    const int32_t gown = g[0];
    __syncthreads();
    double f[ROUNDS][NEN][2];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int sl = tid + r * kBlock;
        if (sl < P.n_elem) {
            const int l[4] = {(int)(ec[r] & 0xffffu), (int)((ec[r] >> 16) & 0xffffu), (int)((ec[r] >> 32) & 0xffffu),
                              (int)(ec[r] >> 48)};
            double xe[NEN][2], ue[NEN][2];
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                xe[a][0] = sx[l[a]];
                xe[a][1] = sy[l[a]];
                ue[a][0] = v0[l[a]];
                ue[a][1] = v1[l[a]];
            }
            mf_forces<NEN>(xe, ue, M.mat, f[r]);
        }
    }
    __syncthreads();  // the node arrays are dead: the forces take their place
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int sl = tid + r * kBlock;
        if (sl < P.n_elem) {
#pragma unroll
            for (int a = 0; a < NEN; ++a) {
                fe[(2 * a) * fs + sl] = f[r][a][0];
                fe[(2 * a + 1) * fs + sl] = f[r][a][1];
            }
        }
    }
    if (tid < 2 * NEN) fe[tid * fs + P.n_elem] = 0.0;
    __syncthreads();
    double dot = 0.0;
    if (owner) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int k = 0; k < MAXA; ++k) {
            const uint32_t ck = (code2[k / 2] >> (16 * (k & 1))) & 0xffffu;
            const uint32_t sl = ck >> 2, cr = ck & 3u;  // the sentinel reads the zero slot
            a0 += fe[(2 * cr) * fs + sl];
            a1 += fe[(2 * cr + 1) * fs + sl];
        }
        const double q0 = (mk & 1u) ? own0 : a0;
        const double q1 = (mk & 2u) ? own1 : a1;
        nxt[2 * gown] = own0;
        nxt[2 * gown + 1] = own1;
        q[2 * gown] = q0;
        q[2 * gown + 1] = q1;
        dot = own0 * q0 + own1 * q1;
    }
    return dot;
}

// The pending direction step's scalars for the matrix-free instances, one
// workgroup per instance (its leader tile): r.r and r.z of the last update
// over the instance's row tiles, convergence and budget (solver.rs:126-131,
// 144-148), beta. The product kernel then reads beta and the status instead
// of reducing the partials in every patch tile: the reduction's latency is
// paid once per instance rather than at the start of every patch.
__global__ __launch_bounds__(kBlock) void cg64_mf_step_kernel(const Tile* __restrict__ leaders,
                                                              const double* __restrict__ partials, int pstride) {
    const Tile t = leaders[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || (S->flags & FL_FRESH) || S->dir) return;
    const double snap_rz = S->snap_rz;
    const int64_t snap_it = S->snap_it;
    const int sl[2] = {P_RR, P_RZ};
    double rr[2];
    reduce_tiles<2>(partials, pstride, t.tile_begin, t.tile_count, sl, rr);
    if (threadIdx.x != 0) return;
    const double res = sqrt(rr[0]) / S->bnorm;
    const int64_t it = snap_it + 1;
    S->it = it;
    S->rr = rr[0];
    S->res = res;
    if (res <= S->tol) {
        S->status = ST_CONVERGED;
    } else {
        S->beta = rr[1] / snap_rz;
        S->rz = rr[1];
        if (it >= S->max_iter) S->status = ST_MAXITER;
    }
}

// Direction step and q = K p of one FP64 iteration on the matrix-free
// operator (cg64_spmv_kernel with node patches for row tiles), after
// cg64_mf_step_kernel has tested convergence and set beta. Tile t is
// mf_tiles[base + xcd_tile(xcd)]; its p.q partial goes to slot P_MPQ, column
// base + xcd_tile(xcd). NT: the patch lists loaded with the nontemporal
// hint (KrylovSet::ntv).
template <int NEN, bool NT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void cg64_mf_spmv_kernel(const Tile* __restrict__ mf_tiles, int base,
                                                              const InstDesc* __restrict__ descs,
                                                              double* __restrict__ partials, int pstride, int par,
                                                              int force_dir, int cap_nodes, int cap_elem, int xcd) {
    const Tile t = mf_tiles[base + xcd_tile(xcd)];
    // the instance's state, its descriptor and the patch record are read
    // together, before the status test, so that their latencies overlap
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags, dir = S->dir;
    const double sbeta = S->beta;
    const InstDesc& dd = descs[t.inst];
    const Dev D = view(dd);
    const MfDev M = mf_view(dd);
    MfPatch P;
    {
        const gptr<const MfPatch> pp = M.patch + t.row_begin;
        P.node_off = pp->node_off, P.n_nodes = pp->n_nodes, P.n_owned = pp->n_owned, P.elem_off = pp->elem_off;
        P.n_elem = pp->n_elem, P.adj_off = pp->adj_off, P.n_adj = pp->n_adj;
    }
    issue_before_test(D.p, D.p_alt, D.z, D.q, M.nodes, M.econn, M.adj, M.xy, M.mask, P.node_off, P.n_nodes, P.n_owned,
                      P.elem_off, P.n_elem, P.adj_off, P.n_adj, par, force_dir, dir, sbeta, cap_nodes,
                      cap_elem, M.mat.c11, M.mat.c12, M.mat.c33, M.mat.t);
    if (status != ST_RUNNING || (flags & FL_FRESH)) return;
    const gptr<const double> cur = par ? D.p_alt : D.p;
    const gptr<double> nxt = par ? D.p : D.p_alt;
    double v[1];
    if (!force_dir && dir) {
        v[0] = mf_patch_spmv<NEN, NT>(M, P, Plain<double>{cur}, nxt, D.q, cap_nodes, cap_elem);
    } else {
        const double beta = force_dir ? 1.0 : sbeta;
        v[0] = mf_patch_spmv<NEN, NT>(M, P, Step<double>{D.z, cur, beta}, nxt, D.q, cap_nodes, cap_elem);
    }
    const int slot[1] = {P_MPQ};
    write_partials<1>(v, slot, partials, pstride, base + xcd_tile(xcd));
}

// ------------------------------------------------ matrix-free C3D8
// q = K p' for meshes of C3D8 elements of one material, in two passes:
//  1. cg64_mf3_elem_kernel: kMf3Elems elements per workgroup, eight threads
//     per element, one per Gauss point; the element's node coordinates and
//     p' = z + beta p staged in LDS; each thread evaluates its point's
//     detJ B^T D B p'_e (mf_hex_gp_forces), the element's nodal forces are
//     the sums over the points (a fixed cross-lane tree), stored node-major
//     in `fe` (node, rank of the element among the node's elements);
//  2. cg64_mf3_node_kernel, over the instance's row tiles: each node sums the
//     forces of its incident elements in element order (the element pass
//     stores each force in the node's slot for that element, so this pass
//     reads one contiguous run per node), constrained dofs take q = p', and
//     the tile stores p' and q and its p'.q partial in the slot the update
//     reads.
// Every element is evaluated once; the sums are in fixed orders, so an
// instance gives the same bits alone and in any batch.
__global__ __launch_bounds__(kMf3Elems * 8) void cg64_mf3_elem_kernel(const Tile* __restrict__ etiles,
                                                                      const InstDesc* __restrict__ descs, int par,
                                                                      int force_dir) {
    __shared__ double sx[kMf3Elems][8][3];
    __shared__ double su[kMf3Elems][8][3];
    const Tile t = etiles[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags, dir = S->dir;
    const double sbeta = S->beta;
    const InstDesc& dd = descs[t.inst];
    const gptr<const double> cur = global(par ? dd.p_alt : dd.p), z = global(dd.z), X = global(dd.mf_xy);
    if (status != ST_RUNNING || (flags & FL_FRESH)) return;
    const bool step = force_dir || !dir;
    const double beta = force_dir ? 1.0 : (step ? sbeta : 0.0);
    const int el = threadIdx.x >> 3, k = threadIdx.x & 7;
    const int32_t e = t.row_begin + el;
    const bool valid = e < t.row_end;
    int32_t g = 0;
    uint32_t rank = 0;
    if (valid) {  // every load of the node issued before the first LDS store
        g = global(dd.mf_conn)[8 * (int64_t)e + k];
        rank = global(dd.mf_rank)[8 * (int64_t)e + k];
        double xv[3], pv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int32_t i = 3 * g + c;
            xv[c] = X[i];
            pv[c] = cur[i];
            zv[c] = z[i];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sx[el][k][c] = xv[c];
            su[el][k][c] = step ? fmadd(beta, pv[c], zv[c]) : pv[c];
        }
    }
    __syncthreads();
    double f[8][3];
    if (valid) {
        mf_hex_gp_forces(k, sx[el], su[el], dd.mf_c11, dd.mf_c12, dd.mf_c33, f);
    } else {
#pragma unroll
        for (int n = 0; n < 8; ++n) f[n][0] = f[n][1] = f[n][2] = 0.0;
    }
    // Sum over the element's eight point threads, reduce-scatter across the
    // eight lanes (partners k ^ 4, k ^ 2, k ^ 1): lane k ends with node k's
    // forces, summed in a fixed tree that depends on the node only.
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double mine = (k & 4) ? f[n ^ 4][c] : f[n][c], other = (k & 4) ? f[n][c] : f[n ^ 4][c];
            f[n][c] = mine + __shfl_xor(other, 4, 8);
        }
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double mine = (k & 2) ? f[n ^ 2][c] : f[n][c], other = (k & 2) ? f[n][c] : f[n ^ 2][c];
            f[n][c] = mine + __shfl_xor(other, 2, 8);
        }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double mine = (k & 1) ? f[1][c] : f[0][c], other = (k & 1) ? f[0][c] : f[1][c];
        f[0][c] = mine + __shfl_xor(other, 1, 8);
    }
    if (valid) {  // node g's slot for this element
        const gptr<double> fe = global(dd.mf_fe) + ((int64_t)g * dd.mf_width + rank) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) fe[c] = f[0][c];
    }
}

__global__ __launch_bounds__(kBlock) void cg64_mf3_node_kernel(const Tile* __restrict__ tiles, int first,
                                                               const InstDesc* __restrict__ descs,
                                                               double* __restrict__ partials, int pstride, int par,
                                                               int force_dir) {
    const Tile t = tiles[first + blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags, dir = S->dir;
    const double sbeta = S->beta;
    const InstDesc& dd = descs[t.inst];
    const gptr<const double> cur = global(par ? dd.p_alt : dd.p), z = global(dd.z), fe = global(dd.mf_fe);
    if (status != ST_RUNNING || (flags & FL_FRESH)) return;
    const bool step = force_dir || !dir;
    const double beta = force_dir ? 1.0 : (step ? sbeta : 0.0);
    const gptr<double> nxt = global(par ? dd.p : dd.p_alt), q = global(dd.q);
    const gptr<const uint8_t> mask = global(dd.mf_mask);
    const int W = dd.mf_width;
    double v[1] = {0.0};
    for (int32_t i = t.row_begin / 3 + (int32_t)threadIdx.x; i < t.row_end / 3; i += kBlock) {
        double a[3] = {0.0, 0.0, 0.0};
        const gptr<const double> fi = fe + (int64_t)i * W * 3;
        for (int w = 0; w < W; ++w) {  // element order; unused slots hold zeros
            a[0] += fi[3 * w];
            a[1] += fi[3 * w + 1];
            a[2] += fi[3 * w + 2];
        }
        const uint32_t mk = mask[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int32_t r = 3 * i + c;
            const double pv = step ? fmadd(beta, cur[r], z[r]) : cur[r];
            const double qv = (mk >> c) & 1u ? pv : a[c];  // constrained dofs: identity rows
            nxt[r] = pv;
            q[r] = qv;
            v[0] += pv * qv;
        }
    }
    const int slot[1] = {P_PQ};
    write_partials<1>(v, slot, partials, pstride, first + blockIdx.x);
}

// u += alpha p, r -= alpha q, z = D^-1 r over [b, e); this thread's r.r, r.z.
template <bool NT>
__device__ __forceinline__ void cg64_update_rows(int32_t b, int32_t e, double alpha, gptr<const double> __restrict__ p,
                                                 gptr<const double> __restrict__ q,
                                                 gptr<const double> __restrict__ dinv, gptr<double> __restrict__ u,
                                                 gptr<double> __restrict__ r, gptr<double> __restrict__ z,
                                                 double (&v)[2]) {
    for (int32_t i = b + threadIdx.x; i < e; i += kBlock) {
        const double pi = load_hint<NT>(p + i), qi = load_hint<NT>(q + i), di = load_hint<NT>(dinv + i);
        const double ui = load_hint<NT>(gptr<const double>(u + i)), r0 = load_hint<NT>(gptr<const double>(r + i));
        store_hint<NT>(u + i, ui + alpha * pi);
        const double ri = r0 - alpha * qi;
        store_hint<NT>(r + i, ri);
        const double zi = ri * di;
        store_hint<NT>(z + i, zi);
        v[0] += ri * ri;
        v[1] += ri * zi;
    }
}

// alpha of one FP64 iteration for the instances with many p.q partials, one
// workgroup per instance (its leader tile): the p.q reduction of the
// instance's SpMV tiles and the breakdown test (solver.rs:116-123), with the
// scalars the update's leader tile records otherwise. The instance's update
// tiles then read alpha instead of each reducing the partials, so the
// reduction's cost no longer grows with the square of the tile count.
__global__ __launch_bounds__(kBlock) void cg64_alpha_kernel(const Tile* __restrict__ leaders,
                                                            const double* __restrict__ partials, int pstride) {
    const Tile t = leaders[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || (S->flags & FL_FRESH)) return;
    const double rz = S->rz;
    const int sp[1] = {t.pq_slot};
    double pq[1];
    reduce_tiles<1>(partials, pstride, t.pq_begin, t.pq_count, sp, pq);
    if (threadIdx.x != 0) return;
    S->pap = pq[0];
    if (pq[0] <= 0.0 || pq[0] != pq[0]) {  // solver.rs:118-123
        S->it += 1;
        S->res = sqrt(S->rr) / S->bnorm;
        S->status = ST_BREAKDOWN;
        return;
    }
    S->alpha = rz / pq[0];
    S->snap_rz = rz;
    S->snap_it = S->it;
    S->dir = 0;  // the next SpMV takes the direction step
}

// One row tile of the FP64 update; TAKEN: cg64_alpha_kernel took alpha (and
// the breakdown test) for the instance, otherwise the tile reduces the p.q
// partials itself. Two instantiations, so each path keeps its own schedule
// (every preload ahead of the reduction). NT: the vectors are loaded and
// stored with the nontemporal hint (KrylovSet::ntv).
template <bool TAKEN, bool NT>
__device__ __forceinline__ void update_tile(const Tile& t, gptr<InstScal> S, const Dev& D,
                                            double* __restrict__ partials, int pstride, int par) {
    const gptr<double> pd = par ? D.p : D.p_alt;
    const int32_t b = t.row_begin, e = t.row_end;
    double p0[kPre], q0[kPre], d0[kPre], u0[kPre], r0[kPre];
    preload<double, NT>(pd, b, e, p0);
    preload<double, NT>(D.q, b, e, q0);
    preload<double, NT>(D.dinv, b, e, d0);
    preload<double, NT>(D.u, b, e, u0);
    preload<double, NT>(D.r, b, e, r0);
    double alpha;
    if constexpr (TAKEN) {
        alpha = S->alpha;
    } else {
        const double rz = S->rz;  // written by the previous phases' leaders only
        const int sp[1] = {t.pq_slot};  // P_PQ over the row tiles, or P_MPQ over the matrix-free tiles
        double pq[1];
        reduce_tiles<1>(partials, pstride, t.pq_begin, t.pq_count, sp, pq);
        if (pq[0] <= 0.0 || pq[0] != pq[0]) {  // solver.rs:118-123
            if (is_leader(t)) {
                S->pap = pq[0];
                S->it += 1;
                S->res = sqrt(S->rr) / S->bnorm;
                S->status = ST_BREAKDOWN;
            }
            return;
        }
        alpha = rz / pq[0];
        if (is_leader(t)) {
            S->pap = pq[0];
            S->alpha = alpha;
            S->snap_rz = rz;
            S->snap_it = S->it;
            S->dir = 0;  // the next SpMV takes the direction step
        }
    }
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
        const int32_t i = b + (int32_t)threadIdx.x + j * kBlock;
        if (i < e) {
            store_hint<NT>(D.u + i, u0[j] + alpha * p0[j]);
            const double ri = r0[j] - alpha * q0[j];
            store_hint<NT>(D.r + i, ri);
            const double zi = ri * d0[j];
            store_hint<NT>(D.z + i, zi);
            v[0] += ri * ri;
            v[1] += ri * zi;
        }
    }
    cg64_update_rows<NT>(b + kPre * kBlock, e, alpha, pd, D.q, D.dinv, D.u, D.r, D.z, v);
    const int slot[2] = {P_RR, P_RZ};
    write_partials<2>(v, slot, partials, pstride, blockIdx.x);
}

template <bool NT>
__global__ __launch_bounds__(kBlock) void cg64_update_kernel(const Tile* __restrict__ tiles,
                                                             const InstDesc* __restrict__ descs,
                                                             double* __restrict__ partials, int pstride, int par) {
    const Tile t = tiles[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags;
    const Dev D = view(descs[t.inst]);  // read with the state, so the loads overlap
    issue_before_test(D.p, D.p_alt, D.q, D.dinv, D.u, D.r, D.z, par, D.alphak);
    if (status != ST_RUNNING || (flags & FL_FRESH)) return;
    if (D.alphak)
        update_tile<true, NT>(t, S, D, partials, pstride, par);
    else
        update_tile<false, NT>(t, S, D, partials, pstride, par);
}

// End of a chunk (an even number of iterations, so the current direction is
// in P[0] = p): the pending direction step of the SpMV (convergence, budget,
// p = z + beta p) done in place, so that no partial sum and no direction
// step is pending at a chunk boundary: the tile tables may then be rebuilt
// (finished instances dropped, new ones added) before the next chunk.
__global__ __launch_bounds__(kBlock) void cg64_chunk_end_kernel(const Tile* __restrict__ tiles,
                                                                const InstDesc* __restrict__ descs,
                                                                const double* __restrict__ partials, int pstride) {
    const Tile t = tiles[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    // A running instance always has its direction step pending here (every
    // chunk ends with an update). S->dir is not read: this kernel's leader
    // sets it, and a tile that read the new value would skip its rows.
    const int32_t status = S->status, flags = S->flags;
    const Dev D = view(descs[t.inst]);  // read with the state, so the loads overlap
    if (status != ST_RUNNING || (flags & FL_FRESH)) return;
    const int32_t b = t.row_begin, e = t.row_end;
    double z0[kPre], p0[kPre];
    preload<double>(D.z, b, e, z0);
    preload<double>(D.p, b, e, p0);
    const double snap_rz = S->snap_rz;
    const int64_t snap_it = S->snap_it;
    const int sl[2] = {P_RR, P_RZ};
    double v[2];
    reduce_tiles<2>(partials, pstride, t.tile_begin, t.tile_count, sl, v);
    const double res = sqrt(v[0]) / S->bnorm;
    const bool conv = res <= S->tol;
    const int64_t it = snap_it + 1;
    const bool budget = it >= S->max_iter;
    const double beta = v[1] / snap_rz;
    if (is_leader(t)) {
        S->it = it;
        S->rr = v[0];
        S->res = res;
        if (conv) {
            S->status = ST_CONVERGED;
        } else {
            S->beta = beta;
            S->rz = v[1];
            S->dir = 1;
            if (budget) S->status = ST_MAXITER;
        }
    }
    if (conv || budget) return;
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
        const int32_t i = b + (int32_t)threadIdx.x + j * kBlock;
        if (i < e) D.p[i] = fmadd(beta, p0[j], z0[j]);
    }
    for (int32_t i = b + kPre * kBlock + (int32_t)threadIdx.x; i < e; i += kBlock) D.p[i] = fmadd(beta, D.p[i], D.z[i]);
}

// ------------------------------------------------------------ FP32-IR
// Direction step and q32 = K32 p32 of one FP32-IR inner iteration: the FP64
// PCG scheme above on the FP32 vectors and the FP32 copy of K (ELL, or the
// FP32 CSR values of a raw CSR instance), FP64 accumulation. The pending step
// also takes the refresh decision of the reliable-update scheme: refresh at
// the end of the period when the recursive residual fell below ru_delta
// times its maximum since the last refresh, reached tol, or went
// kForcedRefresh iterations unchecked; at the budget the instance holds
// (FL_HOLD) for the final FP64 check.
template <int BS, int W, bool NT>
__global__ __launch_bounds__(kBlock) void ir32_spmv_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs,
                                                           double* __restrict__ partials, int pstride, int par,
                                                           int force_dir, int xcd) {
    const Tile t = tiles[xcd_tile(xcd)];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags;
    const Dev D = view(descs[t.inst]);  // read with the state, so the loads overlap
    if (status != ST_RUNNING || (flags & (FL_HOLD | FL_FRESH))) return;
    const gptr<const float> cur = par ? D.p32_alt : D.p32;
    const gptr<float> nxt = par ? D.p32 : D.p32_alt;
    double v[1];
    if (!force_dir && S->dir) {
        v[0] = tile_spmv<BS, W, float, NT>(t, D, D.ell_val32, D.bval, Plain<float>{cur}, nxt, D.q32);
    } else {
        float beta = 1.0f;
        if (!force_dir) {
            const int64_t snap_it = S->snap_it;
            const double snap_rmax = S->snap_rmax, snap_rz = S->snap_rz;
            const int32_t snap_flags = S->snap_flags;
            const int sl[2] = {P_RR, P_RZ};
            double rr[2];
            reduce_tiles<2>(partials, pstride, t.tile_begin, t.tile_count, sl, rr);
            const int64_t it = snap_it + 1;
            const double est = sqrt(rr[0]);
            const double rmax = fmax(snap_rmax, est);
            const double beta_d = rr[1] / snap_rz;
            const bool hold = it >= S->max_iter;
            const bool refresh = hold || (snap_flags & FL_REFRESH) || est / S->bnorm <= S->tol ||
                                 est <= S->ru_delta * rmax || it - S->it_refresh >= kForcedRefresh;
            if (is_leader(t)) {
                S->it = it;
                S->rr = rr[0];
                S->rz = rr[1];
                S->beta = beta_d;
                S->rmax = rmax;
                S->flags = snap_flags | (refresh ? FL_REFRESH : 0) | (hold ? FL_HOLD : 0);
            }
            if (hold) return;
            beta = (float)beta_d;
        }
        v[0] = tile_spmv<BS, W, float, NT>(t, D, D.ell_val32, D.bval, Step<float>{D.z32, cur, beta}, nxt, D.q32);
    }
    const int slot[1] = {P_PQ};
    write_partials<1>(v, slot, partials, pstride, xcd_tile(xcd));
}

template <bool NT>
__device__ __forceinline__ void ir32_update_rows(int32_t b, int32_t e, float alpha, gptr<const float> __restrict__ p,
                                                 gptr<const float> __restrict__ q,
                                                 gptr<const float> __restrict__ dinv, gptr<float> __restrict__ x,
                                                 gptr<float> __restrict__ r, gptr<float> __restrict__ z,
                                                 double (&v)[2]) {
    for (int32_t i = b + threadIdx.x; i < e; i += kBlock) {
        const float pi = load_hint<NT>(p + i), qi = load_hint<NT>(q + i), di = load_hint<NT>(dinv + i);
        const float xi = load_hint<NT>(gptr<const float>(x + i)), r0 = load_hint<NT>(gptr<const float>(r + i));
        store_hint<NT>(x + i, xi + alpha * pi);
        const float ri = r0 - alpha * qi;
        store_hint<NT>(r + i, ri);
        const float zi = ri * di;
        store_hint<NT>(z + i, zi);
        v[0] += (double)ri * (double)ri;
        v[1] += (double)ri * (double)zi;
    }
}

__device__ __forceinline__ void direction32_rows(int32_t b, int32_t e, float beta, gptr<const float> __restrict__ z,
                                                 gptr<float> __restrict__ p) {
    for (int32_t i = b + threadIdx.x; i < e; i += kBlock) {
        const float zi = z[i], pi = p[i];
        p[i] = fmadd(beta, pi, zi);
    }
}

__device__ __forceinline__ void fold_rows(int32_t b, int32_t e, gptr<float> __restrict__ x,
                                          gptr<double> __restrict__ u) {
    for (int32_t i = b + threadIdx.x; i < e; i += kBlock) {
        const float xi = x[i];
        const double ui = u[i];
        u[i] = ui + (double)xi;
        x[i] = 0.0f;
    }
}

// x32 += alpha p32, r32 -= alpha q32, z32 = D^-1 r32 and the partials, with
// p32 the direction the SpMV of this iteration stored (P[1 - par]). NT: the
// vectors are loaded and stored with the nontemporal hint (KrylovSet::ntv).
template <bool NT>
__global__ __launch_bounds__(kBlock) void ir32_update_kernel(const Tile* __restrict__ tiles,
                                                             const InstDesc* __restrict__ descs,
                                                             double* __restrict__ partials, int pstride, int par) {
    const Tile t = tiles[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags;
    const Dev D = view(descs[t.inst]);  // read with the state, so the loads overlap
    if (status != ST_RUNNING || (flags & (FL_HOLD | FL_FRESH))) return;
    const gptr<float> pd = par ? D.p32 : D.p32_alt;
    const int32_t b = t.row_begin, e = t.row_end;
    float p0[kPre], q0[kPre], d0[kPre], x0[kPre], r0[kPre];
    preload<float, NT>(pd, b, e, p0);
    preload<float, NT>(D.q32, b, e, q0);
    preload<float, NT>(D.dinv32, b, e, d0);
    preload<float, NT>(D.x32, b, e, x0);
    preload<float, NT>(D.r32, b, e, r0);
    const double rz = S->rz;  // written by the previous phases' leaders only
    const int sp[1] = {P_PQ};
    double pq[1];
    reduce_tiles<1>(partials, pstride, t.tile_begin, t.tile_count, sp, pq);
    if (pq[0] <= 0.0 || pq[0] != pq[0]) {
        if (is_leader(t)) {
            S->pap = pq[0];
            S->it += 1;
            S->res = sqrt(S->rr) / S->bnorm;
            S->status = ST_BREAKDOWN;
        }
        return;
    }
    const double alpha_d = rz / pq[0];
    if (is_leader(t)) {
        S->pap = pq[0];
        S->alpha = alpha_d;
        S->snap_rz = rz;
        S->snap_rmax = S->rmax;
        S->snap_it = S->it;
        S->snap_flags = S->flags;
        S->dir = 0;  // the next SpMV (or the period end) takes the direction step
    }
    const float alpha = (float)alpha_d;
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
        const int32_t i = b + (int32_t)threadIdx.x + j * kBlock;
        if (i < e) {
            store_hint<NT>(D.x32 + i, x0[j] + alpha * p0[j]);
            const float ri = r0[j] - alpha * q0[j];
            store_hint<NT>(D.r32 + i, ri);
            const float zi = ri * d0[j];
            store_hint<NT>(D.z32 + i, zi);
            v[0] += (double)ri * (double)ri;
            v[1] += (double)ri * (double)zi;
        }
    }
    ir32_update_rows<NT>(b + kPre * kBlock, e, alpha, pd, D.q32, D.dinv32, D.x32, D.r32, D.z32, v);
    const int slot[2] = {P_RR, P_RZ};
    write_partials<2>(v, slot, partials, pstride, blockIdx.x);
}

// End of a period (an even number of iterations, so the current direction is
// in P[0] = p32): the pending step's bookkeeping and refresh decision as in
// the SpMV; an instance flagged for refresh keeps its direction (it is
// rebuilt from the replaced residual), any other takes p32 = z32 + beta p32
// in place and marks it complete.
__global__ __launch_bounds__(kBlock) void ir32_period_end_kernel(const Tile* __restrict__ tiles,
                                                                 const InstDesc* __restrict__ descs,
                                                                 const double* __restrict__ partials, int pstride) {
    const Tile t = tiles[blockIdx.x];
    const gptr<InstScal> S = global(t.s);
    const int32_t status = S->status, flags = S->flags;
    const Dev D = view(descs[t.inst]);  // read with the state, so the loads overlap
    if (status != ST_RUNNING || (flags & (FL_HOLD | FL_FRESH))) return;
    const int32_t b = t.row_begin, e = t.row_end;
    float z0[kPre], p0[kPre];
    preload<float>(D.z32, b, e, z0);
    preload<float>(D.p32, b, e, p0);
    const int64_t snap_it = S->snap_it;
    const double snap_rmax = S->snap_rmax, snap_rz = S->snap_rz;
    const int32_t snap_flags = S->snap_flags;
    const int sl[2] = {P_RR, P_RZ};
    double v[2];
    reduce_tiles<2>(partials, pstride, t.tile_begin, t.tile_count, sl, v);
    const int64_t it = snap_it + 1;
    const double est = sqrt(v[0]);
    const double rmax = fmax(snap_rmax, est);
    const double beta = v[1] / snap_rz;
    const bool hold = it >= S->max_iter;
    const bool refresh = hold || (snap_flags & FL_REFRESH) || est / S->bnorm <= S->tol ||
                         est <= S->ru_delta * rmax || it - S->it_refresh >= kForcedRefresh;
    if (is_leader(t)) {
        S->it = it;
        S->rr = v[0];
        S->rz = v[1];
        S->beta = beta;
        S->rmax = rmax;
        S->flags = snap_flags | (refresh ? FL_REFRESH : 0) | (hold ? FL_HOLD : 0);
        if (!refresh) S->dir = 1;
    }
    if (refresh) return;
    const float bf = (float)beta;
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
        const int32_t i = b + (int32_t)threadIdx.x + j * kBlock;
        if (i < e) D.p32[i] = fmadd(bf, p0[j], z0[j]);
    }
    direction32_rows(b + kPre * kBlock, e, bf, D.z32, D.p32);
}

// Refresh step 1: fold the FP32 correction into the FP64 solution.
__global__ __launch_bounds__(kBlock) void ir32_fold_kernel(const Tile* __restrict__ tiles,
                                                           const InstDesc* __restrict__ descs) {
    const Tile t = tiles[blockIdx.x];
    const Dev D = view(descs[t.inst]);
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || !(S->flags & FL_REFRESH)) return;
    if (is_leader(t)) {
        S->snap_flags = S->flags;
        S->snap_stall = S->stall;
        S->snap_res_best = S->res_best;
    }
    fold_rows(t.row_begin, t.row_end, D.x32, D.u);
}

// Refresh step 2: r = f - K u in FP64 with its evaluation floor; r32, z32.
template <int LANES>
__device__ __forceinline__ void residual_rows(int32_t b, int32_t e, gptr<const int32_t> __restrict__ rp,
                                              gptr<const int32_t> __restrict__ col,
                                              gptr<const double> __restrict__ val, gptr<const double> __restrict__ u,
                                              gptr<const double> __restrict__ f,
                                              gptr<const float> __restrict__ dinv32, gptr<float> __restrict__ r32,
                                              gptr<float> __restrict__ z32, double (&v)[3]) {
    const int lane = threadIdx.x % LANES, grp = threadIdx.x / LANES;
    constexpr int G = kBlock / LANES;
    for (int32_t row = b + grp; row < e; row += G) {
        const double fi = f[row];
        const float di = dinv32[row];
        double ku = 0.0, aku = 0.0;
        for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += LANES) {
            const double w = val[k] * u[col[k]];
            ku += w;
            aku += fabs(w);
        }
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1) {
            ku += __shfl_xor(ku, off, LANES);
            aku += __shfl_xor(aku, off, LANES);
        }
        if (lane == 0) {
            const double ri = fi - ku;
            const float rs = (float)ri;
            const float zs = rs * di;
            r32[row] = rs;
            z32[row] = zs;
            v[0] += ri * ri;
            v[1] += (double)rs * (double)zs;
            const double a = aku + fabs(fi);
            v[2] += a * a;
        }
    }
}

// The same over the ELL chunks [c0, c1) of a node-block instance.
template <int BS, int UNR>
__device__ __forceinline__ void ell_residual_rows(int32_t c0, int32_t c1, const Dev& D, double (&v)[3]) {
    ell_tile<BS, UNR, double, true, false>(c0, c1, D.n / BS, D.ell_off, D.ell_nb, D.ell_col, D.ell_val, Plain<double>{D.u},
                                    [&](int32_t r, const double (&ku)[BS], const double (&aku)[BS]) {
#pragma unroll
                                        for (int a = 0; a < BS; ++a) {
                                            const int32_t i = r * BS + a;
                                            const double fi = D.f[i];
                                            const double ri = fi - ku[a];
                                            const float rs = (float)ri;
                                            const float zs = rs * D.dinv32[i];
                                            D.r32[i] = rs;
                                            D.z32[i] = zs;
                                            v[0] += ri * ri;
                                            v[1] += (double)rs * (double)zs;
                                            const double aa = aku[a] + fabs(fi);
                                            v[2] += aa * aa;
                                        }
                                    });
}

template <int BS, int W>
__global__ __launch_bounds__(kBlock) void ir32_residual_kernel(const Tile* __restrict__ tiles,
                                                               const InstDesc* __restrict__ descs,
                                                               double* __restrict__ partials, int pstride, int xcd) {
    const Tile t = tiles[xcd_tile(xcd)];
    const Dev D = view(descs[t.inst]);
    const gptr<InstScal> S = global(t.s);
    if (S->status != ST_RUNNING || !(S->flags & FL_REFRESH)) return;
    if (is_leader(t)) S->evals += 1;
    double v[3] = {0.0, 0.0, 0.0};
    if constexpr (BS == 1) {
        residual_rows<W>(t.row_begin, t.row_end, D.row_ptr, D.col, D.val, D.u, D.f, D.dinv32, D.r32, D.z32, v);
    } else {
        int32_t c0, c1;
        tile_chunks<BS>(t, c0, c1);
        ell_residual_rows<BS, W>(c0, c1, D, v);
    }
    const int slot[3] = {P_TRR, P_TRZ, P_TFL};
    write_partials<3>(v, slot, partials, pstride, xcd_tile(xcd));
}

// Refresh step 3: convergence / budget / fallback decision, then
// p = z + beta p from the replaced residual.
__global__ __launch_bounds__(kBlock) void ir32_repoint_kernel(const Tile* __restrict__ tiles,
                                                              const InstDesc* __restrict__ descs,
                                                              const double* __restrict__ partials, int pstride) {
    const Tile t = tiles[blockIdx.x];
    const Dev D = view(descs[t.inst]);
    const gptr<InstScal> S = global(t.s);
    const int sl[3] = {P_TRR, P_TRZ, P_TFL};
    double v[3];
    reduce_tiles<3>(partials, pstride, t.tile_begin, t.tile_count, sl, v);
    if (S->status != ST_RUNNING || !(S->snap_flags & FL_REFRESH)) return;
    const double res = sqrt(v[0]) / S->bnorm;
    const double floor = S->floor_scale * sqrt(v[2]) / S->bnorm;
    const bool improved = res < 0.5 * S->snap_res_best;
    const int32_t stall = improved ? 0 : S->snap_stall + 1;
    // The FP32 phase never ends the solve: at or below tol plus the FP64
    // evaluation floor (kFloorSlack floors once the residual has stopped
    // halving), or once the residual stopped improving above it, FP64 PCG
    // finishes on its recursive residual (the reference's test).
    const bool at_floor = res <= S->tol + floor || (stall >= kStallLimit && res <= S->tol + kFloorSlack * floor);
    const bool hold = (S->snap_flags & FL_HOLD) != 0;
    const bool fallback = !hold && (at_floor || stall >= kStallLimit);
    const bool cont = !hold && !fallback;
    const double beta = v[1] / S->snap_rz;
    if (is_leader(t)) {
        S->res = res;
        S->floor = floor;
        S->rr = v[0];
        S->refreshes += 1;
        S->it_refresh = S->it;
        S->res_best = improved ? res : S->snap_res_best;
        S->stall = stall;
        if (hold) {
            S->status = ST_MAXITER;
        } else if (fallback) {
            S->status = ST_FALLBACK;  // FP64 PCG finishes from the current u
            S->fb_reason = at_floor ? 2 : 1;
            if (at_floor) S->tol = fmin(S->tol, 0.1 * res);  // at least a decade of FP64 recursion
        } else {
            S->rmax = sqrt(v[0]);
            S->rz = v[1];
            S->beta = beta;
            S->flags = 0;
            S->dir = 1;  // the direction below is complete
        }
    }
    if (!cont) return;
    direction32_rows(t.row_begin, t.row_end, (float)beta, D.z32, D.p32);
}

// ------------------------------------------------------------ utilities
// out[0]: running instances; out[1]: tiles of running instances in the
// current tile tables (tile_count of an instance describes its current tiling).
__global__ void count_running_kernel(const InstDesc* __restrict__ descs, int32_t n_inst, int32_t* out) {
    __shared__ int32_t cnt, tiles;
    if (threadIdx.x == 0) cnt = 0, tiles = 0;
    __syncthreads();
    int32_t local = 0, ltiles = 0;
    for (int32_t i = threadIdx.x; i < n_inst; i += blockDim.x) {
        const bool run = descs[i].s->status == ST_RUNNING;
        local += run;
        ltiles += run ? descs[i].tile_count : 0;
    }
    atomicAdd(&cnt, local);
    atomicAdd(&tiles, ltiles);
    __syncthreads();
    if (threadIdx.x == 0) out[0] = cnt, out[1] = tiles;
}

__global__ void promote_fallback_kernel(const InstDesc* __restrict__ descs, int32_t n_inst) {
    for (int32_t i = threadIdx.x; i < n_inst; i += blockDim.x) {
        InstScal* S = descs[i].s;
        if (S->status == ST_FALLBACK) {
            S->status = ST_RUNNING;
            S->flags = FL_FRESH | FL_WARM;
            S->fp64_it = -S->it;
        }
    }
}

__global__ void gather_scalars_kernel(const InstDesc* __restrict__ descs, int32_t n_inst, InstScal* out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) out[i] = *descs[i].s;
}

__global__ void set_running_kernel(const InstDesc* __restrict__ descs, int32_t n_inst, int32_t flags) {
    for (int32_t i = threadIdx.x; i < n_inst; i += blockDim.x) {
        descs[i].s->status = ST_RUNNING;
        descs[i].s->flags = flags;
        descs[i].s->snap_flags = flags;
    }
}

__global__ void carry_partials_kernel(const double* __restrict__ old_p, int old_stride, double* __restrict__ new_p,
                                      int new_stride, const int32_t* __restrict__ src, int n_new) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_new) return;
    const int o = src[g];
    if (o < 0 || o >= old_stride) return;
    for (int k = 0; k < kPartialSlots; ++k) new_p[(size_t)k * new_stride + g] = old_p[(size_t)k * old_stride + o];
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

#define SVX_UNROLL_DISPATCH(unroll, MACRO) \
    switch (unroll) {                      \
        case 1: MACRO(1); break;           \
        case 2: MACRO(2); break;           \
        case 3: MACRO(3); break;           \
        default: MACRO(4); break;          \
    }

static void launch_cg64_update(const KrylovSet& k, hipStream_t s, int par) {
    if (k.ntv)
        cg64_update_kernel<true><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par);
    else
        cg64_update_kernel<false><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par);
}

static void launch_ir32_update(const KrylovSet& k, hipStream_t s, int par) {
    if (k.ntv)
        ir32_update_kernel<true><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par);
    else
        ir32_update_kernel<false><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par);
}

void launch_cg64_spmv(const KrylovSet& k, hipStream_t s, int par, int force_dir) {
    const int g = k.n_ell_tiles;
    // the pending step of the matrix-free and the large assembled instances,
    // before any product that reads it
    if (k.n_mf_inst > 0 && !force_dir)
        cg64_mf_step_kernel<<<k.n_mf_inst, kBlock, 0, s>>>(k.mf_leaders, k.partials, k.pstride);
    if (g > 0) {
        if (k.bs == 3) {
#define L(U) cg64_spmv_kernel<3, U, false><<<g, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
#define LN(U) cg64_spmv_kernel<3, U, true><<<g, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
            if (k.nt) {
                SVX_UNROLL_DISPATCH(k.unroll, LN)
            } else {
                SVX_UNROLL_DISPATCH(k.unroll, L)
            }
#undef LN
#undef L
        } else if (k.bs == 2) {
#define L(U) cg64_spmv_kernel<2, U, false><<<g, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
#define LN(U) cg64_spmv_kernel<2, U, true><<<g, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
            if (k.nt) {
                SVX_UNROLL_DISPATCH(k.unroll, LN)
            } else {
                SVX_UNROLL_DISPATCH(k.unroll, L)
            }
#undef LN
#undef L
        } else {
#define L(LN) cg64_spmv_kernel<1, LN, false><<<g, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
            SVX_LANES_DISPATCH(k.lanes, L)
#undef L
        }
    }
    if (k.n_mf3_tiles > 0) {
        cg64_mf3_elem_kernel<<<k.n_mf3_tiles, kMf3Elems * 8, 0, s>>>(k.mf3_tiles, k.descs, par, force_dir);
        cg64_mf3_node_kernel<<<k.n_tiles - k.n_ell_tiles, kBlock, 0, s>>>(k.tiles, k.n_ell_tiles, k.descs,
                                                                            k.partials, k.pstride, par, force_dir);
    }
#define L(NEN, NT, CNT, BASE, I)                                                                                    \
    cg64_mf_spmv_kernel<NEN, NT><<<CNT, kBlock, mf_lds_bytes(NEN, k.mf_cap_nodes[I], k.mf_cap_elem[I]), s>>>(        \
        k.mf_tiles, BASE, k.descs, k.partials, k.pstride, par, force_dir, k.mf_cap_nodes[I], k.mf_cap_elem[I], k.xcd)
    if (k.n_mf4 > 0) {
        if (k.ntv)
            L(4, true, k.n_mf4, 0, 0);
        else
            L(4, false, k.n_mf4, 0, 0);
    }
    if (k.n_mf3 > 0) {
        if (k.ntv)
            L(3, true, k.n_mf3, k.n_mf4, 1);
        else
            L(3, false, k.n_mf3, k.n_mf4, 1);
    }
#undef L
}

void launch_ir32_spmv(const KrylovSet& k, hipStream_t s, int par, int force_dir) {
    if (k.bs == 3) {
#define L(U) \
    ir32_spmv_kernel<3, U, false><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
#define LN(U) \
    ir32_spmv_kernel<3, U, true><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
        if (k.nt32) {
            SVX_UNROLL_DISPATCH(k.unroll, LN)
        } else {
            SVX_UNROLL_DISPATCH(k.unroll, L)
        }
#undef LN
#undef L
    } else if (k.bs == 2) {
#define L(U) \
    ir32_spmv_kernel<2, U, false><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
#define LN(U) \
    ir32_spmv_kernel<2, U, true><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
        if (k.nt32) {
            SVX_UNROLL_DISPATCH(k.unroll, LN)
        } else {
            SVX_UNROLL_DISPATCH(k.unroll, L)
        }
#undef LN
#undef L
    } else {
#define L(LN) \
    ir32_spmv_kernel<1, LN, false><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, par, force_dir, k.xcd)
        SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    }
}

}  // namespace

void launch_spmv_ordered(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* x,
                         double* y, hipStream_t s) {
    if (n <= 0) return;
    spmv_ordered_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, row_ptr, col, val, x, y);
    SVX_HIP(hipGetLastError());
}

namespace {
constexpr int kTrLanes = 16;  // lanes per CSR row of the true-residual pass

// One kTrLanes-lane group per row; block partials of (f - K u)^2,
// (|K||u| + |f|)^2 and f^2, summed in a fixed wave / LDS order.
__global__ __launch_bounds__(kBlock) void true_residual_kernel(int64_t n, const int32_t* __restrict__ rp,
                                                               const int32_t* __restrict__ col,
                                                               const double* __restrict__ val,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ f,
                                                               double* __restrict__ partials) {
    const int lane = threadIdx.x % kTrLanes;
    const int64_t row = (int64_t)blockIdx.x * (kBlock / kTrLanes) + threadIdx.x / kTrLanes;
    double v[3] = {0.0, 0.0, 0.0};
    if (row < n) {
        double ku = 0.0, aku = 0.0;
        for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += kTrLanes) {
            const double w = val[k] * u[col[k]];
            ku += w;
            aku += fabs(w);
        }
#pragma unroll
        for (int off = kTrLanes / 2; off > 0; off >>= 1) {
            ku += __shfl_xor(ku, off, kTrLanes);
            aku += __shfl_xor(aku, off, kTrLanes);
        }
        if (lane == 0) {
            const double fi = f[row], ri = fi - ku, a = aku + fabs(fi);
            v[0] = ri * ri;
            v[1] = a * a;
            v[2] = fi * fi;
        }
    }
    __shared__ double sh[3][kBlock];
#pragma unroll
    for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// One block: each thread sums blocks j, j + kBlock, ... in order, then a
// fixed tree over the threads.
__global__ __launch_bounds__(kBlock) void true_residual_finish_kernel(int64_t n_blocks,
                                                                      const double* __restrict__ partials,
                                                                      double* __restrict__ out) {
    __shared__ double sh[3][kBlock];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int64_t j = threadIdx.x; j < n_blocks; j += kBlock) acc += partials[(size_t)c * n_blocks + j];
        sh[c][threadIdx.x] = acc;
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = sh[threadIdx.x][0];
}
}  // namespace

int64_t true_residual_blocks(int64_t n) { return (n + kBlock / kTrLanes - 1) / (kBlock / kTrLanes); }

void launch_true_residual(int64_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* u,
                          const double* f, double* partials, double* out, hipStream_t s) {
    if (n <= 0) return;
    const int64_t nb = true_residual_blocks(n);
    true_residual_kernel<<<(unsigned)nb, kBlock, 0, s>>>(n, row_ptr, col, val, u, f, partials);
    true_residual_finish_kernel<<<1, kBlock, 0, s>>>(nb, partials, out);
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

void cg64_init(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    cg64_init_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    if (k.bs == 3) {
#define L(U) cg64_init_warm_kernel<3, U><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, k.xcd)
        SVX_UNROLL_DISPATCH(k.unroll, L)
#undef L
    } else if (k.bs == 2) {
#define L(U) cg64_init_warm_kernel<2, U><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, k.xcd)
        SVX_UNROLL_DISPATCH(k.unroll, L)
#undef L
    } else {
#define L(LN) cg64_init_warm_kernel<1, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, k.xcd)
        SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    }
    init_finish_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    SVX_HIP(hipGetLastError());
}

void cg64_iteration(const KrylovSet& k, hipStream_t s, int par) {
    if (k.n_tiles <= 0) return;
    launch_cg64_spmv(k, s, par, 0);
    if (k.n_alpha > 0) cg64_alpha_kernel<<<k.n_alpha, kBlock, 0, s>>>(k.alpha_leaders, k.partials, k.pstride);
    launch_cg64_update(k, s, par);
    SVX_HIP(hipGetLastError());
}

void cg64_chunk_end(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    cg64_chunk_end_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    SVX_HIP(hipGetLastError());
}

void ir32_init(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    ir32_init_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    init_finish_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    SVX_HIP(hipGetLastError());
}

void ir32_iteration(const KrylovSet& k, hipStream_t s, int par) {
    if (k.n_tiles <= 0) return;
    launch_ir32_spmv(k, s, par, 0);
    launch_ir32_update(k, s, par);
    SVX_HIP(hipGetLastError());
}

void ir32_period_end(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    ir32_period_end_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    SVX_HIP(hipGetLastError());
}

void ir32_refresh(const KrylovSet& k, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    ir32_fold_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs);
    if (k.bs == 3) {
#define L(U) ir32_residual_kernel<3, U><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, k.xcd)
        SVX_UNROLL_DISPATCH(k.unroll, L)
#undef L
    } else if (k.bs == 2) {
#define L(U) ir32_residual_kernel<2, U><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, k.xcd)
        SVX_UNROLL_DISPATCH(k.unroll, L)
#undef L
    } else {
#define L(LN) ir32_residual_kernel<1, LN><<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride, k.xcd)
        SVX_LANES_DISPATCH(k.lanes, L)
#undef L
    }
    ir32_repoint_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    SVX_HIP(hipGetLastError());
}

void launch_promote_fallback(const InstDesc* descs, int32_t n_inst, hipStream_t s) {
    promote_fallback_kernel<<<1, kBlock, 0, s>>>(descs, n_inst);
    SVX_HIP(hipGetLastError());
}

void launch_gather_scalars(const InstDesc* descs, int32_t n_inst, InstScal* out, hipStream_t s) {
    if (n_inst <= 0) return;
    gather_scalars_kernel<<<grid_for(n_inst, kBlock), kBlock, 0, s>>>(descs, n_inst, out);
    SVX_HIP(hipGetLastError());
}

void launch_set_running(const InstDesc* descs, int32_t n_inst, hipStream_t s, int32_t flags) {
    set_running_kernel<<<1, kBlock, 0, s>>>(descs, n_inst, flags);
    SVX_HIP(hipGetLastError());
}

void update_phase(const KrylovSet& k, int precision, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    if (precision == SVX_FP64) {
        if (k.n_alpha > 0) cg64_alpha_kernel<<<k.n_alpha, kBlock, 0, s>>>(k.alpha_leaders, k.partials, k.pstride);
        launch_cg64_update(k, s, 0);
    }
    else
        launch_ir32_update(k, s, 0);
    SVX_HIP(hipGetLastError());
}

void period_end_phase(const KrylovSet& k, int precision, hipStream_t s) {
    if (k.n_tiles <= 0 || precision == SVX_FP64) return;
    ir32_period_end_kernel<<<k.n_tiles, kBlock, 0, s>>>(k.tiles, k.descs, k.partials, k.pstride);
    SVX_HIP(hipGetLastError());
}

void spmv_phase(const KrylovSet& k, int precision, hipStream_t s) {
    if (k.n_tiles <= 0) return;
    if (precision == SVX_FP64)
        launch_cg64_spmv(k, s, 0, 1);
    else
        launch_ir32_spmv(k, s, 0, 1);
    SVX_HIP(hipGetLastError());
}

void launch_carry_partials(const double* old_p, int old_stride, double* new_p, int new_stride, const int32_t* src,
                           int n_new, hipStream_t s) {
    if (n_new <= 0 || old_stride <= 0) return;
    carry_partials_kernel<<<grid_for(n_new, kBlock), kBlock, 0, s>>>(old_p, old_stride, new_p, new_stride, src, n_new);
    SVX_HIP(hipGetLastError());
}

void launch_count_running(const InstDesc* descs, int32_t n_inst, int32_t* out, hipStream_t s) {
    count_running_kernel<<<1, kBlock, 0, s>>>(descs, n_inst, out);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
