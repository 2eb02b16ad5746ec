// Global stiffness assembly on gfx950.
//
// Element stiffness (element.rs:10-211, 427-561, 829-851) is computed with
// the reference's operation order and compiled with -ffp-contract=off, so
// each Ke entry equals the CPU restatement's bit for bit. The element
// matrices are then gathered into the CSR values row by row: every CSR entry
// sums its contributions in element order, local row then local column —
// the order in which csr_from_triplets (assembly.rs:87-116) meets them after
// scatter (assembly.rs:120-133) — so the result is deterministic and needs
// no atomics.
#include "svx_internal.hpp"

namespace svx {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ void record_bad(uint8_t* reason, unsigned long long* first_bad, int64_t e, int why) {
    reason[e] = (uint8_t)why;
    atomicMin(first_bad, (unsigned long long)e);
}

// Material checks of element_entry (assembly.rs:145-160); NaN E = no material.
__device__ __forceinline__ int material_reason(double E, double nu, bool solid) {
    if (E != E) return SVX_ELEM_NO_MATERIAL;
    if (E <= 0.0 || fabs(nu) >= 1.0) return SVX_ELEM_BAD_MATERIAL;
    if (solid && nu >= 0.5) return SVX_ELEM_BAD_POISSON_SOLID;
    return SVX_ELEM_OK;
}

// d_matrix (element.rs:10-17), row-major 3x3.
__device__ __forceinline__ void d_matrix(double e, double nu, double* d) {
    double c = e / (1.0 - nu * nu);
    d[0] = c, d[1] = nu * c, d[2] = 0.0;
    d[3] = nu * c, d[4] = c, d[5] = 0.0;
    d[6] = 0.0, d[7] = 0.0, d[8] = c * (1.0 - nu) / 2.0;
}

// btdb (element.rs:829-851) for S = 3: ke += w B^T D B.
template <int N>
__device__ __forceinline__ void btdb3(double* ke, const double* b, const double* d, double w) {
    double db[3 * N];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
            double s = -0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += d[r * 3 + k] * b[k * N + c];
            db[r * N + c] = s;
        }
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
        for (int c = a; c < N; ++c) {
            double s = -0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r) s += b[r * N + a] * db[r * N + c];
            double v = w * s;
            ke[a * N + c] += v;
            if (a != c) ke[c * N + a] += v;
        }
}

// quad_b (element.rs:147-171); returns det J (<= 0 means rejected).
__device__ __forceinline__ double quad_b(const double (*x)[2], double xi, double eta, double* bm) {
    const double cx[4] = {-1.0, 1.0, 1.0, -1.0};
    const double cy[4] = {-1.0, -1.0, 1.0, 1.0};
    double dxi[4], deta[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        dxi[i] = 0.25 * cx[i] * (1.0 + eta * cy[i]);
        deta[i] = 0.25 * cy[i] * (1.0 + xi * cx[i]);
    }
    double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        j00 += dxi[i] * x[i][0];
        j01 += dxi[i] * x[i][1];
        j10 += deta[i] * x[i][0];
        j11 += deta[i] * x[i][1];
    }
    double detj = j00 * j11 - j01 * j10;
    if (detj <= 0.0) return detj;
    double inv = 1.0 / detj;
#pragma unroll
    for (int i = 0; i < 24; ++i) bm[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double dnx = inv * (j11 * dxi[i] - j01 * deta[i]);
        double dny = inv * (j00 * deta[i] - j10 * dxi[i]);
        bm[0 * 8 + 2 * i] = dnx;
        bm[1 * 8 + 2 * i + 1] = dny;
        bm[2 * 8 + 2 * i] = dny;
        bm[2 * 8 + 2 * i + 1] = dnx;
    }
    return detj;
}

// CPS3 (element.rs:91-129), one thread per element.
__global__ void ke_cps3_kernel(int64_t n_sel, const int32_t* __restrict__ sel, const int64_t* __restrict__ conn_ptr,
                               const int32_t* __restrict__ conn, const double* __restrict__ X,
                               const double* __restrict__ E, const double* __restrict__ NU, double t,
                               double* __restrict__ ke_all, const int64_t* __restrict__ ke_off,
                               uint8_t* __restrict__ reason, unsigned long long* first_bad) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_sel) return;
    int64_t e = sel[q];
    int why = material_reason(E[e], NU[e], false);
    if (why) return record_bad(reason, first_bad, e, why);
    const int32_t* cn = conn + conn_ptr[e];
    double x[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k][0] = X[cn[k] * 3], x[k][1] = X[cn[k] * 3 + 1];
    double b[3] = {x[1][1] - x[2][1], x[2][1] - x[0][1], x[0][1] - x[1][1]};
    double c[3] = {x[2][0] - x[1][0], x[0][0] - x[2][0], x[1][0] - x[0][0]};
    double two_a = b[0] * c[1] - b[1] * c[0];
    if (two_a == 0.0) return record_bad(reason, first_bad, e, SVX_ELEM_DEGENERATE_TRIANGLE);
    double inv = 1.0 / two_a;
    double bm[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) bm[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        bm[0 * 6 + 2 * i] = b[i] * inv;
        bm[1 * 6 + 2 * i + 1] = c[i] * inv;
        bm[2 * 6 + 2 * i] = c[i] * inv;
        bm[2 * 6 + 2 * i + 1] = b[i] * inv;
    }
    double d[9];
    d_matrix(E[e], NU[e], d);
    double ke[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) ke[i] = 0.0;
    btdb3<6>(ke, bm, d, t * fabs(two_a) / 2.0);
    double* out = ke_all + ke_off[e];
#pragma unroll
    for (int i = 0; i < 36; ++i) out[i] = ke[i];
    reason[e] = 0;
}

// CPS4 (element.rs:186-211), 2x2 Gauss, one thread per element.
__global__ void ke_cps4_kernel(int64_t n_sel, const int32_t* __restrict__ sel, const int64_t* __restrict__ conn_ptr,
                               const int32_t* __restrict__ conn, const double* __restrict__ X,
                               const double* __restrict__ E, const double* __restrict__ NU, double t, double g,
                               double* __restrict__ ke_all, const int64_t* __restrict__ ke_off,
                               uint8_t* __restrict__ reason, unsigned long long* first_bad) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_sel) return;
    int64_t e = sel[q];
    int why = material_reason(E[e], NU[e], false);
    if (why) return record_bad(reason, first_bad, e, why);
    const int32_t* cn = conn + conn_ptr[e];
    double x[4][2];
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k][0] = X[cn[k] * 3], x[k][1] = X[cn[k] * 3 + 1];
    double d[9];
    d_matrix(E[e], NU[e], d);
    double ke[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) ke[i] = 0.0;
    const double pts[2] = {-g, g};
    double bm[24];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            double detj = quad_b(x, pts[a], pts[b], bm);
            if (detj <= 0.0) return record_bad(reason, first_bad, e, SVX_ELEM_NONPOSITIVE_JACOBIAN);
            btdb3<8>(ke, bm, d, t * detj);
        }
    if (quad_b(x, 0.0, 0.0, bm) <= 0.0) return record_bad(reason, first_bad, e, SVX_ELEM_NONPOSITIVE_JACOBIAN);
    double* out = ke_all + ke_off[e];
#pragma unroll
    for (int i = 0; i < 64; ++i) out[i] = ke[i];
    reason[e] = 0;
}

constexpr int kHexPerBlock = 8;  // 8 elements x 24 lanes = 192 threads

// Jacobian of the trilinear map at (xi, eta, zeta) and its inverse as the
// reference computes it (element.rs:440-521): out[i][j] = cof(j, i) / det
// with cof(r, c) = (-1)^(r+c) times the cyclic-index minor. The cyclic minor
// already carries the cofactor sign, so the four odd-parity entries come out
// negated; for diagonal Jacobians (axis-aligned bricks) they vanish. The
// device reproduces the reference bit for bit, sign pattern included.
// Returns det J.
__device__ __forceinline__ double hex_jac(const double (*x)[3], double xi, double eta, double zeta,
                                          double dn[3][8], double inv[3][3]) {
    const double ca[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
    const double cb[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
    const double cc[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        dn[0][i] = 0.125 * ca[i] * (1.0 + eta * cb[i]) * (1.0 + zeta * cc[i]);
        dn[1][i] = 0.125 * cb[i] * (1.0 + xi * ca[i]) * (1.0 + zeta * cc[i]);
        dn[2][i] = 0.125 * cc[i] * (1.0 + xi * ca[i]) * (1.0 + eta * cb[i]);
    }
    double m[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) m[a][b] += dn[a][i] * x[i][b];
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                 m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (det <= 0.0) return det;
    double id = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // out[i][j] = cof(j, i) * id with cof(r,c) = sign * minor
            int r = j, c = i;
            int ra = (r + 1) % 3, rb = (r + 2) % 3, cd = (c + 1) % 3, ce = (c + 2) % 3;
            double sign = ((r + c) % 2 == 0) ? 1.0 : -1.0;
            inv[i][j] = sign * (m[ra][cd] * m[rb][ce] - m[ra][ce] * m[rb][cd]) * id;
        }
    return det;
}

// Column t (node t/3, component t%3) of the 6x24 B matrix (element.rs:499-519).
__device__ __forceinline__ void hex_b_column(const double dn[3][8], const double inv[3][3], int t, double col[6]) {
    int i = t / 3, comp = t % 3;
    double dnv[3] = {dn[0][i], dn[1][i], dn[2][i]};
    double g[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) g[b] += inv[b][a] * dnv[a];
#pragma unroll
    for (int r = 0; r < 6; ++r) col[r] = 0.0;
    if (comp == 0) {
        col[0] = g[0], col[3] = g[1], col[5] = g[2];
    } else if (comp == 1) {
        col[1] = g[1], col[3] = g[0], col[4] = g[2];
    } else {
        col[2] = g[2], col[4] = g[1], col[5] = g[0];
    }
}

// C3D8 (element.rs:535-561): 24 lanes per element, lane a owns Ke row a.
// Per Gauss point every lane builds its B column and D*B column in LDS; lane
// a then adds w * sum_r B[r][lo] * DB[r][hi] to Ke[a][c] with (lo, hi) the
// ordered pair (min(a,c), max(a,c)), which is exactly the value the
// reference computes once for a <= c and mirrors.
__global__ __launch_bounds__(kHexPerBlock * 24) void ke_c3d8_kernel(
    int64_t n_sel, const int32_t* __restrict__ sel, const int64_t* __restrict__ conn_ptr,
    const int32_t* __restrict__ conn, const double* __restrict__ X, const double* __restrict__ E,
    const double* __restrict__ NU, double g, double* __restrict__ ke_all, const int64_t* __restrict__ ke_off,
    uint8_t* __restrict__ reason, unsigned long long* first_bad) {
    __shared__ double sB[kHexPerBlock][6][24];
    __shared__ double sDB[kHexPerBlock][6][24];
    __shared__ int sBad[kHexPerBlock];
    const int slot = threadIdx.x / 24, a = threadIdx.x % 24;
    int64_t q = (int64_t)blockIdx.x * kHexPerBlock + slot;
    const bool live = q < n_sel;
    int64_t e = live ? sel[q] : 0;
    if (a == 0) sBad[slot] = 0;
    double x[8][3];
    double d[6][6];
    if (live) {
        const int32_t* cn = conn + conn_ptr[e];
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int b = 0; b < 3; ++b) x[k][b] = X[cn[k] * 3 + b];
        double Ee = E[e], nu = NU[e];
        // d3_matrix (element.rs:33-47)
        double lam = Ee * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double mu = Ee / (2.0 * (1.0 + nu));
        double diag = lam + 2.0 * mu;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) d[r][c] = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) d[r][c] = (r == c) ? diag : lam;
        d[3][3] = mu, d[4][4] = mu, d[5][5] = mu;
        int why = material_reason(Ee, nu, true);
        if (why) {
            if (a == 0) record_bad(reason, first_bad, e, why);
            sBad[slot] = 1;  // all lanes agree
        }
    }
    __syncthreads();
    bool active = live && !sBad[slot];
    double ke[24];
#pragma unroll
    for (int c = 0; c < 24; ++c) ke[c] = 0.0;
    const double pts[2] = {-g, g};
    for (int gp = 0; gp < 8; ++gp) {
        double xi = pts[gp >> 2], eta = pts[(gp >> 1) & 1], zeta = pts[gp & 1];
        double w = 0.0;
        if (active) {
            double dn[3][8], inv[3][3];
            w = hex_jac(x, xi, eta, zeta, dn, inv);
            if (w <= 0.0) {
                active = false;  // every lane of the element sees the same det
                if (a == 0) {
                    record_bad(reason, first_bad, e, SVX_ELEM_NONPOSITIVE_JACOBIAN);
                }
            } else {
                double col[6];
                hex_b_column(dn, inv, a, col);
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    double s = -0.0;
#pragma unroll
                    for (int k = 0; k < 6; ++k) s += d[r][k] * col[k];
                    sB[slot][r][a] = col[r];
                    sDB[slot][r][a] = s;
                }
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int c = 0; c < 24; ++c) {
                int lo = a < c ? a : c, hi = a < c ? c : a;
                double s = -0.0;
#pragma unroll
                for (int r = 0; r < 6; ++r) s += sB[slot][r][lo] * sDB[slot][r][hi];
                ke[c] += w * s;
            }
        }
        __syncthreads();
    }
    if (active) {
        double dn[3][8], inv[3][3];
        if (hex_jac(x, 0.0, 0.0, 0.0, dn, inv) <= 0.0) {
            if (a == 0) record_bad(reason, first_bad, e, SVX_ELEM_NONPOSITIVE_JACOBIAN);
            return;
        }
        double* out = ke_all + ke_off[e] + a * 24;
#pragma unroll
        for (int c = 0; c < 24; ++c) out[c] = ke[c];
        if (a == 0) reason[e] = 0;
    }
}

// One wavefront per CSR row; lanes walk the row's entries. The row, and so
// its node's incident elements, their connectivity and Ke row offsets, is
// uniform over the wavefront: taken through readfirstlane, those reads are
// scalar loads, and each lane loads only the Ke entries its column matches.
// Contributions are added in incidence order, then corner order (a node
// listed twice in a degenerate element contributes twice).
__global__ void gather_assembly_kernel(int64_t n_rows, int ndof, const int32_t* __restrict__ row_ptr,
                                       const int32_t* __restrict__ col, const int64_t* __restrict__ inc_ptr,
                                       const int32_t* __restrict__ inc_elem, const int8_t* __restrict__ inc_slot,
                                       const int64_t* __restrict__ conn_ptr, const int32_t* __restrict__ conn,
                                       const uint8_t* __restrict__ kinds, const double* __restrict__ ke,
                                       const int64_t* __restrict__ ke_off, double* __restrict__ vals) {
    // n_rows < 2^31 (int32 CSR indices)
    const int32_t r = __builtin_amdgcn_readfirstlane(
        (int32_t)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave));
    const int lane = threadIdx.x % kWave;
    if (r >= n_rows) return;
    const int32_t i = r / ndof;
    const int a = r % ndof;
    const int64_t ib = inc_ptr[i], ie = inc_ptr[i + 1];
    for (int32_t k = row_ptr[r] + lane; k < row_ptr[r + 1]; k += kWave) {
        const int32_t c = col[k];
        const int32_t j = c / ndof;
        const int b = c % ndof;
        double v = 0.0;
        for (int64_t qq = ib; qq < ie; ++qq) {
            const int32_t e = inc_elem[qq];
            const int li = inc_slot[qq];
            const int64_t cb = conn_ptr[e];
            const int nen = (int)(conn_ptr[e + 1] - cb);
            const int N = nen * ndof;
            const double* kr = ke + ke_off[e] + (int64_t)(li * ndof + a) * N;
            for (int lj = 0; lj < nen; ++lj)
                if (conn[cb + lj] == j) v += kr[lj * ndof + b];
        }
        vals[k] = v;
    }
}

// apply_dirichlet (solver.rs:45-62): one thread per row.
__global__ void dirichlet_kernel(int64_t n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                 const uint8_t* __restrict__ mask, double* __restrict__ vals, double* __restrict__ f) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    bool mr = mask[r] != 0;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
        int32_t c = col[k];
        if (mr || mask[c]) vals[k] = (r == c) ? 1.0 : 0.0;
    }
    if (mr && f) f[r] = 0.0;
}

__global__ void gather_rows_kernel(int64_t n_sel, const int64_t* __restrict__ rows, const int32_t* __restrict__ row_ptr,
                                   const double* __restrict__ vals, const int64_t* __restrict__ off,
                                   double* __restrict__ out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_sel) return;
    int64_t r = rows[q];
    int64_t o = off[q];
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) out[o + (k - row_ptr[r])] = vals[k];
}

// Jacobi inverse diagonal (solver.rs:96-107) and the extreme |K_ii| of
// jacobi_cond (lib.rs:634-644) as ordered bit patterns (|x| >= 0).
__global__ void diag_inv_kernel(int64_t n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                const double* __restrict__ vals, double* __restrict__ dinv, float* __restrict__ dinv32,
                                unsigned long long* absmax, unsigned long long* absmin) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double d = 0.0;
    bool found = false;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
        if (col[k] == r) {
            d = vals[k];
            found = true;
            break;
        }
    double inv = (d == 0.0) ? 0.0 : 1.0 / d;
    if (dinv) dinv[r] = inv;
    if (dinv32) dinv32[r] = (float)inv;
    if (found && absmax) {
        unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(d));
        atomicMax(absmax, bits);
        atomicMin(absmin, bits);
    }
}

// FP32 node-block copy of a CSR matrix whose rows hold full bs x bs node
// blocks (or a lone diagonal for a node outside every element). Values are
// stored component-major, bval[(a*bs + b) * nblk + block], so that lanes
// walking consecutive blocks of a block row read consecutive words.
__global__ void fill_bsr_kernel(int64_t n, int bs, const int32_t* __restrict__ row_ptr,
                                const int32_t* __restrict__ col, const double* __restrict__ vals,
                                const int32_t* __restrict__ brow_ptr, float* __restrict__ bval) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int64_t node = r / bs;
    int a = (int)(r % bs);
    const int64_t nblk = brow_ptr[n / bs];
    int32_t rb = row_ptr[r], re = row_ptr[r + 1];
    int64_t first = brow_ptr[node];
    if (bs > 1 && re - rb == 1) {  // isolated node: one block holding the diagonal
        for (int b = 0; b < bs; ++b) bval[(int64_t)(a * bs + b) * nblk + first] = 0.0f;
    }
    for (int32_t k = rb; k < re; ++k) {
        int32_t nb = (k - rb) / bs;
        int b = col[k] % bs;
        bval[(int64_t)(a * bs + b) * nblk + first + nb] = (float)vals[k];
    }
}

// Node-row sliced ELL copy of a node-block CSR matrix: thread (chunk c, lane
// l) packs node row 64c + l. Value (block k, entry a*bs+b) of that row goes
// to out[(ell_off[c] + k*64) * bs*bs + (a*bs + b)*64 + l], i.e. per chunk
// [block][entry][lane], so that a wavefront reading entry (k, a, b) of its 64
// rows reads 64 consecutive words. Padding blocks and lanes past the last
// node get zeros; a lone node (its rows hold only their diagonal) becomes one
// diagonal block.
template <int BS, typename V>
__global__ void ell_pack_kernel(int64_t n_nodes, int64_t n_chunks, const int32_t* __restrict__ ell_off,
                                const int32_t* __restrict__ ell_nb, const int32_t* __restrict__ row_ptr,
                                const int32_t* __restrict__ brow_ptr, const double* __restrict__ vals,
                                V* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = g / 64;
    const int l = (int)(g % 64);
    if (c >= n_chunks) return;
    const int64_t r = c * 64 + l;
    int32_t nb = 0, r0 = 0;
    bool lone = false;
    if (r < n_nodes) {
        nb = brow_ptr[r + 1] - brow_ptr[r];
        r0 = row_ptr[r * BS];
        lone = row_ptr[r * BS + 1] - r0 != BS * nb;
    }
    V* o = out + (int64_t)ell_off[c] * (BS * BS) + l;
    for (int32_t k = 0; k < ell_nb[c]; ++k)
        for (int a = 0; a < BS; ++a)
            for (int b = 0; b < BS; ++b) {
                double v = 0.0;
                if (k < nb) v = lone ? (a == b ? vals[row_ptr[r * BS + a]] : 0.0) : vals[r0 + a * BS * nb + BS * k + b];
                o[((int64_t)k * BS * BS + a * BS + b) * 64] = (V)v;
            }
}

// (K_orig u)[row] with the row's original values (reactions, stress.rs:103-118).
__global__ void reaction_rows_kernel(int64_t n_sel, const int64_t* __restrict__ rows,
                                     const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                     const int64_t* __restrict__ off, const double* __restrict__ orig,
                                     const double* __restrict__ u, double* __restrict__ out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_sel) return;
    int64_t r = rows[q];
    double o = 0.0;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) o += orig[off[q] + (k - row_ptr[r])] * u[col[k]];
    out[q] = o;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

void launch_element_ke(int kind, int64_t n_sel, const int32_t* sel, const int64_t* conn_ptr, const int32_t* conn,
                       const double* coords, const double* E, const double* nu, double thickness, int ndof,
                       double* ke, const int64_t* ke_off, uint8_t* reason, unsigned long long* first_bad,
                       hipStream_t s) {
    (void)ndof;
    if (n_sel <= 0) return;
    const double g = 1.0 / std::sqrt(3.0);
    if (kind == SVX_CPS3) {
        ke_cps3_kernel<<<grid_for(n_sel, 128), 128, 0, s>>>(n_sel, sel, conn_ptr, conn, coords, E, nu, thickness, ke,
                                                             ke_off, reason, first_bad);
    } else if (kind == SVX_CPS4) {
        ke_cps4_kernel<<<grid_for(n_sel, 128), 128, 0, s>>>(n_sel, sel, conn_ptr, conn, coords, E, nu, thickness, g,
                                                             ke, ke_off, reason, first_bad);
    } else {
        ke_c3d8_kernel<<<grid_for(n_sel, kHexPerBlock), kHexPerBlock * 24, 0, s>>>(
            n_sel, sel, conn_ptr, conn, coords, E, nu, g, ke, ke_off, reason, first_bad);
    }
    SVX_HIP(hipGetLastError());
}

void launch_gather_assembly(int64_t n_rows, int ndof, const int32_t* row_ptr, const int32_t* col,
                            const int64_t* inc_ptr, const int32_t* inc_elem, const int8_t* inc_slot,
                            const int64_t* conn_ptr, const int32_t* conn, const uint8_t* kinds, const double* ke,
                            const int64_t* ke_off, double* vals, hipStream_t s) {
    if (n_rows <= 0) return;
    gather_assembly_kernel<<<grid_for(n_rows * kWave, kBlock), kBlock, 0, s>>>(
        n_rows, ndof, row_ptr, col, inc_ptr, inc_elem, inc_slot, conn_ptr, conn, kinds, ke, ke_off, vals);
    SVX_HIP(hipGetLastError());
}

void launch_dirichlet(int64_t n_rows, const int32_t* row_ptr, const int32_t* col, const uint8_t* mask, double* vals,
                      double* f, hipStream_t s) {
    if (n_rows <= 0) return;
    dirichlet_kernel<<<grid_for(n_rows, kBlock), kBlock, 0, s>>>(n_rows, row_ptr, col, mask, vals, f);
    SVX_HIP(hipGetLastError());
}

void launch_gather_rows(int64_t n_sel, const int64_t* rows, const int32_t* row_ptr, const double* vals,
                        const int64_t* out_off, double* out, hipStream_t s) {
    if (n_sel <= 0) return;
    gather_rows_kernel<<<grid_for(n_sel, kBlock), kBlock, 0, s>>>(n_sel, rows, row_ptr, vals, out_off, out);
    SVX_HIP(hipGetLastError());
}

void launch_diag_inv(int64_t n_rows, const int32_t* row_ptr, const int32_t* col, const double* vals, double* dinv,
                     float* dinv32, unsigned long long* absmax, unsigned long long* absmin, hipStream_t s) {
    if (n_rows <= 0) return;
    diag_inv_kernel<<<grid_for(n_rows, kBlock), kBlock, 0, s>>>(n_rows, row_ptr, col, vals, dinv, dinv32, absmax,
                                                                 absmin);
    SVX_HIP(hipGetLastError());
}

void launch_fill_bsr(int64_t n_rows, int bs, const int32_t* row_ptr, const int32_t* col, const double* vals,
                     const int32_t* brow_ptr, float* bval, hipStream_t s) {
    if (n_rows <= 0) return;
    fill_bsr_kernel<<<grid_for(n_rows, kBlock), kBlock, 0, s>>>(n_rows, bs, row_ptr, col, vals, brow_ptr, bval);
    SVX_HIP(hipGetLastError());
}

void launch_ell_pack(int bs, int64_t n_nodes, const int32_t* ell_off, const int32_t* ell_nb, const int32_t* row_ptr,
                     const int32_t* brow_ptr, const double* vals, double* out64, float* out32, hipStream_t s) {
    if (n_nodes <= 0) return;
    const int64_t n_chunks = (n_nodes + 63) / 64;
    const unsigned grid = grid_for(n_chunks * 64, kBlock);
    if (bs == 2) {
        if (out64) ell_pack_kernel<2, double><<<grid, kBlock, 0, s>>>(n_nodes, n_chunks, ell_off, ell_nb, row_ptr, brow_ptr, vals, out64);
        if (out32) ell_pack_kernel<2, float><<<grid, kBlock, 0, s>>>(n_nodes, n_chunks, ell_off, ell_nb, row_ptr, brow_ptr, vals, out32);
    } else if (bs == 3) {
        if (out64) ell_pack_kernel<3, double><<<grid, kBlock, 0, s>>>(n_nodes, n_chunks, ell_off, ell_nb, row_ptr, brow_ptr, vals, out64);
        if (out32) ell_pack_kernel<3, float><<<grid, kBlock, 0, s>>>(n_nodes, n_chunks, ell_off, ell_nb, row_ptr, brow_ptr, vals, out32);
    } else {
        throw std::invalid_argument("ELL layout needs node blocks of 2 or 3 dofs");
    }
    SVX_HIP(hipGetLastError());
}

void launch_reaction_rows(int64_t n_sel, const int64_t* rows, const int32_t* row_ptr, const int32_t* col,
                          const int64_t* off, const double* orig_vals, const double* u, double* out, hipStream_t s) {
    if (n_sel <= 0) return;
    reaction_rows_kernel<<<grid_for(n_sel, kBlock), kBlock, 0, s>>>(n_sel, rows, row_ptr, col, off, orig_vals, u, out);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
