// Stress recovery on gfx950 (built with -ffp-contract=off, reference order).
//
// element_stress_kernel: per element, the centroid stress and its von Mises
// value (stress.rs:22-42 with element.rs:853-943), and the Gauss-point
// stresses extrapolated to the element's corners (element.rs:945-1053).
// nodal_average_kernel: per node, the unweighted average of the corner
// stresses of its incident elements, summed in element order, then von Mises
// (stress.rs:53-98). Strain sums skip the structural zeros of B, which only
// add signed zeros.
#include "svx_internal.hpp"

namespace svx {
namespace {

__device__ __forceinline__ double von_mises(const double* s) {
    return sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) + (s[2] - s[0]) * (s[2] - s[0])) +
                3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

// plane_stress (element.rs:866-880) from the strain (e0, e1, e2).
__device__ __forceinline__ void plane_stress(const double* d, const double e[3], double* s) {
    s[0] = d[0] * e[0] + d[1] * e[1];
    s[1] = d[3] * e[0] + d[4] * e[1];
    s[2] = 0.0;
    s[3] = d[8] * e[2];
    s[4] = 0.0;
    s[5] = 0.0;
}

__device__ __forceinline__ void d_matrix(double e, double nu, double* d) {
    double c = e / (1.0 - nu * nu);
    d[0] = c, d[1] = nu * c, d[2] = 0.0;
    d[3] = nu * c, d[4] = c, d[5] = 0.0;
    d[6] = 0.0, d[7] = 0.0, d[8] = c * (1.0 - nu) / 2.0;
}

// Strain of the bilinear quad at (xi, eta) (quad_b, element.rs:147-171).
__device__ __forceinline__ void quad_strain(const double (*x)[2], const double* ue, double xi, double eta, double e[3]) {
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
    double inv = 1.0 / detj;
    e[0] = 0.0, e[1] = 0.0, e[2] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double dnx = inv * (j11 * dxi[i] - j01 * deta[i]);
        double dny = inv * (j00 * deta[i] - j10 * dxi[i]);
        // column 2i: rows 0 (dnx) and 2 (dny); column 2i+1: rows 1 (dny), 2 (dnx)
        e[0] += dnx * ue[2 * i];
        e[2] += dny * ue[2 * i];
        e[1] += dny * ue[2 * i + 1];
        e[2] += dnx * ue[2 * i + 1];
    }
}

// Strain of the trilinear hex at a natural point (hex_b, element.rs:479-521),
// with the reference's inverse (sign pattern of element.rs:447-462).
__device__ __forceinline__ void hex_strain(const double (*x)[3], const double* ue, double xi, double eta, double zeta,
                                           double e[6]) {
    const double ca[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
    const double cb[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
    const double cc[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
    double dn[3][8];
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
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double id = 1.0 / det;
    double inv[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int r = j, c = i;
            int ra = (r + 1) % 3, rb = (r + 2) % 3, cd = (c + 1) % 3, ce = (c + 2) % 3;
            double sign = ((r + c) % 2 == 0) ? 1.0 : -1.0;
            inv[i][j] = sign * (m[ra][cd] * m[rb][ce] - m[ra][ce] * m[rb][cd]) * id;
        }
#pragma unroll
    for (int r = 0; r < 6; ++r) e[r] = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double g[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) g[b] += inv[b][a] * dn[a][i];
        double ux = ue[3 * i], uy = ue[3 * i + 1], uz = ue[3 * i + 2];
        // column 3i: rows 0 g0, 3 g1, 5 g2
        e[0] += g[0] * ux;
        e[3] += g[1] * ux;
        e[5] += g[2] * ux;
        // column 3i+1: rows 1 g1, 3 g0, 4 g2
        e[1] += g[1] * uy;
        e[3] += g[0] * uy;
        e[4] += g[2] * uy;
        // column 3i+2: rows 2 g2, 4 g1, 5 g0
        e[2] += g[2] * uz;
        e[4] += g[1] * uz;
        e[5] += g[0] * uz;
    }
}

template <int KIND>
__global__ void element_stress_kernel(int64_t n_sel, const int32_t* __restrict__ sel,
                                      const int64_t* __restrict__ conn_ptr, const int32_t* __restrict__ conn,
                                      const double* __restrict__ X, const double* __restrict__ E,
                                      const double* __restrict__ NU, int ndof, const double* __restrict__ u, double g,
                                      double sqrt3, double* __restrict__ vm_elem, double* __restrict__ nodal,
                                      const int64_t* __restrict__ ns_off) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_sel) return;
    int64_t e = sel[q];
    const int32_t* cn = conn + conn_ptr[e];
    double* out = nodal + ns_off[e];
    auto h = [&](double s) { return 0.5 * (1.0 + s * sqrt3); };
    if (KIND == SVX_CPS3) {
        double x[3][2], ue[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k][0] = X[cn[k] * 3], x[k][1] = X[cn[k] * 3 + 1];
            ue[2 * k] = u[(int64_t)cn[k] * ndof], ue[2 * k + 1] = u[(int64_t)cn[k] * ndof + 1];
        }
        double b[3] = {x[1][1] - x[2][1], x[2][1] - x[0][1], x[0][1] - x[1][1]};
        double c[3] = {x[2][0] - x[1][0], x[0][0] - x[2][0], x[1][0] - x[0][0]};
        double two_a = b[0] * c[1] - b[1] * c[0];
        double inv = 1.0 / two_a;
        double ep[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ep[0] += (b[i] * inv) * ue[2 * i];
            ep[2] += (c[i] * inv) * ue[2 * i];
            ep[1] += (c[i] * inv) * ue[2 * i + 1];
            ep[2] += (b[i] * inv) * ue[2 * i + 1];
        }
        double d[9], s[6];
        d_matrix(E[e], NU[e], d);
        plane_stress(d, ep, s);
        vm_elem[e] = von_mises(s);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int c2 = 0; c2 < 6; ++c2) out[k * 6 + c2] = s[c2];
    } else if (KIND == SVX_CPS4) {
        double x[4][2], ue[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k][0] = X[cn[k] * 3], x[k][1] = X[cn[k] * 3 + 1];
            ue[2 * k] = u[(int64_t)cn[k] * ndof], ue[2 * k + 1] = u[(int64_t)cn[k] * ndof + 1];
        }
        double d[9], s[6], ep[3];
        d_matrix(E[e], NU[e], d);
        quad_strain(x, ue, 0.0, 0.0, ep);
        plane_stress(d, ep, s);
        vm_elem[e] = von_mises(s);
        const double pts[2] = {-g, g};
        double gp[4][6];
        for (int j = 0; j < 4; ++j) {
            quad_strain(x, ue, pts[j >> 1], pts[j & 1], ep);
            plane_stress(d, ep, gp[j]);
        }
        const double cx[4] = {-1.0, 1.0, 1.0, -1.0};
        const double cy[4] = {-1.0, -1.0, 1.0, 1.0};
        const double gx[4] = {-1.0, -1.0, 1.0, 1.0};
        const double gy[4] = {-1.0, 1.0, -1.0, 1.0};
        for (int i = 0; i < 4; ++i) {
            double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int j = 0; j < 4; ++j) {
                double w = h(cx[i] * gx[j]) * h(cy[i] * gy[j]);
#pragma unroll
                for (int k = 0; k < 6; ++k) o[k] += w * gp[j][k];
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) out[i * 6 + k] = o[k];
        }
    } else {
        double x[8][3], ue[24];
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                x[k][b] = X[cn[k] * 3 + b];
                ue[3 * k + b] = u[(int64_t)cn[k] * 3 + b];
            }
        double Ee = E[e], nu = NU[e];
        double lam = Ee * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double mu = Ee / (2.0 * (1.0 + nu));
        double diag = lam + 2.0 * mu;
        double d[6][6];
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c2 = 0; c2 < 6; ++c2) d[r][c2] = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c2 = 0; c2 < 3; ++c2) d[r][c2] = (r == c2) ? diag : lam;
        d[3][3] = mu, d[4][4] = mu, d[5][5] = mu;
        auto solid = [&](const double ep[6], double* s) {
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                double acc = -0.0;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc += d[r][k] * ep[k];
                s[r] = acc;
            }
        };
        double ep[6], s[6];
        hex_strain(x, ue, 0.0, 0.0, 0.0, ep);
        solid(ep, s);
        vm_elem[e] = von_mises(s);
        const double pts[2] = {-g, g};
        double o[8][6];
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int k = 0; k < 6; ++k) o[i][k] = 0.0;
        const double ca[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
        const double cb[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
        const double cc[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
        // Gauss points in stiffness-loop order; o[i] accumulates over j in
        // the same order as extrapolate_to_corners3's inner loop.
        for (int j = 0; j < 8; ++j) {
            double sx = (j >> 2) ? 1.0 : -1.0, sy = ((j >> 1) & 1) ? 1.0 : -1.0, sz = (j & 1) ? 1.0 : -1.0;
            hex_strain(x, ue, pts[j >> 2], pts[(j >> 1) & 1], pts[j & 1], ep);
            solid(ep, s);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                double w = h(ca[i] * sx) * h(cb[i] * sy) * h(cc[i] * sz);
#pragma unroll
                for (int k = 0; k < 6; ++k) o[i][k] += w * s[k];
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int k = 0; k < 6; ++k) out[i * 6 + k] = o[i][k];
    }
}

__global__ void nodal_average_kernel(int64_t n_nodes, const int64_t* __restrict__ inc_ptr,
                                     const int32_t* __restrict__ inc_elem, const int8_t* __restrict__ inc_slot,
                                     const int64_t* __restrict__ ns_off, const double* __restrict__ nodal,
                                     double* __restrict__ vm_node) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    double sum[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int64_t c = inc_ptr[i + 1] - inc_ptr[i];
    if (c == 0) {
        vm_node[i] = 0.0;
        return;
    }
    for (int64_t q = inc_ptr[i]; q < inc_ptr[i + 1]; ++q) {
        const double* s = nodal + ns_off[inc_elem[q]] + inc_slot[q] * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) sum[k] += s[k];
    }
    double cd = (double)c, avg[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) avg[k] = sum[k] / cd;
    vm_node[i] = von_mises(avg);
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

void launch_element_stress(int kind, int64_t n_sel, const int32_t* sel, const int64_t* conn_ptr, const int32_t* conn,
                           const double* coords, const double* E, const double* nu, int ndof, const double* u,
                           double* vm_elem, double* nodal_stress, const int64_t* ns_off, hipStream_t s) {
    if (n_sel <= 0) return;
    const double g = 1.0 / std::sqrt(3.0), sqrt3 = std::sqrt(3.0);
    if (kind == SVX_CPS3)
        element_stress_kernel<SVX_CPS3><<<grid_for(n_sel, 128), 128, 0, s>>>(
            n_sel, sel, conn_ptr, conn, coords, E, nu, ndof, u, g, sqrt3, vm_elem, nodal_stress, ns_off);
    else if (kind == SVX_CPS4)
        element_stress_kernel<SVX_CPS4><<<grid_for(n_sel, 128), 128, 0, s>>>(
            n_sel, sel, conn_ptr, conn, coords, E, nu, ndof, u, g, sqrt3, vm_elem, nodal_stress, ns_off);
    else
        element_stress_kernel<SVX_C3D8><<<grid_for(n_sel, 64), 64, 0, s>>>(
            n_sel, sel, conn_ptr, conn, coords, E, nu, ndof, u, g, sqrt3, vm_elem, nodal_stress, ns_off);
    SVX_HIP(hipGetLastError());
}

void launch_nodal_average(int64_t n_nodes, const int64_t* inc_ptr, const int32_t* inc_elem, const int8_t* inc_slot,
                          const int64_t* ns_off, const double* nodal_stress, double* vm_node, hipStream_t s) {
    if (n_nodes <= 0) return;
    nodal_average_kernel<<<grid_for(n_nodes, kBlock), kBlock, 0, s>>>(n_nodes, inc_ptr, inc_elem, inc_slot, ns_off,
                                                                       nodal_stress, vm_node);
    SVX_HIP(hipGetLastError());
}

}  // namespace svx
