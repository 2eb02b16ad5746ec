// Element forces of the matrix-free plane-stress operator, shared by the
// HIP kernel and the host check of the patch planner.
//
// f_e = K_e u_e without forming K_e: the element formulas of element.rs
// (B of cps3_b_matrix / quad_b, D of d_matrix, the 2x2 Gauss points of
// cps4_stiffness, weights t detJ and t |2A| / 2) applied to the element's
// nodal values, the same linear map as the stiffness btdb assembles, up to
// rounding.
#pragma once

namespace svx {

// Plane-stress D (element.rs:10-17: c = E / (1 - nu^2), [c, nu c, 0; nu c, c,
// 0; 0, 0, c (1 - nu) / 2]) and the model thickness.
struct MfMaterial {
    double c11, c12, c33, t;
};

// CPS4: sum over the Gauss points of t detJ B^T D B u with B = B~ / detJ
// (quad_b), evaluated in contracted form: with the shape-function
// derivatives dN_i = (dxi_i, deta_i), the Jacobian J = sum_i dN_i x_i^T and
// the natural-coordinate gradient Gn = sum_i dN_i u_i^T, the rows of B~ are
// adj(J) dN_i, so B~ u = adj(J) Gn (the strains times detJ) and the nodal
// forces are dN_i^T (adj(J)^T S) with S the stress times t / detJ. The same
// linear map as the stiffness btdb assembles, up to rounding.
__host__ __device__ inline void mf_cps4_forces(const double (&x)[4][2], const double (&u)[4][2], const MfMaterial& m,
                                               double (&f)[4][2]) {
    const double g = 1.0 / __builtin_sqrt(3.0);
    const double cx[4] = {-1.0, 1.0, 1.0, -1.0}, cy[4] = {-1.0, -1.0, 1.0, 1.0};
    for (int a = 0; a < 4; ++a) f[a][0] = f[a][1] = 0.0;
#pragma unroll
    for (int gp = 0; gp < 4; ++gp) {
        const double xi = (gp & 2) ? g : -g, eta = (gp & 1) ? g : -g;
        double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0, gx0 = 0.0, gx1 = 0.0, ge0 = 0.0, ge1 = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double dxi = 0.25 * cx[i] * (1.0 + eta * cy[i]), deta = 0.25 * cy[i] * (1.0 + xi * cx[i]);
            j00 += dxi * x[i][0];
            j01 += dxi * x[i][1];
            j10 += deta * x[i][0];
            j11 += deta * x[i][1];
            gx0 += dxi * u[i][0];
            gx1 += dxi * u[i][1];
            ge0 += deta * u[i][0];
            ge1 += deta * u[i][1];
        }
        const double det = j00 * j11 - j01 * j10;
        // B~ u: rows of B~ are (j11 dxi - j01 deta, j00 deta - j10 dxi)
        const double exx = j11 * gx0 - j01 * ge0;
        const double eyy = j00 * ge1 - j10 * gx1;
        const double gxy = (j00 * ge0 - j10 * gx0) + (j11 * gx1 - j01 * ge1);
        const double sc = m.t / det;
        const double sxx = sc * (m.c11 * exx + m.c12 * eyy);
        const double syy = sc * (m.c12 * exx + m.c11 * eyy);
        const double sxy = sc * (m.c33 * gxy);
        // f_i += B~_i^T s = dxi (j11 sxx - j10 sxy, j11 sxy - j10 syy) + deta (j00 sxy - j01 sxx, j00 syy - j01 sxy)
        const double t00 = j11 * sxx - j10 * sxy, t01 = j11 * sxy - j10 * syy;
        const double t10 = j00 * sxy - j01 * sxx, t11 = j00 * syy - j01 * sxy;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double dxi = 0.25 * cx[i] * (1.0 + eta * cy[i]), deta = 0.25 * cy[i] * (1.0 + xi * cx[i]);
            f[i][0] += dxi * t00 + deta * t10;
            f[i][1] += dxi * t01 + deta * t11;
        }
    }
}

// CPS3: (t |2A| / 2) B^T D B u with B = B~ / 2A (cps3_b_matrix), evaluated as
// (t |2A| / 2 / (2A)^2) B~^T D (B~ u).
__host__ __device__ inline void mf_cps3_forces(const double (&x)[3][2], const double (&u)[3][2], const MfMaterial& m,
                                               double (&f)[3][2]) {
    const double b[3] = {x[1][1] - x[2][1], x[2][1] - x[0][1], x[0][1] - x[1][1]};
    const double c[3] = {x[2][0] - x[1][0], x[0][0] - x[2][0], x[1][0] - x[0][0]};
    const double two_a = b[0] * c[1] - b[1] * c[0];
    double exx = 0.0, eyy = 0.0, gxy = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        exx += b[i] * u[i][0];
        eyy += c[i] * u[i][1];
        gxy += c[i] * u[i][0] + b[i] * u[i][1];
    }
    const double area2 = two_a < 0.0 ? -two_a : two_a;
    const double sc = m.t * area2 * 0.5 / (two_a * two_a);
    const double sxx = sc * (m.c11 * exx + m.c12 * eyy);
    const double syy = sc * (m.c12 * exx + m.c11 * eyy);
    const double sxy = sc * (m.c33 * gxy);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        f[i][0] = b[i] * sxx + c[i] * sxy;
        f[i][1] = c[i] * syy + b[i] * sxy;
    }
}

template <int NEN>
__host__ __device__ inline void mf_forces(const double (&x)[NEN][2], const double (&u)[NEN][2], const MfMaterial& m,
                                          double (&f)[NEN][2]) {
    if constexpr (NEN == 4)
        mf_cps4_forces(x, u, m, f);
    else
        mf_cps3_forces(x, u, m, f);
}

}  // namespace svx

namespace svx {

// One Gauss point of a C3D8 element (c3d8_stiffness, element.rs:535-561:
// 2x2x2 points, xi outer, zeta inner, unit weights): adds
// detJ B^T D (B u) for the point gp (bits 2, 1, 0: the signs of xi, eta,
// zeta) to f, where B is hex_b's (element.rs:466-521) with the reference's
// J^-1 (inv3, element.rs:447-462, whose odd-parity entries carry an extra
// sign) and D is d3_matrix (element.rs:33-47: diag = lambda + 2 mu, off =
// lambda, shear = mu; Voigt order xx, yy, zz, xy, yz, xz). x and u are the
// element's nodal coordinates and values, [node][3]; f receives [node][3].
template <typename In, typename Out>
__host__ __device__ inline void mf_hex_gp_forces(int gp, const In& x, const In& u, double diag, double lam, double mu,
                                                 Out& f) {
    const double g = 1.0 / __builtin_sqrt(3.0);
    const double xi = (gp & 4) ? g : -g, eta = (gp & 2) ? g : -g, zeta = (gp & 1) ? g : -g;
    const double ca[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
    const double cb[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
    const double cc[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
    // dN_i/d(xi, eta, zeta) = 1/8 (a (1 + eta b)(1 + zeta c), ...): with the
    // corner signs a, b, c = +-1 each derivative takes one of four magnitudes
    const double x8[2] = {0.125 * (1.0 - xi), 0.125 * (1.0 + xi)};
    const double e1[2] = {1.0 - eta, 1.0 + eta}, z1[2] = {1.0 - zeta, 1.0 + zeta};
    const double e8[2] = {0.125 * e1[0], 0.125 * e1[1]};
    const double dxi[2][2] = {{e8[0] * z1[0], e8[0] * z1[1]}, {e8[1] * z1[0], e8[1] * z1[1]}};
    const double deta[2][2] = {{x8[0] * z1[0], x8[0] * z1[1]}, {x8[1] * z1[0], x8[1] * z1[1]}};
    const double dzeta[2][2] = {{x8[0] * e1[0], x8[0] * e1[1]}, {x8[1] * e1[0], x8[1] * e1[1]}};
    auto dn = [&](int i, double (&d)[3]) {
        const int ia = ca[i] > 0.0, ib = cb[i] > 0.0, ic = cc[i] > 0.0;
        d[0] = ca[i] > 0.0 ? dxi[ib][ic] : -dxi[ib][ic];
        d[1] = cb[i] > 0.0 ? deta[ia][ic] : -deta[ia][ic];
        d[2] = cc[i] > 0.0 ? dzeta[ia][ib] : -dzeta[ia][ib];
    };
    double m[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double d[3];
        dn(i, d);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) m[a][b] += d[a] * x[i][b];
    }
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    const double id = 1.0 / det;
    double inv[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int r = j, c = i;
            const int ra = (r + 1) % 3, rb = (r + 2) % 3, cd = (c + 1) % 3, ce = (c + 2) % 3;
            const double sign = ((r + c) % 2 == 0) ? 1.0 : -1.0;
            inv[i][j] = sign * (m[ra][cd] * m[rb][ce] - m[ra][ce] * m[rb][cd]) * id;
        }
    // B's node columns are g_i = inv dN_i, so with G = sum_i dN_i u_i^T the
    // displacement gradient is H = inv G (H[b][c] = sum_i g_i[b] u_i[c]) and
    // the strain (engineering shear xy, yz, xz) comes from H
    double G[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double d[3];
        dn(i, d);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) G[a][c] += d[a] * u[i][c];
    }
    double H[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) H[b][c] = inv[b][0] * G[0][c] + inv[b][1] * G[1][c] + inv[b][2] * G[2][c];
    const double exx = H[0][0], eyy = H[1][1], ezz = H[2][2];
    const double gxy = H[1][0] + H[0][1], gyz = H[2][1] + H[1][2], gxz = H[2][0] + H[0][2];
    const double s0 = det * (diag * exx + lam * (eyy + ezz));
    const double s1 = det * (diag * eyy + lam * (exx + ezz));
    const double s2 = det * (diag * ezz + lam * (exx + eyy));
    const double s3 = det * (mu * gxy), s4 = det * (mu * gyz), s5 = det * (mu * gxz);
    // f_i = B_i^T sigma = sum_b g_i[b] S[b][.] = dN_i^T T with T = inv^T S
    const double S[3][3] = {{s0, s3, s5}, {s3, s1, s4}, {s5, s4, s2}};
    double T[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) T[a][c] = inv[0][a] * S[0][c] + inv[1][a] * S[1][c] + inv[2][a] * S[2][c];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double d[3];
        dn(i, d);
#pragma unroll
        for (int c = 0; c < 3; ++c) f[i][c] = d[0] * T[0][c] + d[1] * T[1][c] + d[2] * T[2][c];
    }
}

}  // namespace svx
