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
// (quad_b), evaluated as (t / detJ) B~^T D (B~ u).
__host__ __device__ inline void mf_cps4_forces(const double (&x)[4][2], const double (&u)[4][2], const MfMaterial& m,
                                               double (&f)[4][2]) {
    const double g = 1.0 / __builtin_sqrt(3.0);
    const double cx[4] = {-1.0, 1.0, 1.0, -1.0}, cy[4] = {-1.0, -1.0, 1.0, 1.0};
    for (int a = 0; a < 4; ++a) f[a][0] = f[a][1] = 0.0;
#pragma unroll
    for (int gp = 0; gp < 4; ++gp) {
        const double xi = (gp & 2) ? g : -g, eta = (gp & 1) ? g : -g;
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
        const double det = j00 * j11 - j01 * j10;
        double bx[4], by[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            bx[i] = j11 * dxi[i] - j01 * deta[i];
            by[i] = j00 * deta[i] - j10 * dxi[i];
        }
        double exx = 0.0, eyy = 0.0, gxy = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            exx += bx[i] * u[i][0];
            eyy += by[i] * u[i][1];
            gxy += by[i] * u[i][0] + bx[i] * u[i][1];
        }
        const double sc = m.t / det;
        const double sxx = sc * (m.c11 * exx + m.c12 * eyy);
        const double syy = sc * (m.c12 * exx + m.c11 * eyy);
        const double sxy = sc * (m.c33 * gxy);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            f[i][0] += bx[i] * sxx + by[i] * sxy;
            f[i][1] += by[i] * syy + bx[i] * sxy;
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
