// Test infrastructure only — NOT part of the product.
//
// CPU restatement of Solvix's linear-static path (typezo/Solvix,
// crates/solvix-core/src), used as the parity oracle and as the CPU baseline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Followed, function by function (file:line in the reference):
//   element.rs:10-17    d_matrix              element.rs:33-47   d3_matrix
//   element.rs:91-129   cps3_b_matrix / cps3_stiffness
//   element.rs:131-211  quad_shape_derivs / quad_b / cps4_stiffness
//   element.rs:427-561  HEX_CORNERS / det3 / inv3 / hex_shape_derivs / hex_b /
//                       c3d8_stiffness
//   element.rs:829-851  btdb                  element.rs:853-923 strain /
//                       plane_stress / solid_stress / von_mises
//   element.rs:945-1053 nodal_stresses / extrapolate_to_corners(3)
//   assembly.rs:14-72   worker_threads / Csr::mat_vec_with
//   assembly.rs:85-133  csr_from_triplets / scatter
//   assembly.rs:138-386 element_entry / assemble_with (CLOADs only)
//   solver.rs:15-149    constrained_dofs / apply_dirichlet / pcg_with_threads
//   stress.rs:11-118    gather_ue / von_mises_at_centroids / nodal_von_mises /
//                       reactions
//   lib.rs:106-108      Mesh::node_index      lib.rs:357-367 material_for_element
//   lib.rs:406-437      Model::validate       lib.rs:593-716 solve_with_thermal
//   lib.rs:725-752      ndof_per_node
//
// Floating point: built with -O2 -ffp-contract=off (rustc never fuses a*b+c),
// every sum is the reference's sequential left fold in its written order.
// One deliberate difference: csr_from_triplets sums duplicate (row, col)
// contributions in triplet order (a stable sort), where the reference's
// sort_unstable_by_key leaves that order to Rust's sort implementation;
// the two agree to a few ULP per entry. oracle_opts::dup_seed != 0 shuffles
// each row's triplets (seeded) before the stable sort, so tests can measure
// how far the duplicate order alone moves iteration counts and residuals.
// Stress recovery recomputes the per-element B matrices from the coordinates
// instead of storing them (identical arithmetic, less memory).

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/svx.h"

namespace {

using Clock = std::chrono::steady_clock;
static double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// ---------------------------------------------------------------- elements
using D3 = std::array<std::array<double, 3>, 3>;
using D6 = std::array<std::array<double, 6>, 6>;

// element.rs:10-17
D3 d_matrix(double e, double nu) {
    double c = e / (1.0 - nu * nu);
    D3 d{};
    d[0] = {c, nu * c, 0.0};
    d[1] = {nu * c, c, 0.0};
    d[2] = {0.0, 0.0, c * (1.0 - nu) / 2.0};
    return d;
}

// element.rs:33-47
D6 d3_matrix(double e, double nu) {
    double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = e / (2.0 * (1.0 + nu));
    double diag = lam + 2.0 * mu;
    D6 d{};
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) d[r][c] = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d[r][c] = (r == c) ? diag : lam;
    d[3][3] = mu;
    d[4][4] = mu;
    d[5][5] = mu;
    return d;
}

// element.rs:829-851 — ke += w * B^T D B, B is S x N (row-major b[r*N+c]).
template <int N, int S>
void btdb(double* ke, const double* b, const double* d /*SxS*/, double w) {
    double db[S * N];
    for (int r = 0; r < S; ++r)
        for (int c = 0; c < N; ++c) {
            double s = -0.0;
            for (int k = 0; k < S; ++k) s += d[r * S + k] * b[k * N + c];
            db[r * N + c] = s;
        }
    for (int a = 0; a < N; ++a)
        for (int c = a; c < N; ++c) {
            double s = -0.0;
            for (int r = 0; r < S; ++r) s += b[r * N + a] * db[r * N + c];
            double v = w * s;
            ke[a * N + c] += v;
            if (a != c) ke[c * N + a] += v;
        }
}

// element.rs:91-115. Returns false for a zero signed area.
bool cps3_b_matrix(const double (*x)[2], double* bm /*3x6*/, double* two_a_out) {
    double b[3] = {x[1][1] - x[2][1], x[2][1] - x[0][1], x[0][1] - x[1][1]};
    double c[3] = {x[2][0] - x[1][0], x[0][0] - x[2][0], x[1][0] - x[0][0]};
    double two_a = b[0] * c[1] - b[1] * c[0];
    if (two_a == 0.0) return false;
    double inv = 1.0 / two_a;
    for (int i = 0; i < 18; ++i) bm[i] = 0.0;
    for (int i = 0; i < 3; ++i) {
        bm[0 * 6 + 2 * i] = b[i] * inv;
        bm[1 * 6 + 2 * i + 1] = c[i] * inv;
        bm[2 * 6 + 2 * i] = c[i] * inv;
        bm[2 * 6 + 2 * i + 1] = b[i] * inv;
    }
    *two_a_out = two_a;
    return true;
}

// element.rs:120-129
bool cps3_stiffness(const double (*x)[2], const D3& d, double t, double* ke /*6x6*/) {
    double bm[18], two_a;
    if (!cps3_b_matrix(x, bm, &two_a)) return false;
    for (int i = 0; i < 36; ++i) ke[i] = 0.0;
    btdb<6, 3>(ke, bm, &d[0][0], t * std::fabs(two_a) / 2.0);
    return true;
}

const double QUAD_CORNERS[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};

// element.rs:134-171
bool quad_b(const double (*x)[2], double xi, double eta, double* bm /*3x8*/, double* detj_out) {
    double dxi[4], deta[4];
    for (int i = 0; i < 4; ++i) {
        double xi_i = QUAD_CORNERS[i][0], eta_i = QUAD_CORNERS[i][1];
        dxi[i] = 0.25 * xi_i * (1.0 + eta * eta_i);
        deta[i] = 0.25 * eta_i * (1.0 + xi * xi_i);
    }
    double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0;
    for (int i = 0; i < 4; ++i) {
        j00 += dxi[i] * x[i][0];
        j01 += dxi[i] * x[i][1];
        j10 += deta[i] * x[i][0];
        j11 += deta[i] * x[i][1];
    }
    double detj = j00 * j11 - j01 * j10;
    if (detj <= 0.0) return false;
    double inv = 1.0 / detj;
    for (int i = 0; i < 24; ++i) bm[i] = 0.0;
    for (int i = 0; i < 4; ++i) {
        double dnx = inv * (j11 * dxi[i] - j01 * deta[i]);
        double dny = inv * (j00 * deta[i] - j10 * dxi[i]);
        bm[0 * 8 + 2 * i] = dnx;
        bm[1 * 8 + 2 * i + 1] = dny;
        bm[2 * 8 + 2 * i] = dny;
        bm[2 * 8 + 2 * i + 1] = dnx;
    }
    *detj_out = detj;
    return true;
}

double gauss_g() { return 1.0 / std::sqrt(3.0); }

// element.rs:186-211 (stiffness part; B matrices are recomputed at recovery)
bool cps4_stiffness(const double (*x)[2], const D3& d, double t, double* ke /*8x8*/) {
    const double g = gauss_g();
    const double pts[2] = {-g, g};
    for (int i = 0; i < 64; ++i) ke[i] = 0.0;
    double bm[24], detj;
    for (double xi : pts)
        for (double eta : pts) {
            if (!quad_b(x, xi, eta, bm, &detj)) return false;
            btdb<8, 3>(ke, bm, &d[0][0], t * detj);
        }
    return quad_b(x, 0.0, 0.0, bm, &detj);
}

const double HEX_CORNERS[8][3] = {{-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0},
                                  {-1.0, 1.0, -1.0},  {-1.0, -1.0, 1.0}, {1.0, -1.0, 1.0},
                                  {1.0, 1.0, 1.0},    {-1.0, 1.0, 1.0}};

// element.rs:440-444
double det3(const double m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
           m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// element.rs:447-462
void inv3(const double m[3][3], double det, double out[3][3]) {
    auto cof = [&](int r, int c) {
        int a = (r + 1) % 3, b = (r + 2) % 3;
        int d = (c + 1) % 3, e = (c + 2) % 3;
        double sign = ((r + c) % 2 == 0) ? 1.0 : -1.0;
        return sign * (m[a][d] * m[b][e] - m[a][e] * m[b][d]);
    };
    double id = 1.0 / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[i][j] = cof(j, i) * id;
}

// element.rs:466-521
bool hex_b(const double (*x)[3], double xi, double eta, double zeta, double* bm /*6x24*/,
           double* detj_out) {
    double dn[3][8];
    for (int i = 0; i < 8; ++i) {
        double a = HEX_CORNERS[i][0], b = HEX_CORNERS[i][1], c = HEX_CORNERS[i][2];
        dn[0][i] = 0.125 * a * (1.0 + eta * b) * (1.0 + zeta * c);
        dn[1][i] = 0.125 * b * (1.0 + xi * a) * (1.0 + zeta * c);
        dn[2][i] = 0.125 * c * (1.0 + xi * a) * (1.0 + eta * b);
    }
    double j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) j[a][b] += dn[a][i] * x[i][b];
    double detj = det3(j);
    if (detj <= 0.0) return false;
    double inv[3][3];
    inv3(j, detj, inv);
    for (int i = 0; i < 144; ++i) bm[i] = 0.0;
    for (int i = 0; i < 8; ++i) {
        double dnv[3] = {dn[0][i], dn[1][i], dn[2][i]};
        double g[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) g[b] += inv[b][a] * dnv[a];
        int c = 3 * i;
        bm[0 * 24 + c] = g[0];
        bm[1 * 24 + c + 1] = g[1];
        bm[2 * 24 + c + 2] = g[2];
        bm[3 * 24 + c] = g[1];
        bm[3 * 24 + c + 1] = g[0];
        bm[4 * 24 + c + 1] = g[2];
        bm[4 * 24 + c + 2] = g[1];
        bm[5 * 24 + c] = g[2];
        bm[5 * 24 + c + 2] = g[0];
    }
    *detj_out = detj;
    return true;
}

// element.rs:535-561
bool c3d8_stiffness(const double (*x)[3], const D6& d, double* ke /*24x24*/) {
    const double g = gauss_g();
    const double pts[2] = {-g, g};
    for (int i = 0; i < 576; ++i) ke[i] = 0.0;
    double bm[144], detj;
    for (double xi : pts)
        for (double eta : pts)
            for (double zeta : pts) {
                if (!hex_b(x, xi, eta, zeta, bm, &detj)) return false;
                btdb<24, 6>(ke, bm, &d[0][0], detj);
            }
    return hex_b(x, 0.0, 0.0, 0.0, bm, &detj);
}

// element.rs:854-862
template <int N, int S>
void strain(const double* b, const double* ue, double* e) {
    for (int r = 0; r < S; ++r) {
        e[r] = 0.0;
        for (int c = 0; c < N; ++c) e[r] += b[r * N + c] * ue[c];
    }
}

// element.rs:866-880
template <int N>
void plane_stress(const double* b, const D3& d, const double* ue, double* s) {
    double e[3];
    strain<N, 3>(b, ue, e);
    s[0] = d[0][0] * e[0] + d[0][1] * e[1];
    s[1] = d[1][0] * e[0] + d[1][1] * e[1];
    s[2] = 0.0;
    s[3] = d[2][2] * e[2];
    s[4] = 0.0;
    s[5] = 0.0;
}

// element.rs:902-913
void solid_stress(const double* b, const D6& d, const double* ue, double* s) {
    double e[6];
    strain<24, 6>(b, ue, e);
    for (int r = 0; r < 6; ++r) {
        double acc = -0.0;
        for (int k = 0; k < 6; ++k) acc += d[r][k] * e[k];
        s[r] = acc;
    }
}

// element.rs:917-923
double von_mises(const double* s) {
    return std::sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                            (s[2] - s[0]) * (s[2] - s[0])) +
                     3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
}

const double GP_SIGNS[4][2] = {{-1.0, -1.0}, {-1.0, 1.0}, {1.0, -1.0}, {1.0, 1.0}};
const double GP3_SIGNS[8][3] = {{-1.0, -1.0, -1.0}, {-1.0, -1.0, 1.0}, {-1.0, 1.0, -1.0},
                                {-1.0, 1.0, 1.0},   {1.0, -1.0, -1.0}, {1.0, -1.0, 1.0},
                                {1.0, 1.0, -1.0},   {1.0, 1.0, 1.0}};

// element.rs:1005-1018
void extrapolate_to_corners(const double gp[4][6], double out[4][6]) {
    const double sqrt3 = std::sqrt(3.0);
    auto h = [&](double s) { return 0.5 * (1.0 + s * sqrt3); };
    for (int i = 0; i < 4; ++i) {
        for (int k = 0; k < 6; ++k) out[i][k] = 0.0;
        for (int j = 0; j < 4; ++j) {
            double w = h(QUAD_CORNERS[i][0] * GP_SIGNS[j][0]) * h(QUAD_CORNERS[i][1] * GP_SIGNS[j][1]);
            for (int k = 0; k < 6; ++k) out[i][k] += w * gp[j][k];
        }
    }
}

// element.rs:1040-1053
void extrapolate_to_corners3(const double gp[8][6], double out[8][6]) {
    const double sqrt3 = std::sqrt(3.0);
    auto h = [&](double s) { return 0.5 * (1.0 + s * sqrt3); };
    for (int i = 0; i < 8; ++i) {
        for (int k = 0; k < 6; ++k) out[i][k] = 0.0;
        for (int j = 0; j < 8; ++j) {
            double w = h(HEX_CORNERS[i][0] * GP3_SIGNS[j][0]) * h(HEX_CORNERS[i][1] * GP3_SIGNS[j][1]) *
                       h(HEX_CORNERS[i][2] * GP3_SIGNS[j][2]);
            for (int k = 0; k < 6; ++k) out[i][k] += w * gp[j][k];
        }
    }
}

// --------------------------------------------------------------- CSR / SpMV
struct Csr {
    std::vector<int64_t> row_ptr;
    std::vector<int64_t> col_idx;
    std::vector<double> vals;
    int64_t n() const { return (int64_t)row_ptr.size() - 1; }
};

const int64_t PARALLEL_ASSEMBLY_MIN_ELEMS = 256;  // assembly.rs:25
const int64_t PARALLEL_SPMV_MIN_ROWS = 20000;     // assembly.rs:27

// assembly.rs:17-21
int worker_threads() {
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    return (int)std::min(n, 8u);
}

void spmv_rows(const int64_t* rp, const int64_t* ci, const double* v, const double* x, double* y,
               int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r) {
        double o = 0.0;
        for (int64_t idx = rp[r]; idx < rp[r + 1]; ++idx) o += v[idx] * x[ci[idx]];
        y[r] = o;
    }
}

// assembly.rs:46-72 — threads spawned per call, row chunks of ceil(n/threads).
void mat_vec_with(int64_t n, const int64_t* rp, const int64_t* ci, const double* v, const double* x,
                  double* y, int threads) {
    if (threads > 1 && n > PARALLEL_SPMV_MIN_ROWS) {
        int64_t chunk = (n + threads - 1) / threads;
        std::vector<std::thread> ts;
        for (int64_t start = 0; start < n; start += chunk) {
            int64_t end = std::min(n, start + chunk);
            ts.emplace_back([=] { spmv_rows(rp, ci, v, x, y, start, end); });
        }
        for (auto& t : ts) t.join();
    } else {
        spmv_rows(rp, ci, v, x, y, 0, n);
    }
}

struct Trip {
    int64_t r, c;
    double v;
};

// SplitMix64 step (the decks' generator), for the duplicate-order shuffle.
static uint64_t mix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// assembly.rs:87-116 (stable order for duplicates, see header; dup_seed != 0:
// each row's triplets shuffled first, a different duplicate order)
Csr csr_from_triplets(int64_t n, const std::vector<Trip>& trips, uint64_t dup_seed = 0) {
    std::vector<int64_t> cnt(n + 1, 0);
    for (const auto& t : trips) cnt[t.r + 1]++;
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<std::pair<int64_t, double>> rows(trips.size());
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (const auto& t : trips) rows[pos[t.r]++] = {t.c, t.v};
    }
    Csr k;
    k.row_ptr.reserve(n + 1);
    k.row_ptr.push_back(0);
    for (int64_t r = 0; r < n; ++r) {
        auto b = rows.begin() + cnt[r], e = rows.begin() + cnt[r + 1];
        if (dup_seed) {  // Fisher-Yates with a per-row stream
            uint64_t st = dup_seed ^ (0x5DEECE66Dull * (uint64_t)(r + 1));
            for (int64_t i = (e - b) - 1; i > 0; --i) std::swap(b[i], b[(int64_t)(mix64(st) % (uint64_t)(i + 1))]);
        }
        std::stable_sort(b, e, [](const auto& x, const auto& y) { return x.first < y.first; });
        auto it = b;
        while (it != e) {
            int64_t c = it->first;
            double v = 0.0;
            while (it != e && it->first == c) {
                v += it->second;
                ++it;
            }
            k.col_idx.push_back(c);
            k.vals.push_back(v);
        }
        k.row_ptr.push_back((int64_t)k.col_idx.size());
    }
    return k;
}

// assembly.rs:120-133
void scatter(std::vector<Trip>& trips, const int64_t* node_idx, const double* ke, int N, int ndof) {
    for (int a = 0; a < N; ++a) {
        int64_t gr = node_idx[a / ndof] * ndof + a % ndof;
        for (int c = 0; c < N; ++c) {
            int64_t gc = node_idx[c / ndof] * ndof + c % ndof;
            trips.push_back({gr, gc, ke[a * N + c]});
        }
    }
}

// ----------------------------------------------------------------- model
struct Model {
    const svx_model_desc* m;
    bool faithful;
    std::unordered_map<uint32_t, int64_t> node_pos;  // first position of an id
    std::vector<std::unordered_map<uint32_t, bool>> elset_sets;

    void init() {
        if (!faithful) {
            node_pos.reserve(m->n_nodes * 2);
            for (int64_t i = m->n_nodes - 1; i >= 0; --i) node_pos[m->node_ids[i]] = i;
            elset_sets.resize(m->n_elsets);
            for (int s = 0; s < m->n_elsets; ++s)
                for (int64_t k = m->elset_ptr[s]; k < m->elset_ptr[s + 1]; ++k)
                    elset_sets[s][m->elset_members[k]] = true;
        }
    }
    // lib.rs:106-108 — a linear scan in faithful mode.
    int64_t node_index(uint32_t id) const {
        if (faithful) {
            for (int64_t i = 0; i < m->n_nodes; ++i)
                if (m->node_ids[i] == id) return i;
            return -1;
        }
        auto it = node_pos.find(id);
        return it == node_pos.end() ? -1 : it->second;
    }
    // lib.rs:357-367 — returns the material index or -1.
    int material_for_element(int64_t ei) const {
        uint32_t id = m->elem_ids[ei];
        for (int a = 0; a < m->n_assign; ++a) {
            int s = m->assign_elset[a];
            if (s < 0) continue;
            bool contains = false;
            if (faithful) {
                for (int64_t k = m->elset_ptr[s]; k < m->elset_ptr[s + 1]; ++k)
                    if (m->elset_members[k] == id) {
                        contains = true;
                        break;
                    }
            } else {
                contains = elset_sets[s].count(id) != 0;
            }
            if (contains) return m->assign_material[a];
        }
        return -1;
    }
    int node_count(uint8_t kind) const { return kind == SVX_CPS3 ? 3 : kind == SVX_CPS4 ? 4 : 8; }
};

struct ElemInfo {
    int64_t node_idx[8];
    int mat;
};

struct Err {
    int code = SVX_OK;
    std::string msg;
    int64_t iterations = 0;
    double residual = 0.0;
};

std::string fmt_elem(uint32_t id, const char* what) {
    char buf[256];
    snprintf(buf, sizeof buf, "element %u: %s", id, what);
    return buf;
}

// assembly.rs:138-276 for CPS3 / CPS4 / C3D8. Returns "" or the message.
std::string element_entry(std::vector<Trip>& trips, const Model& M, int64_t ei, int ndof,
                          ElemInfo& info) {
    const svx_model_desc* m = M.m;
    uint32_t id = m->elem_ids[ei];
    uint8_t kind = m->kinds[ei];
    int mat = M.material_for_element(ei);
    if (mat < 0) return fmt_elem(id, "no material assigned");
    double E = m->youngs[mat], nu = m->poisson[mat];
    if (E <= 0.0 || std::fabs(nu) >= 1.0) return fmt_elem(id, "invalid material constants");
    if (kind == SVX_C3D8 && nu >= 0.5)
        return fmt_elem(id, "poisson ratio must be < 0.5 for solids and plane strain");
    int nen = M.node_count(kind);
    for (int k = 0; k < nen; ++k) info.node_idx[k] = M.node_index(m->conn_ids[m->conn_ptr[ei] + k]);
    info.mat = mat;
    const double* X = m->coords;
    if (kind == SVX_CPS3) {
        double x[3][2];
        for (int k = 0; k < 3; ++k) {
            x[k][0] = X[info.node_idx[k] * 3];
            x[k][1] = X[info.node_idx[k] * 3 + 1];
        }
        double ke[36];
        if (!cps3_stiffness(x, d_matrix(E, nu), m->thickness, ke))
            return fmt_elem(id, "degenerate (zero-area) triangle");
        scatter(trips, info.node_idx, ke, 6, ndof);
    } else if (kind == SVX_CPS4) {
        double x[4][2];
        for (int k = 0; k < 4; ++k) {
            x[k][0] = X[info.node_idx[k] * 3];
            x[k][1] = X[info.node_idx[k] * 3 + 1];
        }
        double ke[64];
        if (!cps4_stiffness(x, d_matrix(E, nu), m->thickness, ke))
            return fmt_elem(id, "non-positive Jacobian determinant at a Gauss point");
        scatter(trips, info.node_idx, ke, 8, ndof);
    } else {
        double x[8][3];
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 3; ++b) x[k][b] = X[info.node_idx[k] * 3 + b];
        double ke[576];
        if (!c3d8_stiffness(x, d3_matrix(E, nu), ke))
            return fmt_elem(id, "non-positive Jacobian determinant at a Gauss point");
        scatter(trips, info.node_idx, ke, 24, ndof);
    }
    return "";
}

struct Assembled {
    Csr k;
    std::vector<double> f;
    std::vector<ElemInfo> elems;
};

// assembly.rs:291-386 (CLOADs only: the path's decks carry no DLOAD/gravity)
bool assemble_with(const Model& M, int ndof, int threads, Assembled& out, Err& err, uint64_t dup_seed = 0) {
    const svx_model_desc* m = M.m;
    int64_t ndof_total = m->n_nodes * ndof;
    int64_t n_elem = m->n_elem;
    std::vector<Trip> trips;
    trips.reserve(ndof_total + n_elem * 24);
    for (int64_t i = 0; i < ndof_total; ++i) trips.push_back({i, i, 0.0});
    out.f.assign(ndof_total, 0.0);
    out.elems.resize(n_elem);

    int n_chunks = (threads > 1 && n_elem >= PARALLEL_ASSEMBLY_MIN_ELEMS) ? 2 * threads : 1;
    int64_t chunk_len = (n_elem + n_chunks - 1) / n_chunks;
    struct Chunk {
        std::vector<Trip> t;
        std::string err;
    };
    std::vector<Chunk> chunks(n_chunks);
    auto run_chunk = [&](int c) {
        int64_t a = c * chunk_len, b = std::min(a + chunk_len, n_elem);
        if (a > b) a = b;
        chunks[c].t.reserve((b - a) * 24);
        for (int64_t ei = a; ei < b; ++ei) {
            std::string e = element_entry(chunks[c].t, M, ei, ndof, out.elems[ei]);
            if (!e.empty()) {
                chunks[c].err = e;
                return;
            }
        }
    };
    if (n_chunks == 1) {
        run_chunk(0);
    } else {
        std::vector<std::thread> ts;
        for (int i = 0; i < threads; ++i)
            ts.emplace_back([&, i] {
                for (int a = i; a < n_chunks; a += threads) run_chunk(a);
            });
        for (auto& t : ts) t.join();
    }
    for (auto& c : chunks) {  // merge in chunk order: first bad element wins
        if (!c.err.empty()) {
            err.code = SVX_E_INVALID_MODEL;
            err.msg = c.err;
            return false;
        }
        trips.insert(trips.end(), c.t.begin(), c.t.end());
        std::vector<Trip>().swap(c.t);
    }
    for (int64_t l = 0; l < m->n_cloads; ++l) {  // assembly.rs:365-380
        int dof = m->cload_dof[l];
        uint32_t nid = m->cload_node[l];
        char buf[256];
        if (dof <= 0 || dof > ndof) {
            snprintf(buf, sizeof buf, "load on node %u has unsupported dof %d", nid, dof);
            err.code = SVX_E_INVALID_MODEL;
            err.msg = buf;
            return false;
        }
        int64_t idx = M.node_index(nid);
        if (idx < 0) {
            snprintf(buf, sizeof buf, "load on unknown node %u", nid);
            err.code = SVX_E_INVALID_MODEL;
            err.msg = buf;
            return false;
        }
        out.f[idx * ndof + dof - 1] += m->cload_value[l];
    }
    out.k = csr_from_triplets(ndof_total, trips, dup_seed);
    return true;
}

struct ConstrainedDof {
    int64_t row;
    uint32_t node_id;
    int dof;
};

// solver.rs:15-41
bool constrained_dofs(const Model& M, int ndof, std::vector<ConstrainedDof>& out, Err& err) {
    const svx_model_desc* m = M.m;
    for (int b = 0; b < m->n_bcs; ++b)
        for (int64_t k = m->bc_node_ptr[b]; k < m->bc_node_ptr[b + 1]; ++k) {
            uint32_t nid = m->bc_nodes[k];
            int64_t idx = M.node_index(nid);
            char buf[256];
            if (idx < 0) {
                snprintf(buf, sizeof buf, "boundary condition on unknown node %u", nid);
                err.code = SVX_E_INVALID_MODEL;
                err.msg = buf;
                return false;
            }
            for (int64_t q = m->bc_dof_ptr[b]; q < m->bc_dof_ptr[b + 1]; ++q) {
                int dof = m->bc_dofs[q];
                if (dof <= 0 || dof > ndof) {
                    snprintf(buf, sizeof buf, "boundary condition on node %u has unsupported dof %d", nid, dof);
                    err.code = SVX_E_INVALID_MODEL;
                    err.msg = buf;
                    return false;
                }
                out.push_back({idx * ndof + dof - 1, nid, dof});
            }
        }
    return true;
}

// solver.rs:45-62
void apply_dirichlet(Csr& k, std::vector<double>& f, const std::vector<ConstrainedDof>& cons) {
    int64_t n = (int64_t)f.size();
    std::vector<char> is_c(n, 0);
    for (const auto& c : cons) is_c[c.row] = 1;
    for (int64_t r = 0; r < n; ++r)
        for (int64_t idx = k.row_ptr[r]; idx < k.row_ptr[r + 1]; ++idx) {
            int64_t c = k.col_idx[idx];
            if (is_c[r] || is_c[c]) k.vals[idx] = (r == c) ? 1.0 : 0.0;
        }
    for (const auto& c : cons) f[c.row] = 0.0;
}

double dot(const double* a, const double* b, int64_t n) {
    double s = -0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// solver.rs:82-149. Returns SVX_OK or SVX_E_NOT_CONVERGED (iterations/residual in err).
int pcg_with_threads(int64_t n, const int64_t* rp, const int64_t* ci, const double* kv, const double* f,
                     double tol, int64_t max_iter, int threads, double* u, int64_t* its, double* res_out,
                     Err& err) {
    for (int64_t i = 0; i < n; ++i) u[i] = 0.0;
    std::vector<double> r(f, f + n);
    double b_norm = std::sqrt(dot(f, f, n));
    if (b_norm == 0.0) {
        *its = 0;
        *res_out = 0.0;
        return SVX_OK;
    }
    std::vector<double> inv_diag(n);
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t idx = rp[i]; idx < rp[i + 1]; ++idx)
            if (ci[idx] == i) {
                d = kv[idx];
                break;
            }
        inv_diag[i] = (d == 0.0) ? 0.0 : 1.0 / d;
    }
    auto rel_res = [&](const std::vector<double>& rr) { return std::sqrt(dot(rr.data(), rr.data(), n)) / b_norm; };
    std::vector<double> z(n), p(n), kp(n);
    for (int64_t i = 0; i < n; ++i) z[i] = r[i] * inv_diag[i];
    p = z;
    double rz = dot(r.data(), z.data(), n);
    for (int64_t it = 1; it <= max_iter; ++it) {
        mat_vec_with(n, rp, ci, kv, p.data(), kp.data(), threads);
        double pap = dot(p.data(), kp.data(), n);
        if (pap <= 0.0 || std::isnan(pap)) {
            err.code = SVX_E_NOT_CONVERGED;
            err.iterations = it;
            err.residual = rel_res(r);
            *its = it;
            *res_out = err.residual;
            return SVX_E_NOT_CONVERGED;
        }
        double alpha = rz / pap;
        for (int64_t i = 0; i < n; ++i) u[i] += alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] -= alpha * kp[i];
        double res = rel_res(r);
        if (res <= tol) {
            *its = it;
            *res_out = res;
            return SVX_OK;
        }
        for (int64_t i = 0; i < n; ++i) z[i] = r[i] * inv_diag[i];
        double rz_new = dot(r.data(), z.data(), n);
        double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    err.code = SVX_E_NOT_CONVERGED;
    err.iterations = max_iter;
    err.residual = rel_res(r);
    *its = max_iter;
    *res_out = err.residual;
    return SVX_E_NOT_CONVERGED;
}

// ------------------------------------------------------------ recovery
// Per-element stress at the centroid and at the element's nodes
// (element.rs:926-994 for Tri/Quad/Hex with plane stress / solid D).
void element_stresses(const svx_model_desc* m, const ElemInfo& info, uint8_t kind, int ndof,
                      const double* u, double centroid[6], double nodal[8][6]) {
    double E = m->youngs[info.mat], nu = m->poisson[info.mat];
    int nen = kind == SVX_CPS3 ? 3 : kind == SVX_CPS4 ? 4 : 8;
    double ue[24];
    for (int a = 0; a < nen; ++a)
        for (int d = 0; d < ndof; ++d) ue[a * ndof + d] = u[info.node_idx[a] * ndof + d];
    const double* X = m->coords;
    if (kind == SVX_CPS3) {
        double x[3][2], bm[18], two_a;
        for (int k = 0; k < 3; ++k) {
            x[k][0] = X[info.node_idx[k] * 3];
            x[k][1] = X[info.node_idx[k] * 3 + 1];
        }
        cps3_b_matrix(x, bm, &two_a);
        D3 d = d_matrix(E, nu);
        plane_stress<6>(bm, d, ue, centroid);
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 6; ++c) nodal[k][c] = centroid[c];
    } else if (kind == SVX_CPS4) {
        double x[4][2], bm[24], detj;
        for (int k = 0; k < 4; ++k) {
            x[k][0] = X[info.node_idx[k] * 3];
            x[k][1] = X[info.node_idx[k] * 3 + 1];
        }
        D3 d = d_matrix(E, nu);
        quad_b(x, 0.0, 0.0, bm, &detj);
        plane_stress<8>(bm, d, ue, centroid);
        const double g = gauss_g();
        const double pts[2] = {-g, g};
        double gp[4][6];
        int j = 0;
        for (double xi : pts)
            for (double eta : pts) {
                quad_b(x, xi, eta, bm, &detj);
                plane_stress<8>(bm, d, ue, gp[j++]);
            }
        double out[4][6];
        extrapolate_to_corners(gp, out);
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 6; ++c) nodal[k][c] = out[k][c];
    } else {
        double x[8][3], bm[144], detj;
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 3; ++b) x[k][b] = X[info.node_idx[k] * 3 + b];
        D6 d = d3_matrix(E, nu);
        hex_b(x, 0.0, 0.0, 0.0, bm, &detj);
        solid_stress(bm, d, ue, centroid);
        const double g = gauss_g();
        const double pts[2] = {-g, g};
        double gp[8][6];
        int j = 0;
        for (double xi : pts)
            for (double eta : pts)
                for (double zeta : pts) {
                    hex_b(x, xi, eta, zeta, bm, &detj);
                    solid_stress(bm, d, ue, gp[j++]);
                }
        double out[8][6];
        extrapolate_to_corners3(gp, out);
        for (int k = 0; k < 8; ++k)
            for (int c = 0; c < 6; ++c) nodal[k][c] = out[k][c];
    }
}

}  // namespace

// ===================================================================== C API
extern "C" {

typedef struct oracle_opts {
    int32_t threads;   // <= 0: worker_threads() = min(nproc, 8)
    int32_t faithful;  // 1: the reference's linear-scan lookups; 0: hashed (same numbers)
    double tol;        // <= 0: 1e-10 (lib.rs:666)
    int64_t max_iter;  // <= 0: max(1000, 10 n) (lib.rs:632)
    int32_t stop_after_assembly;
    uint64_t dup_seed;  // != 0: duplicate triplets summed in a seeded shuffled order
} oracle_opts;

struct oracle_result {
    Err err;
    int ndof = 0;
    Csr k_orig;
    std::vector<double> f_orig, u, disp, vm, nodal_vm, react_val;
    std::vector<ConstrainedDof> cons;
    int64_t iterations = 0;
    double residual = 0.0, cond = 0.0;
    double t_assemble = 0, t_pcg = 0, t_recover = 0, t_total = 0;
};

// lib.rs:593-716 without thermal loads and without the LDLT fallback.
oracle_result* oracle_solve(const svx_model_desc* m, const oracle_opts* o) {
    auto* R = new oracle_result();
    auto t_all = Clock::now();
    int threads = (o && o->threads > 0) ? o->threads : worker_threads();
    Model M{m, o ? o->faithful != 0 : true, {}, {}};
    Err& err = R->err;
    auto fail = [&](int code, const std::string& msg) {
        err.code = code;
        err.msg = msg;
        return R;
    };
    if (m->n_bcs == 0) return fail(SVX_E_NO_CONSTRAINTS, "NoConstraints");
    auto t0 = Clock::now();
    M.init();
    // lib.rs:406-437 validate()
    if (m->n_nodes == 0 || m->n_elem == 0) return fail(SVX_E_INVALID_MODEL, "EmptyMesh");
    if (m->n_materials == 0) return fail(SVX_E_INVALID_MODEL, "NoMaterial");
    for (int64_t e = 0; e < m->n_elem; ++e) {
        int64_t len = m->conn_ptr[e + 1] - m->conn_ptr[e];
        if (len != M.node_count(m->kinds[e]))
            return fail(SVX_E_INVALID_MODEL, "MalformedElement(" + std::to_string(m->elem_ids[e]) + ")");
        for (int64_t k = m->conn_ptr[e]; k < m->conn_ptr[e + 1]; ++k)
            if (M.node_index(m->conn_ids[k]) < 0)
                return fail(SVX_E_INVALID_MODEL, "UnknownNode(" + std::to_string(m->conn_ids[k]) + ")");
    }
    if (m->thickness <= 0.0 || std::isnan(m->thickness))
        return fail(SVX_E_INVALID_MODEL, "thickness must be positive");
    // lib.rs:725-752
    bool plane = false, solid = false;
    for (int64_t e = 0; e < m->n_elem; ++e) (m->kinds[e] == SVX_C3D8 ? solid : plane) = true;
    if (plane && solid) return fail(SVX_E_INVALID_MODEL, "mixed dimensionality");
    int ndof = solid ? 3 : 2;
    R->ndof = ndof;

    Assembled A;
    if (!assemble_with(M, ndof, threads, A, err, o ? o->dup_seed : 0)) return R;
    std::vector<ConstrainedDof>& cons = R->cons;
    if (!constrained_dofs(M, ndof, cons, err)) return R;
    R->k_orig = A.k;  // lib.rs:623-628
    R->f_orig = A.f;
    R->t_assemble = ms_since(t0);
    if (o && o->stop_after_assembly) {
        R->t_total = ms_since(t_all);
        return R;
    }
    auto t1 = Clock::now();
    Csr& k = A.k;
    std::vector<double>& f = A.f;
    apply_dirichlet(k, f, cons);
    int64_t n = (int64_t)f.size();
    int64_t max_iter = (o && o->max_iter > 0) ? o->max_iter : std::max<int64_t>(1000, 10 * n);
    double tol = (o && o->tol > 0) ? o->tol : 1e-10;
    R->u.assign(n, 0.0);
    int st = pcg_with_threads(n, k.row_ptr.data(), k.col_idx.data(), k.vals.data(), f.data(), tol, max_iter,
                              threads, R->u.data(), &R->iterations, &R->residual, err);
    R->t_pcg = ms_since(t1);
    if (st != SVX_OK) {
        char buf[256];
        snprintf(buf, sizeof buf, "NotConverged { iterations: %lld, residual: %.6e }", (long long)err.iterations,
                 err.residual);
        err.msg = buf;
        R->t_total = ms_since(t_all);
        return R;
    }
    auto t2 = Clock::now();
    // lib.rs:634-644 jacobi_cond on the constrained K
    double hi = 0.0, lo = INFINITY;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t idx = k.row_ptr[i]; idx < k.row_ptr[i + 1]; ++idx)
            if (k.col_idx[idx] == i) {
                double a = std::fabs(k.vals[idx]);
                hi = std::max(hi, a);
                lo = std::min(lo, a);
                break;
            }
    R->cond = hi / lo;
    // lib.rs:689-698
    R->disp.assign(m->n_nodes * 3, 0.0);
    for (int64_t i = 0; i < m->n_nodes; ++i) {
        R->disp[i * 3] = R->u[i * ndof];
        R->disp[i * 3 + 1] = R->u[i * ndof + 1];
        R->disp[i * 3 + 2] = ndof == 3 ? R->u[i * ndof + 2] : 0.0;
    }
    // stress.rs:22-98
    R->vm.assign(m->n_elem, 0.0);
    std::vector<std::array<double, 6>> sum(m->n_nodes);
    for (auto& s : sum) s.fill(0.0);
    std::vector<uint32_t> count(m->n_nodes, 0);
    for (int64_t e = 0; e < m->n_elem; ++e) {
        double cs[6], ns[8][6];
        element_stresses(m, A.elems[e], m->kinds[e], ndof, R->u.data(), cs, ns);
        R->vm[e] = von_mises(cs);
        int nen = M.node_count(m->kinds[e]);
        for (int i = 0; i < nen; ++i) {
            int64_t node = A.elems[e].node_idx[i];
            for (int c = 0; c < 6; ++c) sum[node][c] += ns[i][c];
            count[node] += 1;
        }
    }
    R->nodal_vm.assign(m->n_nodes, 0.0);
    for (int64_t i = 0; i < m->n_nodes; ++i) {
        if (count[i] == 0) continue;
        double c = (double)count[i];
        double avg[6];
        for (int q = 0; q < 6; ++q) avg[q] = sum[i][q] / c;
        R->nodal_vm[i] = von_mises(avg);
    }
    // stress.rs:103-118
    std::vector<double> ku(n);
    mat_vec_with(n, R->k_orig.row_ptr.data(), R->k_orig.col_idx.data(), R->k_orig.vals.data(), R->u.data(),
                 ku.data(), threads);
    R->react_val.resize(cons.size());
    for (size_t i = 0; i < cons.size(); ++i) R->react_val[i] = ku[cons[i].row] - R->f_orig[cons[i].row];
    R->t_recover = ms_since(t2);
    R->t_total = ms_since(t_all);
    return R;
}

void oracle_free(oracle_result* r) { delete r; }
int oracle_status(const oracle_result* r) { return r->err.code; }
const char* oracle_message(const oracle_result* r) { return r->err.msg.c_str(); }
int oracle_ndof(const oracle_result* r) { return r->ndof; }
int64_t oracle_rows(const oracle_result* r) { return r->k_orig.row_ptr.empty() ? 0 : r->k_orig.n(); }
int64_t oracle_nnz(const oracle_result* r) { return (int64_t)r->k_orig.col_idx.size(); }
int64_t oracle_n_reactions(const oracle_result* r) { return (int64_t)r->cons.size(); }
int64_t oracle_iterations(const oracle_result* r) { return r->err.code == SVX_E_NOT_CONVERGED ? r->err.iterations : r->iterations; }
double oracle_residual(const oracle_result* r) { return r->err.code == SVX_E_NOT_CONVERGED ? r->err.residual : r->residual; }
double oracle_cond(const oracle_result* r) { return r->cond; }
void oracle_times(const oracle_result* r, double* out4) {
    out4[0] = r->t_assemble;
    out4[1] = r->t_pcg;
    out4[2] = r->t_recover;
    out4[3] = r->t_total;
}
void oracle_get_csr(const oracle_result* r, int64_t* rp, int64_t* ci, double* v) {
    const Csr& k = r->k_orig;
    if (rp) std::copy(k.row_ptr.begin(), k.row_ptr.end(), rp);
    if (ci) std::copy(k.col_idx.begin(), k.col_idx.end(), ci);
    if (v) std::copy(k.vals.begin(), k.vals.end(), v);
}
void oracle_get_f(const oracle_result* r, double* f) { std::copy(r->f_orig.begin(), r->f_orig.end(), f); }
void oracle_get_u(const oracle_result* r, double* u) { std::copy(r->u.begin(), r->u.end(), u); }
void oracle_get_solution(const oracle_result* r, double* disp, double* vm, double* nodal_vm) {
    if (disp) std::copy(r->disp.begin(), r->disp.end(), disp);
    if (vm) std::copy(r->vm.begin(), r->vm.end(), vm);
    if (nodal_vm) std::copy(r->nodal_vm.begin(), r->nodal_vm.end(), nodal_vm);
}
void oracle_get_reactions(const oracle_result* r, int64_t* rows, uint32_t* node, int32_t* dof, double* val) {
    for (size_t i = 0; i < r->cons.size(); ++i) {
        if (rows) rows[i] = r->cons[i].row;
        if (node) node[i] = r->cons[i].node_id;
        if (dof) dof[i] = r->cons[i].dof;
        if (val && i < r->react_val.size()) val[i] = r->react_val[i];
    }
}

// Csr::mat_vec_with (assembly.rs:46-72); threads <= 0 picks worker_threads().
void oracle_spmv(int64_t n, const int64_t* rp, const int64_t* ci, const double* v, const double* x, double* y,
                 int threads) {
    mat_vec_with(n, rp, ci, v, x, y, threads > 0 ? threads : worker_threads());
}

// solver::pcg_with_threads (solver.rs:82-149).
int oracle_pcg(int64_t n, const int64_t* rp, const int64_t* ci, const double* v, const double* f, double tol,
               int64_t max_iter, int threads, double* u, int64_t* its, double* res) {
    Err err;
    return pcg_with_threads(n, rp, ci, v, f, tol, max_iter, threads > 0 ? threads : worker_threads(), u, its, res,
                            err);
}

// Element stiffness of one element: kind SVX_CPS3/CPS4/C3D8, coords [nen*3].
// Returns 0, or 4 (degenerate triangle) / 5 (non-positive Jacobian).
int oracle_element_stiffness(int kind, const double* coords, double E, double nu, double t, double* ke) {
    if (kind == SVX_CPS3) {
        double x[3][2];
        for (int k = 0; k < 3; ++k) x[k][0] = coords[k * 3], x[k][1] = coords[k * 3 + 1];
        return cps3_stiffness(x, d_matrix(E, nu), t, ke) ? 0 : SVX_ELEM_DEGENERATE_TRIANGLE;
    }
    if (kind == SVX_CPS4) {
        double x[4][2];
        for (int k = 0; k < 4; ++k) x[k][0] = coords[k * 3], x[k][1] = coords[k * 3 + 1];
        return cps4_stiffness(x, d_matrix(E, nu), t, ke) ? 0 : SVX_ELEM_NONPOSITIVE_JACOBIAN;
    }
    double x[8][3];
    for (int k = 0; k < 8; ++k)
        for (int b = 0; b < 3; ++b) x[k][b] = coords[k * 3 + b];
    return c3d8_stiffness(x, d3_matrix(E, nu), ke) ? 0 : SVX_ELEM_NONPOSITIVE_JACOBIAN;
}

// The deterministic LCG vector of assembly.rs:800-808, values in [-0.5, 0.5).
void oracle_lcg_vector(int64_t n, double* out) {
    uint64_t s = 0x853c49e6748fea9bULL;
    for (int64_t i = 0; i < n; ++i) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        out[i] = ((double)(s >> 11) / (double)(1ULL << 53)) - 0.5;
    }
}

int oracle_worker_threads(void) { return worker_threads(); }

}  // extern "C"
