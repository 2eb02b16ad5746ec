// Host construction of the per-mesh structures, built once per mesh:
//  - node -> (element, local slot) incidence, in element order;
//  - the CSR pattern of csr_from_triplets (assembly.rs:87-116): row
//    node*ndof + a holds every column j*ndof + b for j in the node's
//    element-neighbourhood (sorted), or only its diagonal when the node
//    belongs to no element (assembly.rs:300-302);
//  - the node-block (BSR) pattern with the same neighbourhoods, and its
//    node-row sliced ELL form for the Krylov SpMV.
#include "pattern.hpp"

#include <algorithm>
#include <stdexcept>
#include <thread>

namespace svx {

static void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int T = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (n < 4096 || T <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> ts;
    int64_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; ++t) {
        int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        ts.emplace_back(fn, a, b);
    }
    for (auto& t : ts) t.join();
}

void build_pattern(int ndof, int64_t n_nodes, int64_t n_elem, const int64_t* conn_ptr, const int32_t* conn,
                   HostPattern& P) {
    P.ndof = ndof;
    P.n_nodes = n_nodes;
    // incidence (counting sort by node; stable in (element, slot) order)
    P.inc_ptr.assign(n_nodes + 1, 0);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k) P.inc_ptr[conn[k] + 1]++;
    for (int64_t i = 0; i < n_nodes; ++i) P.inc_ptr[i + 1] += P.inc_ptr[i];
    P.inc_elem.resize(P.inc_ptr[n_nodes]);
    P.inc_slot.resize(P.inc_ptr[n_nodes]);
    {
        std::vector<int64_t> pos(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
        for (int64_t e = 0; e < n_elem; ++e)
            for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k) {
                int64_t q = pos[conn[k]]++;
                P.inc_elem[q] = (int32_t)e;
                P.inc_slot[q] = (int8_t)(k - conn_ptr[e]);
            }
    }
    // node neighbourhoods
    std::vector<int32_t> deg(n_nodes, 0);
    std::vector<std::vector<int32_t>> nb(n_nodes);
    parallel_for(n_nodes, [&](int64_t a, int64_t b) {
        std::vector<int32_t> tmp;
        for (int64_t i = a; i < b; ++i) {
            tmp.clear();
            for (int64_t q = P.inc_ptr[i]; q < P.inc_ptr[i + 1]; ++q) {
                int32_t e = P.inc_elem[q];
                for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k) tmp.push_back(conn[k]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            if (tmp.empty()) tmp.push_back((int32_t)i);  // lone diagonal
            nb[i] = tmp;
            deg[i] = (int32_t)tmp.size();
        }
    });
    // block pattern
    P.brow_ptr.assign(n_nodes + 1, 0);
    for (int64_t i = 0; i < n_nodes; ++i) P.brow_ptr[i + 1] = P.brow_ptr[i] + deg[i];
    if (P.brow_ptr[n_nodes] > INT32_MAX) throw std::runtime_error("block pattern exceeds 2^31 entries");
    P.bcol.resize(P.brow_ptr[n_nodes]);
    // scalar pattern
    const int64_t n = n_nodes * ndof;
    if (n >= INT32_MAX) throw std::runtime_error("more than 2^31 rows in one mesh");
    int64_t total = 0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        bool lone = P.inc_ptr[i + 1] == P.inc_ptr[i];
        total += ndof * (lone ? 1 : (int64_t)deg[i] * ndof);
    }
    if (total > INT32_MAX) throw std::runtime_error("CSR pattern exceeds 2^31 entries");
    P.row_ptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n_nodes; ++i) {
        bool lone = P.inc_ptr[i + 1] == P.inc_ptr[i];
        int32_t len = lone ? 1 : deg[i] * ndof;
        for (int a = 0; a < ndof; ++a) P.row_ptr[i * ndof + a + 1] = len;
    }
    for (int64_t r = 0; r < n; ++r) P.row_ptr[r + 1] += P.row_ptr[r];
    P.col.resize(P.row_ptr[n]);
    parallel_for(n_nodes, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
            const auto& L = nb[i];
            std::copy(L.begin(), L.end(), P.bcol.begin() + P.brow_ptr[i]);
            bool lone = P.inc_ptr[i + 1] == P.inc_ptr[i];
            for (int ad = 0; ad < ndof; ++ad) {
                int64_t r = i * ndof + ad;
                int32_t* out = P.col.data() + P.row_ptr[r];
                if (lone) {
                    out[0] = (int32_t)r;
                    continue;
                }
                for (size_t q = 0; q < L.size(); ++q)
                    for (int bd = 0; bd < ndof; ++bd) *out++ = L[q] * ndof + bd;
            }
        }
    });
    // sliced ELL of the block pattern
    const int64_t nch = (n_nodes + kEllRows - 1) / kEllRows;
    P.ell_off.assign(nch + 1, 0);
    P.ell_nb.assign(nch, 0);
    int64_t ne = 0;
    for (int64_t c = 0; c < nch; ++c) {
        int32_t m = 0;
        for (int64_t r = c * kEllRows; r < std::min(n_nodes, (c + 1) * kEllRows); ++r) m = std::max(m, deg[r]);
        P.ell_nb[c] = m;
        ne += (int64_t)m * kEllRows;
        if (ne > INT32_MAX / (ndof * ndof)) throw std::runtime_error("ELL pattern exceeds 2^31 entries");
        P.ell_off[c + 1] = (int32_t)ne;
    }
    P.ell_col.resize(ne);
    parallel_for(nch, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; ++c)
            for (int l = 0; l < kEllRows; ++l) {
                const int64_t r = c * kEllRows + l;
                for (int32_t k = 0; k < P.ell_nb[c]; ++k) {
                    int32_t v = 0;
                    if (r < n_nodes) v = k < deg[r] ? P.bcol[P.brow_ptr[r] + k] : (int32_t)r;
                    P.ell_col[P.ell_off[c] + (int64_t)k * kEllRows + l] = v;
                }
            }
    });
}

}  // namespace svx
