// Host construction of the per-mesh structures, built once per mesh:
//  - node -> (element, local slot) incidence, in element order;
//  - the CSR pattern of csr_from_triplets (assembly.rs:87-116): row
//    node*ndof + a holds every column j*ndof + b for j in the node's
//    element-neighbourhood (sorted), or only its diagonal when the node
//    belongs to no element (assembly.rs:300-302);
//  - the node-block (BSR) pattern with the same neighbourhoods, and its
//    node-row sliced ELL form for the Krylov SpMV;
//  - on request, the node patches of the matrix-free plane-stress operator.
#include "pattern.hpp"

#include <cmath>

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "mf_element.hpp"

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

namespace {
uint64_t spread3(uint64_t v) {  // bit i -> bit 3i, 21 bits
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
}  // namespace

void build_blk_plan(int64_t n_nodes, const double* coords, const HostPattern& hp, int group, BlkPlan& B) {
    B = BlkPlan{};
    B.group = group;
    if (n_nodes <= 0) {
        B.why = "empty mesh";
        return;
    }
    if (group < 1 || group > 8) {
        B.why = "group size";
        return;
    }
    const int BS = hp.ndof;
    // Morton order of (x, y, z), 21 bits per axis over the bounding cube
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = coords[a];
    for (int64_t i = 0; i < n_nodes; ++i)
        for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], coords[3 * i + a]), hi[a] = std::max(hi[a], coords[3 * i + a]);
    double span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(span > 0.0)) span = 1.0;
    const double scale = 2097151.0 / span;
    std::vector<uint64_t> key(n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) {
        uint64_t k = 0;
        for (int a = 0; a < 3; ++a) {
            const double qv = std::min(2097151.0, std::max(0.0, (coords[3 * i + a] - lo[a]) * scale));
            k |= spread3((uint64_t)qv) << a;
        }
        key[i] = k;
    }
    std::vector<int32_t> order(n_nodes);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    const int64_t ng = (n_nodes + group - 1) / group;
    B.nodes.assign(ng * group, -1);
    for (int64_t i = 0; i < n_nodes; ++i) B.nodes[i] = order[i];
    B.grp.resize(ng * 4);
    B.pk_src.reserve(hp.bcol.size());
    B.pk_stride.reserve(hp.bcol.size());
    std::vector<std::pair<int32_t, int32_t>> cm;  // (column, member)
    for (int64_t g = 0; g < ng; ++g) {
        cm.clear();
        for (int m = 0; m < group; ++m) {
            const int32_t v = B.nodes[g * group + m];
            if (v < 0) break;
            for (int32_t k = hp.brow_ptr[v]; k < hp.brow_ptr[v + 1]; ++k) cm.push_back({hp.bcol[k], m});
        }
        std::sort(cm.begin(), cm.end());
        B.grp[g * 4 + 0] = (int32_t)(B.ucol.size() / 2);
        B.grp[g * 4 + 2] = (int32_t)B.pk_src.size();
        for (size_t q = 0; q < cm.size();) {
            const int32_t col = cm[q].first;
            int32_t mask = 0;
            for (; q < cm.size() && cm[q].first == col; ++q) {
                const int m = cm[q].second;
                const int32_t v = B.nodes[g * group + m];
                mask |= 1 << m;
                // block k of node v: row a of v holds it at row_ptr[v*BS + a] + k*BS (full blocks);
                // a diagonal-only row holds its diagonal at row_ptr[v*BS + a]
                const int32_t kb = hp.brow_ptr[v], nb = hp.brow_ptr[v + 1] - kb;
                const int32_t k = (int32_t)(std::lower_bound(hp.bcol.begin() + kb, hp.bcol.begin() + kb + nb, col) -
                                            (hp.bcol.begin() + kb));
                const int32_t rs = hp.row_ptr[(int64_t)v * BS];
                const bool lone = hp.row_ptr[(int64_t)v * BS + BS] - rs != BS * BS * nb;
                B.pk_src.push_back(lone ? rs : rs + k * BS);
                B.pk_stride.push_back(lone ? 0 : BS * nb);
            }
            B.ucol.push_back(col);
            B.ucol.push_back(mask);
        }
        B.grp[g * 4 + 3] = (int32_t)B.pk_src.size();
        B.max_cols = std::max(B.max_cols, (int32_t)(B.ucol.size() / 2) - B.grp[g * 4 + 0]);
        B.max_blocks = std::max(B.max_blocks, (int32_t)B.pk_src.size() - B.grp[g * 4 + 2]);
        B.grp[g * 4 + 1] = (int32_t)(B.ucol.size() / 2);
        if (B.ucol.size() / 2 > (size_t)INT32_MAX) throw std::runtime_error("block SpMV plan exceeds 2^31 entries");
    }
    B.cols = (int64_t)(B.ucol.size() / 2);
}

double verify_blk_plan(const BlkPlan& B, const HostPattern& hp, int R) {
    auto fail = [](const char* what) { throw std::logic_error(std::string("block SpMV plan: ") + what); };
    const int BS = hp.ndof, G = B.group;
    const int64_t nn = hp.n_nodes, n = nn * BS;
    const int64_t ng = (int64_t)B.grp.size() / 4;
    if ((int64_t)B.nodes.size() != ng * G || ng * G < nn) fail("group count");
    if (B.pk_src.size() != hp.bcol.size() || B.pk_stride.size() != hp.bcol.size()) fail("packed block count");
    std::vector<int> member(nn, 0);
    for (int64_t i = 0; i < ng * G; ++i) {
        const int32_t v = B.nodes[i];
        if (v >= nn) fail("node index");
        if (v >= 0) member[v]++;
    }
    for (int64_t v = 0; v < nn; ++v)
        if (member[v] != 1) fail("a node is not in exactly one group");
    // the CSR entries the packed blocks read: every entry exactly once
    std::vector<int> hit(hp.col.size(), 0);
    for (size_t t = 0; t < B.pk_src.size(); ++t)
        for (int a = 0; a < BS; ++a)
            for (int b = 0; b < BS; ++b) {
                if (B.pk_stride[t] == 0 && a != b) continue;
                const int64_t e = B.pk_stride[t] == 0 ? B.pk_src[t] + a : B.pk_src[t] + (int64_t)a * B.pk_stride[t] + b;
                if (e < 0 || e >= (int64_t)hp.col.size()) fail("packed entry out of range");
                hit[e]++;
            }
    for (int h : hit)
        if (h != 1) fail("a CSR entry is not packed exactly once");
    uint64_t st = 0x9e3779b97f4a7c15ull;
    auto next = [&]() {
        st += 0x9e3779b97f4a7c15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return (double)((z ^ (z >> 31)) >> 11) / 9007199254740992.0 - 0.5;
    };
    std::vector<double> val(hp.col.size()), x((size_t)n * R), ref((size_t)n * R, 0.0), out((size_t)n * R, 0.0);
    for (double& v : val) v = next();
    for (double& v : x) v = next();
    for (int64_t r = 0; r < n; ++r)
        for (int32_t k = hp.row_ptr[r]; k < hp.row_ptr[r + 1]; ++k)
            for (int c = 0; c < R; ++c) ref[(size_t)r * R + c] += val[k] * x[(size_t)hp.col[k] * R + c];
    // packed values as the packing kernel writes them, then the wavefront's loop
    std::vector<double> pv(B.pk_src.size() * BS * BS);
    for (size_t t = 0; t < B.pk_src.size(); ++t)
        for (int a = 0; a < BS; ++a)
            for (int b = 0; b < BS; ++b)
                pv[(t * BS + a) * BS + b] = B.pk_stride[t] == 0 ? (a == b ? val[B.pk_src[t] + a] : 0.0)
                                                                : val[B.pk_src[t] + (int64_t)a * B.pk_stride[t] + b];
    std::vector<double> acc((size_t)G * BS * R);
    for (int64_t g = 0; g < ng; ++g) {
        std::fill(acc.begin(), acc.end(), 0.0);
        int32_t prev = -1;
        int64_t t = B.grp[g * 4 + 2];
        const int32_t nc = B.grp[g * 4 + 1] - B.grp[g * 4];
        if (nc > B.max_cols) fail("union column cap");
        for (int32_t c = B.grp[g * 4]; c < B.grp[g * 4 + 1]; ++c) {
            const int32_t u = B.ucol[2 * c], m = B.ucol[2 * c + 1];
            int members = 0;
            while (members < G && B.nodes[g * G + members] >= 0) ++members;
            if (u <= prev || u < 0 || u >= nn || m == 0 || (m >> members) != 0) fail("union column");
            prev = u;
            for (int mb = 0; mb < G; ++mb) {
                if (!((m >> mb) & 1)) continue;
                for (int a = 0; a < BS; ++a)
                    for (int b = 0; b < BS; ++b)
                        for (int c2 = 0; c2 < R; ++c2)
                            acc[((size_t)mb * BS + a) * R + c2] +=
                                pv[(t * BS + a) * BS + b] * x[((size_t)u * BS + b) * R + c2];
                ++t;
            }
        }
        const int64_t t_end = g + 1 < ng ? B.grp[(g + 1) * 4 + 2] : (int64_t)B.pk_src.size();
        if (t != t_end || t != B.grp[g * 4 + 3] || t - B.grp[g * 4 + 2] > B.max_blocks) fail("packed block run");
        for (int mb = 0; mb < G; ++mb) {
            const int32_t v = B.nodes[g * G + mb];
            if (v < 0) break;
            for (int a = 0; a < BS; ++a)
                for (int c2 = 0; c2 < R; ++c2)
                    out[((size_t)v * BS + a) * R + c2] = acc[((size_t)mb * BS + a) * R + c2];
        }
    }
    double dmax = 0.0, qmax = 0.0;
    for (size_t i = 0; i < ref.size(); ++i) {
        dmax = std::max(dmax, std::fabs(out[i] - ref[i]));
        qmax = std::max(qmax, std::fabs(ref[i]));
    }
    return qmax > 0.0 ? dmax / qmax : dmax;
}

namespace {
uint64_t spread_bits(uint64_t v) {  // bit i -> bit 2i, 32 bits
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}
}  // namespace

void build_mf_plan(int ndof, int64_t n_nodes, const double* coords, int64_t n_elem, const uint8_t* kinds,
                   const int64_t* conn_ptr, const int32_t* conn, const HostPattern& hp, MfPlan& M) {
    M = MfPlan{};
    if (ndof != 2 || n_elem == 0 || n_nodes == 0) {
        M.why = "not a plane mesh";
        return;
    }
    const int kind = kinds[0] == SVX_CPS4 ? 4 : kinds[0] == SVX_CPS3 ? 3 : 0;
    if (kind == 0) {
        M.why = "element kind";
        return;
    }
    for (int64_t e = 0; e < n_elem; ++e)
        if (kinds[e] != kinds[0] || conn_ptr[e + 1] - conn_ptr[e] != kind) {
            M.why = "mixed element kinds";
            return;
        }
    const int max_adj = mf_max_adj(kind), cap_e = mf_max_elem(kind);
    for (int64_t i = 0; i < n_nodes; ++i)
        if (hp.inc_ptr[i + 1] - hp.inc_ptr[i] > max_adj) {
            M.why = "too many elements at a node";
            return;
        }
    // Morton order of (x, y), 22 bits per axis over the bounding square
    double lo[2] = {coords[0], coords[1]}, hi[2] = {coords[0], coords[1]};
    for (int64_t i = 0; i < n_nodes; ++i)
        for (int a = 0; a < 2; ++a) lo[a] = std::min(lo[a], coords[3 * i + a]), hi[a] = std::max(hi[a], coords[3 * i + a]);
    double span = std::max(hi[0] - lo[0], hi[1] - lo[1]);
    if (!(span > 0.0)) span = 1.0;
    const double scale = 4194303.0 / span;
    std::vector<uint64_t> key(n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) {
        const uint64_t qx = (uint64_t)std::min(4194303.0, std::max(0.0, (coords[3 * i] - lo[0]) * scale));
        const uint64_t qy = (uint64_t)std::min(4194303.0, std::max(0.0, (coords[3 * i + 1] - lo[1]) * scale));
        key[i] = spread_bits(qx) | (spread_bits(qy) << 1);
    }
    std::vector<int32_t> order(n_nodes);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    std::vector<int64_t> estamp(n_elem, -1), nstamp(n_nodes, -1);
    std::vector<int32_t> loc(n_nodes, 0), elems, owned, halo;
    std::vector<uint16_t> lists;
    std::vector<int32_t> lcount;
    int64_t pos = 0, p = 0;
    while (pos < n_nodes) {
        // grow the patch along the Morton order while its elements and nodes fit
        elems.clear(), owned.clear();
        int32_t n_nodes_p = 0;
        while (pos < n_nodes && (int)owned.size() < kBlock) {
            const int32_t g = order[pos];
            int add_e = 0, add_n = nstamp[g] != p ? 1 : 0;
            for (int64_t q = hp.inc_ptr[g]; q < hp.inc_ptr[g + 1]; ++q) {
                const int32_t e = hp.inc_elem[q];
                if (estamp[e] == p) continue;
                add_e++;
                for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k) add_n += nstamp[conn[k]] != p && conn[k] != g;
            }
            // add_n over-counts nodes shared by two new elements: an upper bound
            if (!owned.empty() && ((int)elems.size() + add_e > cap_e || n_nodes_p + add_n > kMfMaxNodes)) break;
            if (nstamp[g] != p) nstamp[g] = p, n_nodes_p++;
            for (int64_t q = hp.inc_ptr[g]; q < hp.inc_ptr[g + 1]; ++q) {
                const int32_t e = hp.inc_elem[q];
                if (estamp[e] == p) continue;
                estamp[e] = p;
                elems.push_back(e);
                for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k)
                    if (nstamp[conn[k]] != p) nstamp[conn[k]] = p, n_nodes_p++;
            }
            owned.push_back(g);
            ++pos;
        }
        MfPatch pt{};
        pt.n_owned = (int32_t)owned.size();
        std::sort(elems.begin(), elems.end());
        // node list: owned nodes in patch order, then the halo ascending
        const int64_t tag = -2 - p;  // second stamp of this patch: node placed
        pt.node_off = (int32_t)M.nodes.size();
        for (int32_t k = 0; k < pt.n_owned; ++k) {
            nstamp[owned[k]] = tag;
            loc[owned[k]] = k;
            M.nodes.push_back(owned[k]);
        }
        halo.clear();
        for (int32_t e : elems)
            for (int64_t k = conn_ptr[e]; k < conn_ptr[e + 1]; ++k)
                if (nstamp[conn[k]] != tag) nstamp[conn[k]] = tag, halo.push_back(conn[k]);
        std::sort(halo.begin(), halo.end());
        for (size_t h = 0; h < halo.size(); ++h) {
            loc[halo[h]] = pt.n_owned + (int32_t)h;
            M.nodes.push_back(halo[h]);
        }
        pt.n_nodes = pt.n_owned + (int32_t)halo.size();
        if (pt.n_nodes > kMfMaxNodes || (int)elems.size() > cap_e) throw std::logic_error("matrix-free patch overflow");
        pt.elem_off = (int32_t)(M.econn.size() / 4);
        pt.n_elem = (int32_t)elems.size();
        for (int32_t e : elems)
            for (int a = 0; a < 4; ++a) M.econn.push_back(a < kind ? (uint16_t)loc[conn[conn_ptr[e] + a]] : 0);
        // incidences of the owned nodes, in element order
        lcount.assign(pt.n_owned, 0);
        lists.assign((size_t)pt.n_owned * max_adj, 0);
        for (size_t sl = 0; sl < elems.size(); ++sl)
            for (int a = 0; a < kind; ++a) {
                const int32_t l = loc[conn[conn_ptr[elems[sl]] + a]];
                if (l < pt.n_owned) lists[(size_t)l * max_adj + lcount[l]++] = (uint16_t)(sl * 4 + a);
            }
        int32_t A = 0;
        for (int32_t l = 0; l < pt.n_owned; ++l) A = std::max(A, lcount[l]);
        pt.n_adj = A;
        pt.adj_off = (int32_t)M.adj.size();
        M.adj.resize(M.adj.size() + (size_t)A * kBlock, (uint16_t)(pt.n_elem * 4));  // padding: the zero slot
        for (int32_t l = 0; l < pt.n_owned; ++l)
            for (int32_t k = 0; k < lcount[l]; ++k) M.adj[pt.adj_off + (size_t)k * kBlock + l] = lists[(size_t)l * max_adj + k];
        M.cap_nodes = std::max(M.cap_nodes, pt.n_nodes);
        M.cap_elem = std::max(M.cap_elem, pt.n_elem);
        M.elem_slots += pt.n_elem;
        M.patches.push_back(pt);
        ++p;
        if (M.nodes.size() > (size_t)INT32_MAX || M.adj.size() > (size_t)INT32_MAX || M.econn.size() / 4 > (size_t)INT32_MAX)
            throw std::runtime_error("matrix-free plan exceeds 2^31 entries");
    }
    M.kind = kind;
}

namespace {
template <int NEN>
void mf_element(const double* coords, const int64_t* conn_ptr, const int32_t* conn, int64_t e,
                const std::vector<double>& x, const MfMaterial& mat, double (&f)[NEN][2]) {
    double xe[NEN][2], ue[NEN][2];
    for (int a = 0; a < NEN; ++a) {
        const int32_t g = conn[conn_ptr[e] + a];
        xe[a][0] = coords[3 * g], xe[a][1] = coords[3 * g + 1];
        ue[a][0] = x[2 * g], ue[a][1] = x[2 * g + 1];
    }
    mf_forces<NEN>(xe, ue, mat, f);
}

template <int NEN>
double verify_plan(const MfPlan& M, int64_t n_nodes, const double* coords, int64_t n_elem, const int64_t* conn_ptr,
                   const int32_t* conn) {
    auto fail = [](const char* what) { throw std::logic_error(std::string("matrix-free plan: ") + what); };
    const int max_adj = mf_max_adj(NEN), max_elem = mf_max_elem(NEN);
    const double c = 1.0 / (1.0 - 0.3 * 0.3);
    const MfMaterial mat{c, 0.3 * c, c * (1.0 - 0.3) / 2.0, 1.0};
    std::vector<double> x(2 * n_nodes);
    uint64_t st = 0x5317000000000007ull;
    for (auto& v : x) {  // SplitMix64, (x >> 11) 2^-53 - 0.5
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        v = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    // element by element, in element order
    std::vector<double> ref(2 * n_nodes, 0.0);
    for (int64_t e = 0; e < n_elem; ++e) {
        double f[NEN][2];
        mf_element<NEN>(coords, conn_ptr, conn, e, x, mat, f);
        for (int a = 0; a < NEN; ++a) {
            const int32_t g = conn[conn_ptr[e] + a];
            ref[2 * g] += f[a][0], ref[2 * g + 1] += f[a][1];
        }
    }
    // patch by patch
    std::vector<double> q(2 * n_nodes, 0.0);
    std::vector<int> owned_by(n_nodes, 0);
    std::vector<double> fe;
    for (const MfPatch& P : M.patches) {
        if (P.n_owned < 1 || P.n_owned > 256 || P.n_nodes < P.n_owned || P.n_nodes > kMfMaxNodes) fail("node counts");
        if (P.n_elem < 0 || P.n_elem > max_elem || P.n_adj < 0 || P.n_adj > max_adj) fail("element counts");
        if ((size_t)P.node_off + P.n_nodes > M.nodes.size() || (size_t)(P.elem_off + P.n_elem) * 4 > M.econn.size() ||
            (size_t)P.adj_off + (size_t)P.n_adj * 256 > M.adj.size())
            fail("offsets");
        const int32_t* nl = M.nodes.data() + P.node_off;
        for (int j = 0; j < P.n_nodes; ++j)
            if (nl[j] < 0 || nl[j] >= n_nodes) fail("node index");
        for (int j = 0; j < P.n_owned; ++j) owned_by[nl[j]]++;
        fe.assign((size_t)2 * NEN * (P.n_elem + 1), 0.0);  // the last slot stays zero
        for (int sl = 0; sl < P.n_elem; ++sl) {
            const uint16_t* ec = M.econn.data() + (size_t)(P.elem_off + sl) * 4;
            double xe[NEN][2], ue[NEN][2], f[NEN][2];
            for (int a = 0; a < NEN; ++a) {
                if (ec[a] >= P.n_nodes) fail("local node");
                const int32_t g = nl[ec[a]];
                xe[a][0] = coords[3 * g], xe[a][1] = coords[3 * g + 1];
                ue[a][0] = x[2 * g], ue[a][1] = x[2 * g + 1];
            }
            mf_forces<NEN>(xe, ue, mat, f);
            for (int a = 0; a < NEN; ++a) fe[(2 * a) * (P.n_elem + 1) + sl] = f[a][0], fe[(2 * a + 1) * (P.n_elem + 1) + sl] = f[a][1];
        }
        for (int l = 0; l < P.n_owned; ++l) {
            double a0 = 0.0, a1 = 0.0;
            for (int k = 0; k < P.n_adj; ++k) {
                const uint32_t code = M.adj[P.adj_off + (size_t)k * 256 + l], sl = code >> 2, cr = code & 3u;
                if ((int)sl > P.n_elem || (int)cr >= NEN) fail("incidence code");
                a0 += fe[(2 * cr) * (P.n_elem + 1) + sl];
                a1 += fe[(2 * cr + 1) * (P.n_elem + 1) + sl];
            }
            q[2 * nl[l]] = a0, q[2 * nl[l] + 1] = a1;
        }
    }
    for (int64_t i = 0; i < n_nodes; ++i)
        if (owned_by[i] != 1) fail("a node is not owned by exactly one patch");
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < 2 * n_nodes; ++i) num = std::max(num, std::fabs(q[i] - ref[i])), den = std::max(den, std::fabs(ref[i]));
    return den > 0.0 ? num / den : num;
}
}  // namespace

double verify_mf_plan(const MfPlan& M, int64_t n_nodes, const double* coords, int64_t n_elem, const int64_t* conn_ptr,
                      const int32_t* conn) {
    if (M.kind == 4) return verify_plan<4>(M, n_nodes, coords, n_elem, conn_ptr, conn);
    if (M.kind == 3) return verify_plan<3>(M, n_nodes, coords, n_elem, conn_ptr, conn);
    throw std::logic_error("matrix-free plan: no plan");
}

}  // namespace svx
