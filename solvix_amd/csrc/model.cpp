// Host-side model preparation for svx_solve_model, mirroring the checks and
// their order in solve_with_thermal (crates/solvix-core/src/lib.rs:600-632):
// NoConstraints, validate() (lib.rs:406-437), thickness, ndof_per_node
// (lib.rs:725-752); then the element pass (on the device), CLOAD rows
// (assembly.rs:365-380) and constrained_dofs (solver.rs:15-41). Node and
// material lookups are hashed instead of the reference's linear scans
// (lib.rs:106-108, 357-367) with the same first-match results.
#include <cmath>
#include <cstdio>
#include <limits>
#include <unordered_map>
#include <unordered_set>

#include "api_internal.hpp"

namespace svx {

namespace {
int fail(svx_error* err, int code, const std::string& msg) {
    if (err) {
        err->code = code;
        err->reason = 0;
        err->index = -1;
        snprintf(err->message, sizeof err->message, "%s", msg.c_str());
    }
    return code;
}
}  // namespace

int prepare_model(const svx_model_desc* m, PreparedModel& P, svx_error* err) {
    if (m->n_bcs == 0) return fail(err, SVX_E_NO_CONSTRAINTS, "NoConstraints");
    if (m->n_nodes <= 0 || m->n_elem <= 0) return fail(err, SVX_E_INVALID_MODEL, "EmptyMesh");
    if (m->n_materials <= 0) return fail(err, SVX_E_INVALID_MODEL, "NoMaterial");
    std::unordered_map<uint32_t, int64_t> pos;
    pos.reserve((size_t)m->n_nodes * 2);
    for (int64_t i = 0; i < m->n_nodes; ++i) pos.emplace(m->node_ids[i], i);  // first position wins
    for (int64_t e = 0; e < m->n_elem; ++e)
        if (m->kinds[e] > SVX_C3D8) return fail(err, SVX_E_ARGUMENT, "unsupported element kind");
    P.conn_pos.resize(m->conn_ptr[m->n_elem]);
    for (int64_t e = 0; e < m->n_elem; ++e) {
        int64_t len = m->conn_ptr[e + 1] - m->conn_ptr[e];
        if (len != nodes_of(m->kinds[e]))
            return fail(err, SVX_E_INVALID_MODEL, "MalformedElement(" + std::to_string(m->elem_ids[e]) + ")");
        for (int64_t k = m->conn_ptr[e]; k < m->conn_ptr[e + 1]; ++k) {
            auto it = pos.find(m->conn_ids[k]);
            if (it == pos.end())
                return fail(err, SVX_E_INVALID_MODEL, "UnknownNode(" + std::to_string(m->conn_ids[k]) + ")");
            P.conn_pos[k] = (int32_t)it->second;
        }
    }
    if (!(m->thickness > 0.0)) return fail(err, SVX_E_INVALID_MODEL, "thickness must be positive");
    bool plane = false, solid = false;
    for (int64_t e = 0; e < m->n_elem; ++e) (m->kinds[e] == SVX_C3D8 ? solid : plane) = true;
    if (plane && solid) return fail(err, SVX_E_INVALID_MODEL, "mixed dimensionality");
    P.ndof = solid ? 3 : 2;
    const int ndof = P.ndof;

    // material_for_element: first assignment whose elset contains the id
    std::vector<std::unordered_set<uint32_t>> sets(m->n_elsets);
    for (int32_t a = 0; a < m->n_assign; ++a) {
        int32_t s = m->assign_elset[a];
        if (s < 0 || s >= m->n_elsets || !sets[s].empty()) continue;
        for (int64_t k = m->elset_ptr[s]; k < m->elset_ptr[s + 1]; ++k) sets[s].insert(m->elset_members[k]);
    }
    P.E.assign(m->n_elem, std::numeric_limits<double>::quiet_NaN());
    P.nu.assign(m->n_elem, 0.0);
    for (int64_t e = 0; e < m->n_elem; ++e) {
        for (int32_t a = 0; a < m->n_assign; ++a) {
            int32_t s = m->assign_elset[a];
            if (s < 0 || s >= m->n_elsets) continue;
            if (sets[s].count(m->elem_ids[e])) {
                int32_t mat = m->assign_material[a];
                if (mat >= 0 && mat < m->n_materials) {
                    P.E[e] = m->youngs[mat];
                    P.nu[e] = m->poisson[mat];
                }
                break;
            }
        }
    }
    // load vector (errors are reported after the element pass)
    P.f.assign((size_t)m->n_nodes * ndof, 0.0);
    for (int64_t l = 0; l < m->n_cloads && P.load_error.empty(); ++l) {
        int dof = m->cload_dof[l];
        uint32_t nid = m->cload_node[l];
        char buf[256];
        if (dof <= 0 || dof > ndof) {
            snprintf(buf, sizeof buf, "load on node %u has unsupported dof %d", nid, dof);
            P.load_error = buf;
            break;
        }
        auto it = pos.find(nid);
        if (it == pos.end()) {
            snprintf(buf, sizeof buf, "load on unknown node %u", nid);
            P.load_error = buf;
            break;
        }
        P.f[it->second * ndof + dof - 1] += m->cload_value[l];
    }
    // constrained_dofs (solver.rs:15-41)
    for (int32_t b = 0; b < m->n_bcs && P.bc_error.empty(); ++b)
        for (int64_t k = m->bc_node_ptr[b]; k < m->bc_node_ptr[b + 1] && P.bc_error.empty(); ++k) {
            uint32_t nid = m->bc_nodes[k];
            char buf[256];
            auto it = pos.find(nid);
            if (it == pos.end()) {
                snprintf(buf, sizeof buf, "boundary condition on unknown node %u", nid);
                P.bc_error = buf;
                break;
            }
            for (int64_t q = m->bc_dof_ptr[b]; q < m->bc_dof_ptr[b + 1]; ++q) {
                int dof = m->bc_dofs[q];
                if (dof <= 0 || dof > ndof) {
                    snprintf(buf, sizeof buf, "boundary condition on node %u has unsupported dof %d", nid, dof);
                    P.bc_error = buf;
                    break;
                }
                P.cons_rows.push_back(it->second * ndof + dof - 1);
                P.cons_node.push_back(nid);
                P.cons_dof.push_back(dof);
            }
        }
    return SVX_OK;
}

}  // namespace svx
