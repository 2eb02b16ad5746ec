// A model held by the library (the caller's arrays may go away): the arrays
// of an svx_model_desc and a descriptor pointing at them. Built from a
// caller's descriptor (server submission) or by the .inp reader.
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/svx.h"

namespace svx {

struct OwnedModel {
    std::vector<uint32_t> node_ids, elem_ids, conn_ids, elset_members, bc_nodes, cload_node;
    std::vector<double> coords, youngs, poisson, cload_value;
    std::vector<uint8_t> kinds;
    std::vector<int64_t> conn_ptr{0}, elset_ptr{0}, bc_node_ptr{0}, bc_dof_ptr{0};
    std::vector<int32_t> assign_elset, assign_material, bc_dofs, cload_dof;
    double thickness = 1.0;
    svx_model_desc desc{};

    OwnedModel() { finalize(); }
    explicit OwnedModel(const svx_model_desc& m) {
        copy(node_ids, m.node_ids, m.n_nodes);
        copy(coords, m.coords, m.n_nodes * 3);
        copy(elem_ids, m.elem_ids, m.n_elem);
        copy(kinds, m.kinds, m.n_elem);
        copy(conn_ptr, m.conn_ptr, m.n_elem + 1);
        copy(conn_ids, m.conn_ids, m.conn_ptr ? m.conn_ptr[m.n_elem] : 0);
        copy(youngs, m.youngs, m.n_materials);
        copy(poisson, m.poisson, m.n_materials);
        copy(elset_ptr, m.elset_ptr, m.n_elsets + 1);
        copy(elset_members, m.elset_members, m.elset_ptr ? m.elset_ptr[m.n_elsets] : 0);
        copy(assign_elset, m.assign_elset, m.n_assign);
        copy(assign_material, m.assign_material, m.n_assign);
        copy(bc_node_ptr, m.bc_node_ptr, m.n_bcs + 1);
        copy(bc_nodes, m.bc_nodes, m.bc_node_ptr ? m.bc_node_ptr[m.n_bcs] : 0);
        copy(bc_dof_ptr, m.bc_dof_ptr, m.n_bcs + 1);
        copy(bc_dofs, m.bc_dofs, m.bc_dof_ptr ? m.bc_dof_ptr[m.n_bcs] : 0);
        copy(cload_node, m.cload_node, m.n_cloads);
        copy(cload_dof, m.cload_dof, m.n_cloads);
        copy(cload_value, m.cload_value, m.n_cloads);
        thickness = m.thickness;
        finalize();
    }
    OwnedModel(const OwnedModel&) = delete;
    OwnedModel& operator=(const OwnedModel&) = delete;

    // Point the descriptor at the arrays (after any change to them).
    void finalize() {
        desc = svx_model_desc{};
        desc.n_nodes = (int64_t)node_ids.size();
        desc.node_ids = node_ids.data();
        desc.coords = coords.data();
        desc.n_elem = (int64_t)elem_ids.size();
        desc.elem_ids = elem_ids.data();
        desc.kinds = kinds.data();
        desc.conn_ptr = conn_ptr.data();
        desc.conn_ids = conn_ids.data();
        desc.n_materials = (int32_t)youngs.size();
        desc.youngs = youngs.data();
        desc.poisson = poisson.data();
        desc.n_elsets = (int32_t)elset_ptr.size() - 1;
        desc.elset_ptr = elset_ptr.data();
        desc.elset_members = elset_members.data();
        desc.n_assign = (int32_t)assign_elset.size();
        desc.assign_elset = assign_elset.data();
        desc.assign_material = assign_material.data();
        desc.thickness = thickness;
        desc.n_bcs = (int32_t)bc_node_ptr.size() - 1;
        desc.bc_node_ptr = bc_node_ptr.data();
        desc.bc_nodes = bc_nodes.data();
        desc.bc_dof_ptr = bc_dof_ptr.data();
        desc.bc_dofs = bc_dofs.data();
        desc.n_cloads = (int64_t)cload_node.size();
        desc.cload_node = cload_node.data();
        desc.cload_dof = cload_dof.data();
        desc.cload_value = cload_value.data();
    }

   private:
    template <typename T>
    static void copy(std::vector<T>& v, const T* p, int64_t n) {
        v.assign(p ? p : nullptr, p ? p + std::max<int64_t>(0, n) : nullptr);
    }
};

// Reads the .inp subset of the linear path into `out`. Returns the 1-based
// line of the first error (message in *msg, "line N: ..." without the
// prefix), or 0.
int64_t parse_inp(const char* text, size_t len, OwnedModel& out, std::string* msg);

}  // namespace svx
