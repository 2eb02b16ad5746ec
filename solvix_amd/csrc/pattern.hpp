#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "svx_internal.hpp"  // kEllRows

namespace svx {

struct HostPattern {
    int ndof = 0;
    int64_t n_nodes = 0;
    std::vector<int64_t> inc_ptr;   // [n_nodes + 1]
    std::vector<int32_t> inc_elem;  // incident elements, ascending
    std::vector<int8_t> inc_slot;   // local node position in that element
    std::vector<int32_t> row_ptr;   // [n + 1]
    std::vector<int32_t> col;       // [nnz]
    std::vector<int32_t> brow_ptr;  // [n_nodes + 1]
    std::vector<int32_t> bcol;      // node-block columns
    // Node-row sliced ELL (the Krylov SpMV layout): chunk c holds node rows
    // [64c, 64c + 64), one per lane, each padded to the chunk's longest block
    // row, block k of lane l at ell_off[c] + 64 k + l. Padding blocks point at
    // the row's own node (node 0 beyond the last node) and hold zeros.
    std::vector<int32_t> ell_off;  // [n_chunks + 1], in blocks
    std::vector<int32_t> ell_nb;   // [n_chunks] padded blocks per row
    std::vector<int32_t> ell_col;  // [ell_off[n_chunks]] node-block columns
};

void build_pattern(int ndof, int64_t n_nodes, int64_t n_elem, const int64_t* conn_ptr, const int32_t* conn,
                   HostPattern& out);

// Node patches of the matrix-free plane-stress operator (krylov.hip): the
// nodes in Morton order of their (x, y) coordinates, cut greedily into runs
// of at most kBlock owned nodes whose incident elements fit one (CPS4) or two
// (CPS3) element slots per thread and whose nodes fit kMfMaxNodes.
struct MfPlan {
    int kind = 0;  // 4 (CPS4) or 3 (CPS3); 0: the mesh has no matrix-free operator (see `why`)
    std::vector<MfPatch> patches;
    std::vector<int32_t> nodes;   // per patch: owned nodes (Morton order), then halo nodes (ascending)
    std::vector<uint16_t> econn;  // [element slot][4] patch-local node positions (CPS3: fourth unused)
    std::vector<uint16_t> adj;    // per patch [k][kBlock]: (slot << 2 | corner) of owned node lane's k-th incidence
    int32_t cap_nodes = 0, cap_elem = 0;
    int64_t elem_slots = 0;  // element evaluations per product (patch overlap included)
    const char* why = "";
};
// Plane meshes (ndof 2) of one element kind, CPS3 or CPS4, with at most 8
// (CPS3) or 4 (CPS4) elements per node; `hp` supplies the incidence lists.
void build_mf_plan(int ndof, int64_t n_nodes, const double* coords, int64_t n_elem, const uint8_t* kinds,
                   const int64_t* conn_ptr, const int32_t* conn, const HostPattern& hp, MfPlan& out);
// Node groups of the multi-RHS block SpMV (block.hip): runs of `group`
// consecutive nodes in Morton order of their coordinates (a 2x2x1 patch of a
// structured solid mesh for group 4). One wavefront multiplies a group's
// block rows: it gathers the direction row of each column in the union of the
// members' block columns once and applies it to every member block in that
// column, instead of gathering it once per block. The matrix values are
// packed group by group in that order (union column ascending, then member),
// so a group's blocks are one contiguous run.
struct BlkPlan {
    int group = 0;
    std::vector<int32_t> nodes;      // [n_groups * group] members in Morton order, -1 past the last node
    std::vector<int32_t> grp;        // [n_groups][4]: first union column, end, first packed block, end
    std::vector<int32_t> ucol;       // [n_cols][2]: column node (ascending within a group), member mask
    std::vector<int32_t> pk_src;     // [nblk] CSR offset of packed block t's (0, 0) entry
    std::vector<int32_t> pk_stride;  // [nblk] CSR offset between its rows (0: a diagonal-only row)
    int64_t cols = 0;                // union columns over all groups
    int32_t max_cols = 0, max_blocks = 0;  // the most union columns and blocks of one group
    const char* why = "";            // non-empty: no plan
};
void build_blk_plan(int64_t n_nodes, const double* coords, const HostPattern& hp, int group, BlkPlan& out);

// Host check of a block SpMV plan: every node a member of exactly one group,
// every block (and CSR entry) packed exactly once, and the product of a seeded matrix on the
// pattern with `R` seeded columns evaluated group by group the way the
// kernel indexes it (union columns, member masks, packed values read through
// pk_src / pk_stride) against the plain CSR product. Returns max |difference|
// / max |product|; throws std::logic_error naming a violated invariant.
double verify_blk_plan(const BlkPlan& plan, const HostPattern& hp, int R);

// Host check of a plan: its structural limits, every node owned by exactly
// one patch, and the product q = K x evaluated patch by patch the way the
// kernel does (LDS staging replaced by arrays) against the element-by-element
// product, for a seeded x and unit material. Returns max |difference| /
// max |q|; throws std::logic_error naming a violated limit.
double verify_mf_plan(const MfPlan& plan, int64_t n_nodes, const double* coords, int64_t n_elem,
                      const int64_t* conn_ptr, const int32_t* conn);

}  // namespace svx
