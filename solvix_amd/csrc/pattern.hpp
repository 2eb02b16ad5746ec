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

}  // namespace svx
