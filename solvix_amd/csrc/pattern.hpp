#pragma once
#include <cstdint>
#include <functional>
#include <vector>

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
};

void build_pattern(int ndof, int64_t n_nodes, int64_t n_elem, const int64_t* conn_ptr, const int32_t* conn,
                   HostPattern& out);

}  // namespace svx
