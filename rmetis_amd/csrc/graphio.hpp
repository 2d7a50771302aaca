// graphio.hpp -- METIS .graph files to and from 0-based CSR (host only).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rk {

struct GraphFile {
    int32_t n = 0, ncon = 1;
    std::vector<int32_t> xadj, adjncy, vwgt, vsize, adjwgt;
    bool has_vwgt = false, has_adjwgt = false, has_vsize = false;
};

// Throws rk::Error(-2) on a malformed file (with the line), -4 on I/O errors.
GraphFile read_graph_file(const std::string& path);
// vwgt / vsize / adjwgt may be null (not written).
void write_graph_file(const std::string& path, int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                      const int32_t* vwgt, const int32_t* vsize, const int32_t* adjwgt);

}  // namespace rk
