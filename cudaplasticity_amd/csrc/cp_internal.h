// cp_internal.h -- declarations shared by the library's translation units.
#pragma once

#include <string>
#include <vector>

namespace cp {

// Records the calling thread's last error message and returns `code`.
int fail(int code, const char* msg);
int failf(int code, const char* fmt, ...);

// Node adjacency in the reference CSR order (self first, then first
// appearance while scanning elements) plus the element-ordered incidence
// list node -> (element << 2 | local index).
struct Adjacency {
    std::vector<int> ptr, node;
    std::vector<int> incPtr, inc;
};

int build_adjacency(int nn, int ne, const int* conn, Adjacency& adj);
void export_csr(const Adjacency& adj, int nn, int* rows, int* cols);

}  // namespace cp
