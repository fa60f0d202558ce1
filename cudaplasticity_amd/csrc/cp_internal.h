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
// list node -> element * npe + local index (npe = 4 or 8 nodes per element).
struct Adjacency {
    std::vector<int> ptr, node;
    std::vector<int> incPtr, inc;
    int npe = 4;
};

int build_adjacency(int nn, int ne, int npe, const int* conn, Adjacency& adj);
void export_csr(const Adjacency& adj, int nn, int* rows, int* cols);

}  // namespace cp
