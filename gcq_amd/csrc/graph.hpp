// graph.hpp — host CSR container and synthetic generators (graph.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gcqg {

struct Csr {
  int32_t nv = 0;
  std::vector<int32_t> offsets;  // nv + 1
  std::vector<int32_t> indices;  // offsets[nv]
  int64_t ne() const { return offsets.empty() ? 0 : offsets.back(); }
};

bool lattice(int32_t R, int32_t C, double keep, double diag, uint64_t seed, Csr* g);
bool rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed, Csr* g);
int32_t pick_source(const Csr& g, uint64_t seed, uint64_t k);

enum { kFormatAuto = 0, kFormatMtx = 1, kFormatSnap = 2 };
// 0 on success, -1 with *err set otherwise.
int load_edge_file(const char* path, int format, bool undirected, Csr* g, std::string* err);

}  // namespace gcqg
