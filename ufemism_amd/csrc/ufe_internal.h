// ufe_internal.h -- shared declarations of libufe_dyn (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ufe_dyn.h"

namespace ufe {

constexpr int SELL = 64;          // rows per SELL chunk = one wavefront
constexpr int BLOCK = 256;        // threads per block for row kernels
constexpr int MAX_NZ = 32;

// Host-side mesh builders (mesh_build.hip).  All take/return reference-layout
// arrays through the descriptor.
int build_Ac_mesh(ufe_mesh_desc *m, std::string &err);
int build_AaAc_connectivity(ufe_mesh_desc *m, std::string &err);
int five_colouring(ufe_mesh_desc *m, std::string &err);

// Column-major 1-based access helpers for reference-layout host arrays.
template <typename T>
inline T &at(T *a, int64_t ld, int i, int j) {
  return a[(int64_t)(j - 1) * ld + (i - 1)];
}
template <typename T>
inline const T &at(const T *a, int64_t ld, int i, int j) {
  return a[(int64_t)(j - 1) * ld + (i - 1)];
}

}  // namespace ufe
