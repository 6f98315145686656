// Host side of the aggregation multigrid: greedy aggregation, the Galerkin
// product for piecewise-constant prolongation and the transposed couplings.
// Plain C++ (no HIP types) so that the CPU tests compile it on their own.
#pragma once

#include <cstdint>
#include <vector>

namespace ufe {

struct AmgPair {                // {u-u, v-v} coefficient pair; the layout of a double2
  double x, y;
};

struct AmgHostOp {              // a point operator: off-diagonal pairs, fixed diagonal pairs, u-v coupling
  int n = 0;
  std::vector<int32_t> ptr, col;
  std::vector<AmgPair> cf, fd;
  std::vector<double> S;
};

// Greedy aggregation: a point whose neighbours are all free takes them into a
// new aggregate; every point left joins the aggregate of its first aggregated
// neighbour (or forms its own).
inline std::vector<int32_t> amg_aggregate(const AmgHostOp &A, int *na) {
  std::vector<int32_t> agg(A.n, -1);
  int k = 0;
  for (int i = 0; i < A.n; i++) {
    if (agg[i] >= 0) continue;
    bool free = true;
    for (int q = A.ptr[i]; q < A.ptr[i + 1] && free; q++) free = agg[A.col[q]] < 0;
    if (!free) continue;
    agg[i] = k;
    for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++) agg[A.col[q]] = k;
    k++;
  }
  for (int i = 0; i < A.n; i++) {
    if (agg[i] >= 0) continue;
    int a = -1;
    for (int q = A.ptr[i]; q < A.ptr[i + 1] && a < 0; q++) a = agg[A.col[q]];
    agg[i] = a >= 0 ? a : k++;
  }
  *na = k;
  return agg;
}

// Galerkin product P^T A P for piecewise-constant P, and the member lists.
// C.fd holds the whole fixed diagonal part (the members' fixed parts plus the
// couplings inside each aggregate); intra the couplings inside alone, which is
// what a level adds to the sum of its members' full diagonals.
inline void amg_galerkin(const AmgHostOp &A, const std::vector<int32_t> &agg, int na, AmgHostOp &C,
                         std::vector<int32_t> &mptr, std::vector<int32_t> &mem, std::vector<AmgPair> &intra) {
  mptr.assign(na + 1, 0);
  for (int i = 0; i < A.n; i++) mptr[agg[i] + 1]++;
  for (int I = 0; I < na; I++) mptr[I + 1] += mptr[I];
  mem.resize(A.n);
  {
    std::vector<int32_t> cur(mptr.begin(), mptr.end() - 1);
    for (int i = 0; i < A.n; i++) mem[cur[agg[i]]++] = i;
  }
  C.n = na;
  C.ptr.assign(1, 0);
  C.col.clear();
  C.cf.clear();
  C.fd.assign(na, AmgPair{0.0, 0.0});
  C.S.assign(na, 0.0);
  intra.assign(na, AmgPair{0.0, 0.0});
  std::vector<int64_t> mark(na, -1);
  for (int I = 0; I < na; I++) {
    const int64_t row0 = (int64_t)C.col.size();
    for (int k = mptr[I]; k < mptr[I + 1]; k++) {
      const int i = mem[k];
      C.fd[I].x += A.fd[i].x;
      C.fd[I].y += A.fd[i].y;
      C.S[I] += A.S[i];
      for (int q = A.ptr[i]; q < A.ptr[i + 1]; q++) {
        const int J = agg[A.col[q]];
        const AmgPair c = A.cf[q];
        if (J == I) {
          C.fd[I].x += c.x;
          C.fd[I].y += c.y;
          intra[I].x += c.x;
          intra[I].y += c.y;
        } else if (mark[J] >= row0) {
          C.cf[mark[J]].x += c.x;
          C.cf[mark[J]].y += c.y;
        } else {
          mark[J] = (int64_t)C.col.size();
          C.col.push_back(J);
          C.cf.push_back(c);
        }
      }
    }
    C.ptr.push_back((int32_t)C.col.size());
  }
}

// The transposed couplings of C: the coefficient of I in J's row becomes J's
// in I's row (CSR pT, cT, fT; the diagonal parts are unchanged)
template <typename T>
inline void amg_transpose(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &col,
                          const std::vector<T> &cf, std::vector<int32_t> &pT, std::vector<int32_t> &cT,
                          std::vector<T> &fT) {
  pT.assign(n + 1, 0);
  cT.assign(col.size(), 0);
  fT.assign(col.size(), T{});
  for (size_t k = 0; k < col.size(); k++) pT[col[k] + 1]++;
  for (int i = 0; i < n; i++) pT[i + 1] += pT[i];
  std::vector<int32_t> cur(pT.begin(), pT.end() - 1);
  for (int i = 0; i < n; i++)
    for (int k = ptr[i]; k < ptr[i + 1]; k++) {
      const int q = cur[col[k]]++;
      cT[q] = i;
      fT[q] = cf[k];
    }
}

}  // namespace ufe
