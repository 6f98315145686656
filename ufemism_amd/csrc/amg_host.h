// Host side of the aggregation multigrid: greedy aggregation, the Galerkin
// product for piecewise-constant prolongation and the transposed couplings.
// Plain C++ (no HIP types) so that the CPU tests compile it on their own.
#pragma once

#include <algorithm>
#include <cstdint>
#include <system_error>
#include <thread>
#include <utility>
#include <vector>

namespace ufe {

// Host threads for a set-up pass over n rows: one below 65,536 rows, else the
// hardware's count (at most 16).
inline int amg_threads(int64_t n) {
  if (n < 65536) return 1;
  const unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 16u) : 1;
}

// body(t, begin, end) on `parts` contiguous ranges of [0, n), one thread each;
// a part whose thread cannot be started runs on the calling thread (no
// exception leaves, since callers sit behind a C interface)
template <typename F>
inline void amg_parts(int n, int parts, F body) {
  if (parts <= 1) {
    body(0, 0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(parts);
  for (int t = 0; t < parts; t++) {
    const int b = (int)((int64_t)n * t / parts), e = (int)((int64_t)n * (t + 1) / parts);
    try {
      th.emplace_back(body, t, b, e);
    } catch (const std::system_error &) {
      body(t, b, e);
    }
  }
  for (auto &x : th) x.join();
}

// Rows built in parts: part t holds the rows of its range one after another,
// rowlen[i] the length of row i.  Joins them into CSR ptr, col, cf.
template <typename T>
inline void amg_join_rows(const std::vector<int32_t> &rowlen, std::vector<std::vector<int32_t>> &pc,
                          std::vector<std::vector<T>> &pf, std::vector<int32_t> &ptr, std::vector<int32_t> &col,
                          std::vector<T> &cf) {
  const int n = (int)rowlen.size();
  ptr.assign(n + 1, 0);
  for (int i = 0; i < n; i++) ptr[i + 1] = ptr[i] + rowlen[i];
  col.clear();
  cf.clear();
  col.reserve(ptr[n]);
  cf.reserve(ptr[n]);
  for (size_t t = 0; t < pc.size(); t++) {
    col.insert(col.end(), pc[t].begin(), pc[t].end());
    cf.insert(cf.end(), pf[t].begin(), pf[t].end());
    std::vector<int32_t>().swap(pc[t]);
    std::vector<T>().swap(pf[t]);
  }
}

struct AmgPair {                // {u-u, v-v} coefficient pair; the layout of a double2
  double x, y;
};

struct AmgHostOp {              // a point operator: off-diagonal pairs, fixed diagonal pairs, u-v coupling
  int n = 0;
  std::vector<int32_t> ptr, col;
  std::vector<AmgPair> cf, fd;
  std::vector<double> S;
};

// A point's value as a weighted sum of fine points (a Neumann row eliminated)
using AmgExpansion = std::vector<std::pair<int32_t, double>>;

// The Neumann rows as weighted sums of fine points.  pid[p] is the fine point
// of row p (-1 when p is not an SOR row).  Edge row erow[i] is the mean of its
// listed neighbours eptr/ecol (SOR rows); corner crow[i] is the mean of its
// listed neighbours cptr/ccol, edge rows among them expanded.  exi[p] becomes
// p's index in exl (-1 for rows that are neither).
inline void amg_neumann_expansions(const std::vector<int32_t> &pid, const std::vector<int32_t> &erow,
                                   const std::vector<int32_t> &eptr, const std::vector<int32_t> &ecol,
                                   const std::vector<int32_t> &crow, const std::vector<int32_t> &cptr,
                                   const std::vector<int32_t> &ccol, std::vector<int32_t> &exi,
                                   std::vector<AmgExpansion> &exl) {
  exi.assign(pid.size(), -1);
  exl.clear();
  for (size_t i = 0; i < erow.size(); i++) {
    const double w = 1.0 / (double)(eptr[i + 1] - eptr[i]);
    AmgExpansion e;
    for (int k = eptr[i]; k < eptr[i + 1]; k++)
      if (pid[ecol[k]] >= 0) e.push_back({pid[ecol[k]], w});
    exi[erow[i]] = (int32_t)exl.size();
    exl.push_back(std::move(e));
  }
  for (size_t i = 0; i < crow.size(); i++) {
    const double w = 1.0 / (double)(cptr[i + 1] - cptr[i]);
    AmgExpansion e;
    for (int k = cptr[i]; k < cptr[i + 1]; k++) {
      const int q = ccol[k];
      if (pid[q] >= 0) e.push_back({pid[q], w});
      else if (exi[q] >= 0)
        for (auto &pr : exl[exi[q]]) e.push_back({pr.first, pr.second * w});
    }
    exi[crow[i]] = (int32_t)exl.size();
    exl.push_back(std::move(e));
  }
}

// The operator on the fine points with the Neumann rows eliminated.  Fine
// point i is row rows[i] of a sliced-ELLPACK matrix: slot s of row p holds
// column col[e] and pair cU[e] at e = cptr[p / sell] + p % sell + s * sell,
// for s < deg[p].  A coupling to a Neumann row is spread over its expansion,
// a coupling that comes back to the row itself goes into the fixed diagonal
// part fd, and S = 3 Sxy.  Rows are built in parts on host threads (parts 0:
// amg_threads); a row's entries keep the order of first appearance, found by a
// search of the row so far (short), so the result does not depend on parts.
inline void amg_fine_operator(const std::vector<int32_t> &rows, int sell, const std::vector<int64_t> &cptr,
                              const std::vector<uint8_t> &deg, const std::vector<int32_t> &col, const AmgPair *cU,
                              const std::vector<double> &Sxy, const std::vector<int32_t> &pid,
                              const std::vector<int32_t> &exi, const std::vector<AmgExpansion> &exl, AmgHostOp &A,
                              int parts = 0) {
  const int n0 = (int)rows.size();
  A.n = n0;
  A.fd.assign(n0, AmgPair{0.0, 0.0});
  A.S.resize(n0);
  if (parts <= 0) parts = amg_threads(n0);
  parts = std::max(1, std::min(parts, std::max(n0, 1)));
  std::vector<int32_t> rowlen(n0);
  std::vector<std::vector<int32_t>> pc(parts);
  std::vector<std::vector<AmgPair>> pf(parts);
  amg_parts(n0, parts, [&](int t, int i0, int i1) {
    std::vector<int32_t> lc;             // local, not pc[t]: the parts' vector headers share cache lines
    std::vector<AmgPair> lf;
    for (int i = i0; i < i1; i++) {
      const int p = rows[i];
      const int64_t base = cptr[p / sell] + (p % sell);
      const size_t row0 = lc.size();
      auto add = [&](int j, double cx, double cy) {
        if (j == i) {
          A.fd[i].x += cx;
          A.fd[i].y += cy;
          return;
        }
        for (size_t k = row0; k < lc.size(); k++)
          if (lc[k] == j) {
            lf[k].x += cx;
            lf[k].y += cy;
            return;
          }
        lc.push_back(j);
        lf.push_back(AmgPair{cx, cy});
      };
      for (int s = 0; s < deg[p]; s++) {
        const int64_t e = base + (int64_t)s * sell;
        const int q = col[e];
        const AmgPair c = cU[e];
        if (pid[q] >= 0) add(pid[q], c.x, c.y);
        else if (exi[q] >= 0)
          for (auto &pr : exl[exi[q]]) add(pr.first, c.x * pr.second, c.y * pr.second);
      }
      rowlen[i] = (int32_t)(lc.size() - row0);
      A.S[i] = 3.0 * Sxy[p];
    }
    pc[t] = std::move(lc);
    pf[t] = std::move(lf);
  });
  amg_join_rows(rowlen, pc, pf, A.ptr, A.col, A.cf);
}

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
// parts: host threads for the coarse rows (0: amg_threads); the result does
// not depend on it.
inline void amg_galerkin(const AmgHostOp &A, const std::vector<int32_t> &agg, int na, AmgHostOp &C,
                         std::vector<int32_t> &mptr, std::vector<int32_t> &mem, std::vector<AmgPair> &intra,
                         int parts = 0) {
  mptr.assign(na + 1, 0);
  for (int i = 0; i < A.n; i++) mptr[agg[i] + 1]++;
  for (int I = 0; I < na; I++) mptr[I + 1] += mptr[I];
  mem.resize(A.n);
  {
    std::vector<int32_t> cur(mptr.begin(), mptr.end() - 1);
    for (int i = 0; i < A.n; i++) mem[cur[agg[i]]++] = i;
  }
  C.n = na;
  C.fd.assign(na, AmgPair{0.0, 0.0});
  C.S.assign(na, 0.0);
  intra.assign(na, AmgPair{0.0, 0.0});
  // coarse rows in parts; within a row the entries keep the order of first
  // appearance and the sums the members' order, whatever the number of parts
  if (parts <= 0) parts = amg_threads(A.n);
  parts = std::max(1, std::min(parts, na));
  std::vector<int32_t> rowlen(na);
  std::vector<std::vector<int32_t>> pc(parts);
  std::vector<std::vector<AmgPair>> pf(parts);
  amg_parts(na, parts, [&](int t, int I0, int I1) {
    std::vector<int32_t> col;            // local, not pc[t]: the parts' vector headers share cache lines
    std::vector<AmgPair> cf;
    std::vector<int32_t> mark(na, -1);   // position in this part's col of J in the current row
    for (int I = I0; I < I1; I++) {
      const int32_t row0 = (int32_t)col.size();
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
            cf[mark[J]].x += c.x;
            cf[mark[J]].y += c.y;
          } else {
            mark[J] = (int32_t)col.size();
            col.push_back(J);
            cf.push_back(c);
          }
        }
      }
      rowlen[I] = (int32_t)col.size() - row0;
    }
    pc[t] = std::move(col);
    pf[t] = std::move(cf);
  });
  amg_join_rows(rowlen, pc, pf, C.ptr, C.col, C.cf);
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

// Blocks of points with their neighbourhoods up to `depth` couplings away,
// for kernels in which each block recomputes its halo rather than wait for the
// other blocks.  Block b's list ext[ext_ptr[b], ext_ptr[b+1]) holds its own
// points, then the points at distance 1, 2, ..., depth (each distance in order
// of discovery); cnt[b (depth+1) + d] counts the points within distance d.
// The rows of the points within distance depth-1 are kept with their columns
// renumbered into the block's list: local row q of block b is
// rptr[rstart[b] + q] .. rptr[rstart[b] + q + 1] in lcol, its couplings in the
// original order, and lsrc gives each local coupling's index in col.
//
// The blocks are whole aggregates of the next level (amg_blocks_by_aggregates):
// block b owns the aggregates gagg[gptr[b], gptr[b+1]), whose members, in the
// order of the member lists, are its own points; aggregate entry g has its
// members at local positions [glo[g], glo[g+1]) (glo has one more entry per
// block, so glo[gptr[b+1] + b] closes block b's last aggregate: the entries of
// block b are glo[gptr[b] + b .. gptr[b+1] + b]).
//
// Given the transposed couplings too (ptrT/colT), the same rows of the
// transposed operator are kept alike in rptrT/lcolT/lsrcT (rstart serves
// both); transposed_ok is false when a transposed row reaches outside its
// block's list (an operator that is not structurally symmetric).
struct AmgHalos {
  int depth = 0, nblocks = 0, max_ext = 0;
  std::vector<int32_t> ext_ptr, ext, cnt, rstart, rptr, lcol;
  std::vector<int64_t> lsrc;
  std::vector<int32_t> gptr, gagg, glo;
  bool transposed_ok = false;
  std::vector<int32_t> rptrT, lcolT;
  std::vector<int64_t> lsrcT;
};

// Aggregates of the next level (nc of them, members mptr/mem; graph cptr/ccol)
// gathered into blocks of at most fb members: breadth-first from the
// lowest-numbered aggregate not yet taken, through aggregates not yet taken,
// adding each one whose members still fit (a first aggregate always fits).
inline void amg_blocks_by_aggregates(int nc, const std::vector<int32_t> &cptr, const std::vector<int32_t> &ccol,
                                     const std::vector<int32_t> &mptr, int fb, std::vector<int32_t> &gptr,
                                     std::vector<int32_t> &gagg) {
  gptr.assign(1, 0);
  gagg.clear();
  std::vector<char> taken(nc, 0);
  std::vector<int32_t> queue;
  int seed = 0;
  while (true) {
    while (seed < nc && taken[seed]) seed++;
    if (seed >= nc) break;
    int members = mptr[seed + 1] - mptr[seed];
    taken[seed] = 1;
    gagg.push_back(seed);
    queue.assign(1, seed);
    for (size_t h = 0; h < queue.size(); h++)
      for (int k = cptr[queue[h]]; k < cptr[queue[h] + 1]; k++) {
        const int J = ccol[k];
        const int m = mptr[J + 1] - mptr[J];
        if (taken[J] || members + m > fb) continue;
        taken[J] = 1;
        members += m;
        gagg.push_back(J);
        queue.push_back(J);
      }
    gptr.push_back((int32_t)gagg.size());
  }
}

inline void amg_block_halos(int n, const std::vector<int32_t> &ptr, const std::vector<int32_t> &col,
                            const std::vector<int32_t> &gptr, const std::vector<int32_t> &gagg,
                            const std::vector<int32_t> &mptr, const std::vector<int32_t> &mem, int depth,
                            AmgHalos &H, const std::vector<int32_t> *ptrT = nullptr,
                            const std::vector<int32_t> *colT = nullptr) {
  H = AmgHalos{};
  H.transposed_ok = ptrT && colT;
  if (H.transposed_ok) H.rptrT.assign(1, 0);
  H.depth = depth;
  H.gptr = gptr;
  H.gagg = gagg;
  H.nblocks = (int)gptr.size() - 1;
  H.ext_ptr.assign(1, 0);
  H.rstart.assign(1, 0);
  H.rptr.assign(1, 0);
  H.cnt.assign((size_t)H.nblocks * (depth + 1), 0);
  std::vector<int32_t> local(n, -1);
  for (int b = 0; b < H.nblocks; b++) {
    const size_t e0 = H.ext.size();
    int m = 0;
    for (int g = gptr[b]; g < gptr[b + 1]; g++) {
      H.glo.push_back(m);
      const int I = gagg[g];
      for (int k = mptr[I]; k < mptr[I + 1]; k++) {
        local[mem[k]] = m++;
        H.ext.push_back(mem[k]);
      }
    }
    H.glo.push_back(m);
    int32_t *cb = &H.cnt[(size_t)b * (depth + 1)];
    cb[0] = m;
    int lo = 0;
    for (int d = 1; d <= depth; d++) {
      const int hi = m;
      for (int q = lo; q < hi; q++) {
        const int p = H.ext[e0 + q];
        for (int k = ptr[p]; k < ptr[p + 1]; k++) {
          const int j = col[k];
          if (local[j] < 0) {
            local[j] = m++;
            H.ext.push_back(j);
          }
        }
      }
      cb[d] = m;
      lo = hi;
    }
    const int rows = depth > 0 ? cb[depth - 1] : 0;
    for (int q = 0; q < rows; q++) {
      const int p = H.ext[e0 + q];
      for (int k = ptr[p]; k < ptr[p + 1]; k++) {
        H.lcol.push_back(local[col[k]]);
        H.lsrc.push_back(k);
      }
      H.rptr.push_back((int32_t)H.lcol.size());
      if (H.transposed_ok) {
        for (int k = (*ptrT)[p]; k < (*ptrT)[p + 1]; k++) {
          const int j = local[(*colT)[k]];
          if (j < 0) H.transposed_ok = false;
          H.lcolT.push_back(j);
          H.lsrcT.push_back(k);
        }
        H.rptrT.push_back((int32_t)H.lcolT.size());
      }
    }
    H.rstart.push_back((int32_t)H.rptr.size() - 1);
    for (size_t q = e0; q < H.ext.size(); q++) local[H.ext[q]] = -1;
    H.ext_ptr.push_back((int32_t)H.ext.size());
    H.max_ext = std::max(H.max_ext, m);
  }
}

}  // namespace ufe
