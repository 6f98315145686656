"""The host side of the aggregation multigrid (ufemism_amd/csrc/amg_host.h),
compiled on its own with g++ and checked on random grid operators with
unequal couplings in the two directions:

* every point lies in exactly one aggregate, every aggregate is connected in
  the operator's graph, and the member lists match the aggregate map;
* the Galerkin product is exact: the coarse operator applied to a coarse
  vector equals the fine operator applied to its prolongation, then summed
  over the members (piecewise-constant P);
* the fixed diagonal parts add up (members' parts plus the couplings inside
  the aggregate) and the u-v couplings are the members' sums;
* the transposed couplings give y.(A x) = x.(A^T y);
* a second level coarsens the first in the same way;
* on a grid large enough for the threaded Galerkin product, 1, 3 and 7 host
  threads give the same coarse operator bit for bit;
* the fine operator with the Neumann rows eliminated (a grid with an edge
  ring and four corners in sliced-ELLPACK storage) equals the full rows
  applied to the expanded vector, with 1 and 5 host threads;
* the blocks of the fused coarse-level kernels are whole aggregates of the
  next level, each aggregate in one block with its members in list order,
  connected and of at most the block size; their halos list every point
  within the halo depth at its exact graph distance (breadth-first search),
  and their local rows equal the original rows with renumbered columns, for
  the operator and, when its pattern is symmetric, for its transpose.
"""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ufemism_amd", "csrc")

DRIVER = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "amg_host.h"
using namespace ufe;

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { std::printf("FAIL: " __VA_ARGS__); std::printf("\n"); fails++; } } while (0)

static AmgHostOp grid(int nx, int ny, std::mt19937 &g) {
  std::uniform_real_distribution<double> u(0.1, 1.0), s(-1.0, 1.0);
  AmgHostOp A;
  A.n = nx * ny;
  A.ptr.assign(1, 0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if ((di == 0 && dj == 0) || (di != 0 && dj != 0 && (i + j) % 3)) continue;   // 4- and some 8-links
          const int ii = i + di, jj = j + dj;
          if (ii < 0 || jj < 0 || ii >= nx || jj >= ny) continue;
          A.col.push_back(jj * nx + ii);
          A.cf.push_back(AmgPair{u(g), u(g)});
        }
      A.ptr.push_back((int32_t)A.col.size());
    }
  for (int p = 0; p < A.n; p++) {
    A.fd.push_back(AmgPair{s(g), s(g)});
    A.S.push_back(s(g));
  }
  return A;
}

static void apply(const AmgHostOp &A, const std::vector<AmgPair> &x, std::vector<AmgPair> &y) {
  y.assign(A.n, AmgPair{0.0, 0.0});
  for (int i = 0; i < A.n; i++) {
    double su = A.fd[i].x * x[i].x + A.S[i] * x[i].y, sv = A.fd[i].y * x[i].y + A.S[i] * x[i].x;
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) {
      su += A.cf[k].x * x[A.col[k]].x;
      sv += A.cf[k].y * x[A.col[k]].y;
    }
    y[i] = AmgPair{su, sv};
  }
}

static void check_level(const AmgHostOp &A, std::mt19937 &g, int level, AmgHostOp &C) {
  int na = 0;
  const std::vector<int32_t> agg = amg_aggregate(A, &na);
  CHECK(na > 0 && na < A.n, "level %d: %d aggregates of %d points", level, na, A.n);
  std::vector<int> size(na, 0);
  for (int i = 0; i < A.n; i++) {
    CHECK(agg[i] >= 0 && agg[i] < na, "level %d: point %d aggregate %d", level, i, agg[i]);
    size[agg[i]]++;
  }
  // each aggregate connected (breadth-first search inside it)
  std::vector<int> seen(A.n, 0);
  int ncomp = 0;
  for (int i = 0; i < A.n; i++) {
    if (seen[i]) continue;
    ncomp++;
    std::vector<int> q{i};
    seen[i] = 1;
    for (size_t h = 0; h < q.size(); h++)
      for (int k = A.ptr[q[h]]; k < A.ptr[q[h] + 1]; k++) {
        const int j = A.col[k];
        if (!seen[j] && agg[j] == agg[i]) { seen[j] = 1; q.push_back(j); }
      }
  }
  CHECK(ncomp == na, "level %d: %d connected pieces for %d aggregates", level, ncomp, na);
  std::vector<int32_t> mptr, mem;
  std::vector<AmgPair> intra;
  amg_galerkin(A, agg, na, C, mptr, mem, intra);
  CHECK(C.n == na && (int)mptr.size() == na + 1 && (int)mem.size() == A.n, "level %d: member list sizes", level);
  for (int I = 0; I < na; I++)
    for (int k = mptr[I]; k < mptr[I + 1]; k++) CHECK(agg[mem[k]] == I, "level %d: member %d of %d", level, mem[k], I);
  // exact Galerkin product on random coarse vectors
  std::uniform_real_distribution<double> s(-1.0, 1.0);
  for (int t = 0; t < 3; t++) {
    std::vector<AmgPair> xc(na), xf(A.n), yf, yc;
    for (auto &v : xc) v = AmgPair{s(g), s(g)};
    for (int i = 0; i < A.n; i++) xf[i] = xc[agg[i]];
    apply(A, xf, yf);
    apply(C, xc, yc);
    double err = 0.0, scale = 0.0;
    std::vector<AmgPair> rf(na, AmgPair{0.0, 0.0});
    for (int i = 0; i < A.n; i++) { rf[agg[i]].x += yf[i].x; rf[agg[i]].y += yf[i].y; }
    for (int I = 0; I < na; I++) {
      err = std::fmax(err, std::fmax(std::fabs(rf[I].x - yc[I].x), std::fabs(rf[I].y - yc[I].y)));
      scale = std::fmax(scale, std::fmax(std::fabs(rf[I].x), std::fabs(rf[I].y)));
    }
    CHECK(err <= 1e-12 * scale, "level %d: Galerkin error %.3e of %.3e", level, err, scale);
  }
  // fixed diagonal parts and u-v couplings
  for (int I = 0; I < na; I++) {
    double fx = intra[I].x, fy = intra[I].y, sS = 0.0;
    for (int k = mptr[I]; k < mptr[I + 1]; k++) { fx += A.fd[mem[k]].x; fy += A.fd[mem[k]].y; sS += A.S[mem[k]]; }
    CHECK(std::fabs(fx - C.fd[I].x) <= 1e-12 * (1 + std::fabs(fx)) && std::fabs(fy - C.fd[I].y) <= 1e-12 * (1 + std::fabs(fy)),
          "level %d: fixed diagonal of %d", level, I);
    CHECK(std::fabs(sS - C.S[I]) <= 1e-12 * (1 + std::fabs(sS)), "level %d: u-v coupling of %d", level, I);
  }
  // transposed couplings: y.(A x) = x.(A^T y) on the off-diagonal part
  std::vector<int32_t> pT, cT;
  std::vector<AmgPair> fT;
  amg_transpose(C.n, C.ptr, C.col, C.cf, pT, cT, fT);
  std::vector<AmgPair> x(C.n), y(C.n);
  for (int i = 0; i < C.n; i++) { x[i] = AmgPair{s(g), s(g)}; y[i] = AmgPair{s(g), s(g)}; }
  double lhs = 0.0, rhs = 0.0;
  for (int i = 0; i < C.n; i++) {
    for (int k = C.ptr[i]; k < C.ptr[i + 1]; k++)
      lhs += y[i].x * C.cf[k].x * x[C.col[k]].x + y[i].y * C.cf[k].y * x[C.col[k]].y;
    for (int k = pT[i]; k < pT[i + 1]; k++)
      rhs += x[i].x * fT[k].x * y[cT[k]].x + x[i].y * fT[k].y * y[cT[k]].y;
  }
  CHECK(std::fabs(lhs - rhs) <= 1e-12 * (1 + std::fabs(lhs)), "level %d: transpose %.6e vs %.6e", level, lhs, rhs);
  std::printf("level %d: %d points -> %d aggregates (largest %d)\n", level, A.n, na,
              *std::max_element(size.begin(), size.end()));
}

// The fine operator of a grid with a Neumann ring: a sliced-ELLPACK matrix (slice
// 4) over all points, SOR rows inside, edge rows the mean of their inner
// neighbours, the four corners the mean of their (edge) neighbours.  For a
// random fine vector x, expanded to every point, the full rows equal the
// eliminated operator's rows (off-diagonal and fixed diagonal parts).
static void check_fine_operator(int nx, int ny, std::mt19937 &g, int parts) {
  std::uniform_real_distribution<double> u(0.1, 1.0), s(-1.0, 1.0);
  const int N = nx * ny, sell = 4;
  auto edge = [&](int p) { int i = p % nx, j = p / nx; return i == 0 || j == 0 || i == nx - 1 || j == ny - 1; };
  auto corner = [&](int p) { int i = p % nx, j = p / nx; return (i == 0 || i == nx - 1) && (j == 0 || j == ny - 1); };
  std::vector<std::vector<int>> nb(N);
  for (int p = 0; p < N; p++) {
    const int i = p % nx, j = p / nx;
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        if ((di == 0 && dj == 0) || (di != 0 && dj != 0 && (i + j) % 3)) continue;
        const int ii = i + di, jj = j + dj;
        if (ii >= 0 && jj >= 0 && ii < nx && jj < ny) nb[p].push_back(jj * nx + ii);
      }
  }
  std::vector<int64_t> cptr(1, 0);
  std::vector<uint8_t> deg(N);
  for (int c = 0; c * sell < N; c++) {
    int w = 0;
    for (int p = c * sell; p < std::min(N, (c + 1) * sell); p++) w = std::max(w, (int)nb[p].size());
    cptr.push_back(cptr.back() + (int64_t)w * sell);
  }
  std::vector<int32_t> col(cptr.back(), 0);
  std::vector<AmgPair> cU(cptr.back(), AmgPair{0.0, 0.0});
  std::vector<double> Sxy(N);
  for (int p = 0; p < N; p++) {
    deg[p] = (uint8_t)nb[p].size();
    for (size_t k = 0; k < nb[p].size(); k++) {
      const int64_t e = cptr[p / sell] + p % sell + (int64_t)k * sell;
      col[e] = nb[p][k];
      cU[e] = AmgPair{u(g), u(g)};
    }
    Sxy[p] = s(g);
  }
  std::vector<int32_t> rows, pid(N, -1), erow, eptr{0}, ecol, crow, cptr2{0}, ccol;
  for (int p = 0; p < N; p++)
    if (!edge(p)) { pid[p] = (int32_t)rows.size(); rows.push_back(p); }
  for (int p = 0; p < N; p++) {
    if (!edge(p) || corner(p)) continue;
    erow.push_back(p);
    for (int q : nb[p]) if (!edge(q)) ecol.push_back(q);
    eptr.push_back((int32_t)ecol.size());
  }
  for (int p = 0; p < N; p++) {
    if (!corner(p)) continue;
    crow.push_back(p);
    for (int q : nb[p]) ccol.push_back(q);
    cptr2.push_back((int32_t)ccol.size());
  }
  std::vector<int32_t> exi;
  std::vector<AmgExpansion> exl;
  amg_neumann_expansions(pid, erow, eptr, ecol, crow, cptr2, ccol, exi, exl);
  AmgHostOp A, A1;
  amg_fine_operator(rows, sell, cptr, deg, col, cU.data(), Sxy, pid, exi, exl, A, parts);
  amg_fine_operator(rows, sell, cptr, deg, col, cU.data(), Sxy, pid, exi, exl, A1, 1);
  CHECK(A.n == (int)rows.size() && A.ptr == A1.ptr && A.col == A1.col, "fine operator with %d parts: structure", parts);
  for (size_t k = 0; k < A.cf.size() && k < A1.cf.size(); k++)
    CHECK(A.cf[k].x == A1.cf[k].x && A.cf[k].y == A1.cf[k].y, "fine operator with %d parts: entry %zu", parts, k);
  int self = 0;
  for (int i = 0; i < A.n; i++) self += A.fd[i].x != 0.0;
  CHECK(self > 0, "no self-couplings came back through the Neumann rows");
  std::vector<AmgPair> x(A.n), X(N, AmgPair{0.0, 0.0});
  for (int i = 0; i < A.n; i++) { x[i] = AmgPair{s(g), s(g)}; X[rows[i]] = x[i]; }
  for (size_t i = 0; i < erow.size(); i++) {
    double a = 0.0, b = 0.0;
    for (int k = eptr[i]; k < eptr[i + 1]; k++) { a += X[ecol[k]].x; b += X[ecol[k]].y; }
    X[erow[i]] = AmgPair{a / (eptr[i + 1] - eptr[i]), b / (eptr[i + 1] - eptr[i])};
  }
  for (size_t i = 0; i < crow.size(); i++) {
    double a = 0.0, b = 0.0;
    for (int k = cptr2[i]; k < cptr2[i + 1]; k++) { a += X[ccol[k]].x; b += X[ccol[k]].y; }
    X[crow[i]] = AmgPair{a / (cptr2[i + 1] - cptr2[i]), b / (cptr2[i + 1] - cptr2[i])};
  }
  double err = 0.0, scale = 0.0;
  for (int i = 0; i < A.n; i++) {
    const int p = rows[i];
    double fu = 0.0, fv = 0.0;
    for (int k = 0; k < deg[p]; k++) {
      const int64_t e = cptr[p / sell] + p % sell + (int64_t)k * sell;
      fu += cU[e].x * X[col[e]].x;
      fv += cU[e].y * X[col[e]].y;
    }
    double eu = A.fd[i].x * x[i].x, ev = A.fd[i].y * x[i].y;
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) { eu += A.cf[k].x * x[A.col[k]].x; ev += A.cf[k].y * x[A.col[k]].y; }
    err = std::fmax(err, std::fmax(std::fabs(fu - eu), std::fabs(fv - ev)));
    scale = std::fmax(scale, std::fmax(std::fabs(fu), std::fabs(fv)));
    CHECK(A.S[i] == 3.0 * Sxy[p], "S of fine point %d", i);
  }
  CHECK(err <= 1e-12 * scale, "fine operator %dx%d: eliminated rows differ by %.3e of %.3e", nx, ny, err, scale);
  std::printf("fine operator %dx%d: %d points, %zu couplings, %d with self-couplings\n", nx, ny, A.n, A.col.size(), self);
}

static bool same(const AmgHostOp &a, const AmgHostOp &b) {
  auto eq = [](const std::vector<AmgPair> &x, const std::vector<AmgPair> &y) {
    if (x.size() != y.size()) return false;
    for (size_t k = 0; k < x.size(); k++)
      if (x[k].x != y[k].x || x[k].y != y[k].y) return false;
    return true;
  };
  return a.n == b.n && a.ptr == b.ptr && a.col == b.col && a.S == b.S && eq(a.cf, b.cf) && eq(a.fd, b.fd);
}

// Block halos over whole aggregates of the next level: every aggregate in
// exactly one block, a block's own points are its aggregates' members in the
// order of the member lists (the local ranges glo), at most fb of them unless
// the block is one aggregate, each block connected in the operator's graph,
// each listed point at exactly the graph distance of its class, every point
// within `depth` listed (breadth-first search from the block), and the local
// rows equal the original rows.
static void check_halos(const AmgHostOp &A, int fb, int depth) {
  int na = 0;
  const std::vector<int32_t> agg = amg_aggregate(A, &na);
  AmgHostOp C;
  std::vector<int32_t> mptr, mem;
  std::vector<AmgPair> intra;
  amg_galerkin(A, agg, na, C, mptr, mem, intra, 1);
  std::vector<int32_t> gptr, gagg;
  amg_blocks_by_aggregates(na, C.ptr, C.col, mptr, fb, gptr, gagg);
  std::vector<int32_t> pT, cT;
  std::vector<AmgPair> fT;
  amg_transpose(A.n, A.ptr, A.col, A.cf, pT, cT, fT);
  AmgHalos H;
  amg_block_halos(A.n, A.ptr, A.col, gptr, gagg, mptr, mem, depth, H, &pT, &cT);
  bool symmetric = true;   // every coupling i -> j has its j -> i
  for (int i = 0; i < A.n && symmetric; i++)
    for (int k = A.ptr[i]; k < A.ptr[i + 1] && symmetric; k++) {
      bool found = false;
      for (int q = A.ptr[A.col[k]]; q < A.ptr[A.col[k] + 1]; q++) found = found || A.col[q] == i;
      symmetric = found;
    }
  CHECK(!symmetric || H.transposed_ok, "symmetric pattern, transposed rows refused");
  CHECK(H.nblocks == (int)gptr.size() - 1 && H.nblocks > 0, "halo blocks %d", H.nblocks);
  std::vector<int> seen_agg(na, 0), owner(A.n, -1);
  for (int g : gagg) seen_agg[g]++;
  for (int I = 0; I < na; I++) CHECK(seen_agg[I] == 1, "aggregate %d in %d blocks", I, seen_agg[I]);
  CHECK((int)H.glo.size() == (int)gagg.size() + H.nblocks, "glo entries %d", (int)H.glo.size());
  int maxe = 0;
  for (int b = 0; b < H.nblocks; b++) {
    const int e0 = H.ext_ptr[b], m = H.ext_ptr[b + 1] - e0;
    const int32_t *cb = &H.cnt[(size_t)b * (depth + 1)];
    maxe = std::max(maxe, m);
    CHECK(cb[depth] == m, "block %d: count %d of %d listed", b, cb[depth], m);
    const int own = cb[0];
    const int nag = gptr[b + 1] - gptr[b];
    CHECK(own >= 1 && (own <= fb || nag == 1), "block %d: %d own points in %d aggregates", b, own, nag);
    for (int g = gptr[b]; g < gptr[b + 1]; g++) {
      const int I = gagg[g], lo = H.glo[g + b], hi = H.glo[g + b + 1];
      CHECK(hi - lo == mptr[I + 1] - mptr[I], "block %d: aggregate %d range", b, I);
      for (int k = 0; k < hi - lo && k < mptr[I + 1] - mptr[I]; k++)
        CHECK(H.ext[e0 + lo + k] == mem[mptr[I] + k], "block %d: aggregate %d member %d", b, I, k);
    }
    CHECK(H.glo[gptr[b] + b] == 0 && H.glo[gptr[b + 1] + b] == own, "block %d: member ranges", b);
    for (int q = 0; q < own; q++) {
      const int i = H.ext[e0 + q];
      CHECK(owner[i] < 0, "point %d in blocks %d and %d", i, owner[i], b);
      owner[i] = b;
    }
    {   // connected: a search inside the block from its first point reaches all of it
      std::vector<int> found(1, H.ext[e0]), in(A.n, 0);
      for (int q = 0; q < own; q++) in[H.ext[e0 + q]] = 1;
      in[H.ext[e0]] = 2;
      for (size_t h = 0; h < found.size(); h++)
        for (int k = A.ptr[found[h]]; k < A.ptr[found[h] + 1]; k++)
          if (in[A.col[k]] == 1) {
            in[A.col[k]] = 2;
            found.push_back(A.col[k]);
          }
      CHECK((int)found.size() == own, "block %d: %d of %d own points connected", b, (int)found.size(), own);
    }
    std::vector<int> dist(A.n, -1), frontier;
    for (int q = 0; q < own; q++) {
      dist[H.ext[e0 + q]] = 0;
      frontier.push_back(H.ext[e0 + q]);
    }
    for (int d = 1; d <= depth; d++) {
      std::vector<int> next;
      for (int p : frontier)
        for (int k = A.ptr[p]; k < A.ptr[p + 1]; k++)
          if (dist[A.col[k]] < 0) {
            dist[A.col[k]] = d;
            next.push_back(A.col[k]);
          }
      frontier = next;
    }
    int within = 0;
    for (int i = 0; i < A.n; i++) within += dist[i] >= 0;
    CHECK(within == m, "block %d: %d points within distance %d, %d listed", b, within, depth, m);
    for (int d = 0, q = 0; d <= depth; d++)
      for (; q < cb[d]; q++) CHECK(dist[H.ext[e0 + q]] == d, "block %d: point %d listed at distance %d", b, q, d);
    const int rows = cb[depth - 1];
    for (int q = 0; q < rows; q++) {
      const int p = H.ext[e0 + q], r = H.rstart[b] + q;
      CHECK(H.rptr[r + 1] - H.rptr[r] == A.ptr[p + 1] - A.ptr[p], "block %d: row %d length", b, q);
      for (int k = H.rptr[r], j = A.ptr[p]; k < H.rptr[r + 1] && j < A.ptr[p + 1]; k++, j++) {
        CHECK(H.lsrc[k] == j, "block %d: row %d coupling order", b, q);
        CHECK(H.lcol[k] >= 0 && H.lcol[k] < m && H.ext[e0 + H.lcol[k]] == A.col[j], "block %d: row %d column", b, q);
      }
    }
    CHECK(H.rstart[b + 1] - H.rstart[b] == rows, "block %d: %d local rows", b, H.rstart[b + 1] - H.rstart[b]);
    for (int q = 0; H.transposed_ok && q < rows; q++) {
      const int p = H.ext[e0 + q], r = H.rstart[b] + q;
      CHECK(H.rptrT[r + 1] - H.rptrT[r] == pT[p + 1] - pT[p], "block %d: transposed row %d length", b, q);
      for (int k = H.rptrT[r], j = pT[p]; k < H.rptrT[r + 1] && j < pT[p + 1]; k++, j++) {
        CHECK(H.lsrcT[k] == j, "block %d: transposed row %d order", b, q);
        CHECK(H.lcolT[k] >= 0 && H.lcolT[k] < m && H.ext[e0 + H.lcolT[k]] == cT[j], "block %d: transposed column", b);
      }
    }
  }
  CHECK(H.max_ext == maxe, "max_ext %d of %d", H.max_ext, maxe);
  for (int i = 0; i < A.n; i++) CHECK(owner[i] >= 0, "point %d in no block", i);
}

// the operator with every coupling's reverse added (a structurally symmetric pattern)
static AmgHostOp symmetrized(const AmgHostOp &A, std::mt19937 &g) {
  std::uniform_real_distribution<double> u(0.1, 1.0);
  std::vector<std::vector<std::pair<int, AmgPair>>> rows(A.n);
  for (int i = 0; i < A.n; i++)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) rows[i].push_back({A.col[k], A.cf[k]});
  for (int i = 0; i < A.n; i++)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) {
      const int j = A.col[k];
      bool found = false;
      for (auto &e : rows[j]) found = found || e.first == i;
      if (!found) rows[j].push_back({i, AmgPair{u(g), u(g)}});
    }
  AmgHostOp S = A;
  S.ptr.assign(1, 0);
  S.col.clear();
  S.cf.clear();
  for (int i = 0; i < A.n; i++) {
    for (auto &e : rows[i]) {
      S.col.push_back(e.first);
      S.cf.push_back(e.second);
    }
    S.ptr.push_back((int32_t)S.col.size());
  }
  return S;
}

int main() {
  std::mt19937 g(12345);
  {
    AmgHostOp A = grid(41, 23, g), C1;
    check_halos(A, 64, 3);
    check_halos(A, 256, 2);
    check_halos(symmetrized(A, g), 64, 3);
    check_level(A, g, 1, C1);
    check_halos(C1, 16, 3);
    check_halos(symmetrized(C1, g), 16, 2);
  }
  for (int rep = 0; rep < 3; rep++) {
    AmgHostOp A = grid(37 + 11 * rep, 29 + 5 * rep, g), C1, C2;
    check_level(A, g, 1, C1);
    check_level(C1, g, 2, C2);
  }
  check_fine_operator(13, 9, g, 1);
  check_fine_operator(301, 257, g, 5);
  // a grid large enough for the threaded Galerkin product: the identities hold
  // and the coarse operator is the same, bit for bit, for 1, 3 and 7 parts
  {
    AmgHostOp A = grid(331, 277, g), C1;
    check_level(A, g, 1, C1);
    int na = 0;
    const std::vector<int32_t> agg = amg_aggregate(A, &na);
    AmgHostOp ref;
    std::vector<int32_t> mptr0, mem0;
    std::vector<AmgPair> intra0;
    amg_galerkin(A, agg, na, ref, mptr0, mem0, intra0, 1);
    for (int parts : {3, 7}) {
      AmgHostOp C;
      std::vector<int32_t> mptr, mem;
      std::vector<AmgPair> intra;
      amg_galerkin(A, agg, na, C, mptr, mem, intra, parts);
      bool ok = same(C, ref) && mptr == mptr0 && mem == mem0 && intra.size() == intra0.size();
      for (size_t k = 0; ok && k < intra.size(); k++) ok = intra[k].x == intra0[k].x && intra[k].y == intra0[k].y;
      CHECK(ok, "Galerkin product with %d parts differs from one part", parts);
    }
  }
  std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
  return fails ? 1 : 0;
}
"""


def _driver(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "amg_host_check.cpp"
    src.write_text(DRIVER.replace('#include <vector>\n#include "amg_host.h"',
                                  '#include <vector>\n#include <algorithm>\n#include "amg_host.h"'))
    return cxx, src


def test_multigrid_host_logic(tmp_path):
    cxx, src = _driver(tmp_path)
    exe = tmp_path / "amg_host_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-pthread", "-I", CSRC, "-o", str(exe), str(src)], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_multigrid_host_logic_under_sanitizers(tmp_path, sanitizer):
    """The same checks built with a sanitizer (host code, g++): no memory
    error, undefined behaviour or data race in the threaded set-up."""
    cxx, src = _driver(tmp_path)
    exe = tmp_path / "amg_host_check_san"
    b = subprocess.run([cxx, "-O1", "-g", "-std=c++17", "-pthread", "-fsanitize=" + sanitizer,
                        "-fno-sanitize-recover=all", "-I", CSRC, "-o", str(exe), str(src)],
                       capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip(f"g++ cannot build with -fsanitize={sanitizer}: {b.stderr[-300:]}")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-2000:] + r.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
