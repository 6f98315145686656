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
  threads give the same coarse operator bit for bit.
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

static bool same(const AmgHostOp &a, const AmgHostOp &b) {
  auto eq = [](const std::vector<AmgPair> &x, const std::vector<AmgPair> &y) {
    if (x.size() != y.size()) return false;
    for (size_t k = 0; k < x.size(); k++)
      if (x[k].x != y[k].x || x[k].y != y[k].y) return false;
    return true;
  };
  return a.n == b.n && a.ptr == b.ptr && a.col == b.col && a.S == b.S && eq(a.cf, b.cf) && eq(a.fd, b.fd);
}

int main() {
  std::mt19937 g(12345);
  for (int rep = 0; rep < 3; rep++) {
    AmgHostOp A = grid(37 + 11 * rep, 29 + 5 * rep, g), C1, C2;
    check_level(A, g, 1, C1);
    check_level(C1, g, 2, C2);
  }
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


def test_multigrid_host_logic(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "amg_host_check.cpp"
    src.write_text(DRIVER.replace('#include <vector>\n#include "amg_host.h"',
                                  '#include <vector>\n#include <algorithm>\n#include "amg_host.h"'))
    exe = tmp_path / "amg_host_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-pthread", "-I", CSRC, "-o", str(exe), str(src)], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
