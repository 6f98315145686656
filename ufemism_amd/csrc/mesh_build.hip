// mesh_build.hip -- host-side builders of the secondary mesh data.
//
// These restate the combinatorial set-up routines of the reference, which it
// runs serially on the master rank:
//   make_Ac_mesh + find_Ac_edge_indices  src/mesh_ArakawaC_module.f90:24-285
//   make_combined_AaAc_mesh (topology)   src/mesh_ArakawaC_module.f90:286-430
//   calculate_five_colouring_AaAc        src/mesh_five_colour_module.f90:18-735
// The neighbour functions of the AaAc and Aa meshes are computed on the GPU
// (ufe_dyn.hip).  All arrays are in the reference layout (column-major,
// 1-based indices).
//
// Ac numbering.  The reference scans vi = 1..nV, ci = 1..nC(vi) and numbers a
// connection the first time it meets it, marking the reverse slot as it goes.
// A connection (vi, C(vi,ci)) is therefore new exactly when C(vi,ci) > vi, so
// its Ac index is the rank of (vi, ci) among those pairs in scan order: a
// prefix sum, which is how it is computed here.
//
// Left/right vertices.  With C and iTri counter-clockwise and iTri(vi,k)
// between C(vi,k) and C(vi,k+1), the triangle left of vi->C(vi,ci) is
// (vi, C(vi,ci), C(vi,ci+1)) and the one on the right is
// (vi, C(vi,ci-1), C(vi,ci)); the reference finds the same vertices by
// scanning the triangles of vi.
//
// Colouring.  The reference keeps Q4 and Q5 as arrays with a position table
// and shifts them on every removal, which is quadratic on large meshes.  The
// same ordered sequences are kept here as doubly-linked lists, so every queue
// operation is O(1) while the order -- and therefore the colouring -- is the
// reference's.

#include <cmath>
#include <cstring>
#include <vector>

#include "ufe_internal.h"

namespace ufe {

static int is_boundary_segment(const int32_t *edge, int v1, int v2) {
  // src/mesh_help_functions_module.f90:673-747
  const int e1 = edge[v1 - 1], e2 = edge[v2 - 1];
  if (e1 == 0 || e2 == 0) return 0;
  auto any = [e2](std::initializer_list<int> s) {
    for (int x : s)
      if (x == e2) return 1;
    return 0;
  };
  switch (e1) {
    case 1: return any({8, 1, 2});
    case 2: return any({8, 1, 2, 3, 4});
    case 3: return any({2, 3, 4});
    case 4: return any({2, 3, 4, 5, 6});
    case 5: return any({4, 5, 6});
    case 6: return any({4, 5, 6, 7, 8});
    case 7: return any({6, 7, 8});
    case 8: return any({6, 7, 8, 1, 2});
  }
  return 0;
}

int build_Ac_mesh(ufe_mesh_desc *m, std::string &err) {
  const int nV = m->nV, nCm = m->nC_mem;
  const int64_t ldV = nV;
  auto C = [&](int vi, int ci) { return at(m->C, ldV, vi, ci); };
  auto nC = [&](int vi) { return m->nC[vi - 1]; };
  auto X = [&](int vi) { return at(m->V, ldV, vi, 1); };
  auto Y = [&](int vi) { return at(m->V, ldV, vi, 2); };

  // Ac index of every (vi, ci): prefix sum of "new" connections.
  std::vector<int64_t> first(nV + 1, 0);
  for (int vi = 1; vi <= nV; vi++) {
    int cnt = 0;
    for (int ci = 1; ci <= nC(vi); ci++) cnt += C(vi, ci) > vi;
    first[vi] = first[vi - 1] + cnt;
  }
  const int64_t nAc = first[nV];
  if (nAc != m->nAc) {
    err = "nAc does not match (sum of nC)/2: inconsistent connectivity";
    return UFE_BAD_INPUT;
  }
  std::memset(m->iAci, 0, sizeof(int32_t) * (size_t)nV * nCm);
  for (int vi = 1; vi <= nV; vi++) {
    int64_t k = first[vi - 1];
    for (int ci = 1; ci <= nC(vi); ci++) {
      const int vj = C(vi, ci);
      if (vj > vi) at(m->iAci, ldV, vi, ci) = (int32_t)(++k);
    }
  }
  for (int vi = 1; vi <= nV; vi++) {  // reverse slots
    for (int ci = 1; ci <= nC(vi); ci++) {
      const int vj = C(vi, ci);
      if (vj > vi) continue;
      int found = 0;
      for (int cj = 1; cj <= nC(vj); cj++)
        if (C(vj, cj) == vi) {
          at(m->iAci, ldV, vi, ci) = at(m->iAci, ldV, vj, cj);
          found = 1;
          break;
        }
      if (!found) {
        err = "connectivity is not symmetric";
        return UFE_BAD_INPUT;
      }
    }
  }
  const int64_t ldA = nAc;
  for (int vi = 1; vi <= nV; vi++) {
    const int n = nC(vi);
    const int edge_v = m->edge_index[vi - 1] > 0;
    for (int ci = 1; ci <= n; ci++) {
      const int vj = C(vi, ci);
      if (vj < vi) continue;
      const int aci = at(m->iAci, ldV, vi, ci);
      at(m->VAc, ldA, aci, 1) = (X(vi) + X(vj)) / 2.0;
      at(m->VAc, ldA, aci, 2) = (Y(vi) + Y(vj)) / 2.0;
      double Nx[5], Ny[5];
      if (!is_boundary_segment(m->edge_index, vi, vj)) {
        // left: (vi, C(ci), C(ci+1)); right: (vi, C(ci-1), C(ci))
        int cl = ci + 1, cr = ci - 1;
        if (!edge_v) {
          if (cl > n) cl = 1;
          if (cr < 1) cr = n;
        }
        if (cl > n || cr < 1) {
          err = "interior connection of an edge vertex lacks a neighbouring triangle";
          return UFE_BAD_INPUT;
        }
        const int vl = C(vi, cl), vr = C(vi, cr);
        at(m->Aci, ldA, aci, 1) = vi;
        at(m->Aci, ldA, aci, 2) = vj;
        at(m->Aci, ldA, aci, 3) = vl;
        at(m->Aci, ldA, aci, 4) = vr;
        const double Nxl[5] = {0, Y(vl) - Y(vj), Y(vi) - Y(vl), Y(vj) - Y(vi), 0.0};
        const double Nyl[5] = {0, X(vj) - X(vl), X(vl) - X(vi), X(vi) - X(vj), 0.0};
        const double Nxr[5] = {0, Y(vj) - Y(vr), Y(vr) - Y(vi), 0.0, Y(vi) - Y(vj)};
        const double Nyr[5] = {0, X(vr) - X(vj), X(vi) - X(vr), 0.0, X(vj) - X(vi)};
        const double Nzl = ((X(vj) - X(vi)) * (Y(vl) - Y(vi))) - ((Y(vj) - Y(vi)) * (X(vl) - X(vi)));
        const double Nzr = ((X(vr) - X(vi)) * (Y(vj) - Y(vi))) - ((Y(vr) - Y(vi)) * (X(vj) - X(vi)));
        for (int k = 1; k <= 4; k++) {
          Nx[k] = -((Nxl[k] / Nzl) + (Nxr[k] / Nzr)) / 2.0;
          Ny[k] = -((Nyl[k] / Nzl) + (Nyr[k] / Nzr)) / 2.0;
        }
      } else {
        // a boundary segment is first or last in the fan of an edge vertex
        int vl;
        if (ci == 1) vl = C(vi, 2);
        else if (ci == n) vl = C(vi, n - 1);
        else {
          err = "boundary segment in the middle of a vertex fan";
          return UFE_BAD_INPUT;
        }
        at(m->Aci, ldA, aci, 1) = vi;
        at(m->Aci, ldA, aci, 2) = vj;
        at(m->Aci, ldA, aci, 3) = vl;
        at(m->Aci, ldA, aci, 4) = 1;
        const double Nxl[5] = {0, Y(vl) - Y(vj), Y(vi) - Y(vl), Y(vj) - Y(vi), 0.0};
        const double Nyl[5] = {0, X(vj) - X(vl), X(vl) - X(vi), X(vi) - X(vj), 0.0};
        const double Nzl = ((X(vj) - X(vi)) * (Y(vl) - Y(vi))) - ((Y(vj) - Y(vi)) * (X(vl) - X(vi)));
        for (int k = 1; k <= 4; k++) {
          Nx[k] = -(Nxl[k] / Nzl);
          Ny[k] = -(Nyl[k] / Nzl);
        }
      }
      const double Ux = X(vj) - X(vi), Uy = Y(vj) - Y(vi);
      const double U = std::sqrt(Ux * Ux + Uy * Uy);
      for (int k = 1; k <= 4; k++) {
        at(m->Nx_Ac, ldA, aci, k) = Nx[k];
        at(m->Ny_Ac, ldA, aci, k) = Ny[k];
        at(m->No_Ac, ldA, aci, k) = (Ny[k] * Ux - Nx[k] * Uy) / U;
      }
      m->Np_Ac[aci - 1] = 1.0 / U;
    }
  }
  // find_Ac_edge_indices, src/mesh_ArakawaC_module.f90:242-285
  auto in3 = [](int e, int a, int b, int c) { return e == a || e == b || e == c; };
  for (int aci = 1; aci <= nAc; aci++) {
    const int e1 = m->edge_index[at(m->Aci, ldA, aci, 1) - 1];
    const int e2 = m->edge_index[at(m->Aci, ldA, aci, 2) - 1];
    int r = 0;
    if (in3(e1, 8, 1, 2) && in3(e2, 8, 1, 2)) r = 1;
    else if (in3(e1, 2, 3, 4) && in3(e2, 2, 3, 4)) r = 3;
    else if (in3(e1, 4, 5, 6) && in3(e2, 4, 5, 6)) r = 5;
    else if (in3(e1, 6, 7, 8) && in3(e2, 6, 7, 8)) r = 7;
    m->edge_index_Ac[aci - 1] = r;
  }
  return UFE_OK;
}

int build_AaAc_connectivity(ufe_mesh_desc *m, std::string &err) {
  const int nV = m->nV, nAc = m->nAc, nA = nV + nAc, nCm = m->nC_mem;
  const int64_t ldV = nV, ldA = nAc, ldAA = nA;
  std::memset(m->CAaAc, 0, sizeof(int32_t) * (size_t)nA * nCm);
  // Ac index of connection (a, b), looked up in the fan of a
  auto acof = [&](int a, int b) -> int {
    for (int c = 1; c <= m->nC[a - 1]; c++)
      if (at(m->C, ldV, a, c) == b) return at(m->iAci, ldV, a, c);
    return 0;
  };
  for (int vi = 1; vi <= nV; vi++) {
    at(m->VAaAc, ldAA, vi, 1) = at(m->V, ldV, vi, 1);
    at(m->VAaAc, ldAA, vi, 2) = at(m->V, ldV, vi, 2);
    m->nCAaAc[vi - 1] = m->nC[vi - 1];
    for (int ci = 1; ci <= m->nC[vi - 1]; ci++)
      at(m->CAaAc, ldAA, vi, ci) = at(m->iAci, ldV, vi, ci) + nV;
  }
  for (int aci = 1; aci <= nAc; aci++) {
    const int ai = aci + nV;
    at(m->VAaAc, ldAA, ai, 1) = at(m->VAc, ldA, aci, 1);
    at(m->VAaAc, ldAA, ai, 2) = at(m->VAc, ldA, aci, 2);
    int vi = at(m->Aci, ldA, aci, 1), vj = at(m->Aci, ldA, aci, 2);
    const int e = m->edge_index_Ac[aci - 1];
    if (e > 0) {
      // order the two edge vertices anticlockwise around the Ac point (:351-367)
      const double xi = at(m->V, ldV, vi, 1), xj = at(m->V, ldV, vj, 1);
      const double yi = at(m->V, ldV, vi, 2), yj = at(m->V, ldV, vj, 2);
      const bool sw = (e == 1 && xi > xj) || (e == 3 && yi < yj) || (e == 5 && xi < xj) ||
                      (e == 7 && yi > yj);
      if (sw) std::swap(vi, vj);
      const int vk = at(m->Aci, ldA, aci, 3);
      const int a1 = acof(vk, vi), a2 = acof(vk, vj);
      if (!a1 || !a2) {
        err = "make_combined_AaAc_mesh: Ac point of an edge triangle not found";
        return UFE_BAD_INPUT;
      }
      m->nCAaAc[ai - 1] = 4;
      const int row[4] = {vi, a1 + nV, a2 + nV, vj};
      for (int k = 0; k < 4; k++) at(m->CAaAc, ldAA, ai, k + 1) = row[k];
    } else {
      const int vl = at(m->Aci, ldA, aci, 3), vr = at(m->Aci, ldA, aci, 4);
      const int a1 = acof(vr, vi), a2 = acof(vr, vj), a3 = acof(vl, vj), a4 = acof(vl, vi);
      if (!a1 || !a2 || !a3 || !a4) {
        err = "make_combined_AaAc_mesh: Ac point of a neighbouring triangle not found";
        return UFE_BAD_INPUT;
      }
      m->nCAaAc[ai - 1] = 6;
      const int row[6] = {vi, a1 + nV, a2 + nV, vj, a3 + nV, a4 + nV};
      for (int k = 0; k < 6; k++) at(m->CAaAc, ldAA, ai, k + 1) = row[k];
    }
  }
  return UFE_OK;
}

// ---------------------------------------------------------------------------
// Williams (1985) five-colouring with ordered doubly-linked queues.
// ---------------------------------------------------------------------------
namespace {

struct OrderedSet {                 // ordered sequence with O(1) edits
  std::vector<int32_t> prev, next;  // indices 1..n, 0 = none
  int32_t head = 0, tail = 0, size = 0;
  explicit OrderedSet(int n) : prev(n + 1, 0), next(n + 1, 0) {}
  void push_back(int v) {
    prev[v] = tail;
    next[v] = 0;
    if (tail) next[tail] = v;
    else head = v;
    tail = v;
    size++;
  }
  void push_front(int v) {
    next[v] = head;
    prev[v] = 0;
    if (head) prev[head] = v;
    else tail = v;
    head = v;
    size++;
  }
  void erase(int v) {
    if (prev[v]) next[prev[v]] = next[v];
    else head = next[v];
    if (next[v]) prev[next[v]] = prev[v];
    else tail = prev[v];
    prev[v] = next[v] = 0;
    size--;
  }
};

struct Colourer {
  int n, nCm;
  std::vector<int32_t> deg, L;       // L: row-major [v][k], v 1-based
  std::vector<int8_t> where;         // 0 none, 4 in Q4, 5 in Q5
  OrderedSet Q4, Q5;
  std::vector<int32_t> S_vi, S_L;
  int noofvert;
  int err = 0;
  Colourer(int n_, int nCm_) : n(n_), nCm(nCm_), deg(n_ + 1), L((size_t)(n_ + 1) * nCm_),
                               where(n_ + 1, 0), Q4(n_), Q5(n_) {}
  int32_t *row(int v) { return &L[(size_t)v * nCm]; }

  void add4(int v) {
    if (deg[v] > 4 || where[v]) { err = UFE_COLOURING_FAILED; return; }
    Q4.push_back(v);
    where[v] = 4;
  }
  void add5(int v) {
    if (deg[v] != 5 || where[v]) { err = UFE_COLOURING_FAILED; return; }
    Q5.push_back(v);
    where[v] = 5;
  }
  void rem4(int v) {
    if (where[v] != 4) { err = UFE_COLOURING_FAILED; return; }
    Q4.erase(v);
    where[v] = 0;
  }
  void rem5(int v) {
    if (where[v] != 5) { err = UFE_COLOURING_FAILED; return; }
    Q5.erase(v);
    where[v] = 0;
  }
  void check(int v) {                 // src/mesh_five_colour_module.f90:503-534
    if (deg[v] <= 4) {
      if (where[v] == 5) rem5(v);
      if (where[v] != 4) add4(v);
    } else if (deg[v] == 5) {
      if (where[v] == 4) rem4(v);
      if (where[v] != 5) add5(v);
    } else {
      if (where[v] == 4) { err = UFE_COLOURING_FAILED; return; }
      if (where[v] == 5) rem5(v);
    }
  }
  void del(int v) {                   // src/mesh_five_colour_module.f90:345-404
    int32_t *lv = row(v);
    for (int ci = 0; ci < deg[v]; ci++) {
      const int w = lv[ci];
      int32_t *lw = row(w);
      for (int k = 0; k < deg[w]; k++) {
        if (lw[k] == v) {
          for (int j = k; j < deg[w] - 1; j++) lw[j] = lw[j + 1];
          lw[deg[w] - 1] = 0;
          deg[w]--;
          check(w);
          break;
        }
      }
    }
    S_vi.push_back(v);
    S_L.insert(S_L.end(), lv, lv + nCm);
    if (deg[v] <= 4) rem4(v);
    else if (deg[v] == 5) rem5(v);
    else err = UFE_COLOURING_FAILED;
    noofvert--;
  }
};

}  // namespace

int five_colouring(ufe_mesh_desc *m, std::string &err) {
  const int nV = m->nV, nA = nV + m->nAc, nCm = m->nC_mem;
  const int64_t ldAA = nA;
  Colourer c(nA, nCm);
  c.S_vi.reserve(nA);
  c.S_L.reserve((size_t)nA * nCm);
  for (int v = 1; v <= nA; v++) {
    c.deg[v] = m->nCAaAc[v - 1];
    for (int k = 1; k <= nCm; k++) c.row(v)[k - 1] = at(m->CAaAc, ldAA, v, k);
  }
  c.noofvert = nA;
  for (int v = 1; v <= nA; v++) {     // :91-97
    if (c.deg[v] <= 4) c.add4(v);
    else if (c.deg[v] == 5) c.add5(v);
  }
  long it = 0;
  while (c.noofvert > 5) {            // reduce, :139-237
    if (++it > 2L * nA) { err = "five-colouring: reduce got stuck"; return UFE_COLOURING_FAILED; }
    if (c.Q4.size > 0) {
      c.del(c.Q4.tail);
    } else {
      if (c.Q5.size == 0) { err = "five-colouring: both queues empty"; return UFE_COLOURING_FAILED; }
      const int v = c.Q5.tail;
      const int k = nCm;               // C%nconmax
      bool found = false;
      const int32_t *lv = c.row(v);
      for (int a = 0; a < c.deg[v] && !found; a++) {
        const int x = lv[a];
        for (int b = a + 1; b < c.deg[v]; b++) {
          const int y = lv[b];
          bool adj = false;
          const int32_t *ly = c.row(y);
          for (int q = 0; q < c.deg[y]; q++)
            if (ly[q] == x) { adj = true; break; }
          if (adj) continue;
          if (c.deg[x] < k && c.deg[y] < k) { found = true; break; }
        }
      }
      if (found) {
        err = "five-colouring needs identify(), which the reference aborts on "
              "(src/mesh_five_colour_module.f90:452)";
        return UFE_COLOURING_IDENTIFY;
      }
      c.Q5.erase(v);                   // return v to the rear of Q5 (:222-228)
      c.Q5.push_front(v);
    }
    if (c.err) { err = "five-colouring: queue inconsistency"; return c.err; }
  }
  // colour, :238-316
  std::vector<int32_t> colour(nA + 1, 0);
  if (c.Q4.size < 5) { err = "five-colouring: fewer than five vertices left in Q4"; return UFE_COLOURING_FAILED; }
  {
    int v = c.Q4.head;
    for (int k = 1; k <= 5; k++) { colour[v] = k; v = c.Q4.next[v]; }
  }
  for (int64_t s = (int64_t)c.S_vi.size() - 1; s >= 0; s--) {
    const int v = c.S_vi[s];
    const int32_t *lv = &c.S_L[(size_t)s * nCm];
    int used[6] = {0, 0, 0, 0, 0, 0}, sum = 0;
    for (int k = 0; k < nCm; k++) {
      const int u = lv[k];
      if (u == 0) break;
      if (colour[u]) used[colour[u]] = 1;
    }
    for (int k = 1; k <= 5; k++) sum += used[k];
    if (sum > 4) { err = "five-colouring: no free colour"; return UFE_COLOURING_FAILED; }
    for (int k = 1; k <= 5; k++)
      if (!used[k]) { colour[v] = k; break; }
  }
  for (int v = 1; v <= nA; v++) {     // check_solution, :318-343
    if (!colour[v]) { err = "five-colouring: uncoloured vertex"; return UFE_COLOURING_FAILED; }
    for (int k = 1; k <= m->nCAaAc[v - 1]; k++)
      if (colour[at(m->CAaAc, ldAA, v, k)] == colour[v]) {
        err = "five-colouring: adjacent vertices share a colour";
        return UFE_COLOURING_FAILED;
      }
  }
  std::memset(m->colour_vi, 0, sizeof(int32_t) * (size_t)nA * 5);
  for (int k = 0; k < 5; k++) m->colour_nV[k] = 0;
  for (int v = 1; v <= nA; v++) {
    const int ci = colour[v];
    m->colour[v - 1] = ci;
    m->colour_nV[ci - 1]++;
    at(m->colour_vi, ldAA, m->colour_nV[ci - 1], ci) = v;
  }
  return UFE_OK;
}

}  // namespace ufe
