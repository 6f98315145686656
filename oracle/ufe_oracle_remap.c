/* ufe_oracle_remap.c -- CPU oracle (test infrastructure only) for the mesh-to-
 * mesh remapping of src/mesh_mapping_module.f90 and the geometric centres of
 * src/mesh_help_functions_module.f90.  Loaded only by the tests and the
 * bench's CPU baseline; the library never links it.
 *
 * Conservative weights.  The reference integrates x dy, -xy dx and xy dy along
 * the Voronoi boundary lines of both meshes, each piece assigned to the pair
 * (destination cell, source cell) it separates (:810-954, :1221-2347), and
 * never along the domain boundary (:2376-2379).  Summed per pair, those
 * pieces are the boundary of the overlap of the two cells without its
 * domain-boundary edges.  This restatement clips the two convex cells
 * (Sutherland-Hodgman) and sums the reference's line integrals
 * (line_integral_xdy/mxydx/xydy, src/mesh_help_functions_module.f90:2021-2107)
 * over the edges of the overlap that are not on the domain boundary.  Entries
 * of one destination vertex are sorted by source vertex; weights follow
 * calculate_remapping_weights_from_line_integrals (:2348-2415). */
#include "ufe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GV(vi, k) g->V[(size_t)((k) - 1) * g->nV + (vi) - 1]
#define GI(a, ld, i, k) (a)[(size_t)((k) - 1) * (ld) + (i) - 1]

static double rmax(double a, double b) { return a > b ? a : b; }
static double rmin(double a, double b) { return a < b ? a : b; }

/* :2021-2049 */
static double li_xdy(const double p[2], const double q[2], double tol) {
  if (fabs(p[1] - q[1]) < tol) return 0.0;
  const double dx = q[0] - p[0], dy = q[1] - p[1];
  return p[0] * dy - p[1] * dx + (dx / (2.0 * dy)) * (q[1] * q[1] - p[1] * p[1]);
}
/* :2050-2078 */
static double li_mxydx(const double p[2], const double q[2], double tol) {
  if (fabs(p[0] - q[0]) < tol) return 0.0;
  const double dx = q[0] - p[0], dy = q[1] - p[1];
  return (1.0 / 2.0 * (p[0] * dy / dx - p[1]) * (q[0] * q[0] - p[0] * p[0])) -
         (1.0 / 3.0 * dy / dx * (q[0] * q[0] * q[0] - p[0] * p[0] * p[0]));
}
/* :2079-2107 */
static double li_xydy(const double p[2], const double q[2], double tol) {
  if (fabs(p[1] - q[1]) < tol) return 0.0;
  const double dx = q[0] - p[0], dy = q[1] - p[1];
  return (1.0 / 2.0 * (p[0] - p[1] * dx / dy) * (q[1] * q[1] - p[1] * p[1])) +
         (1.0 / 3.0 * dx / dy * (q[1] * q[1] * q[1] - p[1] * p[1] * p[1]));
}

/* find_Voronoi_cell_geometric_centres, src/mesh_help_functions_module.f90:51-100 */
void om_voronoi_geometric_centres(const om_geom *g, const double *A, double *GC) {
  const int mx = 2 * g->nC_mem + 4;
  for (int vi = 1; vi <= g->nV; vi++) {
    double Vor[2 * (2 * 64 + 4)];
    int out = 0;
    const int n = om_find_Voronoi_cell_vertices(g, vi, Vor, mx, &out);
    double mxy = 0.0, xy = 0.0;
    for (int k = 1; k < n && k < mx; k++) {
      const double *p = Vor + 2 * (k - 1), *q = Vor + 2 * k;
      mxy = mxy + li_mxydx(p, q, g->tol_dist);
      xy = xy + li_xydy(p, q, g->tol_dist);
    }
    if (g->edge_index[vi - 1] > 0 && n > 0) {
      const double v[2] = {GV(vi, 1), GV(vi, 2)};
      const double *last = Vor + 2 * (n - 1), *first = Vor;
      mxy = mxy + li_mxydx(last, v, g->tol_dist);
      xy = xy + li_xydy(last, v, g->tol_dist);
      mxy = mxy + li_mxydx(v, first, g->tol_dist);
      xy = xy + li_xydy(v, first, g->tol_dist);
    }
    GC[vi - 1] = mxy / A[vi - 1];
    GC[g->nV + vi - 1] = xy / A[vi - 1];
  }
}

/* The cell of vi as a closed convex polygon (no repeated closing point, no
 * repeated consecutive points): the clamped Voronoi vertices, edge cells
 * closed through the vertex itself, as in the geometric centres above. */
int om_cell_polygon(const om_geom *g, int vi, double *P, int cap) {
  double Vor[2 * (2 * 64 + 4)];
  int out = 0;
  const int mx = 2 * g->nC_mem + 4;
  int n = om_find_Voronoi_cell_vertices(g, vi, Vor, mx, &out);
  if (n > mx) n = mx;
  int m = 0;
  for (int k = 0; k < n; k++) {
    if (m > 0 && Vor[2 * k] == P[2 * (m - 1)] && Vor[2 * k + 1] == P[2 * (m - 1) + 1]) continue;
    if (m >= cap) return -1;
    P[2 * m] = Vor[2 * k];
    P[2 * m + 1] = Vor[2 * k + 1];
    m++;
  }
  if (g->edge_index[vi - 1] > 0) {
    if (m >= cap) return -1;
    P[2 * m] = GV(vi, 1);
    P[2 * m + 1] = GV(vi, 2);
    m++;
  }
  while (m > 1 && P[0] == P[2 * (m - 1)] && P[1] == P[2 * (m - 1) + 1]) m--;
  /* counter-clockwise, as the line integrals and the clipping assume */
  double a = 0.0;
  for (int i = 0; i < m; i++) {
    const int j = (i + 1) % m;
    a = a + (P[2 * i] * P[2 * j + 1] - P[2 * j] * P[2 * i + 1]);
  }
  if (a < 0.0)
    for (int i = 0, j = m - 1; i < j; i++, j--) {
      const double x = P[2 * i], y = P[2 * i + 1];
      P[2 * i] = P[2 * j];
      P[2 * i + 1] = P[2 * j + 1];
      P[2 * j] = x;
      P[2 * j + 1] = y;
    }
  return m;
}

/* Sutherland-Hodgman: the part of polygon S (n points) inside the convex
 * counter-clockwise polygon C (k points); returns the number of points. */
int om_clip_convex(const double *S, int n, const double *C, int k, double *out, int cap) {
  double buf[2][2 * 160];
  if (n > 160 || cap > 160) return -1;
  memcpy(buf[0], S, sizeof(double) * 2 * n);
  int cur = 0, m = n;
  for (int e = 0; e < k && m > 0; e++) {
    const double ax = C[2 * e], ay = C[2 * e + 1];
    const double bx = C[2 * ((e + 1) % k)], by = C[2 * ((e + 1) % k) + 1];
    const double *in = buf[cur];
    double *o = buf[1 - cur];
    int r = 0;
    for (int i = 0; i < m; i++) {
      const double px = in[2 * i], py = in[2 * i + 1];
      const double qx = in[2 * ((i + 1) % m)], qy = in[2 * ((i + 1) % m) + 1];
      const double sp = (bx - ax) * (py - ay) - (by - ay) * (px - ax);
      const double sq = (bx - ax) * (qy - ay) - (by - ay) * (qx - ax);
      if (sp >= 0.0) {
        if (r >= 160) return -1;
        o[2 * r] = px;
        o[2 * r + 1] = py;
        r++;
      }
      if ((sp >= 0.0) != (sq >= 0.0)) {
        const double t = sp / (sp - sq);
        if (r >= 160) return -1;
        o[2 * r] = px + t * (qx - px);
        o[2 * r + 1] = py + t * (qy - py);
        r++;
      }
    }
    cur = 1 - cur;
    m = r;
  }
  if (m > cap) return -1;
  memcpy(out, buf[cur], sizeof(double) * 2 * m);
  return m;
}

static double shoelace(const double *P, int n) {
  double a = 0.0;
  for (int i = 0; i < n; i++) {
    const int j = (i + 1) % n;
    a = a + (P[2 * i] * P[2 * j + 1] - P[2 * j] * P[2 * i + 1]);
  }
  return a / 2.0;
}

static int on_domain_boundary(const om_geom *g, const double p[2], const double q[2]) {
  const double t = g->tol_dist;
  if (fabs(p[0] - g->xmin) < t && fabs(q[0] - g->xmin) < t) return 1;
  if (fabs(p[0] - g->xmax) < t && fabs(q[0] - g->xmax) < t) return 1;
  if (fabs(p[1] - g->ymin) < t && fabs(q[1] - g->ymin) < t) return 1;
  if (fabs(p[1] - g->ymax) < t && fabs(q[1] - g->ymax) < t) return 1;
  return 0;
}

/* find_containing_vertex, src/mesh_help_functions_module.f90:1742-1789: the
 * greedy walk to the nearest vertex from vi (1-based) */
int om_find_containing_vertex(const om_geom *g, const double p[2], int vi) {
  int vi_prev = vi;
  for (int ncycle = 0; ncycle < g->nV; ncycle++) {
    const double dx0 = GV(vi, 1) - p[0], dy0 = GV(vi, 2) - p[1];
    const double d = sqrt(dx0 * dx0 + dy0 * dy0);
    double dcmin = d + 10.0;
    int vcmin = 0;
    for (int ci = 1; ci <= g->nC[vi - 1]; ci++) {
      const int vc = GI(g->C, g->nV, vi, ci);
      if (vc == vi_prev) continue;
      const double dx = GV(vc, 1) - p[0], dy = GV(vc, 2) - p[1];
      const double dc = sqrt(dx * dx + dy * dy);
      if (dc < dcmin) {
        dcmin = dc;
        vcmin = vc;
      }
    }
    if (dcmin < d) {
      vi_prev = vi;
      vi = vcmin;
    } else {
      break;
    }
  }
  return vi;
}

/* The conservative weights from src to dst.  ptr (nV_dst+1, 0-based offsets)
 * is always filled; the entries only when vi != NULL (up to max_total).
 * Returns the number of entries, or -1 on a polygon overflow. */
int om_remap_conservative(const om_geom *src, const double *A_src, const double *GC_src, const om_geom *dst,
                          const double *A_dst, int *ptr, int *vi_out, double *w0, double *w1x, double *w1y,
                          int max_total) {
  const int cap = 2 * 64 + 8;
  int *stamp = (int *)calloc((size_t)src->nV + 1, sizeof(int));
  int *queue = (int *)malloc(sizeof(int) * (size_t)src->nV);
  int total = 0, vs = 1;
  ptr[0] = 0;
  for (int vd = 1; vd <= dst->nV; vd++) {
    double Pd[2 * (2 * 64 + 8)];
    const int nd = om_cell_polygon(dst, vd, Pd, cap);
    if (nd < 0) { total = -1; break; }
    const double p[2] = {dst->V[vd - 1], dst->V[dst->nV + vd - 1]};
    vs = om_find_containing_vertex(src, p, vs);
    int head = 0, tail = 0, count = 0;
    int list[4096];
    double L0[4096], L1[4096], L2[4096];
    queue[tail++] = vs;
    stamp[vs] = vd;
    while (head < tail) {
      const int s = queue[head++];
      double Ps[2 * (2 * 64 + 8)], O[2 * 160];
      const int ns = om_cell_polygon(src, s, Ps, cap);
      if (ns < 0) { free(stamp); free(queue); return -1; }
      const int no = om_clip_convex(Ps, ns, Pd, nd, O, 160);
      if (no < 0) { free(stamp); free(queue); return -1; }
      if (no < 3 || !(shoelace(O, no) > 1e-12 * A_dst[vd - 1])) {
        if (s != vs) continue;
        /* the start cell is searched from even without an overlap */
        for (int ci = 1; ci <= src->nC[s - 1]; ci++) {
          const int vc = GI(src->C, src->nV, s, ci);
          if (stamp[vc] == vd) continue;
          stamp[vc] = vd;
          queue[tail++] = vc;
        }
        continue;
      }
      double a = 0.0, bx = 0.0, by = 0.0;
      for (int i = 0; i < no; i++) {
        const double *P0 = O + 2 * i, *P1 = O + 2 * ((i + 1) % no);
        if (on_domain_boundary(dst, P0, P1)) continue;
        a = a + li_xdy(P0, P1, dst->tol_dist);
        bx = bx + li_mxydx(P0, P1, dst->tol_dist);
        by = by + li_xydy(P0, P1, dst->tol_dist);
      }
      if (count >= 4096) { free(stamp); free(queue); return -1; }
      list[count] = s;
      L0[count] = a;
      L1[count] = bx;
      L2[count] = by;
      count++;
      for (int ci = 1; ci <= src->nC[s - 1]; ci++) {
        const int vc = GI(src->C, src->nV, s, ci);
        if (stamp[vc] == vd) continue;
        stamp[vc] = vd;
        queue[tail++] = vc;
      }
    }
    /* entries sorted by source vertex */
    for (int i = 1; i < count; i++)
      for (int j = i; j > 0 && list[j - 1] > list[j]; j--) {
        int t = list[j]; list[j] = list[j - 1]; list[j - 1] = t;
        double u = L0[j]; L0[j] = L0[j - 1]; L0[j - 1] = u;
        u = L1[j]; L1[j] = L1[j - 1]; L1[j - 1] = u;
        u = L2[j]; L2[j] = L2[j - 1]; L2[j - 1] = u;
      }
    if (vi_out) {
      for (int i = 0; i < count && total + i < max_total; i++) {
        const int s = list[i], e = total + i;
        vi_out[e] = s;
        if (dst->edge_index[vd - 1] > 0) {   /* :2376-2387 */
          w0[e] = w1x[e] = w1y[e] = 0.0;
        } else {
          w0[e] = (L0[i] / A_dst[vd - 1]);
          w1x[e] = (L1[i] / A_dst[vd - 1]) - (GC_src[s - 1] * w0[e]);
          w1y[e] = (L2[i] / A_dst[vd - 1]) - (GC_src[src->nV + s - 1] * w0[e]);
        }
      }
    }
    total += count;
    ptr[vd] = total;
  }
  free(stamp);
  free(queue);
  return total;
}

/* is_in_triangle, src/mesh_help_functions_module.f90:647-672 */
static int in_tri(const double a[2], const double b[2], const double c[2], const double p[2]) {
  const double tol = 1E-8;
  const double as_x = p[0] - a[0], as_y = p[1] - a[1];
  const double s1 = ((b[0] - a[0]) * as_y - (b[1] - a[1]) * as_x);
  const double s2 = ((c[0] - a[0]) * as_y - (c[1] - a[1]) * as_x);
  const double s3 = ((c[0] - b[0]) * (p[1] - b[1]) - (c[1] - b[1]) * (p[0] - b[0]));
  return s1 > -tol && s2 < tol && s3 > -tol;
}

/* lies_on_line_segment, src/mesh_help_functions_module.f90:1828-1867 */
static int on_segment(const double a[2], const double b[2], const double c[2], double tol) {
  const double dx = b[0] - a[0], dy = b[1] - a[1], ex = c[0] - a[0], ey = c[1] - a[1];
  const double nd = sqrt(dx * dx + dy * dy);
  const double nx = dx / nd, ny = dy / nd;
  const double pr = ex * nx + ey * ny;
  const double epx = pr * nx, epy = pr * ny;
  const double ox = ex - epx, oy = ey - epy;
  if (sqrt(ox * ox + oy * oy) > tol) return 0;
  const double ne = sqrt(epx * epx + epy * epy);
  if ((ex * dx + ey * dy) > 0.0) {
    if (ne > nd) return 0;
  } else if (ne > 0.0) {
    return 0;
  }
  return 1;
}

static void tri_pts(const om_geom *g, int t, double a[2], double b[2], double c[2]) {
  const int ia = GI(g->Tri, g->nTri, t, 1), ib = GI(g->Tri, g->nTri, t, 2), ic = GI(g->Tri, g->nTri, t, 3);
  a[0] = GV(ia, 1); a[1] = GV(ia, 2);
  b[0] = GV(ib, 1); b[1] = GV(ib, 2);
  c[0] = GV(ic, 1); c[1] = GV(ic, 2);
}

/* find_containing_triangle, src/mesh_help_functions_module.f90:1605-1741:
 * greedy walk over triangle neighbours TriC (nTri,3) by centroid distance,
 * then a breadth-first flood; returns the triangle (1-based) or 0 */
int om_find_containing_triangle(const om_geom *g, const int *TriC, const double p[2], int t) {
  double a[2], b[2], c[2];
  tri_pts(g, t, a, b, c);
  if (in_tri(a, b, c, p)) return t;
  int t_prev = t;
  for (int ncycle = 0; ncycle < g->nTri; ncycle++) {
    tri_pts(g, t, a, b, c);
    double gx = (a[0] + b[0] + c[0]) / 3.0, gy = (a[1] + b[1] + c[1]) / 3.0;
    const double d = sqrt((gx - p[0]) * (gx - p[0]) + (gy - p[1]) * (gy - p[1]));
    double dcmin = d + 10.0;
    int tcmin = 0;
    for (int n = 1; n <= 3; n++) {
      const int tc = GI(TriC, g->nTri, t, n);
      if (tc == 0 || tc == t_prev) continue;
      tri_pts(g, tc, a, b, c);
      gx = (a[0] + b[0] + c[0]) / 3.0;
      gy = (a[1] + b[1] + c[1]) / 3.0;
      const double dc = sqrt((gx - p[0]) * (gx - p[0]) + (gy - p[1]) * (gy - p[1]));
      if (dc < dcmin) {
        dcmin = dc;
        tcmin = tc;
      }
    }
    if (dcmin < d) {
      t_prev = t;
      t = tcmin;
    } else {
      break;
    }
  }
  tri_pts(g, t, a, b, c);
  if (in_tri(a, b, c, p)) return t;
  if (on_segment(a, b, p, g->tol_dist) || on_segment(b, c, p, g->tol_dist) || on_segment(c, a, p, g->tol_dist))
    return t;
  char *seen = (char *)calloc((size_t)g->nTri + 1, 1);
  int *s1 = (int *)malloc(sizeof(int) * (size_t)g->nTri), *s2 = (int *)malloc(sizeof(int) * (size_t)g->nTri);
  int n1 = 0, found = 0;
  seen[t] = 1;
  for (int n = 1; n <= 3; n++)
    if (GI(TriC, g->nTri, t, n) > 0) s1[n1++] = GI(TriC, g->nTri, t, n);
  while (!found && n1 > 0) {
    int n2 = 0;
    for (int i = 0; i < n1 && !found; i++) {
      const int ti = s1[i];
      tri_pts(g, ti, a, b, c);
      if (in_tri(a, b, c, p)) {
        found = ti;
        break;
      }
      for (int n = 1; n <= 3; n++) {
        const int tin = GI(TriC, g->nTri, ti, n);
        if (tin == 0 || seen[tin]) continue;
        seen[tin] = 1;
        s2[n2++] = tin;
      }
    }
    int *tmp = s1; s1 = s2; s2 = tmp;
    n1 = n2;
  }
  free(seen); free(s1); free(s2);
  return found;
}

static double tri_area(const double a[2], const double b[2], const double c[2]) {
  /* find_triangle_area, :2008-2018 */
  return fabs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])) / 2.0;
}

/* create_remapping_arrays_trilin (src/mesh_mapping_module.f90:718-771) and
 * create_remapping_arrays_nearest_neighbour (:772-809); the walks start where
 * the previous destination vertex ended, as there.  Returns the number of
 * destination vertices without a containing triangle. */
int om_remap_trilin_nearest(const om_geom *src, const int *TriC, const om_geom *dst, int *tri_vi, double *tri_w,
                            int *nn_vi) {
  int t = 1, v = 1, missing = 0;
  for (int vd = 1; vd <= dst->nV; vd++) {
    const double p[2] = {dst->V[vd - 1], dst->V[dst->nV + vd - 1]};
    const int tf = om_find_containing_triangle(src, TriC, p, t);
    if (tf == 0) {
      missing++;
    } else {
      t = tf;
      double a[2], b[2], c[2];
      tri_pts(src, t, a, b, c);
      const double At = tri_area(a, b, c), Aa = tri_area(b, c, p), Ab = tri_area(c, a, p), Ac = tri_area(a, b, p);
      for (int k = 0; k < 3; k++) tri_vi[(size_t)k * dst->nV + vd - 1] = GI(src->Tri, src->nTri, t, k + 1);
      tri_w[vd - 1] = Aa / At;
      tri_w[dst->nV + vd - 1] = Ab / At;
      tri_w[2 * (size_t)dst->nV + vd - 1] = Ac / At;
    }
    v = om_find_containing_vertex(src, p, v);
    nn_vi[vd - 1] = v;
  }
  return missing;
}
