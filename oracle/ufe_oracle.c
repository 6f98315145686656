/*
 * ufe_oracle.c -- CPU restatement of the UFEMISM ice-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY (what the tests compare the HIP library with,
 * and the timed CPU baseline; see ufe_oracle.h).  Each routine restates the reference routine cited above
 * it, in the same array layout (column-major, 1-based indices), with the
 * same loop order and the same order of floating-point operations, so that
 * the per-row arithmetic matches the reference as written.  Compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Parallelism: the reference runs each loop over an MPI rank's index range
 * and barriers after it.  Loops whose iterations are independent (all of
 * them except the Picard RN sums) run here under OpenMP; the RN sums stay
 * sequential over ai = 1..nVAaAc, which is the reference's result with one
 * rank.
 */
#include "ufe_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* physical constants, src/parameters_module.f90:9-20 */
static const double pi_c = 3.141592653589793;
static const double grav = 9.81;
static const double n_flow = 3.0;
static const double ice_density = 910.0;
static const double seawater_density = 1028.0;

/* 1-based, column-major element access */
#define E1(a, i) ((a)[(i) - 1])
#define E2(a, ld, i, j) ((a)[((size_t)((j) - 1)) * (size_t)(ld) + (size_t)((i) - 1)])

/* mesh accessors */
#define nV_ (m->d.nV)
#define nAc_ (m->d.nAc)
#define nVAaAc_ (m->d.nVAaAc)
#define nCm_ (m->d.nC_mem)
#define MV(vi, k) E2(m->f[OM_V], nV_, vi, k)
#define MnC(vi) E1(m->i[OM_nC], vi)
#define MC(vi, ci) E2(m->i[OM_C], nV_, vi, ci)
#define MniTri(vi) E1(m->i[OM_niTri], vi)
#define MiTri(vi, k) E2(m->i[OM_iTri], nV_, vi, k)
#define Medge(vi) E1(m->i[OM_edge_index], vi)
#define MTri(ti, k) E2(m->i[OM_Tri], m->d.nTri, ti, k)
#define MiAci(vi, ci) E2(m->i[OM_iAci], nV_, vi, ci)
#define MAci(aci, k) E2(m->i[OM_Aci], nAc_, aci, k)
#define MedgeAc(aci) E1(m->i[OM_edge_index_Ac], aci)
#define MnCA(ai) E1(m->i[OM_nCAaAc], ai)
#define MCA(ai, ci) E2(m->i[OM_CAaAc], nVAaAc_, ai, ci)
#define MA(vi) E1(m->f[OM_A], vi)
#define MCw(vi, ci) E2(m->f[OM_Cw], nV_, vi, ci)
#define MNx(vi, ci) E2(m->f[OM_Nx], nV_, vi, ci)
#define MNy(vi, ci) E2(m->f[OM_Ny], nV_, vi, ci)
#define MNxx(vi, ci) E2(m->f[OM_Nxx], nV_, vi, ci)
#define MNxy(vi, ci) E2(m->f[OM_Nxy], nV_, vi, ci)
#define MNyy(vi, ci) E2(m->f[OM_Nyy], nV_, vi, ci)
#define MNxAc(a, k) E2(m->f[OM_Nx_Ac], nAc_, a, k)
#define MNyAc(a, k) E2(m->f[OM_Ny_Ac], nAc_, a, k)
#define MNoAc(a, k) E2(m->f[OM_No_Ac], nAc_, a, k)
#define MNpAc(a) E1(m->f[OM_Np_Ac], a)
#define MNxA(ai, ci) E2(m->f[OM_Nx_AaAc], nVAaAc_, ai, ci)
#define MNyA(ai, ci) E2(m->f[OM_Ny_AaAc], nVAaAc_, ai, ci)
#define MNxxA(ai, ci) E2(m->f[OM_Nxx_AaAc], nVAaAc_, ai, ci)
#define MNxyA(ai, ci) E2(m->f[OM_Nxy_AaAc], nVAaAc_, ai, ci)
#define MNyyA(ai, ci) E2(m->f[OM_Nyy_AaAc], nVAaAc_, ai, ci)

/* ice accessors */
#define IF(id) (ice->f[id])
#define II(id) (ice->i[id])

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* field names (bound by name from Python)                                   */
/* ------------------------------------------------------------------------ */

static const char *mesh_d_names[] = {
  "V", "A", "Cw", "Nx", "Ny", "Nxx", "Nxy", "Nyy", "VAc", "Nx_Ac", "Ny_Ac", "No_Ac",
  "Np_Ac", "VAaAc", "Nx_AaAc", "Ny_AaAc", "Nxx_AaAc", "Nxy_AaAc", "Nyy_AaAc"};
static const char *mesh_i_names[] = {
  "nC", "C", "niTri", "iTri", "edge_index", "Tri", "iAci", "Aci", "edge_index_Ac",
  "nCAaAc", "CAaAc", "colour", "colour_vi"};
static const char *ice_d_names[] = {
  "Hi", "Hb", "SL", "Hs", "dHb_dt", "dHi_dt", "dHs_dt", "Hi_prev",
  "A_flow_mean", "dHi_dx", "dHi_dy", "dHs_dx", "dHs_dy",
  "dHs_dx_shelf", "dHs_dy_shelf", "U_SIA", "V_SIA", "D_SIA", "U_SSA", "V_SSA",
  "A_flow", "Ti", "Ti_pmp", "Cpi", "Ki", "U_3D", "V_3D", "W_3D", "dVi_in",
  "Hi_Ac", "Hb_Ac", "SL_Ac", "Hs_Ac",
  "dHi_dx_Ac", "dHi_dy_Ac", "dHi_dp_Ac", "dHi_do_Ac",
  "dHb_dx_Ac", "dHb_dy_Ac", "dHb_dp_Ac", "dHb_do_Ac",
  "dHs_dx_Ac", "dHs_dy_Ac", "dHs_dp_Ac", "dHs_do_Ac",
  "dSL_dx_Ac", "dSL_dy_Ac", "dSL_dp_Ac", "dSL_do_Ac",
  "dHs_dx_shelf_Ac", "dHs_dy_shelf_Ac", "A_flow_mean_Ac",
  "D_SIA_Ac", "Ux_SIA_Ac", "Uy_SIA_Ac", "Up_SIA_Ac", "Uo_SIA_Ac",
  "Ux_SSA_Ac", "Uy_SSA_Ac", "Up_SSA_Ac", "Uo_SSA_Ac",
  "Qabs_GL_Ac", "Qp_GL_Ac",
  "A_flow_Ac", "D_SIA_3D_Ac", "Ti_Ac",
  "Hi_AaAc", "Hb_AaAc", "SL_AaAc", "dHs_dx_shelf_AaAc", "dHs_dy_shelf_AaAc",
  "A_flow_mean_AaAc", "U_SSA_AaAc", "V_SSA_AaAc",
  "dU_SSA_dx_AaAc", "dU_SSA_dy_AaAc", "dV_SSA_dx_AaAc", "dV_SSA_dy_AaAc",
  "eta_AaAc", "N_AaAc", "N_AaAc_prev", "S_AaAc", "phi_fric_AaAc", "tau_c_AaAc",
  "RHSx_AaAc", "RHSy_AaAc", "eu_i_AaAc", "ev_i_AaAc", "LHSx_AaAc", "LHSy_AaAc",
  "resU_AaAc", "resV_AaAc", "tau_c_override_AaAc"};
static const char *ice_i_names[] = {
  "mask_land", "mask_ocean", "mask_lake", "mask_ice", "mask_sheet",
  "mask_shelf", "mask_coast", "mask_margin", "mask_gl", "mask_cf", "mask",
  "mask_land_Ac", "mask_ocean_Ac", "mask_lake_Ac", "mask_ice_Ac", "mask_sheet_Ac",
  "mask_shelf_Ac", "mask_coast_Ac", "mask_margin_Ac", "mask_gl_Ac", "mask_cf_Ac",
  "mask_Ac"};

const char *om_mesh_field_name_d(int id) {
  return (id >= 0 && id < OM_MESH_D_COUNT) ? mesh_d_names[id] : NULL;
}
const char *om_mesh_field_name_i(int id) {
  return (id >= 0 && id < OM_MESH_I_COUNT) ? mesh_i_names[id] : NULL;
}
const char *om_ice_field_name_d(int id) {
  return (id >= 0 && id < OM_ICE_D_COUNT) ? ice_d_names[id] : NULL;
}
const char *om_ice_field_name_i(int id) {
  return (id >= 0 && id < OM_ICE_I_COUNT) ? ice_i_names[id] : NULL;
}

int om_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Neighbour functions, src/mesh_derivatives_module.f90:76-313               */
/* ------------------------------------------------------------------------ */

#define NMAX 64

void om_get_neighbour_functions_vertex_gr(const double *V_vi, int n, const double *V_vc, int ld,
                                          int is_edge, double *Nx, double *Ny, double *Nxx,
                                          double *Nxy, double *Nyy) {
  /* Nx..Nyy are dense local rows of length n+1 (1-based slot k at [k-1]). */
  double NxTri[NMAX][4], NyTri[NMAX][4], NxSub[NMAX][4], NySub[NMAX][4];
  int nTri, nSub, ti, si, ci;
  memset(NxTri, 0, sizeof(NxTri));
  memset(NyTri, 0, sizeof(NyTri));
  memset(NxSub, 0, sizeof(NxSub));
  memset(NySub, 0, sizeof(NySub));
  for (ci = 0; ci <= n; ci++) Nx[ci] = Ny[ci] = Nxx[ci] = Nxy[ci] = Nyy[ci] = 0.0;
#define VC(c, k) V_vc[((size_t)(k) - 1) * (size_t)ld + (size_t)(c) - 1]
  if (is_edge) {
    nTri = n - 1;
    nSub = n - 2;
  } else {
    nTri = n;
    nSub = n;
  }
  const double xi = V_vi[0], yi = V_vi[1];
  for (ti = 1; ti <= nTri; ti++) {             /* Eq. A1, A3 (:116-137) */
    int tip1s = ti + 1;
    if (tip1s > n) tip1s -= n;
    double xt = VC(ti, 1), yt = VC(ti, 2), xtp1s = VC(tip1s, 1), ytp1s = VC(tip1s, 2);
    double nzt = ((xt - xi) * (ytp1s - yi)) - ((yt - yi) * (xtp1s - xi));
    NxTri[ti][1] = (yt - ytp1s) / nzt;
    NxTri[ti][2] = (ytp1s - yi) / nzt;
    NxTri[ti][3] = (yi - yt) / nzt;
    NyTri[ti][1] = (xtp1s - xt) / nzt;
    NyTri[ti][2] = (xi - xtp1s) / nzt;
    NyTri[ti][3] = (xt - xi) / nzt;
  }
  for (si = 1; si <= nSub; si++) {             /* Eq. A11, A13 (:139-165) */
    int sip1s = si + 1;
    if (sip1s > n) sip1s -= n;
    int sip2s = sip1s + 1;
    if (sip2s > n) sip2s -= n;
    double xs = VC(si, 1), ys = VC(si, 2);
    double xsp1s = VC(sip1s, 1), ysp1s = VC(sip1s, 2);
    double xsp2s = VC(sip2s, 1), ysp2s = VC(sip2s, 2);
    double third = 1.0 / 3.0;
    double nzs = ((third * ((xs + xsp1s) - 2.0 * xi)) * ((ysp1s + ysp2s) - 2.0 * yi)) -
                 ((third * ((ys + ysp1s) - 2.0 * yi)) * ((xsp1s + xsp2s) - 2.0 * xi));
    NxSub[si][1] = (ys - ysp2s) / nzs;
    NxSub[si][2] = ((ysp1s + ysp2s) - 2.0 * yi) / nzs;
    NxSub[si][3] = ((2.0 * yi - ys) - ysp1s) / nzs;
    NySub[si][1] = (xsp2s - xs) / nzs;
    NySub[si][2] = ((2.0 * xi - xsp1s) - xsp2s) / nzs;
    NySub[si][3] = ((xs + xsp1s) - 2.0 * xi) / nzs;
  }
#undef VC
  /* first order (:170-209) */
  if (!is_edge) {
    double sx = 0.0, sy = 0.0;
    for (ti = 1; ti <= n; ti++) { sx += NxTri[ti][1]; sy += NyTri[ti][1]; }
    Nx[n] = 1.0 / (double)n * sx;                                  /* Eq. A7a */
    Ny[n] = 1.0 / (double)n * sy;
    for (ci = 1; ci <= n; ci++) {
      int cim1s = ci - 1;
      if (cim1s == 0) cim1s += n;
      Nx[ci - 1] = 1.0 / (double)n * (NxTri[ci][2] + NxTri[cim1s][3]); /* Eq. A7b */
      Ny[ci - 1] = 1.0 / (double)n * (NyTri[ci][2] + NyTri[cim1s][3]);
    }
  } else {
    double sx = 0.0, sy = 0.0;
    for (ti = 1; ti <= n - 1; ti++) { sx += NxTri[ti][1]; sy += NyTri[ti][1]; }
    double r = 1.0 / (double)(n - 1);
    Nx[n] = r * sx;                                                  /* Eq. A9a */
    Ny[n] = r * sy;
    for (ci = 1; ci <= n; ci++) {                                    /* Eq. A9b */
      if (ci == 1) {
        Nx[ci - 1] = r * NxTri[1][2];
        Ny[ci - 1] = r * NyTri[1][2];
      } else if (ci == n) {
        Nx[ci - 1] = r * NxTri[n - 1][3];
        Ny[ci - 1] = r * NyTri[n - 1][3];
      } else {
        Nx[ci - 1] = r * (NxTri[ci][2] + NxTri[ci - 1][3]);
        Ny[ci - 1] = r * (NyTri[ci][2] + NyTri[ci - 1][3]);
      }
    }
  }
  /* second order (:214-311) */
  if (!is_edge) {
    double r = 1.0 / (double)n;
    for (si = 1; si <= n; si++) {                                    /* Eq. A23a-c */
      int sip1s = si + 1;
      if (sip1s > n) sip1s -= n;
      Nxx[n] = Nxx[n] + (r * (((NxSub[si][1] * Nx[n]) + (NxSub[si][2] * NxTri[si][1])) +
                              (NxSub[si][3] * NxTri[sip1s][1])));
      Nxy[n] = Nxy[n] + (r * (((NySub[si][1] * Nx[n]) + (NySub[si][2] * NxTri[si][1])) +
                              (NySub[si][3] * NxTri[sip1s][1])));
      Nyy[n] = Nyy[n] + (r * (((NySub[si][1] * Ny[n]) + (NySub[si][2] * NyTri[si][1])) +
                              (NySub[si][3] * NyTri[sip1s][1])));
    }
    double sumx = 0.0, sumy = 0.0;
    for (si = 1; si <= n; si++) { sumx += NxSub[si][1]; sumy += NySub[si][1]; }
    for (si = 1; si <= n; si++) {                                    /* Eq. A23d-f */
      int sim1s = si - 1;
      if (sim1s < 1) sim1s += n;
      int sim2s = sim1s - 1;
      if (sim2s < 1) sim2s += n;
      Nxx[si - 1] = r * ((NxTri[si][2] * (NxSub[si][2] + NxSub[sim1s][3])) +
                         ((NxTri[sim1s][3] * (NxSub[sim1s][2] + NxSub[sim2s][3])) +
                          (Nx[si - 1] * sumx)));
      Nxy[si - 1] = r * ((NxTri[si][2] * (NySub[si][2] + NySub[sim1s][3])) +
                         ((NxTri[sim1s][3] * (NySub[sim1s][2] + NySub[sim2s][3])) +
                          (Nx[si - 1] * sumy)));
      Nyy[si - 1] = r * ((NyTri[si][2] * (NySub[si][2] + NySub[sim1s][3])) +
                         ((NyTri[sim1s][3] * (NySub[sim1s][2] + NySub[sim2s][3])) +
                          (Ny[si - 1] * sumy)));
    }
  } else {
    double r = 1.0 / (double)(n - 2);
    for (si = 1; si <= n - 2; si++) {                                /* Eq. A24a-c */
      Nxx[n] = Nxx[n] + (r * (((NxSub[si][1] * Nx[n]) + (NxSub[si][2] * NxTri[si][1])) +
                              (NxSub[si][3] * NxTri[si + 1][1])));
      Nxy[n] = Nxy[n] + (r * (((NySub[si][1] * Nx[n]) + (NySub[si][2] * NxTri[si][1])) +
                              (NySub[si][3] * NxTri[si + 1][1])));
      Nyy[n] = Nyy[n] + (r * (((NySub[si][1] * Ny[n]) + (NySub[si][2] * NyTri[si][1])) +
                              (NySub[si][3] * NyTri[si + 1][1])));
    }
    double sumx = 0.0, sumy = 0.0;
    for (si = 1; si <= n - 2; si++) { sumx += NxSub[si][1]; sumy += NySub[si][1]; }
    for (si = 1; si <= n; si++) {                                    /* Eq. A24d-f, A25 */
      double Axx = 0.0, Axy = 0.0, Ayy = 0.0, Bxx = 0.0, Bxy = 0.0, Byy = 0.0;
      double Cxx = 0.0, Cxy = 0.0, Cyy = 0.0;
      if (si < n - 1) {
        Axx = NxSub[si][2] * NxTri[si][2];
        Axy = NySub[si][2] * NxTri[si][2];
        Ayy = NySub[si][2] * NyTri[si][2];
      }
      if (si > 1 && si < n) {
        Bxx = (NxSub[si - 1][2] * NxTri[si - 1][3]) + (NxSub[si - 1][3] * NxTri[si][2]);
        Bxy = (NySub[si - 1][2] * NxTri[si - 1][3]) + (NySub[si - 1][3] * NxTri[si][2]);
        Byy = (NySub[si - 1][2] * NyTri[si - 1][3]) + (NySub[si - 1][3] * NyTri[si][2]);
      }
      if (si > 2) {
        Cxx = NxSub[si - 2][3] * NxTri[si - 1][3];
        Cxy = NySub[si - 2][3] * NxTri[si - 1][3];
        Cyy = NySub[si - 2][3] * NyTri[si - 1][3];
      }
      Nxx[si - 1] = r * (((Axx + Bxx) + Cxx) + (Nx[si - 1] * sumx));
      Nxy[si - 1] = r * (((Axy + Bxy) + Cxy) + (Nx[si - 1] * sumy));
      Nyy[si - 1] = r * (((Ayy + Byy) + Cyy) + (Ny[si - 1] * sumy));
    }
  }
}

/* src/mesh_derivatives_module.f90:14-75 (vertex part) */
void om_get_neighbour_functions(om_mesh *m) {
  const int w = nCm_ + 1;
  int vi;
#pragma omp parallel for schedule(static)
  for (vi = 1; vi <= nV_; vi++) {
    double Vvi[2], Vvc[2 * NMAX], nx[NMAX + 1], ny[NMAX + 1], nxx[NMAX + 1], nxy[NMAX + 1],
        nyy[NMAX + 1];
    int n = MnC(vi), ci;
    Vvi[0] = MV(vi, 1);
    Vvi[1] = MV(vi, 2);
    for (ci = 1; ci <= n; ci++) {
      int vc = MC(vi, ci);
      Vvc[ci - 1] = MV(vc, 1);
      Vvc[NMAX + ci - 1] = MV(vc, 2);
    }
    om_get_neighbour_functions_vertex_gr(Vvi, n, Vvc, NMAX, Medge(vi) > 0, nx, ny, nxx, nxy, nyy);
    for (ci = 1; ci <= w; ci++) {
      int in = ci <= n + 1;
      MNx(vi, ci) = in ? nx[ci - 1] : 0.0;
      MNy(vi, ci) = in ? ny[ci - 1] : 0.0;
      MNxx(vi, ci) = in ? nxx[ci - 1] : 0.0;
      MNxy(vi, ci) = in ? nxy[ci - 1] : 0.0;
      MNyy(vi, ci) = in ? nyy[ci - 1] : 0.0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Ac and AaAc meshes, src/mesh_ArakawaC_module.f90:24-540                   */
/* ------------------------------------------------------------------------ */

int om_count_Ac(const om_mesh *m) {
  long s = 0;
  for (int vi = 1; vi <= nV_; vi++) s += MnC(vi);
  return (int)(s / 2);
}

/* src/mesh_help_functions_module.f90:673-747 */
static int is_boundary_segment(const om_mesh *m, int v1, int v2) {
  int e1 = Medge(v1), e2 = Medge(v2);
  if (e1 == 0 || e2 == 0) return 0;
  switch (e1) {
    case 1: return e2 == 8 || e2 == 1 || e2 == 2;
    case 2: return e2 == 8 || e2 == 1 || e2 == 2 || e2 == 3 || e2 == 4;
    case 3: return e2 == 2 || e2 == 3 || e2 == 4;
    case 4: return e2 == 2 || e2 == 3 || e2 == 4 || e2 == 5 || e2 == 6;
    case 5: return e2 == 4 || e2 == 5 || e2 == 6;
    case 6: return e2 == 4 || e2 == 5 || e2 == 6 || e2 == 7 || e2 == 8;
    case 7: return e2 == 6 || e2 == 7 || e2 == 8;
    case 8: return e2 == 6 || e2 == 7 || e2 == 8 || e2 == 1 || e2 == 2;
  }
  return 0;
}

static int in3(int e, int a, int b, int c) { return e == a || e == b || e == c; }

int om_make_Ac_mesh(om_mesh *m) {
  const int nV = nV_, nCm = nCm_;
  int vi, ci, vj, cj, iti, ti, n1, n2, n3;
  int vl = 0, vr = 0;     /* persist across connections, as in the reference (:54-55) */
  int nAc = 0;
  const int nAc_alloc = m->d.nAc;
  memset(m->i[OM_iAci], 0, sizeof(int) * (size_t)nV * (size_t)nCm);
  /* :88-197 */
  for (vi = 1; vi <= nV; vi++) {
    for (ci = 1; ci <= MnC(vi); ci++) {
      vj = MC(vi, ci);
      if (MiAci(vi, ci) > 0) continue;
      nAc++;
      if (nAc > nAc_alloc) return OM_BAD_INPUT;
      MiAci(vi, ci) = nAc;
      E2(m->f[OM_VAc], nAc_alloc, nAc, 1) = (MV(vi, 1) + MV(vj, 1)) / 2.0;
      E2(m->f[OM_VAc], nAc_alloc, nAc, 2) = (MV(vi, 2) + MV(vj, 2)) / 2.0;
      for (cj = 1; cj <= MnC(vj); cj++) {
        if (MC(vj, cj) == vi) {
          MiAci(vj, cj) = nAc;
          break;
        }
      }
      double Nx[5], Ny[5], No[5];
      if (!is_boundary_segment(m, vi, vj)) {
        for (iti = 1; iti <= MniTri(vi); iti++) {
          ti = MiTri(vi, iti);
          for (n1 = 1; n1 <= 3; n1++) {
            n2 = n1 + 1;
            if (n2 == 4) n2 = 1;
            n3 = n2 + 1;
            if (n3 == 4) n3 = 1;
            if (MTri(ti, n1) == vi && MTri(ti, n2) == vj) vl = MTri(ti, n3);
            else if (MTri(ti, n1) == vj && MTri(ti, n2) == vi) vr = MTri(ti, n3);
          }
        }
        E2(m->i[OM_Aci], nAc_alloc, nAc, 1) = vi;
        E2(m->i[OM_Aci], nAc_alloc, nAc, 2) = vj;
        E2(m->i[OM_Aci], nAc_alloc, nAc, 3) = vl;
        E2(m->i[OM_Aci], nAc_alloc, nAc, 4) = vr;
        if (vl < 1 || vr < 1) return OM_BAD_INPUT;
        double Nxl[5] = {0, MV(vl, 2) - MV(vj, 2), MV(vi, 2) - MV(vl, 2), MV(vj, 2) - MV(vi, 2), 0.0};
        double Nyl[5] = {0, MV(vj, 1) - MV(vl, 1), MV(vl, 1) - MV(vi, 1), MV(vi, 1) - MV(vj, 1), 0.0};
        double Nxr[5] = {0, MV(vj, 2) - MV(vr, 2), MV(vr, 2) - MV(vi, 2), 0.0, MV(vi, 2) - MV(vj, 2)};
        double Nyr[5] = {0, MV(vr, 1) - MV(vj, 1), MV(vi, 1) - MV(vr, 1), 0.0, MV(vj, 1) - MV(vi, 1)};
        double Nzl = ((MV(vj, 1) - MV(vi, 1)) * (MV(vl, 2) - MV(vi, 2))) -
                     ((MV(vj, 2) - MV(vi, 2)) * (MV(vl, 1) - MV(vi, 1)));
        double Nzr = ((MV(vr, 1) - MV(vi, 1)) * (MV(vj, 2) - MV(vi, 2))) -
                     ((MV(vr, 2) - MV(vi, 2)) * (MV(vj, 1) - MV(vi, 1)));
        for (int k = 1; k <= 4; k++) {
          Nx[k] = -((Nxl[k] / Nzl) + (Nxr[k] / Nzr)) / 2.0;
          Ny[k] = -((Nyl[k] / Nzl) + (Nyr[k] / Nzr)) / 2.0;
        }
      } else {
        for (iti = 1; iti <= MniTri(vi); iti++) {
          ti = MiTri(vi, iti);
          for (n1 = 1; n1 <= 3; n1++) {
            n2 = n1 + 1;
            if (n2 == 4) n2 = 1;
            n3 = n2 + 1;
            if (n3 == 4) n3 = 1;
            if ((MTri(ti, n1) == vi && MTri(ti, n2) == vj) || (MTri(ti, n1) == vj && MTri(ti, n2) == vi))
              vl = MTri(ti, n3);
          }
        }
        E2(m->i[OM_Aci], nAc_alloc, nAc, 1) = vi;
        E2(m->i[OM_Aci], nAc_alloc, nAc, 2) = vj;
        E2(m->i[OM_Aci], nAc_alloc, nAc, 3) = vl;
        E2(m->i[OM_Aci], nAc_alloc, nAc, 4) = 1;
        if (vl < 1) return OM_BAD_INPUT;
        double Nxl[5] = {0, MV(vl, 2) - MV(vj, 2), MV(vi, 2) - MV(vl, 2), MV(vj, 2) - MV(vi, 2), 0.0};
        double Nyl[5] = {0, MV(vj, 1) - MV(vl, 1), MV(vl, 1) - MV(vi, 1), MV(vi, 1) - MV(vj, 1), 0.0};
        double Nzl = ((MV(vj, 1) - MV(vi, 1)) * (MV(vl, 2) - MV(vi, 2))) -
                     ((MV(vj, 2) - MV(vi, 2)) * (MV(vl, 1) - MV(vi, 1)));
        for (int k = 1; k <= 4; k++) {
          Nx[k] = -(Nxl[k] / Nzl);
          Ny[k] = -(Nyl[k] / Nzl);
        }
      }
      double Ux = MV(vj, 1) - MV(vi, 1);
      double Uy = MV(vj, 2) - MV(vi, 2);
      double U = sqrt(Ux * Ux + Uy * Uy);
      double Np = 1.0 / U;
      for (int k = 1; k <= 4; k++) No[k] = (Ny[k] * Ux - Nx[k] * Uy) / U;
      for (int k = 1; k <= 4; k++) {
        E2(m->f[OM_Nx_Ac], nAc_alloc, nAc, k) = Nx[k];
        E2(m->f[OM_Ny_Ac], nAc_alloc, nAc, k) = Ny[k];
        E2(m->f[OM_No_Ac], nAc_alloc, nAc, k) = No[k];
      }
      E1(m->f[OM_Np_Ac], nAc) = Np;
    }
  }
  if (nAc != nAc_alloc) return OM_BAD_INPUT;
  m->d.nAc = nAc;
  m->d.nVAaAc = nV + nAc;

  /* find_Ac_edge_indices, :242-285 */
  for (int aci = 1; aci <= nAc; aci++) {
    int e1 = Medge(MAci(aci, 1)), e2 = Medge(MAci(aci, 2));
    int r;
    if (in3(e1, 8, 1, 2) && in3(e2, 8, 1, 2)) r = 1;
    else if (in3(e1, 2, 3, 4) && in3(e2, 2, 3, 4)) r = 3;
    else if (in3(e1, 4, 5, 6) && in3(e2, 4, 5, 6)) r = 5;
    else if (in3(e1, 6, 7, 8) && in3(e2, 6, 7, 8)) r = 7;
    else r = 0;
    MedgeAc(aci) = r;
  }

  /* make_combined_AaAc_mesh, :286-540 */
  const int nA = m->d.nVAaAc;
  memset(m->i[OM_CAaAc], 0, sizeof(int) * (size_t)nA * (size_t)nCm);
  for (vi = 1; vi <= nV; vi++) {                     /* :325-333 */
    int ai = vi;
    E2(m->f[OM_VAaAc], nA, ai, 1) = MV(vi, 1);
    E2(m->f[OM_VAaAc], nA, ai, 2) = MV(vi, 2);
    MnCA(ai) = MnC(vi);
    for (ci = 1; ci <= MnC(vi); ci++) MCA(ai, ci) = MiAci(vi, ci) + nV;
  }
  for (int aci = 1; aci <= nAc; aci++) {             /* :337-430 */
    int ai = aci + nV;
    E2(m->f[OM_VAaAc], nA, ai, 1) = E2(m->f[OM_VAc], nAc, aci, 1);
    E2(m->f[OM_VAaAc], nA, ai, 2) = E2(m->f[OM_VAc], nAc, aci, 2);
    if (MedgeAc(aci) > 0) {
      int a = MAci(aci, 1), b = MAci(aci, 2), sw = 0;
      switch (MedgeAc(aci)) {
        case 1: sw = MV(a, 1) > MV(b, 1); break;
        case 3: sw = MV(a, 2) < MV(b, 2); break;
        case 5: sw = MV(a, 1) < MV(b, 1); break;
        case 7: sw = MV(a, 2) > MV(b, 2); break;
        default: return OM_BAD_INPUT;
      }
      if (sw) { int t = a; a = b; b = t; }
      int vk = MAci(aci, 3), aci1 = 0, aci2 = 0;
      for (ci = 1; ci <= MnC(vk); ci++) {
        if (MC(vk, ci) == a) aci1 = MiAci(vk, ci);
        else if (MC(vk, ci) == b) aci2 = MiAci(vk, ci);
      }
      if (aci1 == 0 || aci2 == 0) return OM_BAD_INPUT;
      MnCA(ai) = 4;
      MCA(ai, 1) = a;
      MCA(ai, 2) = aci1 + nV;
      MCA(ai, 3) = aci2 + nV;
      MCA(ai, 4) = b;
    } else {
      int a = MAci(aci, 1), b = MAci(aci, 2), l = MAci(aci, 3), r = MAci(aci, 4);
      int aci1 = 0, aci2 = 0, aci3 = 0, aci4 = 0;
      for (ci = 1; ci <= MnC(r); ci++) {
        if (MC(r, ci) == a) aci1 = MiAci(r, ci);
        else if (MC(r, ci) == b) aci2 = MiAci(r, ci);
      }
      for (ci = 1; ci <= MnC(l); ci++) {
        if (MC(l, ci) == b) aci3 = MiAci(l, ci);
        else if (MC(l, ci) == a) aci4 = MiAci(l, ci);
      }
      if (aci1 == 0 || aci2 == 0 || aci3 == 0 || aci4 == 0) return OM_BAD_INPUT;
      MnCA(ai) = 6;
      MCA(ai, 1) = a;
      MCA(ai, 2) = aci1 + nV;
      MCA(ai, 3) = aci2 + nV;
      MCA(ai, 4) = b;
      MCA(ai, 5) = aci3 + nV;
      MCA(ai, 6) = aci4 + nV;
    }
  }
  /* neighbour functions on every AaAc point, :511-537 */
  const int w = nCm + 1;
  int ai;
#pragma omp parallel for schedule(static)
  for (ai = 1; ai <= nA; ai++) {
    double Vvi[2], Vvc[2 * NMAX], nx[NMAX + 1], ny[NMAX + 1], nxx[NMAX + 1], nxy[NMAX + 1],
        nyy[NMAX + 1];
    int n = MnCA(ai), c;
    Vvi[0] = E2(m->f[OM_VAaAc], nA, ai, 1);
    Vvi[1] = E2(m->f[OM_VAaAc], nA, ai, 2);
    for (c = 1; c <= n; c++) {
      int aj = MCA(ai, c);
      Vvc[c - 1] = E2(m->f[OM_VAaAc], nA, aj, 1);
      Vvc[NMAX + c - 1] = E2(m->f[OM_VAaAc], nA, aj, 2);
    }
    int is_edge = (ai <= nV) ? (Medge(ai) > 0) : (MedgeAc(ai - nV) > 0);
    om_get_neighbour_functions_vertex_gr(Vvi, n, Vvc, NMAX, is_edge, nx, ny, nxx, nxy, nyy);
    for (c = 1; c <= w; c++) {
      int in = c <= n + 1;
      MNxA(ai, c) = in ? nx[c - 1] : 0.0;
      MNyA(ai, c) = in ? ny[c - 1] : 0.0;
      MNxxA(ai, c) = in ? nxx[c - 1] : 0.0;
      MNxyA(ai, c) = in ? nxy[c - 1] : 0.0;
      MNyyA(ai, c) = in ? nyy[c - 1] : 0.0;
    }
  }
  return OM_OK;
}

/* ------------------------------------------------------------------------ */
/* Five-colouring, src/mesh_five_colour_module.f90:18-735                    */
/* Literal array queues with position table LDP and O(n) shifts.             */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, nCm;
  int *deg, *L, *Q4, *Q5, *LDP, *S_vi, *S_L, *S_u;
  int Q4n, Q5n, Sn, noofvert;
  int err;
} colctx;

#define CL(c, v, k) (c)->L[((size_t)(k) - 1) * (size_t)(c)->n + (size_t)(v) - 1]
#define CSL(c, s, k) (c)->S_L[((size_t)(k) - 1) * (size_t)(c)->n + (size_t)(s) - 1]

static int is_in_Q4(colctx *c, int vi) {            /* :688-708 */
  int qi = c->LDP[vi - 1];
  if (qi == 0) return 0;
  int isso = c->Q4[qi - 1] == vi;
  if (c->Q5[qi - 1] == vi && isso) c->err = OM_COLOURING_FAILED;
  return isso;
}
static int is_in_Q5(colctx *c, int vi) {            /* :709-729 */
  int qi = c->LDP[vi - 1];
  if (qi == 0) return 0;
  int isso = c->Q5[qi - 1] == vi;
  if (c->Q4[qi - 1] == vi && isso) c->err = OM_COLOURING_FAILED;
  return isso;
}
static void add_to_Q4(colctx *c, int vi) {          /* :560-586 */
  if (c->deg[vi - 1] > 4 || c->LDP[vi - 1] > 0) { c->err = OM_COLOURING_FAILED; return; }
  c->Q4n++;
  c->Q4[c->Q4n - 1] = vi;
  c->LDP[vi - 1] = c->Q4n;
}
static void add_to_Q5(colctx *c, int vi) {          /* :587-613 */
  if (c->deg[vi - 1] != 5 || c->LDP[vi - 1] > 0) { c->err = OM_COLOURING_FAILED; return; }
  c->Q5n++;
  c->Q5[c->Q5n - 1] = vi;
  c->LDP[vi - 1] = c->Q5n;
}
static void remove_from_Q4(colctx *c, int vi) {     /* :614-650 */
  int qi = c->LDP[vi - 1];
  if (qi == 0 || c->Q5[qi - 1] == vi || c->Q4[qi - 1] != vi) { c->err = OM_COLOURING_FAILED; return; }
  memmove(&c->Q4[qi - 1], &c->Q4[qi], sizeof(int) * (size_t)(c->Q4n - qi));
  c->Q4[c->Q4n - 1] = 0;
  c->Q4n--;
  c->LDP[vi - 1] = 0;
  for (int q = qi; q <= c->Q4n; q++) c->LDP[c->Q4[q - 1] - 1] = q;
}
static void remove_from_Q5(colctx *c, int vi) {     /* :651-687 */
  int qi = c->LDP[vi - 1];
  if (qi == 0 || c->Q4[qi - 1] == vi || c->Q5[qi - 1] != vi) { c->err = OM_COLOURING_FAILED; return; }
  memmove(&c->Q5[qi - 1], &c->Q5[qi], sizeof(int) * (size_t)(c->Q5n - qi));
  c->Q5[c->Q5n - 1] = 0;
  c->Q5n--;
  c->LDP[vi - 1] = 0;
  for (int q = qi; q <= c->Q5n; q++) c->LDP[c->Q5[q - 1] - 1] = q;
}
static void check(colctx *c, int vi) {              /* :503-534 */
  int d = c->deg[vi - 1];
  if (d <= 4) {
    if (is_in_Q5(c, vi)) remove_from_Q5(c, vi);
    if (!is_in_Q4(c, vi)) add_to_Q4(c, vi);
  } else if (d == 5) {
    if (is_in_Q4(c, vi)) remove_from_Q4(c, vi);
    if (!is_in_Q5(c, vi)) add_to_Q5(c, vi);
  } else {
    if (is_in_Q4(c, vi)) { c->err = OM_COLOURING_FAILED; return; }
    if (is_in_Q5(c, vi)) remove_from_Q5(c, vi);
  }
}
static void cdelete(colctx *c, int vi) {            /* :345-404 */
  for (int ci = 1; ci <= c->deg[vi - 1]; ci++) {
    int wi = CL(c, vi, ci);
    for (int ci2 = 1; ci2 <= c->deg[wi - 1]; ci2++) {
      if (CL(c, wi, ci2) == vi) {
        for (int k = ci2; k <= c->deg[wi - 1] - 1; k++) CL(c, wi, k) = CL(c, wi, k + 1);
        CL(c, wi, c->deg[wi - 1]) = 0;
        c->deg[wi - 1]--;
        check(c, wi);
        break;
      }
    }
  }
  c->Sn++;
  c->S_vi[c->Sn - 1] = vi;
  for (int k = 1; k <= c->nCm; k++) CSL(c, c->Sn, k) = CL(c, vi, k);
  if (c->deg[vi - 1] <= 4) remove_from_Q4(c, vi);
  else if (c->deg[vi - 1] == 5) remove_from_Q5(c, vi);
  else c->err = OM_COLOURING_FAILED;
  c->noofvert--;
}

int om_calculate_five_colouring_AaAc(om_mesh *m) {
  const int n = nVAaAc_, nCm = nCm_;
  colctx c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  c.nCm = nCm;
  c.deg = (int *)calloc((size_t)n, sizeof(int));
  c.L = (int *)calloc((size_t)n * (size_t)nCm, sizeof(int));
  c.Q4 = (int *)calloc((size_t)n, sizeof(int));
  c.Q5 = (int *)calloc((size_t)n, sizeof(int));
  c.LDP = (int *)calloc((size_t)n, sizeof(int));
  c.S_vi = (int *)calloc((size_t)n, sizeof(int));
  c.S_L = (int *)calloc((size_t)n * (size_t)nCm, sizeof(int));
  c.S_u = (int *)calloc((size_t)n, sizeof(int));
  int status = OM_OK;
  int *colour = m->i[OM_colour];
  if (!c.deg || !c.L || !c.Q4 || !c.Q5 || !c.LDP || !c.S_vi || !c.S_L || !c.S_u) {
    status = OM_BAD_INPUT;
    goto done;
  }
  for (int ai = 1; ai <= n; ai++) {
    c.deg[ai - 1] = MnCA(ai);
    for (int k = 1; k <= nCm; k++) CL(&c, ai, k) = MCA(ai, k);
  }
  c.noofvert = n;
  for (int vi = 1; vi <= n; vi++) {                 /* :91-97 */
    if (c.deg[vi - 1] <= 4) add_to_Q4(&c, vi);
    else if (c.deg[vi - 1] == 5) add_to_Q5(&c, vi);
  }
  /* reduce, :139-237 */
  {
    long it = 0;
    while (c.noofvert > 5) {
      it++;
      if (it > 2L * n) { status = OM_COLOURING_FAILED; goto done; }
      if (c.Q4n > 0) {
        cdelete(&c, c.Q4[c.Q4n - 1]);
      } else {
        if (c.Q5n == 0) { status = OM_COLOURING_FAILED; goto done; }
        int vi = c.Q5[c.Q5n - 1], found = 0;
        int k = nCm;                                 /* C%nconmax */
        for (int ci1 = 1; ci1 <= c.deg[vi - 1] && !found; ci1++) {
          int xi = CL(&c, vi, ci1);
          for (int ci2 = ci1 + 1; ci2 <= c.deg[vi - 1]; ci2++) {
            int yi = CL(&c, vi, ci2), adj = 0;
            for (int ci3 = 1; ci3 <= c.deg[yi - 1]; ci3++)
              if (CL(&c, yi, ci3) == xi) { adj = 1; break; }
            if (adj) continue;
            if (c.deg[xi - 1] < k && c.deg[yi - 1] < k) { found = 1; break; }
          }
        }
        if (found) {
          /* delete(v) then identify(x,y); identify() aborts unconditionally
             (src/mesh_five_colour_module.f90:452) */
          status = OM_COLOURING_IDENTIFY;
          goto done;
        }
        /* return v to the rear of Q5 (:222-228) */
        memmove(&c.Q5[1], &c.Q5[0], sizeof(int) * (size_t)(c.Q5n - 1));
        c.Q5[0] = vi;
        for (int q = 1; q <= c.Q5n; q++) c.LDP[c.Q5[q - 1] - 1] = q;
      }
      if (c.err) { status = c.err; goto done; }
    }
  }
  /* colour, :238-316 */
  memset(colour, 0, sizeof(int) * (size_t)n);
  if (c.Q4n < 5) { status = OM_COLOURING_FAILED; goto done; }
  for (int k = 1; k <= 5; k++) colour[c.Q4[k - 1] - 1] = k;
  while (c.Sn > 0) {
    int vi = c.S_vi[c.Sn - 1];
    if (c.S_u[c.Sn - 1] > 0) {
      colour[c.S_u[c.Sn - 1] - 1] = colour[vi - 1];
    } else {
      int Lc[6] = {0, 0, 0, 0, 0, 0}, s = 0;
      for (int ci = 1; ci <= nCm; ci++) {
        int ui = CSL(&c, c.Sn, ci);
        if (ui == 0) break;
        int col = colour[ui - 1];
        if (col == 0) continue;
        Lc[col] = 1;
      }
      for (int k = 1; k <= 5; k++) s += Lc[k];
      if (s > 4) { status = OM_COLOURING_FAILED; goto done; }
      for (int k = 1; k <= 5; k++)
        if (Lc[k] == 0) { colour[vi - 1] = k; break; }
    }
    c.Sn--;
  }
  /* check_solution, :318-343 */
  for (int vi = 1; vi <= n; vi++) {
    if (colour[vi - 1] == 0) { status = OM_COLOURING_FAILED; goto done; }
    for (int ci = 1; ci <= MnCA(vi); ci++)
      if (colour[vi - 1] == colour[MCA(vi, ci) - 1]) { status = OM_COLOURING_FAILED; goto done; }
  }
  /* colour_vi / colour_nV, :106-116 */
  memset(m->i[OM_colour_vi], 0, sizeof(int) * (size_t)n * 5);
  for (int k = 0; k < 5; k++) m->d.colour_nV[k] = 0;
  for (int vi = 1; vi <= n; vi++) {
    int ci = colour[vi - 1];
    m->d.colour_nV[ci - 1]++;
    E2(m->i[OM_colour_vi], n, m->d.colour_nV[ci - 1], ci) = vi;
  }
done:
  free(c.deg); free(c.L); free(c.Q4); free(c.Q5); free(c.LDP);
  free(c.S_vi); free(c.S_L); free(c.S_u);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Aa operators, src/mesh_derivatives_module.f90:315-432, 485-597            */
/* ------------------------------------------------------------------------ */

void om_get_mesh_derivatives(const om_mesh *m, const double *d, double *ddx, double *ddy) {
  int vi;
#pragma omp parallel for schedule(static)
  for (vi = 1; vi <= nV_; vi++) {
    int n = MnC(vi);
    double x = MNx(vi, n + 1) * E1(d, vi), y = MNy(vi, n + 1) * E1(d, vi);
    for (int ci = 1; ci <= n; ci++) {
      x = x + MNx(vi, ci) * E1(d, MC(vi, ci));
      y = y + MNy(vi, ci) * E1(d, MC(vi, ci));
    }
    E1(ddx, vi) = x;
    E1(ddy, vi) = y;
  }
}

void om_get_mesh_curvatures(const om_mesh *m, const double *d, double *ddxx, double *ddxy,
                            double *ddyy) {
  int vi;
#pragma omp parallel for schedule(static)
  for (vi = 1; vi <= nV_; vi++) {
    int n = MnC(vi);
    double xx = MNxx(vi, n + 1) * E1(d, vi), xy = MNxy(vi, n + 1) * E1(d, vi),
           yy = MNyy(vi, n + 1) * E1(d, vi);
    for (int ci = 1; ci <= n; ci++) {
      double dc = E1(d, MC(vi, ci));
      xx = xx + MNxx(vi, ci) * dc;
      xy = xy + MNxy(vi, ci) * dc;
      yy = yy + MNyy(vi, ci) * dc;
    }
    E1(ddxx, vi) = xx;
    E1(ddxy, vi) = xy;
    E1(ddyy, vi) = yy;
  }
}

void om_get_mesh_derivatives_3D(const om_mesh *m, const double *d, double *ddx, double *ddy,
                                int nz) {
  int vi;
#pragma omp parallel for schedule(static)
  for (vi = 1; vi <= nV_; vi++) {
    int n = MnC(vi);
    for (int k = 1; k <= nz; k++) {
      double x = MNx(vi, n + 1) * E2(d, nV_, vi, k), y = MNy(vi, n + 1) * E2(d, nV_, vi, k);
      for (int ci = 1; ci <= n; ci++) {
        x = x + MNx(vi, ci) * E2(d, nV_, MC(vi, ci), k);
        y = y + MNy(vi, ci) * E2(d, nV_, MC(vi, ci), k);
      }
      E2(ddx, nV_, vi, k) = x;
      E2(ddy, nV_, vi, k) = y;
    }
  }
}

void om_apply_Neumann_boundary(const om_mesh *m, double *d) {
  for (int vi = 5; vi <= nV_; vi++) {
    if (Medge(vi) == 0) continue;
    double s = 0.0;
    int nvals = 0;
    for (int ci = 1; ci <= MnC(vi); ci++) {
      int vc = MC(vi, ci);
      if (Medge(vc) > 0) continue;
      nvals++;
      s = s + E1(d, vc);
    }
    E1(d, vi) = s / (double)nvals;
  }
  for (int vi = 1; vi <= 4; vi++) {
    double s = 0.0;
    int nvals = 0;
    for (int ci = 1; ci <= MnC(vi); ci++) {
      nvals++;
      s = s + E1(d, MC(vi, ci));
    }
    E1(d, vi) = s / (double)nvals;
  }
}

void om_apply_Neumann_boundary_3D(const om_mesh *m, double *d, int nz) {
  for (int vi = 5; vi <= nV_; vi++) {
    if (Medge(vi) == 0) continue;
    for (int k = 1; k <= nz; k++) {
      double s = 0.0;
      int nvals = 0;
      for (int ci = 1; ci <= MnC(vi); ci++) {
        int vc = MC(vi, ci);
        if (Medge(vc) > 0) continue;
        nvals++;
        s = s + E2(d, nV_, vc, k);
      }
      E2(d, nV_, vi, k) = s / (double)nvals;
    }
  }
  for (int vi = 1; vi <= 4; vi++) {
    for (int k = 1; k <= nz; k++) {
      double s = 0.0;
      int nvals = 0;
      for (int ci = 1; ci <= MnC(vi); ci++) {
        nvals++;
        s = s + E2(d, nV_, MC(vi, ci), k);
      }
      E2(d, nV_, vi, k) = s / (double)nvals;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Ac / AaAc operators and maps, src/mesh_ArakawaC_module.f90:542-844        */
/* ------------------------------------------------------------------------ */

void om_get_mesh_derivatives_Ac(const om_mesh *m, const double *d, double *ddx, double *ddy,
                                double *ddp, double *ddo) {
  int aci;
#pragma omp parallel for schedule(static)
  for (aci = 1; aci <= nAc_; aci++) {
    double x = 0.0, y = 0.0, o = 0.0;
    for (int k = 1; k <= 4; k++) {
      double dv = E1(d, MAci(aci, k));
      x = x + MNxAc(aci, k) * dv;
      y = y + MNyAc(aci, k) * dv;
      o = o + MNoAc(aci, k) * dv;
    }
    E1(ddx, aci) = x;
    E1(ddy, aci) = y;
    E1(ddo, aci) = o;
    E1(ddp, aci) = MNpAc(aci) * (E1(d, MAci(aci, 2)) - E1(d, MAci(aci, 1)));
  }
}

void om_get_mesh_derivatives_AaAc(const om_mesh *m, const double *d, double *ddx, double *ddy) {
  int ai;
#pragma omp parallel for schedule(static)
  for (ai = 1; ai <= nVAaAc_; ai++) {
    int n = MnCA(ai);
    double x = MNxA(ai, n + 1) * E1(d, ai), y = MNyA(ai, n + 1) * E1(d, ai);
    for (int ci = 1; ci <= n; ci++) {
      x = x + MNxA(ai, ci) * E1(d, MCA(ai, ci));
      y = y + MNyA(ai, ci) * E1(d, MCA(ai, ci));
    }
    E1(ddx, ai) = x;
    E1(ddy, ai) = y;
  }
}

/* The reference multiplies every weight by d(ai), the home value
   (src/mesh_ArakawaC_module.f90:651-656); restated as written. */
static inline void curvatures_vertex_AaAc(const om_mesh *m, const double *d, int ai, double *xx,
                                          double *xy, double *yy) {
  int n = MnCA(ai);
  double h = E1(d, ai);
  double a = h * MNxxA(ai, n + 1), b = h * MNxyA(ai, n + 1), c = h * MNyyA(ai, n + 1);
  for (int ci = 1; ci <= n; ci++) {
    a = a + h * MNxxA(ai, ci);
    b = b + h * MNxyA(ai, ci);
    c = c + h * MNyyA(ai, ci);
  }
  *xx = a;
  *xy = b;
  *yy = c;
}

void om_get_mesh_curvatures_AaAc(const om_mesh *m, const double *d, double *ddxx, double *ddxy,
                                 double *ddyy) {
  int ai;
#pragma omp parallel for schedule(static)
  for (ai = 1; ai <= nVAaAc_; ai++)
    curvatures_vertex_AaAc(m, d, ai, &E1(ddxx, ai), &E1(ddxy, ai), &E1(ddyy, ai));
}

static inline int is_edge_AaAc(const om_mesh *m, int ai) {
  return ai <= nV_ ? Medge(ai) > 0 : MedgeAc(ai - nV_) > 0;
}

void om_apply_Neumann_boundary_AaAc(const om_mesh *m, double *d) {
  int ai;
  /* edge points read only non-edge points, so the pass is order-free */
#pragma omp parallel for schedule(static)
  for (ai = 5; ai <= nVAaAc_; ai++) {
    if (!is_edge_AaAc(m, ai)) continue;
    double s = 0.0;
    int nvals = 0;
    for (int ci = 1; ci <= MnCA(ai); ci++) {
      int ac = MCA(ai, ci);
      if (is_edge_AaAc(m, ac)) continue;
      nvals++;
      s = s + E1(d, ac);
    }
    E1(d, ai) = s / (double)nvals;
  }
  for (ai = 1; ai <= 4; ai++) {
    double s = 0.0;
    int nvals = 0;
    for (int ci = 1; ci <= MnCA(ai); ci++) {
      nvals++;
      s = s + E1(d, MCA(ai, ci));
    }
    E1(d, ai) = s / (double)nvals;
  }
}

void om_map_Aa_to_Ac(const om_mesh *m, const double *d_Aa, double *d_Ac) {
  for (int aci = 1; aci <= nAc_; aci++)
    E1(d_Ac, aci) = (E1(d_Aa, MAci(aci, 1)) + E1(d_Aa, MAci(aci, 2))) / 2.0;
}

void om_map_Aa_to_Ac_3D(const om_mesh *m, const double *d_Aa, double *d_Ac, int nz) {
  for (int aci = 1; aci <= nAc_; aci++)
    for (int k = 1; k <= nz; k++)
      E2(d_Ac, nAc_, aci, k) =
          (E2(d_Aa, nV_, MAci(aci, 1), k) + E2(d_Aa, nV_, MAci(aci, 2), k)) / 2.0;
}

void om_map_Ac_to_Aa(const om_mesh *m, const double *d_Ac, double *d_Aa) {
  for (int vi = 1; vi <= nV_; vi++) {
    double s = 0.0;
    for (int ci = 1; ci <= MnC(vi); ci++) s = s + E1(d_Ac, MiAci(vi, ci)) / (double)MnC(vi);
    E1(d_Aa, vi) = s;
  }
}

void om_map_Ac_to_Aa_3D(const om_mesh *m, const double *d_Ac, double *d_Aa, int nz) {
  for (int vi = 1; vi <= nV_; vi++)
    for (int k = 1; k <= nz; k++) {
      double s = 0.0;
      for (int ci = 1; ci <= MnC(vi); ci++)
        s = s + E2(d_Ac, nAc_, MiAci(vi, ci), k) / (double)MnC(vi);
      E2(d_Aa, nV_, vi, k) = s;
    }
}

void om_rotate_xy_to_po(const om_mesh *m, const double *x, const double *y, double *p,
                        double *o) {
  for (int ci = 1; ci <= nAc_; ci++) {
    int vi = MAci(ci, 1), vj = MAci(ci, 2);
    double Dx = MV(vj, 1) - MV(vi, 1), Dy = MV(vj, 2) - MV(vi, 2);
    double D = sqrt(Dx * Dx + Dy * Dy);
    E1(p, ci) = E1(x, ci) * Dx / D + E1(y, ci) * Dy / D;
    E1(o, ci) = E1(y, ci) * Dx / D - E1(x, ci) * Dy / D;
  }
}

/* ------------------------------------------------------------------------ */
/* General ice-model data, src/general_ice_model_data_module.f90:23-297,464  */
/* ------------------------------------------------------------------------ */

static inline int is_floating(double Hi, double Hb, double SL) {
  return Hi < (SL - Hb) * seawater_density / ice_density;
}

void om_determine_masks(const om_mesh *m, om_ice *ice) {
  const int nV = nV_, nAc = nAc_;
  enum { type_land = 0, type_ocean = 1, type_lake = 2, type_sheet = 3, type_shelf = 4,
         type_coast = 5, type_margin = 6, type_groundingline = 7, type_calvingfront = 8 };
  int *land = II(OI_mask_land), *ocean = II(OI_mask_ocean), *lake = II(OI_mask_lake),
      *icem = II(OI_mask_ice), *sheet = II(OI_mask_sheet), *shelf = II(OI_mask_shelf),
      *coast = II(OI_mask_coast), *margin = II(OI_mask_margin), *gl = II(OI_mask_gl),
      *cf = II(OI_mask_cf), *mask = II(OI_mask);
  int *landA = II(OI_mask_land_Ac), *oceanA = II(OI_mask_ocean_Ac), *lakeA = II(OI_mask_lake_Ac),
      *iceA = II(OI_mask_ice_Ac), *sheetA = II(OI_mask_sheet_Ac), *shelfA = II(OI_mask_shelf_Ac),
      *coastA = II(OI_mask_coast_Ac), *marginA = II(OI_mask_margin_Ac), *glA = II(OI_mask_gl_Ac),
      *cfA = II(OI_mask_cf_Ac), *maskA = II(OI_mask_Ac);
  const double *Hi = IF(OI_Hi), *Hb = IF(OI_Hb), *SL = IF(OI_SL);
  const double *HiA = IF(OI_Hi_Ac), *HbA = IF(OI_Hb_Ac), *SLA = IF(OI_SL_Ac);
  for (int v = 0; v < nV; v++) {
    land[v] = 1; ocean[v] = 0; lake[v] = 0; icem[v] = 0; sheet[v] = 0; shelf[v] = 0;
    coast[v] = 0; margin[v] = 0; gl[v] = 0; cf[v] = 0; mask[v] = type_land;
  }
  for (int a = 0; a < nAc; a++) {
    landA[a] = 1; oceanA[a] = 0; lakeA[a] = 0; iceA[a] = 0; sheetA[a] = 0; shelfA[a] = 0;
    coastA[a] = 0; marginA[a] = 0; glA[a] = 0; cfA[a] = 0; maskA[a] = type_land;
  }
  for (int vi = 1; vi <= nV; vi++) {                 /* :160-187 */
    int v = vi - 1;
    if (is_floating(Hi[v], Hb[v], SL[v])) { ocean[v] = 1; land[v] = 0; mask[v] = type_ocean; }
    if (Hi[v] > 0.0) icem[v] = 1;
    if (icem[v] == 1 && land[v] == 1) { sheet[v] = 1; mask[v] = type_sheet; }
    if (icem[v] == 1 && ocean[v] == 1) { shelf[v] = 1; mask[v] = type_shelf; }
  }
  for (int vi = 1; vi <= nV; vi++) {                 /* :190-240 */
    int v = vi - 1;
    if (land[v] == 1)
      for (int ci = 1; ci <= MnC(vi); ci++)
        if (ocean[MC(vi, ci) - 1] == 1) { mask[v] = type_coast; coast[v] = 1; }
    if (icem[v] == 1)
      for (int ci = 1; ci <= MnC(vi); ci++)
        if (icem[MC(vi, ci) - 1] == 0) { mask[v] = type_margin; margin[v] = 1; }
    if (sheet[v] == 1)
      for (int ci = 1; ci <= MnC(vi); ci++)
        if (shelf[MC(vi, ci) - 1] == 1) { mask[v] = type_groundingline; gl[v] = 1; }
    if (icem[v] == 1)
      for (int ci = 1; ci <= MnC(vi); ci++)
        if (ocean[MC(vi, ci) - 1] == 1) { mask[v] = type_calvingfront; cf[v] = 1; }
  }
  for (int a = 0; a < nAc; a++) {                    /* :245-270 */
    if (is_floating(HiA[a], HbA[a], SLA[a])) { oceanA[a] = 1; landA[a] = 0; maskA[a] = type_ocean; }
    if (HiA[a] > 0.0) iceA[a] = 1;
    if (iceA[a] == 1 && landA[a] == 1) { sheetA[a] = 1; maskA[a] = type_sheet; }
    if (iceA[a] == 1 && oceanA[a] == 1) { shelfA[a] = 1; maskA[a] = type_shelf; }
  }
  for (int aci = 1; aci <= nAc; aci++) {             /* :273-295 */
    int a = aci - 1, vi = MAci(aci, 1) - 1, vj = MAci(aci, 2) - 1;
    if ((land[vi] == 1 && ocean[vj] == 1) || (land[vj] == 1 && ocean[vi] == 1)) {
      maskA[a] = type_coast; coastA[a] = 1;
    }
    if ((icem[vi] == 1 && icem[vj] == 0) || (icem[vj] == 1 && icem[vi] == 0)) {
      maskA[a] = type_margin; marginA[a] = 1;
    }
    if ((sheet[vi] == 1 && shelf[vj] == 1) || (sheet[vj] == 1 && shelf[vi] == 1)) {
      maskA[a] = type_groundingline; glA[a] = 1;
    }
    if ((icem[vi] == 1 && shelf[vj] == 0 && ocean[vj] == 1) ||
        (icem[vj] == 1 && shelf[vi] == 0 && ocean[vi] == 1)) {
      maskA[a] = type_calvingfront; cfA[a] = 1;
    }
  }
}

void om_update_general_ice_model_data(const om_mesh *m, om_ice *ice) {
  const int nV = nV_, nAc = nAc_;
  om_map_Aa_to_Ac(m, IF(OI_Hi), IF(OI_Hi_Ac));       /* :40-43 (Ti not carried) */
  om_map_Aa_to_Ac(m, IF(OI_Hb), IF(OI_Hb_Ac));
  om_map_Aa_to_Ac(m, IF(OI_SL), IF(OI_SL_Ac));
  for (int v = 0; v < nV; v++)                        /* :46-52 */
    IF(OI_Hs)[v] = IF(OI_Hi)[v] +
                   dmax(IF(OI_SL)[v] - ice_density / seawater_density * IF(OI_Hi)[v], IF(OI_Hb)[v]);
  for (int a = 0; a < nAc; a++)
    IF(OI_Hs_Ac)[a] = IF(OI_Hi_Ac)[a] + dmax(IF(OI_SL_Ac)[a] -
                                              ice_density / seawater_density * IF(OI_Hi_Ac)[a],
                                              IF(OI_Hb_Ac)[a]);
  for (int v = 0; v < nV; v++) IF(OI_dHs_dt)[v] = IF(OI_dHb_dt)[v] + IF(OI_dHi_dt)[v];  /* :55 */
  om_determine_masks(m, ice);                                                           /* :58 */
  om_get_mesh_derivatives(m, IF(OI_Hi), IF(OI_dHi_dx), IF(OI_dHi_dy));                  /* :61-66 */
  om_get_mesh_derivatives(m, IF(OI_Hs), IF(OI_dHs_dx), IF(OI_dHs_dy));
  om_get_mesh_derivatives_Ac(m, IF(OI_Hi), IF(OI_dHi_dx_Ac), IF(OI_dHi_dy_Ac), IF(OI_dHi_dp_Ac),
                             IF(OI_dHi_do_Ac));
  om_get_mesh_derivatives_Ac(m, IF(OI_Hb), IF(OI_dHb_dx_Ac), IF(OI_dHb_dy_Ac), IF(OI_dHb_dp_Ac),
                             IF(OI_dHb_do_Ac));
  om_get_mesh_derivatives_Ac(m, IF(OI_Hs), IF(OI_dHs_dx_Ac), IF(OI_dHs_dy_Ac), IF(OI_dHs_dp_Ac),
                             IF(OI_dHs_do_Ac));
  om_get_mesh_derivatives_Ac(m, IF(OI_SL), IF(OI_dSL_dx_Ac), IF(OI_dSL_dy_Ac), IF(OI_dSL_dp_Ac),
                             IF(OI_dSL_do_Ac));
  for (int v = 0; v < nV; v++) {                      /* :70-87 */
    if (II(OI_mask_ocean)[v] == 0) {
      IF(OI_dHs_dx_shelf)[v] = IF(OI_dHs_dx)[v];
      IF(OI_dHs_dy_shelf)[v] = IF(OI_dHs_dy)[v];
    } else {
      IF(OI_dHs_dx_shelf)[v] = (1.0 - ice_density / seawater_density) * IF(OI_dHi_dx)[v];
      IF(OI_dHs_dy_shelf)[v] = (1.0 - ice_density / seawater_density) * IF(OI_dHi_dy)[v];
    }
  }
  for (int a = 0; a < nAc; a++) {
    if (II(OI_mask_ocean_Ac)[a] == 0) {
      IF(OI_dHs_dx_shelf_Ac)[a] = IF(OI_dHs_dx_Ac)[a];
      IF(OI_dHs_dy_shelf_Ac)[a] = IF(OI_dHs_dy_Ac)[a];
    } else {
      IF(OI_dHs_dx_shelf_Ac)[a] = (1.0 - ice_density / seawater_density) * IF(OI_dHi_dx_Ac)[a];
      IF(OI_dHs_dy_shelf_Ac)[a] = (1.0 - ice_density / seawater_density) * IF(OI_dHi_dy_Ac)[a];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Ice physical properties, src/general_ice_model_data_module.f90:298-462    */
/* (the non-benchmark branch), after Ti -> Ti_Ac (:40-43).                   */
/* ------------------------------------------------------------------------ */
static double arrhenius(double T) {                                    /* :379-383 */
  const double A_low_temp = 1.14E-05, A_high_temp = 5.47E+10, Q_low_temp = 6.0E+04,
               Q_high_temp = 13.9E+04, R_gas = 8.314;
  if (T < 263.15) return A_low_temp * exp(-Q_low_temp / (R_gas * T));
  return A_high_temp * exp(-Q_high_temp / (R_gas * T));
}

/* mode 1: the temperature-dependent branch (:370-460); mode 2: the fixed
 * values of the EISMINT / Halfar / Bueler / mesh_generation_test benchmarks
 * (:323-348), which set no Ac temperature and no Ti-dependent field. */
void om_ice_physical_properties(const om_mesh *m, om_ice *ice, const om_params *p, int mode) {
  const int nV = nV_, nAc = nAc_, nZ = p->nZ;
  const double T0 = 273.16, CC = 8.7E-04, SMT = 271.15;               /* src/parameters_module.f90 */
  const double sec_per_year = 31556943.36;
  double *A = IF(OI_A_flow), *Am = IF(OI_A_flow_mean), *Tp = IF(OI_Ti_pmp), *Cp = IF(OI_Cpi),
         *Ki = IF(OI_Ki);
  const double *Ti = IF(OI_Ti), *Hi = IF(OI_Hi);
  if (mode == 2) {
    for (int k = 0; k < nZ; k++)
      for (int v = 0; v < nV; v++) {
        A[(size_t)k * nV + v] = 1.0E-16;
        Ki[(size_t)k * nV + v] = 2.1 * sec_per_year;
        Cp[(size_t)k * nV + v] = 2009.0;
        Tp[(size_t)k * nV + v] = T0 - (p->zeta[k] * Hi[v] * 8.7E-04);
      }
    for (int v = 0; v < nV; v++) Am[v] = 1.0E-16;
    for (size_t i = 0; i < (size_t)nAc * nZ; i++) IF(OI_A_flow_Ac)[i] = 1.0E-16;
    for (int a = 0; a < nAc; a++) IF(OI_A_flow_mean_Ac)[a] = 1.0E-16;
    return;
  }
  om_map_Aa_to_Ac_3D(m, IF(OI_Ti), IF(OI_Ti_Ac), nZ);                  /* :40-43 */
  double prof[32];
  for (int v = 0; v < nV; v++) {                                       /* :370-408 */
    for (int k = 0; k < nZ; k++) {
      const double T = Ti[(size_t)k * nV + v];
      Tp[(size_t)k * nV + v] = T0 - CC * Hi[v] * p->zeta[k];
      A[(size_t)k * nV + v] = arrhenius(T);
      Cp[(size_t)k * nV + v] = 2115.3 + 7.79293 * (T - T0);
      Ki[(size_t)k * nV + v] = 3.101E+08 * exp(-0.0057 * T);
      prof[k] = A[(size_t)k * nV + v];
    }
    if (II(OI_mask_sheet)[v] == 1) Am[v] = om_vertical_average(p, prof);
    else Am[v] = arrhenius((Ti[v] + SMT) / 2.0);
  }
  double *AA = IF(OI_A_flow_Ac), *AmA = IF(OI_A_flow_mean_Ac);
  const double *TA = IF(OI_Ti_Ac);
  for (int a = 0; a < nAc; a++) {                                      /* :413-460 */
    for (int k = 0; k < nZ; k++) {
      AA[(size_t)k * nAc + a] = arrhenius(TA[(size_t)k * nAc + a]);
      prof[k] = AA[(size_t)k * nAc + a];
    }
    if (II(OI_mask_sheet_Ac)[a] == 1) AmA[a] = om_vertical_average(p, prof);
    else AmA[a] = arrhenius((TA[a] + SMT) / 2.0);
  }
}

/* ------------------------------------------------------------------------ */
/* Vertical integrals, src/zeta_module.f90:33-85                             */
/* ------------------------------------------------------------------------ */

double om_vertical_average(const om_params *p, const double *f) {
  double a = 0.0;
  for (int k = 1; k <= p->nZ - 1; k++)
    a = a + 0.5 * (f[k] + f[k - 1]) * (p->zeta[k] - p->zeta[k - 1]);
  return a;
}

void om_vertical_integrate(const om_params *p, const double *f, double *int_f) {
  int_f[p->nZ - 1] = 0.0;
  for (int k = p->nZ - 1; k >= 1; k--)
    int_f[k - 1] = int_f[k] - 0.5 * (f[k] + f[k - 1]) * (p->zeta[k] - p->zeta[k - 1]);
}

/* ------------------------------------------------------------------------ */
/* SIA, src/ice_dynamics_module.f90:240-314                                  */
/* ------------------------------------------------------------------------ */

void om_solve_SIA(const om_mesh *m, om_ice *ice, const om_params *p) {
  const int nAc = nAc_, nZ = p->nZ;
  const double D_uv_3D_cutoff = -1e5;
  double *D3 = IF(OI_D_SIA_3D_Ac);
  const double *Hi = IF(OI_Hi_Ac), *dp = IF(OI_dHs_dp_Ac), *dO = IF(OI_dHs_do_Ac),
               *Af = IF(OI_A_flow_Ac);
  const int *sheet = II(OI_mask_sheet_Ac);
  memset(D3, 0, sizeof(double) * (size_t)nAc * (size_t)nZ);
  memset(IF(OI_D_SIA_Ac), 0, sizeof(double) * (size_t)nAc);
  memset(IF(OI_Ux_SIA_Ac), 0, sizeof(double) * (size_t)nAc);
  memset(IF(OI_Uy_SIA_Ac), 0, sizeof(double) * (size_t)nAc);
  memset(IF(OI_Up_SIA_Ac), 0, sizeof(double) * (size_t)nAc);
  memset(IF(OI_Uo_SIA_Ac), 0, sizeof(double) * (size_t)nAc);
  int aci;
#pragma omp parallel for schedule(static)
  for (aci = 1; aci <= nAc; aci++) {
    int a = aci - 1;
    if (sheet[a] == 1) {
      double f[32], I[32];
      double D_0 = pow(ice_density * grav * Hi[a], n_flow) *
                   pow(dp[a] * dp[a] + dO[a] * dO[a], (n_flow - 1.0) / 2.0);
      for (int k = 0; k < nZ; k++)
        f[k] = p->m_enh_sia * E2(Af, nAc, aci, k + 1) * pow(p->zeta[k], n_flow);
      om_vertical_integrate(p, f, I);
      for (int k = 0; k < nZ; k++) {
        double Dd = 2.0 * Hi[a] * I[k];
        E2(D3, nAc, aci, k + 1) = D_0 * Dd;
      }
    }
    for (int k = 1; k <= nZ; k++)
      if (E2(D3, nAc, aci, k) < D_uv_3D_cutoff) E2(D3, nAc, aci, k) = D_uv_3D_cutoff;
  }
#pragma omp parallel for schedule(static)
  for (aci = 1; aci <= nAc; aci++) {
    int a = aci - 1;
    if (sheet[a] == 1) {
      double prof[32];
      for (int k = 0; k < nZ; k++) prof[k] = E2(D3, nAc, aci, k + 1);
      double D2 = om_vertical_average(p, prof);
      IF(OI_D_SIA_Ac)[a] = Hi[a] * D2;
      IF(OI_Ux_SIA_Ac)[a] = D2 * IF(OI_dHs_dx_Ac)[a];
      IF(OI_Uy_SIA_Ac)[a] = D2 * IF(OI_dHs_dy_Ac)[a];
      IF(OI_Up_SIA_Ac)[a] = D2 * dp[a];
      IF(OI_Uo_SIA_Ac)[a] = D2 * dO[a];
    }
  }
  om_map_Ac_to_Aa(m, IF(OI_Ux_SIA_Ac), IF(OI_U_SIA));
  om_map_Ac_to_Aa(m, IF(OI_Uy_SIA_Ac), IF(OI_V_SIA));
  om_map_Ac_to_Aa(m, IF(OI_D_SIA_Ac), IF(OI_D_SIA));
}

/* ------------------------------------------------------------------------ */
/* SSA, src/ice_dynamics_module.f90:408-844                                  */
/* ------------------------------------------------------------------------ */

/* :780-844 */
void om_basal_yield_stress(const om_mesh *m, om_ice *ice) {
  const double pf1 = -1000.0, pf2 = 0.0, p_min = 5.0, p_max = 20.0;
  double *phi = IF(OI_phi_fric_AaAc), *tau = IF(OI_tau_c_AaAc);
  for (int vi = 1; vi <= nV_; vi++) {
    int v = vi - 1;
    double Hb = IF(OI_Hb)[v], SL = IF(OI_SL)[v], Hi = IF(OI_Hi)[v];
    double lambda_p = dmax(0.0, dmin(1.0, (1.0 - (Hb - SL) / 1000.0)));
    double pwp = 0.96 * ice_density * grav * dmax(0.1, Hi) * lambda_p;
    phi[v] = dmax(p_min, dmin(p_max, (p_min + (p_max - p_min) * (1.0 + (Hb - pf2) / (pf2 - pf1)))));
    tau[v] = tan((pi_c / 180.0) * phi[v]) * (ice_density * grav * dmax(0.1, Hi) - pwp);
  }
  for (int aci = 1; aci <= nAc_; aci++) {
    int a = aci - 1, ai = aci + nV_ - 1;
    double Hb = IF(OI_Hb_Ac)[a], SL = IF(OI_SL_Ac)[a], Hi = IF(OI_Hi_Ac)[a];
    double lambda_p = dmax(0.0, dmin(1.0, (1.0 - (Hb - SL) / 1000.0)));
    double pwp = 0.96 * ice_density * grav * dmax(0.1, Hi) * lambda_p;
    phi[ai] = dmax(p_min, dmin(p_max, (p_min + (p_max - p_min) * (1.0 + (Hb - pf2) / (pf2 - pf1)))));
    tau[ai] = tan((pi_c / 180.0) * phi[ai]) * (ice_density * grav * dmax(0.1, Hi) - pwp);
  }
}

/* :695-726 */
void om_SSA_effective_viscosity(const om_mesh *m, om_ice *ice, const om_params *p) {
  const double epsilon_sq_0 = 1e-12;
  om_get_mesh_derivatives_AaAc(m, IF(OI_U_SSA_AaAc), IF(OI_dU_SSA_dx_AaAc), IF(OI_dU_SSA_dy_AaAc));
  om_get_mesh_derivatives_AaAc(m, IF(OI_V_SSA_AaAc), IF(OI_dV_SSA_dx_AaAc), IF(OI_dV_SSA_dy_AaAc));
  const double *ux = IF(OI_dU_SSA_dx_AaAc), *uy = IF(OI_dU_SSA_dy_AaAc),
               *vx = IF(OI_dV_SSA_dx_AaAc), *vy = IF(OI_dV_SSA_dy_AaAc),
               *A = IF(OI_A_flow_mean_AaAc), *H = IF(OI_Hi_AaAc);
  double *eta = IF(OI_eta_AaAc), *N = IF(OI_N_AaAc);
  int a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < nVAaAc_; a++) {
    double s = uy[a] + vx[a];
    eta[a] = pow(p->m_enh_ssa * 0.5 * A[a], -1.0 / n_flow) *
             pow(ux[a] * ux[a] + vy[a] * vy[a] + ux[a] * vy[a] + 0.25 * (s * s) + epsilon_sq_0,
                 (1.0 - n_flow) / (2.0 * n_flow));
    N[a] = eta[a] * dmax(0.1, H[a]);
  }
}

/* :727-779 (Coulomb_regularised) */
void om_SSA_sliding_term(const om_mesh *m, om_ice *ice) {
  const double delta_v = 1e-3, q_plastic = 0.30, u_threshold = 100.0;
  const double *tau = IF(OI_tau_c_AaAc), *U = IF(OI_U_SSA_AaAc), *V = IF(OI_V_SSA_AaAc);
  double *S = IF(OI_S_AaAc);
  int a;
#pragma omp parallel for schedule(static)
  for (a = 0; a < nVAaAc_; a++)
    S[a] = tau[a] * (pow(delta_v * delta_v + U[a] * U[a] + V[a] * V[a], 0.5 * (q_plastic - 1.0))) /
           pow(u_threshold, q_plastic);
}

/* :558-694 */
int om_solve_SSA_linearised(const om_mesh *m, om_ice *ice, const om_params *p, int *did_reset,
                            double *last_max_residual, int *sweeps) {
  const int nV = nV_, nA = nVAaAc_;
  double *U = IF(OI_U_SSA_AaAc), *V = IF(OI_V_SSA_AaAc);
  double *eu = IF(OI_eu_i_AaAc), *ev = IF(OI_ev_i_AaAc);
  double *RHSx = IF(OI_RHSx_AaAc), *RHSy = IF(OI_RHSy_AaAc);
  double *LHSx = IF(OI_LHSx_AaAc), *LHSy = IF(OI_LHSy_AaAc);
  double *resU = IF(OI_resU_AaAc), *resV = IF(OI_resV_AaAc);
  const double *eta = IF(OI_eta_AaAc), *S = IF(OI_S_AaAc), *H = IF(OI_Hi_AaAc),
               *Hb = IF(OI_Hb_AaAc), *SL = IF(OI_SL_AaAc);
  const int *glA = II(OI_mask_gl_Ac);
  const double omega = p->SSA_SOR_omega;
  int a;
  for (a = 0; a < nA; a++) {                                          /* :581-584 */
    RHSx[a] = ice_density * grav * IF(OI_dHs_dx_shelf_AaAc)[a] / eta[a];
    RHSy[a] = ice_density * grav * IF(OI_dHs_dy_shelf_AaAc)[a] / eta[a];
  }
  for (int ai = 1; ai <= nA; ai++) {                                  /* :587-595 */
    int n = MnCA(ai);
    a = ai - 1;
    if (!is_floating(H[a], Hb[a], SL[a])) {
      eu[a] = (4.0 * MNxxA(ai, n + 1) + MNyyA(ai, n + 1)) - S[a] / (dmax(0.1, H[a]) * eta[a]);
      ev[a] = (4.0 * MNyyA(ai, n + 1) + MNxxA(ai, n + 1)) - S[a] / (dmax(0.1, H[a]) * eta[a]);
    } else {
      eu[a] = (4.0 * MNxxA(ai, n + 1) + MNyyA(ai, n + 1));
      ev[a] = (4.0 * MNyyA(ai, n + 1) + MNxxA(ai, n + 1));
    }
  }
  int has_converged = 0, inner = 0, status = OM_OK;
  double maxres = 0.0;
  while (!has_converged && inner < p->SSA_max_inner_loops) {          /* :601-691 */
    inner++;
    maxres = 0.0;
    for (int fci = 1; fci <= 5; fci++) {
      const int cnt = m->d.colour_nV[fci - 1];
      int fcvi;
      double cmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : cmax)
      for (fcvi = 1; fcvi <= cnt; fcvi++) {
        int ai = E2(m->i[OM_colour_vi], nA, fcvi, fci);
        int i = ai - 1;
        if (is_edge_AaAc(m, ai)) continue;
        if (p->use_analytical_GL_flux && ai > nV && glA[ai - nV - 1] == 1) continue;
        double Uxx, Uxy, Uyy, Vxx, Vxy, Vyy;
        curvatures_vertex_AaAc(m, U, ai, &Uxx, &Uxy, &Uyy);
        curvatures_vertex_AaAc(m, V, ai, &Vxx, &Vxy, &Vyy);
        double sumU = 0.0, sumV = 0.0;
        for (int ci = 1; ci <= MnCA(ai); ci++) {
          int ac = MCA(ai, ci);
          sumU = sumU + U[ac - 1] * (4.0 * MNxxA(ai, ci) + MNyyA(ai, ci));
          sumV = sumV + V[ac - 1] * (4.0 * MNyyA(ai, ci) + MNxxA(ai, ci));
        }
        LHSx[i] = sumU + (3.0 * Vxy) + (eu[i] * U[i]);
        LHSy[i] = sumV + (3.0 * Uxy) + (ev[i] * V[i]);
        resU[i] = (LHSx[i] - RHSx[i]) / eu[i];
        resV[i] = (LHSy[i] - RHSy[i]) / ev[i];
        cmax = dmax(cmax, fabs(resU[i]));
        cmax = dmax(cmax, fabs(resV[i]));
        U[i] = U[i] - omega * resU[i];
        V[i] = V[i] - omega * resV[i];
      }
      maxres = dmax(maxres, cmax);
    }
    om_apply_Neumann_boundary_AaAc(m, U);                              /* :669-670 */
    om_apply_Neumann_boundary_AaAc(m, V);
    if (maxres < p->SSA_max_residual_UV) {                             /* :676-689 */
      *did_reset = 0;
      has_converged = 1;
    } else if (maxres > 1e6) {
      memset(U, 0, sizeof(double) * (size_t)nA);
      memset(V, 0, sizeof(double) * (size_t)nA);
      *did_reset = 1;
      has_converged = 1;
    } else if (inner == p->SSA_max_inner_loops) {
      status = OM_SOR_NOT_CONVERGED;
    }
  }
  if (last_max_residual) *last_max_residual = maxres;
  if (sweeps) *sweeps = inner;
  return status;
}

/* :408-557 */

/* ------------------------------------------------------------------------ */
/* Grounding-line flux, src/ice_dynamics_module.f90:848-949 (the regularised */
/* Coulomb branch: Tsai et al. 2015).  At every GL Ac point the flux follows */
/* the flotation-crossing thickness and points down the thickness-above-     */
/* flotation gradient; it prescribes Ux/Uy_SSA_Ac there.                     */
/* ------------------------------------------------------------------------ */
void om_calculate_GL_flux(const om_mesh *m, om_ice *ice) {
  const int nAc = nAc_;
  const double Q0 = 0.61;
  const double *Hi = IF(OI_Hi), *Hb = IF(OI_Hb), *SL = IF(OI_SL), *phi = IF(OI_phi_fric_AaAc),
               *Am = IF(OI_A_flow_mean);
  const int *glA = II(OI_mask_gl_Ac), *sheet = II(OI_mask_sheet);
  double *Qabs = IF(OI_Qabs_GL_Ac), *Qp = IF(OI_Qp_GL_Ac), *Ux = IF(OI_Ux_SSA_Ac), *Uy = IF(OI_Uy_SSA_Ac);
  for (int a = 0; a < nAc; a++) { Qabs[a] = 0.0; Qp[a] = 0.0; }       /* :870-871 */
  for (int aci = 1; aci <= nAc; aci++) {                                /* :878 */
    const int a = aci - 1;
    if (glA[a] == 0) continue;
    const int vi = MAci(aci, 1), vj = MAci(aci, 2);
    const double TAFi = E1(Hi, vi) - ((E1(SL, vi) - E1(Hb, vi)) * (seawater_density / ice_density));
    const double TAFj = E1(Hi, vj) - ((E1(SL, vj) - E1(Hb, vj)) * (seawater_density / ice_density));
    const double lambda_GL = TAFi / (TAFi - TAFj);
    const double Hi_GL = (E1(Hi, vi) * (1.0 - lambda_GL)) + (E1(Hi, vj) * lambda_GL);
    double phi_GL, A_GL;                                               /* grounded side, :893-899 */
    if (E1(sheet, vi) == 1) { phi_GL = E1(phi, vi); A_GL = E1(Am, vi); }
    else { phi_GL = E1(phi, vj); A_GL = E1(Am, vj); }
    const double factor_Tsai = (8.0 * Q0 * A_GL * pow(ice_density * grav, n_flow) *   /* :908-911 */
                                pow(1.0 - (ice_density / seawater_density), n_flow - 1.0) /
                                pow(4.0, n_flow));
    Qabs[a] = factor_Tsai * pow(Hi_GL, n_flow + 2.0) / tan(phi_GL * (pi_c / 180.0));  /* :917 */
    double Fx = -(IF(OI_dHi_dx_Ac)[a] - ((IF(OI_dSL_dx_Ac)[a] - IF(OI_dHb_dx_Ac)[a]) *
                                         (seawater_density / ice_density)));
    double Fy = -(IF(OI_dHi_dy_Ac)[a] - ((IF(OI_dSL_dy_Ac)[a] - IF(OI_dHb_dy_Ac)[a]) *
                                         (seawater_density / ice_density)));
    const double F = sqrt(Fx * Fx + Fy * Fy);
    Fx = Fx / F;
    Fy = Fy / F;
    Ux[a] = Qabs[a] * Fx / Hi_GL;                                      /* :934-935 */
    Uy[a] = Qabs[a] * Fy / Hi_GL;
    double Dx = MV(vj, 1) - MV(vi, 1), Dy = MV(vj, 2) - MV(vi, 2);
    const double D = sqrt(Dx * Dx + Dy * Dy);
    Dx = Dx / D;
    Dy = Dy / D;
    Qp[a] = Qabs[a] * (Dx * Fx + Dy * Fy);                             /* :944 */
  }
}

int om_solve_SSA(const om_mesh *m, om_ice *ice, const om_params *p, om_ssa_stats *st) {
  const int nV = nV_, nAc = nAc_, nA = nVAaAc_;
  memset(st, 0, sizeof(*st));
  int zero = p->ssa_forced_zero;
  long sheet_sum = 0;
  for (int v = 0; v < nV; v++) sheet_sum += II(OI_mask_sheet)[v];
  if (sheet_sum == 0) zero = 1;
  if (zero) {                                                          /* :451-465 */
    memset(IF(OI_U_SSA_AaAc), 0, sizeof(double) * (size_t)nA);
    memset(IF(OI_V_SSA_AaAc), 0, sizeof(double) * (size_t)nA);
    memset(IF(OI_U_SSA), 0, sizeof(double) * (size_t)nV);
    memset(IF(OI_V_SSA), 0, sizeof(double) * (size_t)nV);
    memset(IF(OI_Ux_SSA_Ac), 0, sizeof(double) * (size_t)nAc);
    memset(IF(OI_Uy_SSA_Ac), 0, sizeof(double) * (size_t)nAc);
    memset(IF(OI_Up_SSA_Ac), 0, sizeof(double) * (size_t)nAc);
    memset(IF(OI_Uo_SSA_Ac), 0, sizeof(double) * (size_t)nAc);
    return OM_OK;
  }
  om_basal_yield_stress(m, ice);                                       /* :468 */
  if (p->use_analytical_GL_flux) om_calculate_GL_flux(m, ice);         /* :471 */
  if (p->use_tau_c_override)
    memcpy(IF(OI_tau_c_AaAc), IF(OI_tau_c_override_AaAc), sizeof(double) * (size_t)nA);
  /* gather, :478-495 */
#define GATHER(dst, srcAa, srcAc)                                           \
  do {                                                                      \
    memcpy(IF(dst), IF(srcAa), sizeof(double) * (size_t)nV);                \
    memcpy(IF(dst) + nV, IF(srcAc), sizeof(double) * (size_t)nAc);         \
  } while (0)
  GATHER(OI_Hi_AaAc, OI_Hi, OI_Hi_Ac);
  GATHER(OI_Hb_AaAc, OI_Hb, OI_Hb_Ac);
  GATHER(OI_SL_AaAc, OI_SL, OI_SL_Ac);
  GATHER(OI_dHs_dx_shelf_AaAc, OI_dHs_dx_shelf, OI_dHs_dx_shelf_Ac);
  GATHER(OI_dHs_dy_shelf_AaAc, OI_dHs_dy_shelf, OI_dHs_dy_shelf_Ac);
  GATHER(OI_A_flow_mean_AaAc, OI_A_flow_mean, OI_A_flow_mean_Ac);
  GATHER(OI_U_SSA_AaAc, OI_U_SSA, OI_Ux_SSA_Ac);
  GATHER(OI_V_SSA_AaAc, OI_V_SSA, OI_Uy_SSA_Ac);
#undef GATHER
  int has_converged = 0, it = 0, did_reset_before = 0, did_reset_now = 0;
  double *N = IF(OI_N_AaAc), *Np = IF(OI_N_AaAc_prev);
  while (!has_converged && it < p->SSA_max_outer_loops) {             /* :501-545 */
    it++;
    memcpy(Np, N, sizeof(double) * (size_t)nA);
    om_SSA_effective_viscosity(m, ice, p);
    double sdn = 0.0, sn = 0.0;
    for (int a = 0; a < nA; a++) {
      double d = N[a] - Np[a];
      sdn = sdn + d * d;
    }
    for (int a = 0; a < nA; a++) sn = sn + N[a] * N[a];
    double RN = sqrt(sdn / sn);
    st->last_RN = RN;
    st->picard_iterations = it;
    if (RN < p->SSA_RN_tol) {
      has_converged = 1;
      st->picard_converged = 1;
      break;
    }
    om_SSA_sliding_term(m, ice);
    double lm = 0.0;
    int sweeps = 0;
    int r = om_solve_SSA_linearised(m, ice, p, &did_reset_now, &lm, &sweeps);
    if (r == OM_SOR_NOT_CONVERGED) st->status = OM_SOR_NOT_CONVERGED;
    if (it <= 64) st->sweeps[it - 1] = sweeps;
    st->total_sweeps += sweeps;
    st->last_max_residual = lm;
    if (did_reset_now) {
      st->resets++;
      if (!did_reset_before) did_reset_before = 1;
      else { st->status = OM_SSA_UNSTABLE; return OM_SSA_UNSTABLE; }
    }
  }
  /* scatter, :548-552 */
  memcpy(IF(OI_U_SSA), IF(OI_U_SSA_AaAc), sizeof(double) * (size_t)nV);
  memcpy(IF(OI_V_SSA), IF(OI_V_SSA_AaAc), sizeof(double) * (size_t)nV);
  memcpy(IF(OI_Ux_SSA_Ac), IF(OI_U_SSA_AaAc) + nV, sizeof(double) * (size_t)nAc);
  memcpy(IF(OI_Uy_SSA_Ac), IF(OI_V_SSA_AaAc) + nV, sizeof(double) * (size_t)nAc);
  om_rotate_xy_to_po(m, IF(OI_Ux_SSA_Ac), IF(OI_Uy_SSA_Ac), IF(OI_Up_SSA_Ac), IF(OI_Uo_SSA_Ac));
  return st->status;
}

/* ------------------------------------------------------------------------ */
/* Mass continuity, src/ice_dynamics_module.f90:31-237                       */
/* ------------------------------------------------------------------------ */

void om_calculate_ice_thickness_change(const om_mesh *m, om_ice *ice, const double *SMB_year,
                                       const double *BMB, double dt, const int *mask_noice,
                                       const om_params *p) {
  const int nV = nV_, nAc = nAc_, nCm = nCm_;
  double *dVi = IF(OI_dVi_in), *Hi = IF(OI_Hi);
  memset(dVi, 0, sizeof(double) * (size_t)nV * (size_t)nCm);
  double *Vi_SMB = (double *)calloc((size_t)nV, sizeof(double));
  for (int aci = 1; aci <= nAc; aci++) {                               /* :64-105 */
    int vi = MAci(aci, 1), vj = MAci(aci, 2), ci = 0, cj = 0;
    for (int c = 1; c <= MnC(vi); c++) if (MC(vi, c) == vj) { ci = c; break; }
    for (int c = 1; c <= MnC(vj); c++) if (MC(vj, c) == vi) { cj = c; break; }
    double Upar = IF(OI_Up_SIA_Ac)[aci - 1] + IF(OI_Up_SSA_Ac)[aci - 1];
    double dV;
    if (Upar > 0.0) dV = Hi[vi - 1] * Upar * MCw(vi, ci) * dt;
    else dV = Hi[vj - 1] * Upar * MCw(vi, ci) * dt;
    E2(dVi, nV, vi, ci) = -dV;
    E2(dVi, nV, vj, cj) = dV;
  }
  for (int v = 0; v < nV; v++) Vi_SMB[v] = (SMB_year[v] + BMB[v]) * MA(v + 1) * dt;  /* :112 */
  for (int vi = 1; vi <= nV; vi++) {                                   /* :115-166 */
    double avail = MA(vi) * Hi[vi - 1], Vin = 0.0, Vout = 0.0;
    for (int ci = 1; ci <= MnC(vi); ci++) {
      double x = E2(dVi, nV, vi, ci);
      if (x > 0.0) Vin = Vin + x;
      else Vout = Vout - x;
    }
    (void)Vin;
    double rescale = 1.0;
    if (-Vi_SMB[vi - 1] >= avail) {
      Vi_SMB[vi - 1] = -avail;
      rescale = 0.0;
    }
    if (Vout > avail + Vi_SMB[vi - 1]) rescale = (avail + Vi_SMB[vi - 1]) / Vout;
    if (rescale < 1.0) {
      for (int ci = 1; ci <= MnC(vi); ci++) {
        int vj = MC(vi, ci);
        if (E2(dVi, nV, vi, ci) < 0.0) {
          E2(dVi, nV, vi, ci) = E2(dVi, nV, vi, ci) * rescale;
          for (int cj = 1; cj <= MnC(vj); cj++)
            if (MC(vj, cj) == vi) { E2(dVi, nV, vj, cj) = -E2(dVi, nV, vi, ci); break; }
        }
      }
    }
  }
  for (int vi = 1; vi <= nV; vi++) {                                   /* :172-176 */
    double s = 0.0;
    for (int ci = 1; ci <= nCm; ci++) s = s + E2(dVi, nV, vi, ci);
    IF(OI_dHi_dt)[vi - 1] = (s + Vi_SMB[vi - 1]) / (MA(vi) * dt);
  }
  if (dt == 0.0) memset(IF(OI_dHi_dt), 0, sizeof(double) * (size_t)nV);  /* :180 */
  for (int v = 0; v < nV; v++) {                                        /* :186-187 */
    IF(OI_Hi_prev)[v] = Hi[v];
    Hi[v] = Hi[v] + (IF(OI_dHi_dt)[v] * dt);
  }
  if (p->zero_Hi_at_edges)                                              /* :189-229 */
    for (int vi = 1; vi <= nV; vi++) if (Medge(vi) > 0) Hi[vi - 1] = 0.0;
  if (mask_noice)                                                       /* :232-235 */
    for (int v = 0; v < nV; v++) if (mask_noice[v] == 1) Hi[v] = 0.0;
  free(Vi_SMB);
}

/* ------------------------------------------------------------------------ */
/* solve_SIA_3D, src/ice_dynamics_module.f90:317-405                         */
/* ------------------------------------------------------------------------ */
static void vertex_deriv_3D(const om_mesh *m, const double *d, int vi, int k, double *ddx, double *ddy) {
  /* get_mesh_derivatives_vertex_3D, src/mesh_derivatives_module.f90:175-190 (1-based vi, k) */
  const int n = MnC(vi);
  double x = E2(m->f[OM_Nx], nV_, vi, n + 1) * E2(d, nV_, vi, k);
  double y = E2(m->f[OM_Ny], nV_, vi, n + 1) * E2(d, nV_, vi, k);
  for (int ci = 1; ci <= n; ci++) {
    const int vc = E2(m->i[OM_C], nV_, vi, ci);
    x = x + E2(m->f[OM_Nx], nV_, vi, ci) * E2(d, nV_, vc, k);
    y = y + E2(m->f[OM_Ny], nV_, vi, ci) * E2(d, nV_, vc, k);
  }
  *ddx = x;
  *ddy = y;
}

void om_solve_SIA_3D(const om_mesh *m, om_ice *ice, const om_params *p) {
  const int nV = nV_, nZ = p->nZ;
  double *U3 = IF(OI_U_3D), *V3 = IF(OI_V_3D), *W3 = IF(OI_W_3D);
  const double *Hi = IF(OI_Hi), *Af = IF(OI_A_flow), *Us = IF(OI_U_SSA), *Vs = IF(OI_V_SSA);
  const double *sx = IF(OI_dHs_dx), *sy = IF(OI_dHs_dy), *ix = IF(OI_dHi_dx), *iy = IF(OI_dHi_dy);
  const double D_uv_3D_cutoff = -1E5;
  for (size_t i = 0; i < (size_t)nV * nZ; i++) U3[i] = V3[i] = W3[i] = 0.0;   /* :339-342 */
  for (int v = 0; v < nV; v++) {                                                /* :345-364 */
    for (int k = 0; k < nZ; k++) {
      U3[(size_t)k * nV + v] = Us[v];
      V3[(size_t)k * nV + v] = Vs[v];
    }
    if (II(OI_mask_shelf)[v] == 1) continue;
    if (Hi[v] == 0.0) continue;
    const double D_0 = pow(ice_density * grav * Hi[v], n_flow) *
                       pow(sx[v] * sx[v] + sy[v] * sy[v], (n_flow - 1.0) / 2.0);
    double prof[32], Ddef[32];
    for (int k = 0; k < nZ; k++) prof[k] = p->m_enh_sia * Af[(size_t)k * nV + v] * pow(p->zeta[k], n_flow);
    om_vertical_integrate(p, prof, Ddef);
    for (int k = 0; k < nZ; k++) {
      Ddef[k] = 2.0 * Hi[v] * Ddef[k];
      const double D3 = dmax(D_0 * Ddef[k], D_uv_3D_cutoff);
      U3[(size_t)k * nV + v] = D3 * sx[v] + Us[v];
      V3[(size_t)k * nV + v] = D3 * sy[v] + Vs[v];
    }
  }
  om_apply_Neumann_boundary_3D(m, U3, nZ);                                     /* :366-367 */
  om_apply_Neumann_boundary_3D(m, V3, nZ);
  for (int vi = 1; vi <= nV; vi++) {                                            /* :369-399 */
    const int v = vi - 1;
    if (Medge(vi) > 0) continue;
    if (II(OI_mask_sheet)[v] == 0) continue;
    const double dHb_dx = sx[v] - ix[v], dHb_dy = sy[v] - iy[v];
    W3[(size_t)(nZ - 1) * nV + v] = IF(OI_dHb_dt)[v] + U3[(size_t)(nZ - 1) * nV + v] * dHb_dx +
                                    V3[(size_t)(nZ - 1) * nV + v] * dHb_dy;
    for (int k = nZ - 1; k >= 1; k--) {   /* 1-based layer k, k+1 */
      double dUdx_k, dUdy_k, dUdx_kp1, dUdy_kp1, dVdx_k, dVdy_k, dVdx_kp1, dVdy_kp1;
      vertex_deriv_3D(m, U3, vi, k, &dUdx_k, &dUdy_k);
      vertex_deriv_3D(m, U3, vi, k + 1, &dUdx_kp1, &dUdy_kp1);
      vertex_deriv_3D(m, V3, vi, k, &dVdx_k, &dVdy_k);
      vertex_deriv_3D(m, V3, vi, k + 1, &dVdx_kp1, &dVdy_kp1);
      const double w1 = (dUdx_k + dUdx_kp1) / 2.0;
      const double w2 = (dVdy_k + dVdy_kp1) / 2.0;
      const double zk = p->zeta[k - 1], zk1 = p->zeta[k];
      const double w3 = ((sx[v] - 0.5 * (zk1 + zk) * ix[v]) / dmax(0.1, Hi[v])) *
                        ((E2(U3, nV, vi, k + 1) - E2(U3, nV, vi, k)) / (zk1 - zk));
      const double w4 = ((sy[v] - 0.5 * (zk1 + zk) * iy[v]) / dmax(0.1, Hi[v])) *
                        ((E2(V3, nV, vi, k + 1) - E2(V3, nV, vi, k)) / (zk1 - zk));
      E2(W3, nV, vi, k) = E2(W3, nV, vi, k + 1) - Hi[v] * (w1 + w2 + w3 + w4) * (zk1 - zk);
    }
  }
  om_apply_Neumann_boundary_3D(m, W3, nZ);                                     /* :403 */
}

/* ------------------------------------------------------------------------ */
/* Critical time steps, src/UFEMISM_main_model.f90:734-775                   */
/* ------------------------------------------------------------------------ */
void om_critical_timesteps(const om_mesh *m, const om_ice *ice, const om_params *p, double out[3]) {
  const double pi = 3.141592653589793;                    /* src/parameters_module.f90:9 */
  const int nV = nV_, nZ = p->nZ;
  double dD = 1000.0, dS = 1000.0, d3 = 1000.0;
  const double *Us = IF(OI_U_SSA), *Vs = IF(OI_V_SSA), *D = IF(OI_D_SIA_Ac), *A = m->f[OM_A];
  for (int aci = 1; aci <= nAc_; aci++) {
    const int vi = MAci(aci, 1), vj = MAci(aci, 2);
    const double dx = E2(m->f[OM_V], nV, vj, 1) - E2(m->f[OM_V], nV, vi, 1);
    const double dy = E2(m->f[OM_V], nV, vj, 2) - E2(m->f[OM_V], nV, vi, 2);
    const double dist = sqrt(dx * dx + dy * dy);
    /* 1E-09 is a default-real literal in the reference */
    dD = dmin(dist * dist / (-6.0 * pi * (D[aci - 1] - (double)1E-09f)), dD);
    dS = dmin(dist / (fabs(Us[vi - 1]) + fabs(Vs[vi - 1])), dS);
    dS = dmin(dist / (fabs(Us[vj - 1]) + fabs(Vs[vj - 1])), dS);
  }
  for (int v = 0; v < nV; v++) {
    const double r = sqrt(A[v] / pi);
    dS = dmin(r / (fabs(Us[v]) + fabs(Vs[v])), dS);
    for (int k = 0; k < nZ; k++)
      d3 = dmin(r / (fabs(IF(OI_U_3D)[(size_t)k * nV + v]) + fabs(IF(OI_V_3D)[(size_t)k * nV + v])), d3);
  }
  const double dt_correction_factor = 0.9;
  out[0] = dD * dt_correction_factor;
  out[1] = dS * dt_correction_factor;
  out[2] = d3 * dt_correction_factor;
}

/* ------------------------------------------------------------------------ */
/* Upwind derivatives, src/mesh_derivatives_module.f90:30-47, 434-483        */
/* ------------------------------------------------------------------------ */
#define MTri(ti, k) E2(m->i[OM_Tri], m->d.nTri, ti, k)
void om_triangle_neighbour_functions(const om_mesh *m, double *NxTri, double *NyTri) {
  const int nT = m->d.nTri;
  for (int ti = 1; ti <= nT; ti++) {
    const double ax = E2(m->f[OM_V], nV_, MTri(ti, 1), 1), ay = E2(m->f[OM_V], nV_, MTri(ti, 1), 2);
    const double bx = E2(m->f[OM_V], nV_, MTri(ti, 2), 1), by = E2(m->f[OM_V], nV_, MTri(ti, 2), 2);
    const double cx = E2(m->f[OM_V], nV_, MTri(ti, 3), 1), cy = E2(m->f[OM_V], nV_, MTri(ti, 3), 2);
    const double D = ax * (by - cy) + bx * (cy - ay) + cx * (ay - by);
    E2(NxTri, nT, ti, 1) = (by - cy) / D;
    E2(NxTri, nT, ti, 2) = (cy - ay) / D;
    E2(NxTri, nT, ti, 3) = (ay - by) / D;
    E2(NyTri, nT, ti, 1) = (cx - bx) / D;
    E2(NyTri, nT, ti, 2) = (ax - cx) / D;
    E2(NyTri, nT, ti, 3) = (bx - ax) / D;
  }
}

/* is_in_triangle, src/mesh_help_functions_module.f90:647-672 */
static int in_triangle(const double pa[2], const double pb[2], const double pc[2], const double p[2]) {
  const double tol = 1E-8;
  const double as_x = p[0] - pa[0], as_y = p[1] - pa[1];
  const double s1 = ((pb[0] - pa[0]) * as_y - (pb[1] - pa[1]) * as_x);
  const double s2 = ((pc[0] - pa[0]) * as_y - (pc[1] - pa[1]) * as_x);
  const double s3 = ((pc[0] - pb[0]) * (p[1] - pb[1]) - (pc[1] - pb[1]) * (p[0] - pb[0]));
  return s1 > -tol && s2 < tol && s3 > -tol;
}

int om_upwind_derivatives_3D(const om_mesh *m, const double *R, const double *NxTri, const double *NyTri,
                             const double *U, const double *V, const double *d, int nz, double *ddx,
                             double *ddy) {
  const int nV = nV_, nT = m->d.nTri;
  int missing = 0;
  for (int vi = 1; vi <= nV; vi++)
    for (int k = 1; k <= nz; k++) {
      const double u = E2(U, nV, vi, k), w = E2(V, nV, vi, k);
      double x = 0.0, y = 0.0;
      if (fabs(u) < 1E-10 && fabs(w) < 1E-10) {
        vertex_deriv_3D(m, d, vi, k, &x, &y);
      } else {
        /* W = [U,V] * R / (4 |[U,V]|), evaluated left to right */
        const double den = 4.0 * sqrt(u * u + w * w);
        const double Wx = u * E1(R, vi) / den, Wy = w * E1(R, vi) / den;
        const double p[2] = {E2(m->f[OM_V], nV, vi, 1) - Wx, E2(m->f[OM_V], nV, vi, 2) - Wy};
        int tup = 0;
        for (int iti = 1; iti <= MniTri(vi); iti++) {
          const int ti = E2(m->i[OM_iTri], nV, vi, iti);
          double pa[2], pb[2], pc[2];
          for (int c = 0; c < 2; c++) {
            pa[c] = E2(m->f[OM_V], nV, MTri(ti, 1), c + 1);
            pb[c] = E2(m->f[OM_V], nV, MTri(ti, 2), c + 1);
            pc[c] = E2(m->f[OM_V], nV, MTri(ti, 3), c + 1);
          }
          if (in_triangle(pa, pb, pc, p)) {
            tup = ti;
            break;
          }
        }
        if (tup == 0) {
          missing++;
        } else {
          x = E2(NxTri, nT, tup, 1) * E2(d, nV, MTri(tup, 1), k) + E2(NxTri, nT, tup, 2) * E2(d, nV, MTri(tup, 2), k) +
              E2(NxTri, nT, tup, 3) * E2(d, nV, MTri(tup, 3), k);
          y = E2(NyTri, nT, tup, 1) * E2(d, nV, MTri(tup, 1), k) + E2(NyTri, nT, tup, 2) * E2(d, nV, MTri(tup, 2), k) +
              E2(NyTri, nT, tup, 3) * E2(d, nV, MTri(tup, 3), k);
        }
      }
      E2(ddx, nV, vi, k) = x;
      E2(ddy, nV, vi, k) = y;
    }
  return missing;
}

/* ------------------------------------------------------------------------ */
/* Primary-mesh geometry, src/mesh_help_functions_module.f90                 */
/* ------------------------------------------------------------------------ */
#define GV(vi, k) E2(g->V, g->nV, vi, k)
#define GTri(ti, k) E2(g->Tri, g->nTri, ti, k)
#define GTricc(ti, k) E2(g->Tricc, g->nTri, ti, k)
#define GiTri(vi, k) E2(g->iTri, g->nV, vi, k)

/* line la*x + lb*y = lc through p and q (:1918-1930) */
static void g_line_from_points(const double p[2], const double q[2], double *la, double *lb, double *lc) {
  *la = q[1] - p[1];
  *lb = p[0] - q[0];
  *lc = (*la) * p[0] + (*lb) * p[1];
}

/* (:1931-1948) */
static void g_perpendicular_bisector(const double p[2], const double q[2], double la1, double lb1,
                                     double *la2, double *lb2, double *lc2) {
  const double m0 = (p[0] + q[0]) / 2.0, m1 = (p[1] + q[1]) / 2.0;
  *lc2 = -lb1 * m0 + la1 * m1;
  *la2 = -lb1;
  *lb2 = la1;
}

/* (:1949-1965); parallel lines give the single-precision literal 1E30 */
static void g_line_line_intersection(double la1, double lb1, double lc1, double la2, double lb2, double lc2,
                                     double out[2]) {
  const double d = la1 * lb2 - la2 * lb1;
  if (d == 0.0) {
    out[0] = out[1] = (double)1E30f;
  } else {
    out[0] = (lb2 * lc1 - lb1 * lc2) / d;
    out[1] = (la1 * lc2 - la2 * lc1) / d;
  }
}

/* (:1966-1996) */
static void g_find_circumcenter(const double p[2], const double q[2], const double r[2], double cc[2]) {
  double la1, lb1, lc1, le1, lf1, lg1, la2, lb2, lc2, le2, lf2, lg2;
  g_line_from_points(p, q, &la1, &lb1, &lc1);
  g_line_from_points(q, r, &le1, &lf1, &lg1);
  g_perpendicular_bisector(p, q, la1, lb1, &la2, &lb2, &lc2);
  g_perpendicular_bisector(q, r, le1, lf1, &le2, &lf2, &lg2);
  g_line_line_intersection(la2, lb2, lc2, le2, lf2, lg2, cc);
}

static void g_tri_cc(const om_geom *g, int ti, double cc[2]) {
  double v1[2] = {GV(GTri(ti, 1), 1), GV(GTri(ti, 1), 2)};
  double v2[2] = {GV(GTri(ti, 2), 1), GV(GTri(ti, 2), 2)};
  double v3[2] = {GV(GTri(ti, 3), 1), GV(GTri(ti, 3), 2)};
  g_find_circumcenter(v1, v2, v3, cc);
}

static void g_rotate_tri(om_geom *g, int ti) {
  const int a = GTri(ti, 1), b = GTri(ti, 2), c = GTri(ti, 3);
  GTri(ti, 1) = b;
  GTri(ti, 2) = c;
  GTri(ti, 3) = a;
}

int om_update_triangle_circumcenters(om_geom *g, int *rotated) {
  const double Lx = g->xmax - g->xmin, Ly = g->ymax - g->ymin;
  int bad = 0, rot = 0;
  for (int ti = 1; ti <= g->nTri; ti++) {
    double cc[2];
    g_tri_cc(g, ti, cc);
    /* the first retry tests y against 1e4 domain widths, the second against 1e5 (:867-885) */
    if (cc[0] > Lx * 100000.0 || cc[1] > Ly * 10000.0) {
      g_rotate_tri(g, ti);
      g_tri_cc(g, ti, cc);
      rot++;
    }
    if (cc[0] > Lx * 100000.0 || cc[1] > Ly * 100000.0) {
      g_rotate_tri(g, ti);
      g_tri_cc(g, ti, cc);
      rot++;
    }
    if (cc[0] > Lx * 100000.0 || cc[1] > Ly * 100000.0) bad++;
    GTricc(ti, 1) = cc[0];
    GTricc(ti, 2) = cc[1];
  }
  if (rotated) *rotated = rot;
  return bad;
}

/* crop the circumcentre of t1 in the direction of t2 (:452-487) */
static void g_crop_circumcenter(const om_geom *g, int t1, int t2, double ccc[2]) {
  double la, lb, lc, le, lf, lg;
  ccc[0] = GTricc(t1, 1);
  ccc[1] = GTricc(t1, 2);
  const int te = E1(g->Tri_edge_index, t1);
  double a[2], b[2];
  if (te == 1 && GTricc(t1, 2) > g->ymax) {
    a[0] = g->xmin; a[1] = g->ymax; b[0] = g->xmax; b[1] = g->ymax;
  } else if (te == 3 && GTricc(t1, 1) > g->xmax) {
    a[0] = g->xmax; a[1] = g->ymax; b[0] = g->xmax; b[1] = g->ymin;
  } else if (te == 5 && GTricc(t1, 2) < g->ymin) {
    a[0] = g->xmin; a[1] = g->ymin; b[0] = g->xmax; b[1] = g->ymin;
  } else if (te == 7 && GTricc(t1, 1) < g->xmin) {
    a[0] = g->xmin; a[1] = g->ymax; b[0] = g->xmin; b[1] = g->ymin;
  } else {
    return;
  }
  g_line_from_points(a, b, &la, &lb, &lc);
  const double p[2] = {GTricc(t1, 1), GTricc(t1, 2)}, q[2] = {GTricc(t2, 1), GTricc(t2, 2)};
  g_line_from_points(p, q, &le, &lf, &lg);
  g_line_line_intersection(la, lb, lc, le, lf, lg, ccc);
}

static void g_push(double *Vor, int max_vor, int *n, double x, double y) {
  if (*n < max_vor) {
    Vor[2 * (*n)] = x;
    Vor[2 * (*n) + 1] = y;
  }
  (*n)++;
}

/* :253-305 */
static int g_vor_free(const om_geom *g, int vi, double *Vor, int mx) {
  int n = 0;
  const int nt = E1(g->niTri, vi);
  for (int iti = 1; iti <= nt; iti++) {
    int ic = iti - 1, ia = iti + 1;
    if (ic == 0) ic = nt;
    if (ia > nt) ia = 1;
    const int ti = GiTri(vi, iti), tc = GiTri(vi, ic), ta = GiTri(vi, ia);
    double cc_c[2], cc_a[2];
    g_crop_circumcenter(g, ti, tc, cc_c);
    g_crop_circumcenter(g, ti, ta, cc_a);
    if (cc_c[0] != cc_a[0] || cc_c[1] != cc_a[1]) {
      g_push(Vor, mx, &n, cc_c[0], cc_c[1]);
      g_push(Vor, mx, &n, cc_a[0], cc_a[1]);
    } else {
      g_push(Vor, mx, &n, GTricc(ti, 1), GTricc(ti, 2));
    }
  }
  g_push(Vor, mx, &n, Vor[0], Vor[1]);
  return n;
}

/* :306-393 */
static int g_vor_edge(const om_geom *g, int vi, double *Vor, int mx) {
  int n = 0;
  const int nt = E1(g->niTri, vi), ei = E1(g->edge_index, vi);
  for (int iti = 1; iti <= nt; iti++) {
    const int ti = GiTri(vi, iti);
    const double cc0 = GTricc(ti, 1), cc1 = GTricc(ti, 2);
    double c[2], c2[2];
    if (iti == 1) {
      if ((ei == 1 || ei == 2) && cc1 < g->ymax) g_push(Vor, mx, &n, cc0, g->ymax);
      else if ((ei == 3 || ei == 4) && cc0 < g->xmax) g_push(Vor, mx, &n, g->xmax, cc1);
      else if ((ei == 5 || ei == 6) && cc1 > g->ymin) g_push(Vor, mx, &n, cc0, g->ymin);
      else if ((ei == 7 || ei == 8) && cc0 > g->xmin) g_push(Vor, mx, &n, g->xmin, cc1);
      g_crop_circumcenter(g, GiTri(vi, 1), GiTri(vi, 2), c);
      g_push(Vor, mx, &n, c[0], c[1]);
    }
    if (iti > 1 && iti < nt) {
      g_crop_circumcenter(g, ti, GiTri(vi, iti + 1), c2);   /* anticlockwise */
      g_crop_circumcenter(g, ti, GiTri(vi, iti - 1), c);    /* clockwise */
      if (c2[0] != c[0] || c2[1] != c[1]) {
        g_push(Vor, mx, &n, c[0], c[1]);
        g_push(Vor, mx, &n, c2[0], c2[1]);
      } else {
        g_push(Vor, mx, &n, cc0, cc1);
      }
    }
    if (iti == nt) {
      g_crop_circumcenter(g, ti, GiTri(vi, iti - 1), c);
      g_push(Vor, mx, &n, c[0], c[1]);
      if ((ei == 1 || ei == 8) && cc1 < g->ymax) g_push(Vor, mx, &n, cc0, g->ymax);
      else if ((ei == 3 || ei == 2) && cc0 < g->xmax) g_push(Vor, mx, &n, g->xmax, cc1);
      else if ((ei == 5 || ei == 4) && cc1 > g->ymin) g_push(Vor, mx, &n, cc0, g->ymin);
      else if ((ei == 7 || ei == 6) && cc0 > g->xmin) g_push(Vor, mx, &n, g->xmin, cc1);
    }
  }
  return n;
}

/* :394-451 */
static int g_vor_corner(const om_geom *g, int vi, double *Vor, int mx) {
  if (E1(g->niTri, vi) > 1) return g_vor_edge(g, vi, Vor, mx);
  int n = 0;
  const int ti = GiTri(vi, 1);
  const double c0 = GTricc(ti, 1), c1 = GTricc(ti, 2);
  switch (E1(g->edge_index, vi)) {
    case 2:
      g_push(Vor, mx, &n, c0, g->ymax); g_push(Vor, mx, &n, c0, c1); g_push(Vor, mx, &n, g->xmax, c1);
      break;
    case 4:
      g_push(Vor, mx, &n, g->xmax, c1); g_push(Vor, mx, &n, c0, c1); g_push(Vor, mx, &n, c0, g->ymin);
      break;
    case 6:
      g_push(Vor, mx, &n, c0, g->ymin); g_push(Vor, mx, &n, c0, c1); g_push(Vor, mx, &n, g->xmin, c1);
      break;
    default:
      g_push(Vor, mx, &n, g->xmin, c1); g_push(Vor, mx, &n, c0, c1); g_push(Vor, mx, &n, c0, g->ymax);
      break;
  }
  return n;
}

/* :219-252 */
int om_find_Voronoi_cell_vertices(const om_geom *g, int vi, double *Vor, int max_vor, int *outside) {
  const int ei = E1(g->edge_index, vi);
  int n;
  if (ei == 0) n = g_vor_free(g, vi, Vor, max_vor);
  else if (ei == 2 || ei == 4 || ei == 6 || ei == 8) n = g_vor_corner(g, vi, Vor, max_vor);
  else n = g_vor_edge(g, vi, Vor, max_vor);
  int out = 0;
  for (int k = 0; k < n && k < max_vor; k++) {
    double x = Vor[2 * k], y = Vor[2 * k + 1];
    if (x < g->xmin - g->tol_dist || x > g->xmax + g->tol_dist || y < g->ymin - g->tol_dist ||
        y > g->ymax + g->tol_dist)
      out++;
    Vor[2 * k] = dmax(dmin(x, g->xmax), g->xmin);
    Vor[2 * k + 1] = dmax(dmin(y, g->ymax), g->ymin);
  }
  if (outside) *outside = out;
  return n;
}

int om_find_Voronoi_cell_areas(const om_geom *g, double *A) {
  const int mx = 2 * g->nC_mem + 4;
  int outside_total = 0;
  if (g->nC_mem > 64) return -1;
#pragma omp parallel for schedule(static) reduction(+ : outside_total)
  for (int vi = 1; vi <= g->nV; vi++) {
    double Vor[2 * (2 * 64 + 4)];
    int out = 0;
    const int n = om_find_Voronoi_cell_vertices(g, vi, Vor, mx, &out);
    outside_total += out + (n > g->nC_mem + 2);   /* the reference's Vor holds nC_mem+2 points */
    const double x0 = GV(vi, 1), y0 = GV(vi, 2);
    double a = 0.0;
    for (int k = 1; k < n && k < mx; k++) {
      const double ax = Vor[2 * k] - x0, ay = Vor[2 * k + 1] - y0;
      const double bx = Vor[2 * k - 2] - x0, by = Vor[2 * k - 1] - y0;
      a = a + fabs((ax * by) - (ay * bx)) / 2.0;
    }
    E1(A, vi) = a;
  }
  return outside_total;
}

int om_find_connection_widths(const om_geom *g, double *Cw) {
  int fails = 0;
  for (size_t k = 0; k < (size_t)g->nV * g->nC_mem; k++) Cw[k] = 0.0;
  for (int v1 = 1; v1 <= g->nV; v1++) {
    for (int nv2 = 1; nv2 <= E1(g->nC, v1); nv2++) {
      const int v2 = E2(g->C, g->nV, v1, nv2);
      int t1 = 0, t2 = 0;
      for (int iti = 1; iti <= E1(g->niTri, v1); iti++) {
        const int ti = GiTri(v1, iti);
        const int has = GTri(ti, 1) == v2 || GTri(ti, 2) == v2 || GTri(ti, 3) == v2;
        if (has) {
          if (t1 == 0) t1 = ti;
          else t2 = ti;
        }
      }
      double w = 0.0;
      if (t1 == 0) {
        fails++;
      } else if (t2 > 0) {
        const double dx = GTricc(t1, 1) - GTricc(t2, 1), dy = GTricc(t1, 2) - GTricc(t2, 2);
        w = sqrt(dx * dx + dy * dy);   /* NORM2 */
      } else {
        const int te = E1(g->Tri_edge_index, t1);
        if (te == 1) w = dmax(0.0, g->ymax - GTricc(t1, 2));
        else if (te == 3) w = dmax(0.0, g->xmax - GTricc(t1, 1));
        else if (te == 5) w = dmax(0.0, GTricc(t1, 2) - g->ymin);
        else if (te == 7) w = dmax(0.0, GTricc(t1, 1) - g->xmin);
        else fails++;
      }
      E2(Cw, g->nV, v1, nv2) = w;
    }
  }
  return fails;
}

void om_find_triangle_areas(const om_geom *g, double *TriA) {
  for (int ti = 1; ti <= g->nTri; ti++) {
    const double qx = GV(GTri(ti, 1), 1), qy = GV(GTri(ti, 1), 2);
    const double ax = GV(GTri(ti, 2), 1) - qx, ay = GV(GTri(ti, 2), 2) - qy;
    const double bx = GV(GTri(ti, 3), 1) - qx, by = GV(GTri(ti, 3), 2) - qy;
    E1(TriA, ti) = fabs((ax * by) - (ay * bx)) / 2.0;
  }
}

void om_determine_mesh_resolution(const om_geom *g, double *R, double *rmin, double *rmax) {
  for (int vi = 1; vi <= g->nV; vi++) {
    double r = g->xmax - g->xmin;
    for (int ci = 1; ci <= E1(g->nC, vi); ci++) {
      const int vj = E2(g->C, g->nV, vi, ci);
      const double dx = GV(vj, 1) - GV(vi, 1), dy = GV(vj, 2) - GV(vi, 2);
      r = dmin(r, sqrt(dx * dx + dy * dy));
    }
    E1(R, vi) = r;
  }
  double lo = R[0], hi = R[0];
  for (int vi = 1; vi < g->nV; vi++) {
    lo = dmin(lo, R[vi]);
    hi = dmax(hi, R[vi]);
  }
  *rmin = lo;
  *rmax = hi;
}
