/*
 * fmm_oracle_lfmm.c -- CPU oracle restatement of the Laplace particle FMM
 * driver lfmm3dparttarg (src/lfmm3dpart.f:138-522) and its schedule
 * lfmm3dparttargmain (src/lfmm3dpart.f:528-1169) + lfmm3d_list2
 * (src/lfmm3drouts.f:148-566).  TEST INFRASTRUCTURE ONLY.
 *
 * Same tree, same expansion orders (l3dterms at every level, list-2
 * truncation from l3dterms_list2), same eight steps with the same pruning
 * rules; expansions carry rscale = 1/box side at each level (the reference
 * uses 1, lfmm3dpart.f:338-341; the scale cancels mathematically).
 */
#include "fmm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(n, m) ((n) * (n) + (n) + (m))

/* lfmm3dpart.f:277-300 */
int orc_iprec_params(int iprec, int ns, int nt, double *eps, int *nbox) {
  static const double e[] = {.5e0, .5e-1, .5e-2, .5e-3, .5e-6, .5e-9, .5e-12, .5e-15, 0.0};
  static const int nb[] = {40, 50, 80, 160, 400, 800, 1200, 1400, 0};
  if (iprec < -2 || iprec > 6) return 0;
  *eps = e[iprec + 2];
  *nbox = (iprec == 6) ? ns + nt : nb[iprec + 2];
  return 1;
}

/* lfmm3dpart.f:1175-1287 lfmm3dpart_direct_self + :1293-1385 _direct_targ:
 * sources [s0,s0+sn) act on sources [t0,t0+tn) (skipping equal indices) and
 * on targets [u0,u0+un). */
static void p2p_block(const double *src, int ifcharge, const zdouble *q, int ifdipole,
                      const zdouble *ds, const double *dv, int s0, int sn,
                      int t0, int tn, int ifpot, zdouble *pot, int iffld, zdouble *fld,
                      const double *trg, int u0, int un, int ifpottarg, zdouble *pottarg,
                      int iffldtarg, zdouble *fldtarg) {
  for (int pass = 0; pass < 2; pass++) {
    int b0 = pass ? u0 : t0, bn = pass ? un : tn;
    int wp = pass ? ifpottarg : ifpot, wf = pass ? iffldtarg : iffld;
    if (!wp && !wf) continue;
    const double *P = pass ? trg : src;
    zdouble *po = pass ? pottarg : pot, *fo = pass ? fldtarg : fld;
    for (int j = b0; j < b0 + bn; j++) {
      zdouble ph = 0, f[3] = {0, 0, 0};
      const double *t = P + 3 * j;
      for (int i = s0; i < s0 + sn; i++) {
        if (!pass && i == j) continue;
        const double *s = src + 3 * i;
        double dx = t[0] - s[0], dy = t[1] - s[1], dz = t[2] - s[2];
        double dd = dx * dx + dy * dy + dz * dz, d = sqrt(dd);
        double dinv = 1.0 / d, dinv2 = dinv * dinv, dinv3 = dinv * dinv2;
        if (ifcharge) {
          ph += q[i] * dinv;
          f[0] += q[i] * dx * dinv3; f[1] += q[i] * dy * dinv3; f[2] += q[i] * dz * dinv3;
        }
        if (ifdipole) {
          const double *v = dv + 3 * i;
          double dot = dx * v[0] + dy * v[1] + dz * v[2], rt = 3.0 * dot * dinv3 * dinv2;
          ph += ds[i] * (dot * dinv3);
          f[0] += ds[i] * (rt * dx - dinv3 * v[0]);
          f[1] += ds[i] * (rt * dy - dinv3 * v[1]);
          f[2] += ds[i] * (rt * dz - dinv3 * v[2]);
        }
      }
      if (wp) po[j] += ph;
      if (wf) for (int k = 0; k < 3; k++) fo[3 * j + k] += f[k];
    }
  }
}

int orc_lfmm3dparttarg(int iprec, int ns, const double *src_in,
                       int ifcharge, const zdouble *q_in, int ifdipole, const zdouble *ds_in,
                       const double *dv_in, int ifpot, zdouble *pot_out, int iffld, zdouble *fld_out,
                       int nt, const double *trg_in, int ifpottarg, zdouble *pottarg_out,
                       int iffldtarg, zdouble *fldtarg_out) {
  double eps;
  int nbox;
  if (!orc_iprec_params(iprec, ns, nt, &eps, &nbox)) return 32;
  if (nt < 0) nt = 0;
  orc_tree t;
  if (ns < 1 || orc_tree_build(src_in, ns, trg_in, nt, nbox, &t) != 0) return 4;
  const int nb = t.nboxes, nlev = t.nlev;

  /* l3dreorder / l3dreordertarg (lfmm3drouts.f:667-708) */
  double *src = malloc(sizeof(double) * 3 * ns), *trg = malloc(sizeof(double) * 3 * (nt + 1));
  zdouble *q = calloc(ns, sizeof(zdouble)), *ds = calloc(ns, sizeof(zdouble));
  double *dv = calloc(3 * ns, sizeof(double));
  for (int i = 0; i < ns; i++) {
    int o = t.isrc[i];
    memcpy(src + 3 * i, src_in + 3 * o, sizeof(double) * 3);
    if (ifcharge) q[i] = q_in[o];
    if (ifdipole) { ds[i] = ds_in[o]; memcpy(dv + 3 * i, dv_in + 3 * o, sizeof(double) * 3); }
  }
  for (int i = 0; i < nt; i++) memcpy(trg + 3 * i, trg_in + 3 * t.itrg[i], sizeof(double) * 3);
  zdouble *pot = calloc(ns, sizeof(zdouble)), *fld = calloc(3 * ns, sizeof(zdouble));
  zdouble *pottarg = calloc(nt + 1, sizeof(zdouble)), *fldtarg = calloc(3 * (nt + 1), sizeof(zdouble));

  int ier;
  int p = orc_l3dterms(eps, &ier);                 /* lfmm3dpart.f:401-407 */
  int itable[343];
  orc_l3dterms_list2(eps, itable, &ier);
  int K = (p + 1) * (p + 1);
  zdouble *mp = calloc((size_t)nb * K, sizeof(zdouble)), *lo = calloc((size_t)nb * K, sizeof(zdouble));
  double *scale = malloc(sizeof(double) * (nlev + 1));
  for (int l = 0; l <= nlev; l++) scale[l] = pow(2.0, l) / t.size;
  const orc_box *B = t.boxes;

  /* STEP 1: P2M at leaves with sources (lfmm3dpart.f:700-769) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    if (B[b].nsrc == 0 || B[b].nchild) continue;
    double sc = scale[B[b].level];
    if (ifcharge) orc_l3dformmp(sc, src + 3 * B[b].isrc, q + B[b].isrc, B[b].nsrc, B[b].center, p, mp + (size_t)b * K);
    if (ifdipole) orc_l3dformmp_dp(sc, src + 3 * B[b].isrc, ds + B[b].isrc, dv + 3 * B[b].isrc, B[b].nsrc, B[b].center, p, mp + (size_t)b * K);
  }
  /* STEP 2: P2L from list 4 (lfmm3dpart.f:788-869) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    double sc = scale[B[b].level];
    for (int i = t.lst_off[3][b]; i < t.lst_off[3][b + 1]; i++) {
      int j = t.lst[3][i];
      if (B[j].nsrc == 0) continue;
      if (ifcharge) orc_l3dformta(sc, src + 3 * B[j].isrc, q + B[j].isrc, B[j].nsrc, B[b].center, p, lo + (size_t)b * K);
      if (ifdipole) orc_l3dformta_dp(sc, src + 3 * B[j].isrc, ds + B[j].isrc, dv + 3 * B[j].isrc, B[j].nsrc, B[b].center, p, lo + (size_t)b * K);
    }
  }
  /* STEP 3: M2M, parents at levels nlev-1..2 (lfmm3drouts.f:195-258) */
  for (int lev = nlev - 1; lev >= 2; lev--) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = t.ladd[lev]; b < t.ladd[lev + 1]; b++) {
      if (B[b].nsrc == 0 || B[b].nchild == 0) continue;
      memset(mp + (size_t)b * K, 0, sizeof(zdouble) * K);
      for (int k = 0; k < B[b].nchild; k++) {
        int c = B[b].child[k];
        orc_l3dmpmp(scale[lev + 1], B[c].center, mp + (size_t)c * K, p, scale[lev], B[b].center, mp + (size_t)b * K, p);
      }
    }
  }
  /* STEP 4: M2L over list 2, levels 2..nlev (lfmm3drouts.f:347-483) */
  int prune = !(ifpot || iffld);
  for (int lev = 2; lev <= nlev; lev++) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = t.ladd[lev]; b < t.ladd[lev + 1]; b++) {
      for (int i = t.lst_off[1][b]; i < t.lst_off[1][b + 1]; i++) {
        int j = t.lst[1][i];
        if (B[j].nsrc == 0) continue;
        if (B[b].ntrg == 0 && prune) continue;
        int di = B[j].ijk[0] - B[b].ijk[0], dj = B[j].ijk[1] - B[b].ijk[1], dk = B[j].ijk[2] - B[b].ijk[2];
        int pt = itable[((di + 3) * 7 + (dj + 3)) * 7 + (dk + 3)];
        if (pt > p) pt = p;
        orc_l3dmploc(scale[lev], B[j].center, mp + (size_t)j * K, p, pt, scale[lev], B[b].center, lo + (size_t)b * K, p);
      }
    }
  }
  /* STEP 5: L2L, parents at levels 2..nlev-1 (lfmm3drouts.f:498-558) */
  for (int lev = 2; lev <= nlev - 1; lev++) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = t.ladd[lev]; b < t.ladd[lev + 1]; b++) {
      for (int k = 0; k < B[b].nchild; k++) {
        int c = B[b].child[k];
        orc_l3dlocloc(scale[lev], B[b].center, lo + (size_t)b * K, p, scale[lev + 1], B[c].center, lo + (size_t)c * K, p);
      }
    }
  }
  /* STEP 6: M2P from list 3 (lfmm3dpart.f:907-984) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    for (int i = t.lst_off[2][b]; i < t.lst_off[2][b + 1]; i++) {
      int j = t.lst[2][i];
      double sc = scale[B[j].level];
      if (B[b].nsrc > 0)
        orc_l3dmpeval(sc, B[j].center, mp + (size_t)j * K, p, src + 3 * B[b].isrc, B[b].nsrc,
                      ifpot, pot + B[b].isrc, iffld, fld + 3 * B[b].isrc);
      if (B[b].ntrg > 0)
        orc_l3dmpeval(sc, B[j].center, mp + (size_t)j * K, p, trg + 3 * B[b].itrg, B[b].ntrg,
                      ifpottarg, pottarg + B[b].itrg, iffldtarg, fldtarg + 3 * B[b].itrg);
    }
  }
  /* STEP 7: L2P at leaves of level >= 2 (lfmm3dpart.f:1003-1056) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    if (B[b].nchild || B[b].level < 2) continue;
    double sc = scale[B[b].level];
    orc_l3dtaeval(sc, B[b].center, lo + (size_t)b * K, p, src + 3 * B[b].isrc, B[b].nsrc,
                  ifpot, pot + B[b].isrc, iffld, fld + 3 * B[b].isrc);
    orc_l3dtaeval(sc, B[b].center, lo + (size_t)b * K, p, trg + 3 * B[b].itrg, B[b].ntrg,
                  ifpottarg, pottarg + B[b].itrg, iffldtarg, fldtarg + 3 * B[b].itrg);
  }
  /* STEP 8: P2P, self + list 1 (lfmm3dpart.f:1076-1141) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    if (B[b].nchild) continue;
    if (B[b].nsrc > 0)
      p2p_block(src, ifcharge, q, ifdipole, ds, dv, B[b].isrc, B[b].nsrc, B[b].isrc, B[b].nsrc,
                ifpot, pot, iffld, fld, trg, B[b].itrg, B[b].ntrg, ifpottarg, pottarg, iffldtarg, fldtarg);
    for (int i = t.lst_off[0][b]; i < t.lst_off[0][b + 1]; i++) {
      int j = t.lst[0][i];
      if (B[j].nsrc == 0) continue;
      p2p_block(src, ifcharge, q, ifdipole, ds, dv, B[j].isrc, B[j].nsrc, B[b].isrc, B[b].nsrc,
                ifpot, pot, iffld, fld, trg, B[b].itrg, B[b].ntrg, ifpottarg, pottarg, iffldtarg, fldtarg);
    }
  }
  /* l3dpsort / l3dfsort (lfmm3drouts.f:583-615): scatter back */
  for (int i = 0; i < ns; i++) {
    int o = t.isrc[i];
    if (ifpot) pot_out[o] = pot[i];
    if (iffld) for (int k = 0; k < 3; k++) fld_out[3 * o + k] = fld[3 * i + k];
  }
  for (int i = 0; i < nt; i++) {
    int o = t.itrg[i];
    if (ifpottarg) pottarg_out[o] = pottarg[i];
    if (iffldtarg) for (int k = 0; k < 3; k++) fldtarg_out[3 * o + k] = fldtarg[3 * i + k];
  }
  free(src); free(trg); free(q); free(ds); free(dv); free(pot); free(fld); free(pottarg); free(fldtarg);
  free(mp); free(lo); free(scale);
  orc_tree_free(&t);
  return 0;
}
