/*
 * fmm_oracle.h -- CPU oracle for the particle FMM path of fmmlib3d.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference algorithm (Laplace lfmm3dparttarg / l3dpartdirect and Helmholtz
 * hfmm3dparttarg / h3dpartdirect, src/lfmm3dpart.f, src/hfmm3dpart.f) used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product (fmmlib3d_amd) never links or calls it.
 *
 * Conventions (identical to the reference's "_trunc" routines, which the
 * reference FMM uses):
 *   Yhat_n^m(theta,phi) = (-1)^m sqrt((n-|m|)!/(n+|m|)!) P_n^|m|(cos theta)
 *                         e^{i m phi},  P without Condon-Shortley phase,
 *   regular   R_n^m(x) = |x|^n Yhat_n^m,   irregular S_n^m(x) = Yhat_n^m/|x|^{n+1},
 *   multipole M_n^m = rscale^n  sum_j q_j conj(R_n^m(y_j - c)),
 *   local     L_n^m = rscale^-n sum_j q_j conj(S_n^m(y_j - c)),
 * evaluated as sum M_n^m S_n^m / rscale^n and sum L_n^m R_n^m rscale^n.
 * Expansions are stored packed: coefficient (n,m) at index n*n+n+m.
 * The reference stores them as mpole(0:p,-p:p); tests convert.
 */
#ifndef FMM_ORACLE_H
#define FMM_ORACLE_H
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex zdouble;

/* ---- expansion orders (src/l3dterms.f, src/h3dterms.f) ---- */
int  orc_l3dterms(double eps, int *ier);
int  orc_l3dterms_eval(int itype, double eps, int *ier);
void orc_l3dterms_list2(double eps, int *itable /* 7*7*7, [i+3][j+3][k+3] */, int *ier);
int  orc_h3dterms(double size, zdouble zk, double eps, int *ier);
void orc_h3dterms_list2(double size, zdouble zk, double eps, int *itable, int *ier);

/* ---- solid harmonics (packed, n <= p) ---- */
void orc_regular(int p, const double *x, zdouble *R);
void orc_irregular(int p, const double *x, zdouble *S);

/* ---- Laplace operators (accumulate into the output expansion) ---- */
void orc_l3dformmp(double rscale, const double *src, const zdouble *q, int ns,
                   const double *center, int p, zdouble *mp);
void orc_l3dformmp_dp(double rscale, const double *src, const zdouble *ds,
                      const double *dv, int ns, const double *center, int p, zdouble *mp);
void orc_l3dformta(double rscale, const double *src, const zdouble *q, int ns,
                   const double *center, int p, zdouble *loc);
void orc_l3dformta_dp(double rscale, const double *src, const zdouble *ds,
                      const double *dv, int ns, const double *center, int p, zdouble *loc);
void orc_l3dmpeval(double rscale, const double *center, const zdouble *mp, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld);
void orc_l3dtaeval(double rscale, const double *center, const zdouble *loc, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld);
void orc_l3dmpmp(double sc1, const double *c1, const zdouble *mp1, int p1,
                 double sc2, const double *c2, zdouble *mp2, int p2);
void orc_l3dmploc(double sc1, const double *c1, const zdouble *mp, int p1, int ptrunc,
                  double sc2, const double *c2, zdouble *loc, int p2);
void orc_l3dlocloc(double sc1, const double *c1, const zdouble *loc1, int p1,
                   double sc2, const double *c2, zdouble *loc2, int p2);

/* ---- Helmholtz special functions and operators (fmm_oracle_helm.c; the
 * coefficient convention is stated there).  Expansions packed as above,
 * operators accumulate into their output. ---- */
void orc_h3dall(int n, zdouble z, double sc, zdouble *hs);      /* h_i(z) sc^i, i<=n */
int  orc_jfuns3d(int n, zdouble z, double sc, zdouble *fj);     /* j_i(z)/sc^i, i<=n */
void orc_h3dformmp(zdouble zk, double sc, const double *src, const zdouble *q, int ns,
                   const double *center, int p, zdouble *mp);
void orc_h3dformmp_dp(zdouble zk, double sc, const double *src, const zdouble *ds, const double *dv,
                      int ns, const double *center, int p, zdouble *mp);
void orc_h3dformta(zdouble zk, double sc, const double *src, const zdouble *q, int ns,
                   const double *center, int p, zdouble *loc);
void orc_h3dformta_dp(zdouble zk, double sc, const double *src, const zdouble *ds, const double *dv,
                      int ns, const double *center, int p, zdouble *loc);
void orc_h3dmpeval(zdouble zk, double sc, const double *center, const zdouble *mp, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld);
void orc_h3dtaeval(zdouble zk, double sc, const double *center, const zdouble *loc, int p,
                   const double *trg, int nt, int ifpot, zdouble *pot, int iffld, zdouble *fld);
void orc_h3dmpmp(zdouble zk, double sc1, const double *c1, const zdouble *mp1, int p1,
                 double sc2, const double *c2, zdouble *mp2, int p2);
void orc_h3dmploc(zdouble zk, double sc1, const double *c1, const zdouble *mp, int p1, int ptrunc,
                  double sc2, const double *c2, zdouble *loc, int p2);
void orc_h3dlocloc(zdouble zk, double sc1, const double *c1, const zdouble *loc1, int p1,
                   double sc2, const double *c2, zdouble *loc2, int p2);

/* rotation about y by theta of a packed expansion (coefficient convention above) */
void orc_rotate_y(int p, double theta, const zdouble *in, zdouble *out);

/* ---- direct sums (src/lfmm3dpart.f:1391, src/hfmm3dpart.f:1268) ---- */
void orc_l3dpartdirect(int ns, const double *src, int ifcharge, const zdouble *q,
                       int ifdipole, const zdouble *ds, const double *dv,
                       int ifpot, zdouble *pot, int iffld, zdouble *fld,
                       int nt, const double *trg, int ifpottarg, zdouble *pottarg,
                       int iffldtarg, zdouble *fldtarg);
/* Hessians (pot_xx, pot_yy, pot_zz, pot_xy, pot_xz, pot_yz) of the direct sum
 * at nt points (laprouts3d.f:5442 lpotfld3dhess, :5619 lpotfld3dhess_dp) */
void orc_l3ddirect_hess_sample(int ns, const double *src, int ifcharge, const zdouble *q,
                               int ifdipole, const zdouble *ds, const double *dv,
                               int nt, const double *trg, const int *self_index, zdouble *hess);
/* ---- quadrupole sources (fmm_oracle_qp.c; laprouts3d.f:8691 lpotfld3dhess_qp,
 * :7914 l3dformmp_qp_trunc, :8261 l3dformta_qp_trunc).  quadvec(6,n) orders
 * (xx, yy, zz, xy, xz, yz). ---- */
void orc_qp_ladder_regular(int n, int m, double *Lj);
void orc_qp_ladder_irregular(int n, int m, double *Lj);
void orc_qp_channels(zdouble quadstr, const double *qv, zdouble *qj);
void orc_l3dform_qp(int irreg, double rscale, const double *src, const zdouble *quadstr, const double *quadvec,
                    int ns, const double *center, int p, zdouble *out);
void orc_l3ddirect_qp_sample(int ns, const double *src, const zdouble *quadstr, const double *quadvec, int npts,
                             const double *pts, const int *self_index, zdouble *pot, zdouble *fld);
/* direct sum evaluated at a subset of targets/sources given by index lists */
void orc_l3ddirect_sample(int ns, const double *src, int ifcharge, const zdouble *q,
                          int ifdipole, const zdouble *ds, const double *dv,
                          int nt, const double *trg, const int *self_index,
                          zdouble *pot, zdouble *fld);
void orc_h3dpartdirect(zdouble zk, int ns, const double *src, int ifcharge, const zdouble *q,
                       int ifdipole, const zdouble *ds, const double *dv,
                       int ifpot, zdouble *pot, int iffld, zdouble *fld,
                       int nt, const double *trg, int ifpottarg, zdouble *pottarg,
                       int iffldtarg, zdouble *fldtarg);
void orc_h3ddirect_sample(zdouble zk, int ns, const double *src, int ifcharge, const zdouble *q,
                          int ifdipole, const zdouble *ds, const double *dv,
                          int nt, const double *trg, const int *self_index,
                          zdouble *pot, zdouble *fld);

/* ---- adaptive octree (src/d3tstrcr.f) ---- */
typedef struct {
  int level, ijk[3], parent, child[8], nchild;
  int isrc, nsrc, itrg, ntrg;      /* 0-based start into the sorted order */
  double center[3];
} orc_box;

typedef struct {
  int nboxes, nlev;
  double center[3], size;
  orc_box *boxes;
  int *isrc, *itrg;                /* sorted -> original index (0-based) */
  int *ladd;                       /* first box of each level, nlev+2 entries */
  int *lst_off[5], *lst[5];        /* CSR lists 1..5 (index 0 = list 1) */
} orc_tree;

int  orc_tree_build(const double *src, int ns, const double *trg, int nt, int nbox,
                    orc_tree *t);
void orc_tree_free(orc_tree *t);
/* flattened accessors for tests: boxes as 20 ints like d3tgetb (1-based ids) */
int  orc_tree_summary(const double *src, int ns, const double *trg, int nt, int nbox,
                      int *nboxes, int *nlev, int *boxes20, int maxboxes,
                      int *isrc, int *itrg, int *list_counts /* 5*maxboxes */,
                      int *lists /* concatenated, sorted ascending per box */, int maxlist);

/* ---- FMM drivers (reference calling sequences, 0 = ok) ---- */
int orc_lfmm3dparttarg(int iprec, int ns, const double *src,
                       int ifcharge, const zdouble *q, int ifdipole, const zdouble *ds,
                       const double *dv, int ifpot, zdouble *pot, int iffld, zdouble *fld,
                       int nt, const double *trg, int ifpottarg, zdouble *pottarg,
                       int iffldtarg, zdouble *fldtarg);
int orc_hfmm3dparttarg(int iprec, zdouble zk, int ns, const double *src,
                       int ifcharge, const zdouble *q, int ifdipole, const zdouble *ds,
                       const double *dv, int ifpot, zdouble *pot, int iffld, zdouble *fld,
                       int nt, const double *trg, int ifpottarg, zdouble *pottarg,
                       int iffldtarg, zdouble *fldtarg);

/* ---- Laplace layer potentials of flat triangles (fmm_oracle_tria.c) ---- */
/* unit-density SLP pot, SLP fld[3], DLP pot, DLP fld[3] of one triangle
 * (vertices xyz, 9 doubles) at x; self != 0 for the triangle's own centroid */
void orc_tri_unit(const double *tri, const double *x, int self, double *out);
void orc_l3dtriadirect(int ntri, const double *tri, const double *cent, int ifcharge,
                       const zdouble *charge, int ifdipole, const zdouble *dipstr, int ifpot,
                       zdouble *pot, int iffld, zdouble *fld, int ntarget, const double *target,
                       int ifpottarg, zdouble *pottarg, int iffldtarg, zdouble *fldtarg);
void orc_l3dtriadirect_sample(int ntri, const double *tri, int ifcharge, const zdouble *charge,
                              int ifdipole, const zdouble *dipstr, int npts, const double *pts,
                              const int *self_index, zdouble *pot, zdouble *fld);

/* ---- Helmholtz layer potentials of flat triangles (fmm_oracle_htria.c) ---- */
/* D3-symmetric rule of degree 1..6 on the reference triangle (weights sum
 * to 1/2); returns the node count (at most 12), 0 for other degrees */
int  orc_tri_symrule(int degree, double *xi, double *eta, double *w);
/* unit-density SLP pot, fld[3], DLP pot, fld[3] (complex) of one triangle at x */
void orc_htri_unit(const double *tri, const double *x, int self, zdouble zk, int nq, const double *rxi,
                   const double *reta, const double *rw, int lap_c, int lap_d, zdouble *out);
void orc_h3dtriadirect_sample(int ntri, const double *tri, zdouble zk, int nqtri, int ifcharge,
                              const zdouble *charge, int ifdipole, const zdouble *dipstr, int npts,
                              const double *pts, const int *self_index, zdouble *pot, zdouble *fld);

#ifdef __cplusplus
}
#endif
#endif
