/*
 * fmm3d_mi355x.h -- C ABI of libfmm3d_mi355x.so, the MI355X (gfx950) particle
 * FMM that replaces fmmlib3d's lfmm3dpart* / hfmm3dpart* / l3dpartdirect /
 * h3dpartdirect.
 *
 * Calling conventions mirror the reference routines one for one:
 *  - arrays use the reference's Fortran layouts: source(3,n) and target(3,n)
 *    are n consecutive (x,y,z) doubles, charge/dipstr/pot are complex*16
 *    (interleaved re,im doubles), dipvec(3,n) doubles, fld(3,n) complex*16
 *    (x,y,z components of a point consecutive);
 *  - flags are 0/1 exactly as in the reference; outputs are written only when
 *    their flag is 1 and are fully overwritten (not accumulated);
 *  - arrays whose flag is 0 may be NULL; a NULL array that the flags ask
 *    for (source, charge with ifcharge=1, dipstr/dipvec with ifdipole=1,
 *    target with ntarget>0, an output whose flag is 1) returns 128;
 *  - the return value is the reference's `ier`: 0 ok, 4 tree failure (also
 *    returned for nsource < 1, as the reference does through d3tstrcr's
 *    ier=128 and ten retries), 8 bulk workspace, 16 expansion workspace.
 *    Extensions: 32 invalid iprec (the reference leaves it undefined),
 *    64 HIP runtime error, 128 invalid argument (negative sizes, zk == 0).
 * All pointers are host pointers unless the name ends in _dev.
 * The functions are thread safe; calls on one device are serialised.
 */
#ifndef FMM3D_MI355X_H
#define FMM3D_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* Precision modes for the *_ex / *_dev entry points (a bit mask).  Point
 * positions, offsets and output accumulation stay FP64 in every mode.  A
 * mode is honoured while FP32 can meet the tolerance of iprec: at iprec >= 3
 * (0.5e-9 and tighter, src/lfmm3dpart.f:277-285) every mode runs FP64
 * throughout; at iprec 2 (0.5e-6) the FP32 expansions of FAR32 and FP32
 * run only on an octree whose leaves all lie on one level, and in FP64 on
 * an adaptive one (boxes of different sizes evaluate expansions near their
 * radius of convergence, where FP32 rounding reached 6e-7).  In the FP32
 * near field the closest pairs (under 1/32 of the leaf side for Laplace,
 * 1/8 for Helmholtz) run in FP64.
 * fmm3d_last_precision() tells which mode a call ran in. */
#define FMM3D_PREC_FP64 0      /* every stage in FP64 (the default)                 */
#define FMM3D_PREC_MIXED 1     /* FP32 near-field pair arithmetic, FP64 far field   */
#define FMM3D_PREC_FAR32 2     /* FP32 expansions and translations, FP64 near field */
#define FMM3D_PREC_FP32 3      /* FP32 near and far field                          */

/* Replaces lfmm3dparttarg, src/lfmm3dpart.f:138-141. */
int fmm3d_lfmm3dparttarg(int iprec, int nsource, const double *source,
                         int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces lfmm3dpartself, src/lfmm3dpart.f:87-89 (and lfmm3dpart, :36-38). */
int fmm3d_lfmm3dpartself(int iprec, int nsource, const double *source,
                         int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld);

/* Replaces l3dpartdirect, src/lfmm3dpart.f:1391-1394 (O(N^2) direct sum). */
void fmm3d_l3dpartdirect(int nsource, const double *source,
                         int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* fmm3d_lfmm3dparttarg with Hessians: hess(6,nsource) and hesstarg(6,ntarget)
 * complex*16, per point (pot_xx, pot_yy, pot_zz, pot_xy, pot_xz, pot_yz) of the
 * same potential, written when ifhess / ifhesstarg is 1 (other outputs as
 * fmm3d_lfmm3dparttarg).  FP64 throughout.  Extension: the reference's
 * particle FMM returns pot and fld only; the quantities and their order are
 * those of its library routines lpotfld3dhess / lpotfld3dhess_dp
 * (src/laprouts3d.f:5442, :5619) and l3dtaevalhess / l3dmpevalhess (:5297,
 * :5229). */
int fmm3d_lfmm3dparttarg_hess(int iprec, int nsource, const double *source,
                              int ifcharge, const double *charge,
                              int ifdipole, const double *dipstr, const double *dipvec,
                              int ifpot, double *pot, int iffld, double *fld, int ifhess, double *hess,
                              int ntarget, const double *target,
                              int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg,
                              int ifhesstarg, double *hesstarg);

/* fmm3d_lfmm3dparttarg with quadrupole sources as well: when ifquad is 1,
 * quadstr(nsource) complex*16 and quadvec(6,nsource) real*8, ordered
 * (xx, yy, zz, xy, xz, yz), add at each point quadstr * sum_k quadvec_k
 * V_k(x - y), V_ab = d_a d_b (1/r) (mixed terms weighed once), and its
 * -gradient to fld.  FP64 throughout.  Extension: the reference's particle
 * FMM takes charges and dipoles only; the source and its expansions are
 * those of its library routines lpotfld3dhess_qp (src/laprouts3d.f:8691),
 * l3dformmp_qp_trunc (:7914) and l3dformta_qp_trunc (:8261). */
int fmm3d_lfmm3dparttarg_qp(int iprec, int nsource, const double *source,
                            int ifcharge, const double *charge,
                            int ifdipole, const double *dipstr, const double *dipvec,
                            int ifquad, const double *quadstr, const double *quadvec,
                            int ifpot, double *pot, int iffld, double *fld,
                            int ntarget, const double *target,
                            int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces lfmm3dtriatarg (src/lfmm3dtria.f:121-141): the triangle FMM.
 * Piecewise-constant single layer (charge) and dipole layer (dipstr) on
 * nsource flat triangles, Green's function 1/r without 1/(4 pi),
 * self-interactions included; pot/fld (= -gradient) at the points `source`
 * (the triangles' centroids, one per triangle, evaluated on their own
 * triangle in the plane) and at ntarget targets.  triaflat is (3,3,nsource)
 * in Fortran order (x,y,z of vertex 1, 2, 3 per triangle); the dipole layer
 * points along the normal of the vertex order, (v2-v1)x(v3-v1) -- the
 * reference requires dipvec (and trianorm) to be that normal and so do
 * these entries, which read neither.  Near field by the analytic
 * quadratures (direct3dtritarglaps2 / lapd2, direct3dtrilaps2 / lapd2,
 * src/triquadflatlib.f:147-818), far field through quadrature points of
 * each triangle.  FP64; leaf size as the reference's (lfmm3dtria.f:297-304).
 * Returns ier as fmm3d_lfmm3dparttarg. */
int fmm3d_lfmm3dtriatarg(int iprec, int nsource, const double *triaflat, const double *trianorm,
                         const double *source, int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces lfmm3dtriaself (src/lfmm3dtria.f:80): no targets. */
int fmm3d_lfmm3dtriaself(int iprec, int nsource, const double *triaflat, const double *trianorm,
                         const double *source, int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld);

/* Replaces lfmm3dtriampftarg (src/lfmm3dtria.f:1566): the triangle FMM with
 * outputs at the targets only. */
int fmm3d_lfmm3dtriampftarg(int iprec, int nsource, const double *triaflat, const double *trianorm,
                            const double *source, int ifcharge, const double *charge,
                            int ifdipole, const double *dipstr, const double *dipvec,
                            int ntarget, const double *target,
                            int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces l3dtriadirect (src/lfmm3dtria.f:1385): the O(N^2) triangle sum. */
void fmm3d_l3dtriadirect(int nsource, const double *triaflat, const double *trianorm,
                         const double *source, int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces hfmm3dtriatarg (src/hfmm3dtria.f:133-144): the Helmholtz
 * triangle FMM, Green's function e^{ikr}/r without 1/(4 pi), zk = (re, im),
 * arguments as fmm3d_lfmm3dtriatarg.  ifcharge / ifdipole = 2 subtract that
 * density's Laplace part (the difference kernel e^{ikr}/r - 1/r), as the
 * reference's flag 2.  Near field: the Laplace analytic quadratures plus a
 * D3-symmetric rule of degree 2, 4 or 6 by iprec on (e^{ikr}-1)/r
 * (direct3dtritarghelms3 / helmd3, direct3dtrihelms2 / helmd2,
 * src/triquadflatlib.f:819-1358; hfmm3dtria.f:730-751); far field through
 * quadrature points of each triangle.  FP64.  Returns ier as
 * fmm3d_hfmm3dparttarg (128 also for a density flag outside 0..2). */
int fmm3d_hfmm3dtriatarg(int iprec, const double *zk, int nsource, const double *triaflat,
                         const double *trianorm, const double *source, int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces hfmm3dtriaself (src/hfmm3dtria.f:92): no targets. */
int fmm3d_hfmm3dtriaself(int iprec, const double *zk, int nsource, const double *triaflat,
                         const double *trianorm, const double *source, int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld);

/* Replaces hfmm3dtriampftarg (src/hfmm3dtria.f:1668): outputs at the
 * targets only. */
int fmm3d_hfmm3dtriampftarg(int iprec, const double *zk, int nsource, const double *triaflat,
                            const double *trianorm, const double *source, int ifcharge, const double *charge,
                            int ifdipole, const double *dipstr, const double *dipvec,
                            int ntarget, const double *target,
                            int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces h3dtriadirect (src/hfmm3dtria.f:1456): the O(N^2) Helmholtz
 * triangle sum with the symmetric rule of degree nqtri (1..6; the
 * reference's triasymq rules, degree 3 being its degree-4 rule).  Like the
 * reference it has no error channel: an unsupported nqtri, a density flag
 * outside 0..2 or a missing array computes nothing. */
void fmm3d_h3dtriadirect(int nqtri, const double *zk, int nsource, const double *triaflat,
                         const double *trianorm, const double *source, int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Same as fmm3d_lfmm3dparttarg with an explicit precision mode (FMM3D_PREC_*).
 * Extension with no reference counterpart: the reference has FP64 only. */
int fmm3d_lfmm3dparttarg_ex(int prec_mode, int iprec, int nsource, const double *source,
                            int ifcharge, const double *charge,
                            int ifdipole, const double *dipstr, const double *dipvec,
                            int ifpot, double *pot, int iffld, double *fld,
                            int ntarget, const double *target,
                            int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Device-resident variant of lfmm3dparttarg: every array is a device pointer
 * in the same layout; `stream` is a hipStream_t (NULL = default stream).
 * The call returns when the results are complete on the device.  The octree
 * is built on the device from the device coordinates (only per-box
 * histograms come back to the host, which decides the boxes and builds the
 * interaction lists); the coordinates themselves never leave HBM.
 * Extension with no reference counterpart, for callers whose data already
 * lives in HBM (benchmarks, the serving loop). */
int fmm3d_lfmm3dparttarg_dev(int prec_mode, int iprec, int nsource, const double *source_dev,
                             int ifcharge, const double *charge_dev,
                             int ifdipole, const double *dipstr_dev, const double *dipvec_dev,
                             int ifpot, double *pot_dev, int iffld, double *fld_dev,
                             int ntarget, const double *target_dev,
                             int ifpottarg, double *pottarg_dev, int iffldtarg, double *fldtarg_dev,
                             void *stream);

/* Size of the device array behind fmm3d_lfmm3dparttarg_dev_noalloc (256 MiB;
 * each GPU the entry point runs on holds one until the process ends). */
#define FMM3D_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* fmm3d_lfmm3dparttarg_dev without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc).  Arguments and
 * results are those of fmm3d_lfmm3dparttarg_dev (and bitwise equal to them);
 * every work buffer of the call (sorted points, expansions, plan, tables,
 * sums) is a slice of one static device array of
 * FMM3D_NOALLOC_SCRATCH_BYTES, so the call makes no hipMalloc, hipFree or
 * hipHostMalloc (host staging is ordinary heap memory, and the near field
 * stays on `stream`).  A problem whose buffers do not fit returns 256; as a
 * rule of thumb (nsource + ntarget) up to about 4e5 fit at iprec 2, and
 * any nsource + ntarget above 1,048,576 is rejected before device work
 * starts.  The array keeps the last call's contents; calls on one GPU are
 * serialised by the library within a process, and calls from two processes
 * on one GPU use separate arrays.  Like the other entry points, the call
 * synchronises with the host (tree decisions and lists are made there), so
 * it cannot be captured in a HIP graph. */
int fmm3d_lfmm3dparttarg_dev_noalloc(int prec_mode, int iprec, int nsource, const double *source_dev,
                                     int ifcharge, const double *charge_dev,
                                     int ifdipole, const double *dipstr_dev, const double *dipvec_dev,
                                     int ifpot, double *pot_dev, int iffld, double *fld_dev,
                                     int ntarget, const double *target_dev,
                                     int ifpottarg, double *pottarg_dev, int iffldtarg,
                                     double *fldtarg_dev, void *stream);

/* Replaces hfmm3dparttarg, src/hfmm3dpart.f:141-144: Helmholtz kernel
 * e^{ik r}/r, zk = complex*16 wavenumber passed by reference as (re, im)
 * (zk = 0 returns 128).  Other arguments as fmm3d_lfmm3dparttarg. */
int fmm3d_hfmm3dparttarg(int iprec, const double *zk, int nsource, const double *source,
                         int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Replaces hfmm3dpartself, src/hfmm3dpart.f:90-92 (and hfmm3dpart, :39-41). */
int fmm3d_hfmm3dpartself(int iprec, const double *zk, int nsource, const double *source,
                         int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld);

/* Replaces h3dpartdirect, src/hfmm3dpart.f:1268-1271 (O(N^2) direct sum). */
void fmm3d_h3dpartdirect(const double *zk, int nsource, const double *source,
                         int ifcharge, const double *charge,
                         int ifdipole, const double *dipstr, const double *dipvec,
                         int ifpot, double *pot, int iffld, double *fld,
                         int ntarget, const double *target,
                         int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* fmm3d_hfmm3dparttarg with a precision mode: FMM3D_PREC_FP64, or
 * FMM3D_PREC_MIXED (FP32 near-field pair arithmetic, FP64 translations);
 * other modes return 128.  Extension with no reference counterpart. */
int fmm3d_hfmm3dparttarg_ex(int prec_mode, int iprec, const double *zk, int nsource,
                            const double *source, int ifcharge, const double *charge,
                            int ifdipole, const double *dipstr, const double *dipvec,
                            int ifpot, double *pot, int iffld, double *fld,
                            int ntarget, const double *target,
                            int ifpottarg, double *pottarg, int iffldtarg, double *fldtarg);

/* Device-resident variant of hfmm3dparttarg (device pointers, hipStream_t
 * `stream`; zk is a host pointer).  Extension with no reference counterpart. */
int fmm3d_hfmm3dparttarg_dev(int prec_mode, int iprec, const double *zk, int nsource,
                             const double *source_dev, int ifcharge, const double *charge_dev,
                             int ifdipole, const double *dipstr_dev, const double *dipvec_dev,
                             int ifpot, double *pot_dev, int iffld, double *fld_dev,
                             int ntarget, const double *target_dev,
                             int ifpottarg, double *pottarg_dev, int iffldtarg, double *fldtarg_dev,
                             void *stream);

/* Replaces d3tstrcr with its accessors d3tgetb/d3tgetl
 * (src/d3tstrcr.f:24-259, :580-689): builds the adaptive octree the FMM
 * uses (host code) and returns it flattened.  boxes20 is the d3tgetb
 * layout (level, i, j, k, parent, 8 children, source start, source count,
 * target start, target count, 3 unused; ids and starts 1-based),
 * isource/itarget the 1-based permutations, list_counts[k*maxboxes+b] the
 * length of list k+1 (k = 0..3) of box b and `lists` their entries (1-based,
 * ascending per box) concatenated list by list, box by box.  Call with
 * maxboxes = 0 to obtain nboxes and the total list length (*nlist) first.
 * Returns 0, 128 (nsource < 1) or 32 (arrays too small). */
int fmm3d_d3tstrcr(int nsource, const double *source, int ntarget, const double *target,
                   int nbox, int *nboxes, int *nlev, double *center, double *size, int *nlist,
                   int maxboxes, int *boxes20, int *isource, int *itarget,
                   int *list_counts, int maxlist, int *lists);

/* ---- Request server (extension with no reference counterpart; not part of
 * the drop-in contract).  A persistent loop that runs several evaluations at
 * once on the GPU: `nworkers` slots, each with its own HIP stream and
 * device arena (trees, expansions and tables stay resident between
 * requests); when a slot is free, the oldest queued request whose
 * device-memory estimate fits what is left of `mem_budget_bytes` starts
 * (smaller requests may pass a large one that does not fit yet, at most 8
 * times, after which the large one is served first).  Inputs are copied at
 * submission.  All functions are thread safe. */
void *fmm3d_server_create(int nworkers, long long mem_budget_bytes);
/* kernel 0 = Laplace (zk ignored), 1 = Helmholtz (zk = (re, im)); other
 * arguments as fmm3d_lfmm3dparttarg_ex.  Returns a request id >= 0, or
 * -ier for an invalid argument. */
long long fmm3d_server_submit(void *server, int kernel, int prec_mode, int iprec, const double *zk,
                              int nsource, const double *source, int ifcharge, const double *charge,
                              int ifdipole, const double *dipstr, const double *dipvec,
                              int ifpot, int iffld, int ntarget, const double *target,
                              int ifpottarg, int iffldtarg);
/* 1 when the request has finished (outputs copied, *ier and the latency from
 * submission to completion set; the id is released), 0 when it is still
 * pending (block = 0), -1 for an unknown id. */
int fmm3d_server_wait(void *server, long long id, int block, double *pot, double *fld,
                      double *pottarg, double *fldtarg, int *ier, double *latency_ms);
/* Queue length, busy workers, device-memory estimate in use, and the number
 * of requests served so far in fused batches (several small Laplace requests
 * with the same tolerance, precision and output flags, run as one forest in
 * one graph-captured launch sequence).  Any pointer may be NULL. */
void fmm3d_server_stats(void *server, long long *queued, int *running, long long *bytes_in_use,
                        long long *fused_requests);
/* Finishes the queued requests, then releases the workers and their memory. */
void fmm3d_server_destroy(void *server);

/* Per-stage timings (milliseconds, device time from HIP events) of the last
 * FMM call on this thread: [0] host time before the first kernel (tree and
 * box plan), [1] gather/upload, [2] P2M, [3] P2L, [4] M2M, [5] M2L, [6] L2L,
 * [7] L2P, [8] M2P, [9] P2P, [10] total, [11] tree build (within [0]),
 * [12] host time of the interaction lists and their plan (overlapped with
 * P2M and M2M, then with P2P, on the device).  P2P runs on a second stream
 * at the same time as the far-field stages, so [9] overlaps [2]-[8], and
 * those stage times include the sharing of the GPU; [10] is the end-to-end
 * time of the call.
 * Mirrors the reference's timeinfo(1..8) (src/lfmm3dpart.f:1155). */
void fmm3d_last_timings(double *ms, int n);

/* The FMM3D_PREC_* mode the last FMM call of the calling thread ran in
 * (FMM3D_PREC_FP64 when the requested mode fell back at iprec >= 3).
 * Extension with no reference counterpart. */
int fmm3d_last_precision(void);

/* Library version string. */
const char *fmm3d_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FMM3D_MI355X_H */
