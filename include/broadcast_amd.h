/*
 * broadcast_amd -- C ABI of the MI355X (gfx950) FP64 residual / JVP path of
 * BROADCAST's DNC5 finite-volume scheme.
 *
 * Every array argument is a DEVICE pointer (HIP global memory) laid out
 * exactly as the reference's Fortran dummies: column-major, lower bound 1-gh
 * on both axes, components as trailing planes.  With ni = im+2gh,
 * nj = jm+2gh:
 *   cell arrays  w, residu, wd, ...   ni x nj x 5      (float64)
 *   node arrays  x0, y0               (ni+1) x (nj+1)
 *   face arrays  nx, ny               (ni+1) x (nj+1) x 2
 *   vol, xc, yc                       ni x nj
 *   volf                              ni x nj x 2
 * Work is enqueued on the context's HIP stream (the Jacobian and VJP calls
 * also run their edge-row face kernels on three streams of the context that
 * wait for it and are joined back into it before the call returns, so every
 * result is ordered on the context's stream, also under stream capture);
 * calls return 0 on success and a negative code on error (bca_last_error()
 * gives the message).  The caller
 * owns every buffer; the context owns its scratch.  Calls on one context are
 * not thread-safe; separate contexts are independent.
 *
 * Side codes (loc): 0 = 'Ilo', 1 = 'Ihi', 2 = 'Jlo', 3 = 'Jhi'.
 * Boxes: {imin, jmin, imax, jmax} as the reference's interf(2,2)
 * (BROADCAST_npz.py:706-731).
 */
#ifndef BROADCAST_AMD_H
#define BROADCAST_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bca_ctx bca_ctx;

/* Thermodynamic and scheme constants in the positional order of
 * flux_num_dnc5_2d (srcfv/rhs/flux_num_dnc5.F90:7-8). */
typedef struct {
  double cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4;
} bca_phys;

/* Metrics produced by bca_geom (nx, ny, vol, volf; device pointers). */
typedef struct {
  const double *nx, *ny, *vol, *volf;
} bca_metrics;

/* One ghost-filling operation of a boundary-condition sequence. */
enum { BCA_BC_INLET = 0, BCA_BC_NOREF = 1, BCA_BC_EXTRAP2 = 2, BCA_BC_WALL_ADIA = 3, BCA_BC_JOIN = 4 };
typedef struct {
  int kind;            /* BCA_BC_* */
  int loc;             /* side code */
  int box[4];          /* interface box; for BCA_BC_JOIN the receiving box prr */
  int box2[4];         /* BCA_BC_JOIN: donor box prd */
  int tr[2];           /* BCA_BC_JOIN: index transform */
  const double *data;  /* INLET: field (lm x gh x 5); NOREF: wbd (lm x 5); device */
  int lm;              /* leading dimension of data */
  double gam;
} bca_bc_op;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart: holds the mesh dimensions, the stream and the
 * scratch planes that the reference keeps as automatic arrays
 * (srcfv/rhs/flux_num_dnc5.F90:29-37). */
int bca_ctx_create(bca_ctx **ctx, int im, int jm, int gh, void *hip_stream);
int bca_ctx_destroy(bca_ctx *ctx);
int bca_ctx_set_stream(bca_ctx *ctx, void *hip_stream);
const char *bca_last_error(void);
/* Number of bytes of device scratch the context currently holds. */
long bca_ctx_scratch_bytes(const bca_ctx *ctx);
/* Kernel timing with HIP events recorded on the context stream around each
 * pipeline launch (no reference counterpart; the reference times with wall
 * clocks, BROADCAST_npz.py:1003-1007).  enable resets the record; read sums
 * the launches of one kernel, named "k_face<f64>", "k_face<dual1>",
 * "k_face<dual5>", "k_prim<..>", "k_grad<..>", "k_balance<..>",
 * "k_balance_scatter<dual5>". */
int bca_profile_enable(bca_ctx *ctx, int on);
int bca_profile_read(bca_ctx *ctx, const char *kernel, double *total_ms, long *launches);

/* ---- f_geom -------------------------------------------------------------- */
/* computegeom_2d(x0,y0,nx,ny,xc,yc,vol,volf,im,jm,gh)
 * replaces srcfv/geom/computegeom.F90:3-104 (x0, y0 ghost nodes are
 * overwritten by extrapolation, as in the reference). */
int bca_geom(bca_ctx *ctx, double *x0, double *y0, double *nx, double *ny, double *xc,
             double *yc, double *vol, double *volf);
/* volf = 2/(vol_L + vol_R) on faces 1..im+1 x 1..jm+1 from the current vol
 * (the O-mesh recomputation of cylinder.py:524-527). */
int bca_volf(bca_ctx *ctx, const double *vol, double *volf);

/* ---- f_sch --------------------------------------------------------------- */
/* flux_num_dnc5_2d(residu,w,x0,y0,nx,ny,xc,yc,vol,volf,gh,cp,...,k4,im,jm)
 * replaces srcfv/rhs/flux_num_dnc5.F90:7-226.  Writes residu on the interior
 * only (ghosts untouched); x0, y0, xc, yc are not needed (never referenced by
 * the reference routine). */
int bca_residual_dnc5(bca_ctx *ctx, double *residu, const double *w, bca_metrics g,
                      const bca_phys *phys);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the residual of bca_residual_dnc5, same kernels and
 * results, without a context, enqueued on `hip_stream`.  Its scratch planes
 * (prim, grad, hn) live in one static device array of
 * BCA_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends; the
 * call allocates and frees nothing.  A mesh needing more scratch
 * (bca_residual_noalloc_bytes) is refused with -4 before any launch.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (two streams, two host threads) race and must be ordered by the caller. */
#define BCA_NOALLOC_SCRATCH_BYTES (64L * 1024L * 1024L)
long bca_residual_noalloc_bytes(int im, int jm, int gh);
int bca_residual_dnc5_noalloc(double *residu, const double *w, bca_metrics g, const bca_phys *phys,
                              int im, int jm, int gh, void *hip_stream);

/* ---- f_lin --------------------------------------------------------------- */
/* flux_num_dnc5_2d_d(residu,residud,w,wd,...)
 * replaces srcfv/tangent/flux_num_dnc5_d.f90:15-3870: residud = dR/dw . wd
 * on the interior (wd's ghosts as given).  The reference's tangent leaves
 * residu as given; here residu may be NULL (the drop-in passes NULL) or
 * receives R(w) from the same pass. */
int bca_jvp_dnc5(bca_ctx *ctx, double *residu, double *residud, const double *w,
                 const double *wd, bca_metrics g, const bca_phys *phys);

/* flux_num_dnc5_nowall_2d(residu,w,...) replaces
 * srcfv/rhs/flux_num_dnc5_nowall.F90:7-176 (the sensitivity card's scheme,
 * card_bl2d_fv_npz_sens.py:87): the DNC5 residual with interior formulas on
 * every face (no wall rows at j = 1), ghost layers closing the stencils. */
int bca_residual_dnc5_nowall(bca_ctx *ctx, double *residu, const double *w, bca_metrics g,
                             const bca_phys *phys);
/* flux_num_dnc5_nowall_2d_d(residu,residud,w,wd,...) replaces
 * srcfv/tangent/flux_num_dnc5_nowall_d.f90:15: residud = dR/dw . wd of the
 * variant above on the interior.  As the reference's tangent does, it leaves
 * residu as given (the argument may be NULL). */
int bca_jvp_dnc5_nowall(bca_ctx *ctx, double *residu, double *residud, const double *w,
                        const double *wd, bca_metrics g, const bca_phys *phys);

/* The 3rd-order scheme, on a context of gh = 2 (the reference's
 * gh = (order+1)/2, BROADCAST_npz.py:501-502): 2nd-order gradients, 4-point
 * centred Euler flux and predictor, 2nd-order viscous flux, one off-centred
 * Euler row next to the wall.
 * flux_num_dnc3_2d(residu,w,...) replaces srcfv/rhs/flux_num_dnc3.F90:7-219;
 * flux_num_dnc3_nowall_2d(residu,w,...) replaces
 * srcfv/rhs/flux_num_dnc3_nowall.F90:7 (interior formulas on every face). */
int bca_residual_dnc3(bca_ctx *ctx, double *residu, const double *w, bca_metrics g,
                      const bca_phys *phys);
int bca_residual_dnc3_nowall(bca_ctx *ctx, double *residu, const double *w, bca_metrics g,
                             const bca_phys *phys);
/* flux_num_dnc3_2d_d / flux_num_dnc3_nowall_2d_d(residu,residud,w,wd,...)
 * replace srcfv/tangent/flux_num_dnc3_d.f90:15 and
 * srcfv/tangent/flux_num_dnc3_nowall_d.f90:15: residud = dR/dw . wd on the
 * interior; residu is left as given, as by the reference (may be NULL). */
int bca_jvp_dnc3(bca_ctx *ctx, double *residu, double *residud, const double *w, const double *wd,
                 bca_metrics g, const bca_phys *phys);
int bca_jvp_dnc3_nowall(bca_ctx *ctx, double *residu, double *residud, const double *w,
                        const double *wd, bca_metrics g, const bca_phys *phys);

/* Any order of the family with or without wall rows, on a context of
 * gh >= (order + 1) / 2 (gh = 3 for order 5):
 * flux_num_dnc{order}[_nowall]_2d(residu,w,...) replaces
 * srcfv/rhs/flux_num_dnc{3,5,7,9}[_nowall].F90 (order 7: flux_num_dnc7.F90:
 * 74-176, 8-point centred Euler flux, 9-point predictor, 6th-order gradients,
 * off-centred Euler rows 2..4 (coefnearbnd_9p.F); order 9: flux_num_dnc9.F90,
 * 10-point Euler, 11-point predictor, 8th-order gradients, rows 2..5
 * (coefnearbnd_11p.F)); the wall variants need jm >= 9 (order 7) or 11
 * (order 9).  bca_jvp_dnc replaces the matching
 * srcfv/tangent/flux_num_dnc*_d.f90: residud only, residu left as given. */
int bca_residual_dnc(bca_ctx *ctx, int order, int nowall, double *residu, const double *w, bca_metrics g,
                     const bca_phys *phys);
int bca_jvp_dnc(bca_ctx *ctx, int order, int nowall, double *residu, double *residud, const double *w,
                const double *wd, bca_metrics g, const bca_phys *phys);

/* ---- pruned CSR of the colored COO ---------------------------------------- */
/* remove_zero_jac (BROADCAST_npz.py:129-135) + scipy.sparse.csr_matrix of
 * createMatPetscCSR (misc/PETSc_func.py:85) on a COO filled by the colored
 * assembly (nzones = 1, withjn as passed to it; nzones = 2 for the two-zone
 * loop; jac holds nbentry >= nzones*im*jm*(2gh+1)^2*25 values): entries with |jac| > mini, duplicate (row, column) pairs summed,
 * columns ascending in each row.  _count fills indptr (5*im*jm + 1 entries)
 * and returns nnz (it synchronizes the stream); _fill then writes indices and
 * data (nnz entries each) from that indptr. */
int bca_csr_colored_count(bca_ctx *ctx, const double *jac, long nbentry, int nzones, int withjn,
                          double mini, int *indptr, long *nnz);
int bca_csr_colored_fill(bca_ctx *ctx, const double *jac, long nbentry, int nzones, int withjn,
                         double mini, const int *indptr, int *indices, double *data);

/* flux_num_dnc5_2d_b(residu,residub,w,wb,...)
 * replaces srcfv/adjoint/flux_num_dnc5_b.f90:15-6676: wb = (dR/dw)^T residub,
 * with w's ghost frame as independent inputs (no BC adjoint), over the whole
 * array including the ghost frame; wb is overwritten, residub is consumed
 * (its interior is zeroed, as the reference does), residu is not touched.
 * The primal is recomputed; nothing is stored between calls. */
int bca_vjp_dnc5(bca_ctx *ctx, double *residub, double *wb, const double *w, bca_metrics g,
                 const bca_phys *phys);

/* ---- f_bnd (wd == NULL) and f_lin (wd != NULL: linearised BC, w and wd
 * both updated, as the TAPENADE _d routines) -------------------------------- */
/* bc_supandsubinlet_2d[_d]  srcfv/borders/bc_supandsubinlet.F90:2-171,
 *                           srcfv/tangent/bc_supandsubinlet_d.f90:11 */
int bca_bc_inlet(bca_ctx *ctx, double *w, double *wd, int loc, const int box[4],
                 const double *field, int lm, const double *nx, const double *ny, double gam);
/* bc_no_reflexion_2d[_d]    srcfv/borders/bc_no_reflexion.F90:8-152,
 *                           srcfv/tangent/bc_no_reflexion_d.f90:16 */
int bca_bc_noref(bca_ctx *ctx, double *w, double *wd, const double *wbd, int lm, int loc,
                 const int box[4], const double *nx, const double *ny, double gam);
/* bc_extrapolate_o2_2d[_d]  srcfv/borders/bc_extrapolate.F90:39-71,
 *                           srcfv/tangent/bc_extrapolateo2_d.f90:18 */
int bca_bc_extrap2(bca_ctx *ctx, double *w, double *wd, int loc, const int box[4], int em);
/* bc_wall_viscous_adia_2d[_d] srcfv/borders/bc_wall_viscous.F90:2-99,
 *                           srcfv/tangent/bc_wall_viscous_d.f90:11 */
int bca_bc_wall_adia(bca_ctx *ctx, double *w, double *wd, int loc, const int box[4], double gam);
/* jn_match_2d               srcfv/borders/jn_match.F90:3-66 (same array as
 * receiver and donor, as the O-mesh periodic join uses it); the array has
 * leading dimensions ni_arr x nj_arr (cells: ni x nj, faces: ni+1 x nj+1) and
 * em planes. */
int bca_bc_join(bca_ctx *ctx, double *a, const int prr[4], const int prd[4], const int tr[2],
                int em, int ni_arr, int nj_arr);

/* ---- f_norm -------------------------------------------------------------- */
/* compute_norml2inf(norm,ninf,rhs,im,jm,gh)  srcfv/norm.F90:34-77
 * compute_norml2(norm,nmoy,rhs,im,jm,gh)     srcfv/norm.F90:2-32
 * out[0..4] = norm, out[5..9] = ninf (or nmoy); out is HOST memory and the
 * call synchronises the context stream. */
int bca_norml2inf(bca_ctx *ctx, const double *rhs, double *out10);
int bca_norml2(bca_ctx *ctx, const double *rhs, double *out10);

/* ---- f_misc -------------------------------------------------------------- */
/* testvector(wd,m,l,k,gh,im,jm)  misc/ComputeJacobian.f90:357-374 */
int bca_seed(bca_ctx *ctx, double *wd, int m, int l, int k);
/* computejacobianfromjv[_relaxed][_withjn](jac,ia,ja,resd,m,l,k,gh,im,jm,
 * nbentry[,coefdiag])  misc/ComputeJacobian.f90:292-355, 503-570, 847-926:
 * coefdiag == NULL selects the plain variant, withjn the join-aware one.
 * jac/ia/ja have nbentry = im*jm*(2gh+1)^2*25 entries; coefdiag is im x jm. */
int bca_scatter_jv(bca_ctx *ctx, double *jac, int *ia, int *ja, const double *resd, int m, int l,
                   int k, const double *coefdiag, int withjn, long nbentry);
/* set_bndbl_2d(w,field,wbd,im,jm,gh)  set_bnd.f90:2-24 */
int bca_set_bndbl(bca_ctx *ctx, const double *w, double *field, double *wbd);

/* testvector_partial(wd,m,l,k,gh,im,jm,istart,iend,jstart,jend)
 * misc/ComputeJacobian.f90:1075-1092: the seeds of (m, l, k) restricted to the
 * 0-based index box [istart, iend] x [jstart, jend]. */
int bca_seed_partial(bca_ctx *ctx, double *wd, int m, int l, int k, int istart, int iend, int jstart,
                     int jend);
/* computejacobianfromjv_relaxed_withjnandcheck(jac,ia,ja,resd,m,l,k,gh,im,jm,
 * nbentry,coefdiag,mini,n)  misc/ComputeJacobian.f90:1095-1204: the two-zone
 * periodic scatter of zone n (0 or 1) into slot block n of a COO of
 * nbentry = 2*im*jm*(2gh+1)^2*25 entries; slots holding |jac| >= mini are kept
 * or reset to the dummy row 5*im*jm-2 as the reference does. */
int bca_scatter_jv_2zones(bca_ctx *ctx, double *jac, int *ia, int *ja, const double *resd, int m, int l,
                          int k, const double *coefdiag, double mini, int n, long nbentry);

/* ---- spanwise operators (3D linearisation) and J(beta) -------------------- */
/* coeffs_5p_dz(dz,w,wd,...) srcfv/dz/coeffs_5p_dz.F90:10-174 and
 * coeffs_5p_dz2(dz2,w,wd,...) srcfv/dz/coeffs_5p_dz2.F90: the interior of dz
 * and/or dz2 (either may be NULL) from the state w and tangent wd (ghost
 * frames as given).  dz2 does not read the metrics. */
int bca_dz_5p(bca_ctx *ctx, double *dz, double *dz2, const double *w, const double *wd, bca_metrics g,
              const bca_phys *phys);
/* computejacobianfromdz(jac,ia,ja,dz,m,l,k,gh,im,jm,nbentry)
 * misc/ComputeJacobian.f90:708-779: +dz into the COO slots of seed (m, l, k)
 * with its symmetric window rules (dummy row 5*im*jm-1). */
int bca_scatter_dz(bca_ctx *ctx, double *jac, int *ia, int *ja, const double *dz, int m, int l, int k,
                   long nbentry);
/* The Dz/Dzz seed loop of BROADCAST_npz.py:1230-1245 in one call: for every
 * seed (m, l, k), testvector + the linearised BC sequence `ops` + coeffs_5p_dz
 * and coeffs_5p_dz2 + computejacobianfromdz, into the COO values jdz and/or
 * jdz2 (either may be NULL; ia/ja are the same for both). */
int bca_dz_colored(bca_ctx *ctx, const double *w, bca_metrics g, const bca_phys *phys, const bca_bc_op *ops,
                   int nops, double *jdz, double *jdz2, int *ia, int *ja, long nbentry, int write_index);
/* J(beta) = Jac/vol(row) - beta^2 Dzz + i beta Dz (BROADCAST_npz.py:1206-1209,
 * biglobal.py:82) for nbeta wavenumbers, from three COOs on the slot layout
 * of the plain/relaxed single-zone colored Jacobian (jac from
 * bca_jacobian_colored with withjn = 0, jdz and jdz2 from bca_dz_colored):
 * each pruned at |x| > mini like remove_zero_jac, summed over the union of
 * their patterns.  _count fills indptr (5*im*jm + 1) and returns nnz (it
 * synchronizes the stream); _fill writes the shared column indices and
 * vals[b][nnz] as (re, im) pairs (complex128). */
int bca_jbeta_count(bca_ctx *ctx, const double *jac, const double *jdz, const double *jdz2, long nbentry,
                    double mini, int *indptr, long *nnz);
int bca_jbeta_fill(bca_ctx *ctx, const double *jac, const double *jdz, const double *jdz2, long nbentry,
                   double mini, const double *vol, const double *beta, int nbeta, const int *indptr,
                   int *indices, double *vals, long nnz);

/* ---- batched colored Jacobian --------------------------------------------- */
/* The whole seed loop of BROADCAST_npz.py:1068-1127 (cylinder.py:944-985 for
 * the O-mesh): for every seed (m, l, k) of testvector, the linearised BC
 * sequence `ops` (applied in order) + the tangent residual + the scatter of
 * computejacobianfromjv[_relaxed][_withjn], into the same COO arrays the
 * reference loop fills (a fresh, zeroed COO).  No dense tangent is formed:
 * face-local Jacobian blocks are computed once and summed per row and seed.
 * write_index = 0 skips ia/ja (they depend on the mesh only).  Every BC box
 * of `ops` must lie on the block edge and every join box in the 7-layer
 * boundary frame; the seeded states are stored for that frame only, and any
 * other box returns an error. */
int bca_jacobian_colored(bca_ctx *ctx, const double *w, bca_metrics g, const bca_phys *phys,
                         const bca_bc_op *ops, int nops, const double *coefdiag, int withjn,
                         double *jac, int *ia, int *ja, long nbentry, int write_index);
/* The two-zone loop of cylinder.py:895-935 (O-mesh with im not a multiple of
 * 2gh+1): zones i in [0, isplit-1] and [isplit, im-1] (0-based; the driver
 * takes isplit = im/2/(2gh+1)*(2gh+1)), seeds of testvector_partial and the
 * scatter of computejacobianfromjv_relaxed_withjnandcheck on a zeroed COO of
 * nbentry >= 2*im*jm*(2gh+1)^2*25 entries. */
int bca_jacobian_colored_2zones(bca_ctx *ctx, const double *w, bca_metrics g, const bca_phys *phys,
                                const bca_bc_op *ops, int nops, const double *coefdiag, int isplit,
                                double *jac, int *ia, int *ja, long nbentry, int write_index);

/* ---- uneven-mesh batch ----------------------------------------------------- */
/* The reference's per-mesh iteration work -- the BC calls in order
 * (BROADCAST_npz.py:1018-1021, tangent :1079-1082) then flux_num_dnc5_2d
 * (srcfv/rhs/flux_num_dnc5.F90:7-226) or flux_num_dnc5_2d_d
 * (srcfv/tangent/flux_num_dnc5_d.f90:15) -- for up to 16 meshes of different
 * sizes (gh = 3), one launch per stage over all meshes.  Per item: the ghost
 * frames of w (and wd) are filled in place by the BC sequence, residu (and
 * residud = dR/dw . wd) are written on the interior.  Items carry wd and
 * residud all or none (residual only).  No reference counterpart as a batch:
 * the reference calls the per-mesh routines once per mesh. */
typedef struct bca_batch bca_batch;
typedef struct {
  int im, jm;
  bca_metrics g;
  bca_phys phys;
  const bca_bc_op *ops; /* host array, applied in order, at most 8 */
  int nops;
  double *w, *wd;        /* device, ni*nj*5 each (wd may be NULL) */
  double *residu, *residud;
} bca_batch_item;
int bca_batch_create(bca_batch **batch, void *hip_stream);
int bca_batch_set_stream(bca_batch *batch, void *hip_stream);
int bca_batch_destroy(bca_batch *batch);
int bca_batch_residual_jvp(bca_batch *batch, const bca_batch_item *items, int n);

/* ---- f_adj boundary conditions --------------------------------------------- */
/* bc_supandsubinlet_2d_b  srcfv/adjoint/bc_supandsubinlet_b.f90:11
 * bc_no_reflexion_2d_b    srcfv/adjoint/bc_no_reflexion_b.f90:16
 * bc_extrapolate_o2_2d_b  srcfv/adjoint/bc_extrapolateo2_b.f90:18
 * bc_wall_viscous_adia_2d_b srcfv/adjoint/bc_wall_viscous_b.f90:11
 * jn_match_2d_b           srcfv/adjoint/jn_match_2d_b.f90:11 (in-block join)
 * One operation `op` (kind as in the BC sequence): wb += (d ghosts / d w)^T
 * wb(ghosts) at the cells the BC reads, then wb(ghosts) = 0; w is unchanged
 * (the reference restores it).  For the join, receiving and donor boxes in
 * one block: wb(prd) += wb(prr), wb(prr) = 0 -- the transpose of the copy
 * (the reference routine takes two arrays and zeroes the donor adjoint
 * first, which in one block would clear wb).  nx, ny: the metrics (inlet and
 * no-reflexion read them).  Except for the join, the box must lie on the
 * block edge (the seeded states cover the 7-layer boundary frame only);
 * other boxes return an error. */
int bca_bc_adjoint(bca_ctx *ctx, const double *w, double *wb, const bca_bc_op *op, const double *nx,
                   const double *ny);

/* ---- fused colored Jacobian -> pruned CSR ---------------------------------- */
/* The seed loop of BROADCAST_npz.py:1068-1127 (nzones = 1, zparam = withjn)
 * or of cylinder.py:895-935 (nzones = 2, zparam = isplit as for
 * bca_jacobian_colored_2zones) followed by remove_zero_jac + csr_matrix
 * (BROADCAST_npz.py:129-135,1136-1150; misc/PETSc_func.py:85), without the
 * COO: indptr (5*im*jm + 1), and when nnz <= cap, indices and data (nnz each),
 * bit for bit what bca_jacobian_colored[_2zones] + bca_csr_colored_count/_fill
 * give.  Returns -5 with *nnz set when cap is too small (call again with
 * arrays of *nnz entries).  Synchronizes the stream once (nnz). */
int bca_jacobian_colored_csr(bca_ctx *ctx, const double *w, bca_metrics g, const bca_phys *phys,
                             const bca_bc_op *ops, int nops, const double *coefdiag, int nzones, int zparam,
                             double mini, int *indptr, int *indices, double *data, long cap, long *nnz);
/* The same CSR for a Newton loop (BROADCAST_npz.py:1068-1157 re-assembles
 * and re-converts every iteration): when indptr, indices and data are the
 * arrays the previous bca_jacobian_colored_csr[_refresh] call on this context
 * filled, with the same nzones, zparam and mini, the kept set of every row
 * is checked against that call's; if no row changed only data is rewritten
 * (indptr and indices already hold this CSR) and *rebuilt = 0, otherwise the
 * whole CSR is rebuilt as bca_jacobian_colored_csr does and *rebuilt = 1.
 * The result is identical to bca_jacobian_colored_csr either way. */
int bca_jacobian_colored_csr_refresh(bca_ctx *ctx, const double *w, bca_metrics g, const bca_phys *phys,
                                     const bca_bc_op *ops, int nops, const double *coefdiag, int nzones,
                                     int zparam, double mini, int *indptr, int *indices, double *data, long cap,
                                     long *nnz, int *rebuilt);

/* ---- Newton linear solve ---------------------------------------------------- */
/* The reference solves each Newton system with MUMPS through PETSc
 * (misc/PETSc_func.py:137-152,247-263; BROADCAST_npz.py:1152-1157).  The
 * project's driver factors the block-banded matrix by block Gaussian
 * elimination; these entries are its dense parts.
 * bca_lu_np: in-place LU without pivoting (L unit lower, U upper) of the
 * n x n row-major block at A (leading dimension ld): 32-column panels, the
 * trailing updates on the FP64 matrix cores.  *bad (device int, set by the
 * caller to 0) becomes 1 if a pivot falls below tiny times its tile's
 * largest entry; work: device scratch of 32 x 32 doubles.
 * bca_inv_np: in-place inverse without pivoting of the n x n row-major block
 * at A (ld) by Gauss-Jordan steps on 32-column panels (each step a product
 * on the FP64 matrix cores over the whole block); the block-banded driver
 * multiplies its off-diagonal blocks and right-hand sides by it.  work:
 * device scratch of n x n doubles; *bad as for bca_lu_np.
 * bca_gemm_f64_sub: C -= A B, row-major, FP64 matrix cores.
 * bca_csr_spmv: y = A x for a CSR matrix (the residual of a refined solve). */
int bca_lu_np(void *hip_stream, double *A, int n, long ld, double tiny, int *bad, double *work);
int bca_inv_np(void *hip_stream, double *A, int n, long ld, double *work, double tiny, int *bad);
int bca_gemm_f64_sub(void *hip_stream, int M, int N, int K, const double *A, long lda, const double *B,
                     long ldb, double *C, long ldc);
int bca_csr_spmv(void *hip_stream, long n, const int *indptr, const int *indices, const double *data,
                 const double *x, double *y);

/* ---- BF16 GEMM of the learned Newton guess ---------------------------------- */
/* No reference counterpart (the reference has no learned part).  C (fp32,
 * M x N) = A (bf16 bit patterns, M x K) * B (bf16, K x N), FP32 accumulation on
 * the matrix cores; element (r, c) of each operand at r * s_r + c * s_c, so
 * transposed operands need no copy.  accumulate = 1 adds to C.  ksplit > 1
 * splits K over ksplit block rows that add their partial sums to C with
 * float atomics (C must hold the initial value, e.g. zeros). */
int bca_gemm_bf16(void *hip_stream, int M, int N, int K, const unsigned short *A, long sam, long sak,
                  const unsigned short *B, long sbk, long sbn, float *C, long scm, long scn, int accumulate,
                  int ksplit);

/* ---- launch-shape autotuner ------------------------------------------------ */
/* No reference counterpart.  Times candidate block shapes of the residual face
 * kernel, the interior face-block kernels and the colored assembly on this
 * context's mesh and state (median of reps launches each; nzones and zparam as
 * for bca_jacobian_colored_csr; jac: a COO buffer of nbentry values used as
 * scratch) and keeps the fastest for the context's later launches.  Results
 * do not depend on the shapes.  report (optional) receives the choices. */
int bca_autotune(bca_ctx *ctx, const double *w, bca_metrics g, const bca_phys *phys, const bca_bc_op *ops,
                 int nops, int nzones, int zparam, double *jac, long nbentry, int reps, char *report,
                 int report_len);

#ifdef __cplusplus
}
#endif
#endif /* BROADCAST_AMD_H */
