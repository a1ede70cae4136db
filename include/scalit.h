/*
 * scalit.h -- C ABI of the MI355X implementation of ScalIT's iterate hot path:
 * the H*X product of a direct-product DVR Hamiltonian, the OSB block
 * preconditioner P*X, and the QMR / PIST / Lanczos iterations that use them.
 *
 * Two layers:
 *
 * 1. Handle API (scalit_*): a context owns the operator on one GPU (it replaces
 *    the reference's module globals H0/RES/DEP/VOSB/EIG0,
 *    src/iterate/module/osb.data.h:28-69).  Vector arguments are DEVICE
 *    pointers; nvec vectors of length N are stored with leading dimension ld.
 *    Kernels are enqueued on the given HIP stream (NULL = the default stream);
 *    the solvers, which return host results, synchronise the device on entry
 *    and run on the context's own stream.
 *
 * 2. Fortran drop-in layer (lower case + trailing underscore, all arguments by
 *    reference, HOST arrays, column-major, 1-based semantics): the callbacks
 *    and external library functions the reference's solvers bind.  They act on
 *    the context made current with scalit_bind().
 *
 * Return codes: the reference's own conventions where a reference function is
 * replaced (QMR: i / -i / 0 / -1; PIST_CONV: cnt / -cnt / 0; ...).  Handle
 * API functions return SCALIT_OK (0) or a negative SCALIT_E* code; they never
 * throw across the ABI.
 */
#ifndef SCALIT_H
#define SCALIT_H

#ifdef __cplusplus
extern "C" {
#endif

#define SCALIT_FMAX 8

#define SCALIT_OK 0
#define SCALIT_EINVAL (-1001)   /* bad argument */
#define SCALIT_EHIP (-1002)     /* HIP runtime error */
#define SCALIT_ENOMEM (-1003)   /* device allocation failed */
#define SCALIT_ESTATE (-1004)   /* operator incomplete (e.g. no preconditioner yet) */
#define SCALIT_EIO (-1005)      /* file missing, short or not in the expected format */

/* H*X shift selector: src/iterate/module/osb.hx.h:82-89 */
#define SCALIT_TA0 0   /* Y = H X */
#define SCALIT_TA1 1   /* Y = (H - E0) X */
/* complex operators (absorbing potentials): src/iterate/module/osb.data.h:18-22 */
#define SCALIT_TAAP0 2   /* (H + i AP) X */
#define SCALIT_TMAP0 3   /* (H - i AP) X */
#define SCALIT_TAAP 100  /* (H + i AP - E0) X */
#define SCALIT_TAAPP 101 /* (H + i APP - E0) X */
#define SCALIT_TAAPR 102 /* (H + i APR - E0) X */
#define SCALIT_TMAP 200  /* (H - i AP - E0) X */
#define SCALIT_TMAPP 201 /* (H - i APP - E0) X */
#define SCALIT_TMAPR 202 /* (H - i APR - E0) X */
/* absorbing-potential arrays: osb.data.h:61 */
#define SCALIT_AP 0
#define SCALIT_APP 1
#define SCALIT_APR 2
/* preconditioner kernels: src/iterate/module/osb.px.h:156-211 */
#define SCALIT_TOSB 0
#define SCALIT_TOSBD1 1
#define SCALIT_TOSBD2 2
#define SCALIT_TOSBW 3   /* the windowed preconditioner; needs scalit_set_osbw */
/* arithmetic of an operator application */
#define SCALIT_FP64 0
#define SCALIT_FP32 1
#define SCALIT_BF16 2  /* vectors stored in bf16, arithmetic in fp32 */

typedef struct scalit_ctx scalit_ctx;
typedef void *scalit_stream; /* hipStream_t */

/* ------------------------------------------------------------------------ */
/* context and operator data                                                 */
/* ------------------------------------------------------------------------ */

/* Replaces initOSB/calOSB (src/iterate/module/osb.f90:17, osb.init.h:2-50):
 * sF levels of sizes sN[0..sF-1] (level 1 fastest), sDEP[l] != 0 when level
 * l+1 is multiplied by the coordinate-dependent factor DEP (sDEP[0] must be 0,
 * src/iterate/module/mosb.print.h:241-242).  Allocates nothing large yet. */
int scalit_create(scalit_ctx **ctx, int sF, const int *sN, const int *sDEP, int device);
int scalit_destroy(scalit_ctx *ctx);
int scalit_get_len(const scalit_ctx *ctx, long *n);

/* Operator pieces, host arrays (replaces loadH0/loadRES/loadDep,
 * src/iterate/module/osb.io.h:2-60):  h0 is n x n column-major for level
 * `level` (1-based); dep has myDim(level) = prod(sN[0..level-2]) entries; res
 * has N entries. */
int scalit_set_h0(scalit_ctx *ctx, int level, const double *h0);
int scalit_set_dep(scalit_ctx *ctx, int level, const double *dep);
int scalit_set_res(scalit_ctx *ctx, const double *res);
/* sNDVR = T outermost layer (OUTH of src/iterate/module/osb.data.h:62, read by
 * loadOutH, osb.io.h:230-235): outh holds myDim(sF) x sN(sF) x sN(sF) values,
 * point index fastest.  Level sF of H*X then applies H3X
 * (src/libutil/serial/base/matrix.f90:188-203, osb.hx.h:93-94) instead of
 * H0 [x DEP], its OSB block is OUTH itself (osb.diag.h:22-23), and its
 * H0 / DEP are not needed (sDEP(sF) is forced off, osb.print.h:204).
 * Requires sF >= 2. */
int scalit_set_outh(scalit_ctx *ctx, const double *outh);
/* sOSBW%mE0 / mDE / mBeta (src/iterate/module/a_osbtype.f90:26-29) */
int scalit_set_shift(scalit_ctx *ctx, double e0, double de, double beta);
/* sOSB, the preconditioner kernel of the solvers' P*X (TOSB, TOSBD1,
 * TOSBD2 or TOSBW; the input read at src/iterate/module/mosb.print.h:245,
 * used by EPX at src/iterate/module/mosb.px.h:40-46): QMR, PIST and the
 * mixed-precision inner solves then divide by that kernel, with DE and beta
 * from scalit_set_shift.  Default TOSB.  TOSBW is FP64 only (the FP32 and
 * BF16 inner solves return SCALIT_EINVAL) and takes the per-step QMR. */
int scalit_set_osb(scalit_ctx *ctx, int nosb);
/* initOSBW (src/iterate/module/osb.osbw.h:2-50) for the TOSBW kernel
 * (src/iterate/module/osb.px.h:191-206): adjustWinSize (osb.osbw.h:117-145)
 * widens or narrows the window from de until exactly cnt OSB states have
 * |EIG0 - e0| <= de (at most 10,000 steps, else SCALIT_ESTATE), getOSBWIndex
 * (osb.osbw.h:79-95) takes them in grid order, getVOSBH0
 * (src/iterate/module/osb.hij.h:133-160) forms H_W(i,j) = v_i . H v_j with
 * EIG0 on the diagonal, and createWH (src/iterate/module/osb.pij.h:6-40)
 * subtracts e0 from the diagonal when npc is SCALIT_TA1.  P*X with TOSBW then
 * divides the rotated vector by EIG0 - shift and replaces the window's
 * components by (H_W - e0)^-1 times them.  cnt 1..1024; needs the
 * preconditioner (scalit_block_diag or scalit_set_precond, which drop the
 * window again) and RES; not for sNDVR operators.  *hwlen gets the window's
 * size and *de_out the width found (either may be NULL). */
int scalit_set_osbw(scalit_ctx *ctx, int npc, double e0, double de, int cnt, int *hwlen, double *de_out);
/* No counterpart in the reference (its QMR loop, qmr.f90:56-175, calls the
 * operators once per step).  FP64 QMR solves on a fused 3-level operator
 * (sF = 3, DEP on level 3 only, levels of at most 48 points, at most 65,536
 * points, n1 n2 <= 1,536) run every iteration in one persistent launch when
 * on = 1 (and PIST runs its Krylov steps block by block the same way), one
 * kernel per step of the iteration when on = 0; -1 (the default) follows the
 * environment variable SCALIT_QMR_PERSIST (unset: 1).  Same results to
 * rounding: only the order of the terms of the dot products differs. */
int scalit_set_qmr_persistent(scalit_ctx *ctx, int on);

/* OSB set-up on the GPU: myDiag (src/iterate/module/osb.diag.h:53-151) with
 * HosbDiag/MJACOBI (src/iterate/libiterate/diag/hosbdiag.f90:5-75,
 * jacobi.f90:2-83).  Leaves VOSB and EIG0 resident. */
int scalit_block_diag(scalit_ctx *ctx, int bjmax, double bjtol);
/* Load / fetch VOSB and EIG0 (host arrays, the sVOSB<0 / >0 file modes,
 * src/iterate/module/osb.progDiag.h:6-25): vosb[l] holds myBlk(l)*n_l^2. */
int scalit_set_precond(scalit_ctx *ctx, const double *const *vosb, const double *eig0);
int scalit_get_precond(scalit_ctx *ctx, double *const *vosb, double *eig0);

/* ------------------------------------------------------------------------ */
/* operator applications (device pointers)                                  */
/* ------------------------------------------------------------------------ */

/* MYHX (src/iterate/module/osb.hx.h:71-110):
 * Y_b = (RES - shift_b) o X_b + sum_l (I x H0_l [x DEP_l] x I) X_b.
 * ntype TA0: shift 0; TA1: shift E0.  `shifts` (device, nvec doubles) when not
 * NULL overrides the shift per vector (used for batched PIST windows). */
int scalit_hx(scalit_ctx *ctx, int ntype, int nvec, const double *x, long ldx, double *y, long ldy,
              const double *shifts, scalit_stream stream);
/* The same product in fp32 (prec SCALIT_FP32: float vectors) or with bf16
 * storage (SCALIT_BF16: unsigned short vectors), fp32 accumulation. */
int scalit_hx_lp(scalit_ctx *ctx, int prec, int ntype, int nvec, const void *x, long ldx, void *y, long ldy,
                 const float *shifts, scalit_stream stream);
/* Adjoint of H*X for gradients: dx_b = H^T g_b (H0 transposed per level; RES
 * and DEP diagonal) and, when dv != NULL, dv += sum_b x_b o g_b. */
int scalit_hx_adjoint(scalit_ctx *ctx, int ntype, int nvec, const double *x, const double *g, long ld,
                      double *dx, double *dv, scalit_stream stream);

/* MYPX (src/iterate/module/osb.px.h:125-219): Y = V (EIG0 - E0)^-1 V^T X with
 * the TOSB / TOSBD1 / TOSBD2 kernels; npc TA0 uses E0 = 0. */
int scalit_px(scalit_ctx *ctx, int nosb, int npc, int nvec, const double *x, long ldx, double *y, long ldy,
              const double *shifts, scalit_stream stream);
int scalit_px_lp(scalit_ctx *ctx, int prec, int npc, int nvec, const void *x, long ldx, void *y, long ldy,
                 const float *shifts, scalit_stream stream);

/* ------------------------------------------------------------------------ */
/* solvers on the context's operator (device vectors)                       */
/* ------------------------------------------------------------------------ */

/* QMR (src/iterate/libiterate/qmr/qmr.f90:5-175) on (H - E0) with the OSB
 * preconditioner, for nvec right-hand sides at once (shifts per vector when
 * not NULL; host array).  ret[b] gets the reference's return code for vector
 * b, res[b] its final residual.  Returns SCALIT_OK or an error code. */
int scalit_qmr(scalit_ctx *ctx, int ntype, int qmrmax, double qmrtol, int nvec, const double *b, double *x,
               long ld, const double *shifts, int *ret, double *res, scalit_stream stream);
/* Mixed-precision QMR: inner iterations in fp32 (prec SCALIT_FP32) or with
 * bf16-stored vectors (SCALIT_BF16), outer fp64 iterative refinement until
 * ||b - (H-E0)x|| <= tol ||b||.  ret[b] as scalit_qmr (total inner steps). */
int scalit_qmr_mixed(scalit_ctx *ctx, int prec, int qmrmax, double qmrtol, int inner_max, double inner_tol, int nvec,
                     const double *b, double *x, long ld, const double *shifts, int *ret, double *res,
                     scalit_stream stream);

/* PIST_CONV (src/iterate/libiterate/pist/pist.f90:7-98) with pistH0
 * (pistHij.f90:328-383), LinSolv = OSB QMR (src/iterate/module/osbw.qmr.h:13),
 * H0X = H (TA0).  x is the device start vector; eig (host, m entries)
 * receives the eigenvalues ascending.  The return value is the reference's
 * (0 for its parameter errors), or a negative SCALIT_E* code when the solve
 * cannot run (e.g. an inner precision that the preconditioner kernel refuses).
 * prec selects the inner solve (SCALIT_FP64 = plain QMR, FP32/BF16 = mixed
 * with fp64 refinement); stats (host, 4 ints, may be NULL): total QMR
 * iterations, linear solves, final Krylov size, H*X applications. */
int scalit_pist_conv(scalit_ctx *ctx, double e0, double etol, int ntype, const double *x, int startcnt, int stepcnt,
                     int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int prec, double *eig,
                     double *res, int *stats);
/* Batched PIST: nwin windows with their own E0 (host array) and start
 * vectors (device, leading dimension ld), solved in lock-step so that every
 * QMR iteration is one batched H*X / P*X launch.  eig is nwin x m (host),
 * ret / res per window. */
int scalit_pist_conv_batched(scalit_ctx *ctx, int nwin, const double *e0, double etol, int ntype, const double *x,
                             long ld, int startcnt, int stepcnt, int stepeig, int m, int m_max, int qmr_type,
                             int qmrmax, double qmrtol, int prec, double *eig, int *ret, double *res, int *stats);
/* Orthogonalisation passes per PIST Krylov vector: 1 (the default) is the
 * reference's MGS_ORTH (src/iterate/libiterate/orth/gsorth.f90:42-72), one
 * modified Gram-Schmidt pass; 2 orthogonalises each new vector a second time
 * against the basis.  No counterpart in the reference: a long PIST solve (a
 * Krylov space of hundreds of vectors, many levels at once) loses
 * orthogonality with one pass and its Rayleigh-Ritz step then returns
 * spurious values; two passes keep the basis orthonormal to rounding. */
int scalit_set_pist_orth(scalit_ctx *ctx, int passes);
/* Ritz vectors of the last PIST solve of window `win` for its m eigenvalues:
 * psi (device, N x m, leading dimension ld), unit norm.  SCALIT_ESTATE when
 * that window's solve did not converge (its return code was <= 0). */
int scalit_pist_vectors(scalit_ctx *ctx, int win, double *psi, long ld);
/* The same for every window of the last (batched) PIST solve in one call:
 * window k's m vectors at psi + (k m + i) ld (device, ncols columns of N,
 * ncols >= nwin m, else SCALIT_EINVAL); SCALIT_ESTATE when any window's
 * solve did not converge. */
int scalit_pist_vectors_all(scalit_ctx *ctx, double *psi, long ld, int ncols);

/* Lan_Conv (src/iterate/libiterate/lanczos/lan.f90:7-105) with lanH0
 * (lanHij.f90:6-45): FP64 Lanczos with full re-orthogonalisation. */
int scalit_lan_conv(scalit_ctx *ctx, double e0, double etol, int ntype, const double *x, int startcnt, int stepcnt,
                    int stepeig, int m, int m_max, double *eig, double *res);

/* ------------------------------------------------------------------------ */
/* absorbing-potential (resonance) jobs: complex vectors in the split layout */
/* (vector v's real part at x + 2 v N, its imaginary part at x + (2 v + 1) N; */
/* the reference's double complex arrays are interleaved)                   */
/* ------------------------------------------------------------------------ */

/* AP / APP / APR (src/iterate/module/osb.data.h:61, loaded by osb.io.h:20-30:
 * a resonance job sets AP = APP; a CRP job AP = APP + APR).  Host, N values. */
int scalit_set_ap(scalit_ctx *ctx, int which, const double *ap);
/* MYHX_DX (src/iterate/module/osb.hxdx.h:177-240): Y = (RES - E0 +- i A) X +
 * the level products, ntype one of SCALIT_TA0 ... SCALIT_TMAPR (E0 = 0 for
 * TA0, TAAP0, TMAP0).  x, y device, nvec split vectors. */
int scalit_hx_cx(scalit_ctx *ctx, int ntype, int nvec, const double *x, double *y, scalit_stream stream);
/* MYPX_DX (src/iterate/module/osb.mypxdx.h): V (kernel)^-1 V^T X with the
 * complex divisor EIG0 - E0 (clamped by TOSBD1 / TOSBD2) +- i A of npc.
 * nosb TOSB, TOSBD1, TOSBD2 or TOSBW (after scalit_set_osbw, real level
 * matrices): the window's components are replaced by HWR^-1 times them,
 * HWR = H_W + diag(-E0 +- i A(WIND)) (createWH_DXCX, osb.pij.h:65-166;
 * ZGETRS / ZHETRS in osb.mypxdx.h:127-170), its inverse formed on the host
 * once per (npc, E0, absorbing potentials, window); TOSBW with npc TAAP0 or
 * TMAP0 is SCALIT_EINVAL (the reference divides by EIG0 - E0 there but
 * leaves E0 out of HWR).  The complex solvers below take sOSB from
 * scalit_set_osb, TOSBW included. */
int scalit_px_cx(scalit_ctx *ctx, int nosb, int npc, int nvec, const double *x, double *y, scalit_stream stream);
/* OSB_QMRDX (src/iterate/module/osbw.qmrdx.h:16-27): QMR_CX
 * (src/iterate/libiterate/qmr/qmrcx.f90:4-179) with HX_DX = MYHX_DX(hc) and
 * PX_DX = MYPX_DX(sOSB, pc).  b, x device (one split vector; x is the start
 * vector when qmr_type > 0); *ret gets QMR_CX's value, *res its residual. */
int scalit_qmr_cx(scalit_ctx *ctx, int qmr_type, int hc, int pc, int qmrmax, double qmrtol, const double *b, double *x,
                  int *ret, double *res);
/* OSB_PISTCONVDX1 (src/iterate/module/osbw.pistconv.h:98-140): PIST_CONV_SX
 * (src/iterate/libiterate/pist/pist_sx.f90:9-110) with H0X = MYHX_DX(TMAP0),
 * LinSolv = OSB_QMRDX with hc = TMAP and pc = SCALIT_TMAP (sAP) or SCALIT_TA1,
 * the operator shift E0 of scalit_set_shift and the window centre e0.  x0
 * device (one split vector); eig host, 2 m doubles (re, im of each energy,
 * ascending real part); *ret gets PIST_CONV_SX's value.  ntype > 0 is
 * SCALIT_EINVAL (the reference passes getSepMax_CX five arguments for four). */
int scalit_pist_conv_sx(scalit_ctx *ctx, double e0, double etol, int ntype, const double *x0, int startcnt,
                        int stepcnt, int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int pc,
                        double *eig, int *ret, double *res, int *stats);
/* OSB_PISTCONVDX2 (src/iterate/module/osbw.pistconv.h:204-244): PIST_CONV_CX
 * (src/iterate/libiterate/pist/pist_cx.f90, pistHij_cx.f90) -- as
 * scalit_pist_conv_sx with the Hermitian products x^H y and real norms of
 * MGS_ORTH_CX (orth/gsorth_cx.f90:51-79) for the Krylov basis and its matrix. */
int scalit_pist_conv_cx(scalit_ctx *ctx, double e0, double etol, int ntype, const double *x0, int startcnt,
                        int stepcnt, int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int pc,
                        double *eig, int *ret, double *res, int *stats);
/* osbCRP (src/iterate/module/osbw.crp.h:7-49): the cumulative reaction
 * probability N(E) = 4 sum of the nm Lanczos eigenvalues (LAN_DX,
 * src/iterate/libiterate/lanczos/lan_dx.f90:299-325, full re-orthogonalisation)
 * of APR^1/2 (H - i AP - E)^-1 APP (H + i AP - E)^-1 APR^1/2 (CRP_HXDX,
 * osbw.crp.h:79-131), each inverse an OSB_QMRDX solve whose preconditioner
 * includes the absorbing term when sap != 0 (sAP).  Needs APP and APR; sets
 * AP = APP + APR as osbCRP does.  x0: device, ncrp split start vectors (one
 * per energy, as the reference draws one per energy); crp host, ncrp values
 * (0 where the Lanczos or its eigensolve fails, as calOSBCRP returns);
 * stats (host, 2 ints, may be NULL): QMR iterations, QMR solves.  The ctx's
 * shift is restored afterwards. */
int scalit_crp(scalit_ctx *ctx, int ncrp, const double *energies, int nm, const double *x0, int qmr_type,
               int qmrmax, double qmrtol, int sap, double *crp, int *stats);
/* ZGEEV('N', 'N') as PIST_CONV_SX calls it (pist_sx.f90:62,71), on the host:
 * the eigenvalues of a general complex n x n matrix (a interleaved re/im,
 * column-major, leading dimension lda) into w (2 n doubles). */
int scalit_small_eig_cx(int n, const double *a, int lda, double *w);

/* Complex level matrices (sCX = T, src/iterate/module/osb.data.h:31,63):
 * H0CX of a level (n x n, column-major; its real and imaginary parts as
 * separate host arrays), DEPCX of a dependent level (myDim values) and, for
 * an sNDVR operator, OUTHCX (N n_sF values, the layout of OUTH).  They
 * replace H0 / DEP / OUTH, as loadH0 / loadDEP / loadOutH do under sCX
 * (src/iterate/module/osb.io.h:45-55,211-236), once scalit_set_scx(ctx, 1)
 * switches them on: the complex operators (scalit_hx_cx and every complex
 * solver) then apply MYHX_CX (src/iterate/module/osb.hxcx.h:191-245; TMAPR
 * takes -APR as mosb3.hxcx.h:92-94 does) and scalit_block_diag runs the sCX
 * branch of myDiag (osb.diag.h:76-101): HosbDiagDX
 * (src/iterate/libiterate/diag/hosbdiag_dx.f90:6-73) with MJACOBI_CX
 * (jacobi_cx.f90:5-96), real VOSB, EIG0 the real parts of the rotated
 * diagonal.  scalit_set_scx(ctx, 1) returns SCALIT_ESTATE until every level
 * has its complex data.  scalit_set_osbw then forms the complex window
 * Hamiltonian of getVOSBH0_CX (src/iterate/module/osb.hijcx.h:135-164:
 * v_i . MYHX_CX(TA0) v_j above the diagonal, its conjugate below, EIG0 on
 * it) for the complex solvers' TOSBW. */
int scalit_set_h0_cx(scalit_ctx *ctx, int level, const double *h0_re, const double *h0_im);
int scalit_set_dep_cx(scalit_ctx *ctx, int level, const double *dep_re, const double *dep_im);
int scalit_set_outh_cx(scalit_ctx *ctx, const double *outh_re, const double *outh_im);
int scalit_set_scx(scalit_ctx *ctx, int on);

/* Hellmann-Feynman gradients: dE_k/dp_j = sum_grid psi_k^2 dV/dp_j.
 * psi device N x nk (ld), dvdp device N x np (lddv); out host nk x np
 * (row-major).  Uses the adjoint product's dV = x o g with x = g = psi_k. */
int scalit_eig_gradients(scalit_ctx *ctx, int nk, const double *psi, long ld, int np, const double *dvdp, long lddv,
                         double *out);

/* scalit_eig_gradients without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as scalit_eig_gradients, with nk * np <= 2048 (else
 * SCALIT_EINVAL): its partial sums live in a fixed 8.4 MB device array per
 * GPU, held until the process ends.  Calls that overlap on one GPU share that
 * array and race: the caller serialises them. */
int scalit_eig_gradients_noalloc(scalit_ctx *ctx, int nk, const double *psi, long ld, int np, const double *dvdp,
                                 long lddv, double *out);

/* ------------------------------------------------------------------------ */
/* Fortran drop-in layer (host arrays, arguments by reference)               */
/* ------------------------------------------------------------------------ */

/* Make ctx the operator behind hx_/px_/... (replaces the module state). */
int scalit_bind(scalit_ctx *ctx);

typedef void (*scalit_fop)(const int *n, const double *x, double *y);
typedef int (*scalit_flinsolv)(const int *n, const double *b, double *x, double *res);

/* HX(N,X,Y) with sHC (src/iterate/module/osb.hx.h:13-20), HX0 (TA0, :41-48),
 * EHX (TA1, :56-63) */
void hx_(const int *n, const double *x, double *y);
void hx0_(const int *n, const double *x, double *y);
void ehx_(const int *n, const double *x, double *y);
/* PX(N,X,Y) with sOSB/sPC (src/iterate/module/osb.px.h:12-19), EPX (TA1) */
void px_(const int *n, const double *x, double *y);
void epx_(const int *n, const double *x, double *y);
/* OSB_QMR(N,B,X,RES): the PIST LinSolv callback (src/iterate/module/osbw.qmr.h:13-23) */
int osb_qmr_(const int *n, const double *b, double *x, double *eres);
/* Set the QMR parameters OSB_QMR uses (sQMRConvType, sQMR%mMax, sQMR%mTol). */
void scalit_set_qmr_params(int qmr_type, int qmrmax, double qmrtol);

/* QMR(nTYPE,QMRMAX,QMRTOL,N,B,HX,PX,X,RES) (src/iterate/libiterate/qmr/qmr.f90:5).
 * When HX/PX are this library's own callbacks the whole solve stays on the
 * GPU; other callbacks are called on host copies of the vectors. */
int qmr_(const int *ntype, const int *qmrmax, const double *qmrtol, const int *n, const double *b, scalit_fop hx,
         scalit_fop px, double *x, double *res);
/* PIST_CONV(E0,ETOL,NTYPE,N,X,STARTCNT,STEPCNT,STEPEIG,M,M_MAX,H0X,LinSolv,EIG,RES)
 * (src/iterate/libiterate/pist/pist.f90:7) */
int pist_conv_(const double *e0, const double *etol, const int *ntype, const int *n, const double *x,
               const int *startcnt, const int *stepcnt, const int *stepeig, const int *m, const int *m_max,
               scalit_fop h0x, scalit_flinsolv linsolv, double *eig, double *res);
/* Lan_Conv(E0,ETOL,nType,N,X,startCnt,stepCnt,stepEig,M,M_MAX,HX,EIG,RES)
 * (src/iterate/libiterate/lanczos/lan.f90:7) */
int lan_conv_(const double *e0, const double *etol, const int *ntype, const int *n, const double *x,
              const int *startcnt, const int *stepcnt, const int *stepeig, const int *m, const int *m_max,
              scalit_fop hx, double *eig, double *res);
/* MGS_ORTH(N,WJ,M,VJ) (src/iterate/libiterate/orth/gsorth.f90:42): returns 1/0 */
int mgs_orth_(const int *n, const double *wj, const int *m, double *vj);

/* ------------------------------------------------------------------------ */
/* test and profiling hooks (no counterpart in the reference)                */
/* ------------------------------------------------------------------------ */
/* Force (on != 0) the generic any-level H*X kernels instead of the fused
 * 3-level path, for parity tests of both. */
int scalit_debug_force_generic(int on);
/* Lane layout of the 16x16x4 MFMA builtins (prec SCALIT_FP64 / SCALIT_FP32):
 * out (host, 256 doubles) receives accumulator q of lane l at 4 l + q for
 * D(i,j) = 100 (i+1) + (j+1). */
int scalit_debug_mfma_probe(int prec, double *out);
/* MFMA throughput probe: `blocks` workgroups of 4 waves, each wave `iters`
 * rounds of 8 independent 16x16x4 chains from registers; *ms = kernel time
 * (FLOP = blocks * 4 * iters * 8 * 2048). */
int scalit_debug_mfma_peak(int prec, int iters, int blocks, float *ms);
/* The small dense symmetric eigensolve of the Rayleigh-Ritz steps on the GPU
 * (host arrays, column-major, upper triangle of a read): w = eigenvalues
 * ascending; z (n x n, ldz) = eigenvectors when not NULL. */
int scalit_debug_sym_eig(int n, const double *a, int lda, double *w, double *z, int ldz);
/* Eigenvalues of several small symmetric matrices in one launch, as the
 * Rayleigh-Ritz steps of all PIST windows are solved: matrix j is the leading
 * ns[j] x ns[j] block of a + j lda lda (host, column-major, leading dimension
 * lda, upper triangle read); its eigenvalues, ascending, go to w + j lda. */
int scalit_debug_sym_eigvals_batch(int njobs, const int *ns, const double *a, int lda, double *w);
/* The fused H*X kernel variants (pass A block shape, pass B width) the
 * autotuner picked for this precision and batch size (tuning them if needed). */
int scalit_debug_hx_variant(scalit_ctx *ctx, int prec, int nvec, int *va, int *vb);
/* Launch one pass of the fused FP64 H*X alone (1: modes 1+2 and the diagonal,
 * 2: mode 3 accumulate) so that each kernel can be timed by itself. */
int scalit_profile_hx_pass(scalit_ctx *ctx, int pass, int nvec, const double *x, long ldx, double *y, long ldy,
                           scalit_stream stream);

/* Several problems (uneven grids) in one call: job j runs
 * scalit_pist_conv_batched on its own context (its windows, start vectors and
 * outputs as there), every job in its own host thread on its context's
 * stream, so the problems' kernels share the device.  No counterpart in the
 * reference, which solves one problem per run (PIST_CONV,
 * src/iterate/libiterate/pist/pist.f90:7-98, per window).  status and stats
 * are per job; the call returns the first job error, else SCALIT_OK.  One job
 * per context. */
typedef struct scalit_pist_job {
    scalit_ctx *ctx;
    int nwin;
    const double *e0;   /* nwin window centres */
    const double *x;    /* device, nwin start vectors, leading dimension ld */
    long ld;
    double *eig;        /* host, nwin * m */
    int *ret;           /* host, nwin */
    double *res;        /* host, nwin, may be NULL */
    int stats[4];
    int status;
} scalit_pist_job;
int scalit_pist_conv_multi(int njob, scalit_pist_job *jobs, double etol, int ntype, int startcnt, int stepcnt,
                           int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int prec);

/* No counterpart in the reference (QMR, qmr.f90:5-175, solves one operator
 * per call).  FP64 QMR solves of several operators (one job per context, all
 * contexts on one device, grids of any sizes) in lock step: one iteration of
 * every job is recorded as one HIP graph, a branch per job, and each graph
 * launch advances all of them.  Results are those of scalit_qmr with the
 * persistent solver switched off (scalit_set_qmr_persistent(ctx, 0)), bit for
 * bit.  status is per job; returns the first error, else SCALIT_OK. */
typedef struct scalit_qmr_job {
    scalit_ctx *ctx;
    int nvec;
    const double *b;      /* device, nvec right-hand sides, leading dimension ld */
    double *x;            /* device, nvec solutions (start vectors when ntype > 0) */
    long ld;
    const double *shifts; /* host, nvec shifts, or NULL for the context's E0 */
    int *ret;             /* host, nvec */
    double *res;          /* host, nvec, may be NULL */
    int status;
} scalit_qmr_job;
int scalit_qmr_multi(int njob, scalit_qmr_job *jobs, int ntype, int qmrmax, double qmrtol);

/* ------------------------------------------------------------------------ */
/* sequential truncation of a DVR basis (host arrays; eigensolves on the     */
/* current GPU)                                                              */
/* ------------------------------------------------------------------------ */

/* DVR_Step2 + DVR_Step3 (src/ham/libham/dvr/dvr.f90:84-160): the
 * potential-optimised DVR of one coordinate.  x: the n primitive grid points,
 * t: their n x n kinetic matrix (column-major), v: the 1-D reference
 * potential at x.  Keeps the m lowest eigenvectors of T + diag(v) and returns
 * the m new grid points x0 (ascending: the eigenvalues of the position
 * operator in that basis), the m x m Hamiltonian h0 = W^T diag(E) W on them
 * (each column of the eigenvectors U of T + diag(v), and then of W, with its
 * largest-magnitude entry positive: the reference leaves these signs to
 * DSYEV, and every choice gives an equivalent operator) and, when e != NULL,
 * the m eigenvalues E.  n <= 1024. */
int scalit_podvr(int n, const double *x, const double *t, const double *v, int m, double *x0, double *h0, double *e);
/* saveH0CBALL (src/ham/ja3/module/ja3.h0:154-199): the energy-cut combined
 * layer of two coordinates.  Keeps the nkeep product points (i, j) with the
 * lowest v1[i] + v2[j] (ties in the reference's loop order, j outer), lists
 * them in that loop order (idx1[a] = i, idx2[a] = j, 0-based) and builds the
 * nkeep x nkeep layer Hamiltonian h0c(a,b) = [j_a == j_b] h1(i_a,i_b)
 * + [i_a == i_b] h2(j_a,j_b).  Host only. */
int scalit_combine_2d(int n1, const double *h1, const double *v1, int n2, const double *h2, const double *v2,
                      int nkeep, int *idx1, int *idx2, double *h0c);

/* ------------------------------------------------------------------------ */
/* on-disk formats (host only, no device needed)                             */
/* ------------------------------------------------------------------------ */

/* saveData / loadData (src/libutil/serial/io/io.f90:9-33, 98-126): bin_mode != 0
 * is Fortran sequential unformatted, one record framed by 4-byte byte counts
 * (at most 2^31-1 bytes); bin_mode == 0 is list-directed text.  Loading reads
 * n values from the first record (text: the first n items). */
int scalit_save_data(const char *path, long n, const double *data, int bin_mode);
int scalit_load_data(const char *path, long n, double *data, int bin_mode);
/* saveDataDir / loadDataDir (src/libutil/serial/io/iodirect.f90:10-36, 68-93):
 * direct access, one record whose length is the array's: the raw n doubles. */
int scalit_save_data_dir(const char *path, long n, const double *data);
int scalit_load_data_dir(const char *path, long n, double *data);
/* P2SaveDataGrid's per-point direct-access records
 * (src/libutil/parallel/p_io/p2io_grid.f90:62-83), the OUTH / H0GM files of
 * ScalIT's ja3 generator: record p holds the ng x ng block of grid point p
 * (column-major, RECL = ng*ng doubles, no record markers).  outh is the
 * operator's layout, npts x ng x ng with the point index fastest (as
 * scalit_set_outh takes it); npts records are read or written. */
int scalit_load_data_grid(const char *path, long npts, int ng, double *outh);
int scalit_save_data_grid(const char *path, long npts, int ng, const double *outh);
/* calOSB / calDataInfo (src/iterate/module/osb.init.h:2-49): 0-based start of
 * each level's piece in the concatenated H0 (sN(l)^2 each, none for the sNDVR
 * layer), DEP (myDim(l) for each dependent level) and VOSB (myBlk(l) sN(l)^2)
 * arrays, any of which may be NULL, and totals[5] = {myLen, myH0%mLen,
 * myDEP%mLen, myVOSB%mLen, outLen}: the sizes of the files the reference's
 * loadH0, loadDEP, loadVOSBEig and loadOutH read (osb.io.h:46-62, 171-235). */
int scalit_layout(int sF, const int *sN, const int *sDEP, int ndvr, long *h0_start, long *dep_start,
                  long *vosb_start, long *totals);

#ifdef __cplusplus
}
#endif
#endif
