/*
 * CPU oracle for the ScalIT H*X / OSB / QMR / PIST / Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, the smoke check and the CPU-baseline leg of the
 * benchmark as the checker.  The product library (scalit_amd) never links,
 * loads or calls it.
 *
 * Conventions follow the reference: all arrays are column-major, FP64, and a
 * vector on the direct-product grid has level 1 as the fastest index
 * (src/iterate/module/osb.init.h:2-50).  Indices in the comments are the
 * reference's 1-based ones; the code is 0-based.
 */
#ifndef SCALIT_ORACLE_H
#define SCALIT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_FMAX 8

/* H*X / P*X shift selector (reference TA0 / TA1, src/iterate/module/osb.hx.h:82-89) */
#define ORC_TA0 0
#define ORC_TA1 1
/* preconditioner kernels (src/iterate/module/osb.px.h:156-211) */
#define ORC_TOSB   0
#define ORC_TOSBD1 1
#define ORC_TOSBD2 2

typedef struct {
    int sF;                   /* number of levels */
    int sN[ORC_FMAX];         /* grid size per level */
    int sDEP[ORC_FMAX];       /* level uses the coordinate-dependent factor */
    int myDim[ORC_FMAX];      /* product of inner sizes (NIN) */
    int myBlk[ORC_FMAX];      /* product of outer sizes (BLK) */
    long myLen;               /* N */
    const double *H0[ORC_FMAX];   /* n x n, column-major */
    const double *DEP[ORC_FMAX];  /* myDim(level) or NULL */
    const double *RES;            /* N */
    double *VOSB[ORC_FMAX];       /* myBlk * n * n  (filled by orc_my_diag) */
    double *EIG0;                 /* N              (filled by orc_my_diag) */
    double E0;                    /* sOSBW%mE0 */
    double DE, beta;              /* sOSBW%mDE / mBeta for the OSBD kernels */
    int osb_kind;                 /* ORC_TOSB / ORC_TOSBD1 / ORC_TOSBD2 */
    int hc, pc;                   /* sHC / sPC */
    /* sNDVR = T: the outermost level is the dense per-point layer OUTH
     * (myDim(sF) x sN(sF) x sN(sF)), src/iterate/module/osb.hx.h:93-94 */
    int ndvr;
    const double *OUTH;
} orc_osb;

/* ---- sizes: src/iterate/module/osb.init.h:2-50 ---- */
void orc_cal_osb(orc_osb *o);

/* ---- dense mode products: src/libutil/serial/base/matrix.f90:91-226 ---- */
void orc_h1x(int nin, int nout, const double *h1, const double *x, double *xout);
void orc_h1tx(int nin, int nout, const double *h1, const double *x, double *xout);
void orc_h2x(int nin, int nout, const double *h1, const double *corx, const double *x, double *xout);
void orc_h3x(int nin, int nout, const double *a, const double *x, double *xout);
/* adjoint of H3X: XOUT(:,i) = sum_j A(:,j,i) X(:,j) */
void orc_h3tx(int nin, int nout, const double *a, const double *x, double *xout);
/* ---- src/iterate/libiterate/hpx/hpx.f90:5-73 ---- */
void orc_h1x_xyz(int nin, int nout, int blk, const double *h1, const double *x, double *y);
void orc_h1x_dep(int nin, int nout, int blk, const double *h1, const double *corx, const double *x, double *y);
void orc_vx(int nin, int nout, int blk, const double *vout, double *x);
void orc_vtx(int nin, int nout, int blk, const double *vout, double *x);

/* ---- operator: src/iterate/module/osb.hx.h:71-110, osb.px.h:125-219 ---- */
void orc_myhx(const orc_osb *o, int ntype, const double *x, double *y);
void orc_mypx(const orc_osb *o, int nosb, int npc, const double *x, double *y);
/* adjoint of MYHX: y = H^T x (H0 transposed per level); V and DEP are diagonal */
void orc_myhx_t(const orc_osb *o, int ntype, const double *x, double *y);

/* ---- OSB set-up: src/iterate/libiterate/diag/{hosbdiag,jacobi,hinit*,vupdate,sumdiag}.f90, src/iterate/module/osb.diag.h:53-151 ---- */
void orc_hosbinit_xyz(int blk, int mak, int ndk, const double *h1, const double *res, double *hosb);
void orc_hosbinit_dep(int blk, int mak, int ndk, const double *h1, const double *dep, const double *res, double *hosb);
double orc_sum_off_diag(int nin, int nout, const double *hout);
double orc_mjacobi(int p, int q, int nout, int nin, double *hout);   /* p,q 1-based */
void orc_vinit(int nout, double *vout);
void orc_vupdate(int p, int q, double phi, int nout, double *vout);  /* p,q 1-based */
void orc_hosb_diag(int bjmax, double bjtol, int blk, int mak, int ndk, double *hosb, double *vosb, double *eigval,
                   double *stats5);
int  orc_my_diag(orc_osb *o, int bjmax, double bjtol);   /* VOSB[l], EIG0 caller-allocated */

/* ---- vector helpers (Fortran dot_product order) ---- */
double orc_dot(long n, const double *a, const double *b);

/* ---- Krylov: src/iterate/libiterate/{qmr,orth,pist,lanczos} ---- */
typedef void (*orc_op_fn)(void *ctx, long n, const double *x, double *y);
typedef int  (*orc_linsolv_fn)(void *ctx, long n, const double *b, double *x, double *res);

/* callback context for the OSB operator (sHC = sPC = TA1) */
typedef struct {
    orc_osb *o;
    int qmr_type, qmr_max;
    double qmr_tol;
    long hx_calls, px_calls, solves, qmr_iters;
    /* nonzero: H*X and P*X run on the host threads (orc_myhx_omp /
     * orc_mypx_omp, results identical to the serial ones) with tmp as their
     * N-value work vector */
    int threaded;
    double *tmp;
} orc_solver_ctx;
void orc_cb_hx_e(void *ctx, long n, const double *x, double *y);   /* (H - E0) x */
void orc_cb_hx0(void *ctx, long n, const double *x, double *y);    /* H x */
void orc_cb_px(void *ctx, long n, const double *x, double *y);     /* V (EIG0 - E0)^-1 V^T x */
int  orc_cb_linsolv(void *ctx, long n, const double *b, double *x, double *res);  /* OSB_QMR */

int orc_qmr(int ntype, int qmrmax, double qmrtol, long n, const double *b,
            orc_op_fn hx, orc_op_fn px, void *ctx, double *x, double *res);
int orc_mgs_orth(long n, const double *wj, int m, double *vj);
int orc_lan_mgs_orth(long n, const double *wj, int m, double *vj, double *beta);
int orc_pist_h0(long n, const double *x, int m_max, int m_low, int m_high,
                orc_op_fn h0x, orc_linsolv_fn linsolv, void *ctx, double *vj, double *hmat, int *qmr_iters);
int orc_pist_conv(double e0, double etol, int ntype, long n, const double *x, int startcnt, int stepcnt,
                  int stepeig, int m, int m_max, orc_op_fn h0x, orc_linsolv_fn linsolv, void *ctx,
                  double *eig, double *res, int *qmr_iters);
int orc_lan_h0(long n, const double *x, double *vj, int m_low, int m_high, orc_op_fn hx, void *ctx,
               double *alpha, double *beta);
int orc_lan_conv(double e0, double etol, int ntype, long n, const double *x, int startcnt, int stepcnt,
                 int stepeig, int m, int m_max, orc_op_fn hx, void *ctx, double *eig, double *res);

/* ---- small dense eigensolvers standing in for DSYEV('N') / DSTEV('N') ---- */
int orc_sym_eigvals(int n, const double *a, int lda, double *w);      /* ascending */
int orc_tridiag_eigvals(int n, const double *d, const double *e, double *w);

/* ---- window helpers: src/libutil/serial/base/util.f90 ---- */
void orc_get_window0(double e0, int n, const double *vec, int m, int *vecind, double *newvec);
double orc_get_diff_max(int n, const double *v1, const double *v2);
double orc_get_sep_max(int n1, const double *v1, int n2, const double *v2);
void orc_reorder_asc(int n, double *vec);

/* ---- DVR closed forms: src/ham/libham/dvr/sinc.f90, sinc_xh.f90, src/ham/libham/poly ---- */
void orc_h1_sinc2(int n, double a0, double b0, double mass, double *h1);
void orc_sinc2_dvr(int n, double a0, double b0, double mass, double *x0, double *h0);
void orc_gau_legendre(double x1, double x2, int n, double *x, double *w);
void orc_yj_polys(int n0, int l, const double *x, double *yj);
void orc_legendre_l2(int n, double *x, double *l2);   /* j(j+1) in the Gauss-Legendre DVR */

/* ---- OpenMP-threaded H*X for the CPU baseline (same loop nests, blocks split over threads) ---- */
void orc_myhx_omp(const orc_osb *o, int ntype, const double *x, double *y, double *tmp);
void orc_mypx_omp(const orc_osb *o, int nosb, int npc, const double *x, double *y);
int  orc_omp_threads(void);

#ifdef __cplusplus
}
#endif
#endif
