/*
 * TEST INFRASTRUCTURE ONLY.  Glue for the reference checker library
 * (_ref/libscalit_ref.so).  The reference's operator callbacks HX/PX and the
 * linear-solver callback live in Fortran modules that pull in its whole I/O
 * and input layer; this file restates those module loops in C
 * (src/iterate/module/osb.hx.h:71-110, osb.px.h:125-219, osbw.qmr.h:13-23,
 * osb.diag.h:111-137) and calls the reference's own compiled library routines
 * (H1X_XYZ, H1X_DEP, H3X, VX, VTX, HOSBINIT_*, HosbDiag, QMR) for all arithmetic.
 */
#include <stdlib.h>
#include <string.h>

#define FMAX 8

extern void h1x_xyz_(const int *nin, const int *nout, const int *blk, const double *h1, const double *x, double *y);
extern void h1x_dep_(const int *nin, const int *nout, const int *blk, const double *h1, const double *corx,
                     const double *x, double *y);
extern void h3x_(const int *nin, const int *nout, const double *a, const double *x, double *xout);
extern void vx_(const int *nin, const int *nout, const int *blk, const double *vout, double *x);
extern void vtx_(const int *nin, const int *nout, const int *blk, const double *vout, double *x);
extern void hosbinit_xyz_(const int *blk, const int *mak, const int *ndk, const double *h1, const double *res,
                          double *hosb);
extern void hosbinit_dep_(const int *blk, const int *mak, const int *ndk, const double *h1, const double *dep,
                          const double *res, double *hosb);
extern void hosbdiag_(const int *bjmax, const double *bjtol, const int *blk, const int *mak, const int *ndk,
                      double *hosb, double *vosb, double *eigval);
typedef void (*fop)(const int *, const double *, double *);
extern int qmr_(const int *ntype, const int *qmrmax, const double *qmrtol, const int *n, const double *b, fop hx,
                fop px, double *x, double *res);

static struct {
    int sF, sN[FMAX], sDEP[FMAX], myDim[FMAX], myBlk[FMAX], myLen;
    const double *H0[FMAX], *DEP[FMAX], *RES;
    const double *OUTH;   /* sNDVR = T outer layer, or NULL */
    const double *VOSB[FMAX];
    const double *EIG0;
    double E0;
    int qmr_type, qmr_max;
    double qmr_tol;
    double *tmp;
} S;

void refdrv_setup(int sF, const int *sN, const int *sDEP, const double **H0, const double **DEP, const double *RES,
                  double E0, const double *OUTH)
{
    S.OUTH = OUTH;
    S.sF = sF;
    for (int i = 0; i < sF; i++) {
        S.sN[i] = sN[i];
        S.sDEP[i] = sDEP[i];
        S.H0[i] = H0[i];
        S.DEP[i] = DEP[i];
    }
    S.myDim[0] = 1;
    for (int i = 0; i < sF - 1; i++) S.myDim[i + 1] = S.myDim[i] * sN[i];
    S.myBlk[sF - 1] = 1;
    for (int i = sF - 1; i >= 1; i--) S.myBlk[i - 1] = S.myBlk[i] * sN[i];
    S.myLen = S.myBlk[0] * sN[0];
    S.RES = RES;
    S.E0 = E0;
    free(S.tmp);
    S.tmp = (double *)malloc(sizeof(double) * S.myLen);
}

void refdrv_set_precond(const double **VOSB, const double *EIG0)
{
    for (int i = 0; i < S.sF; i++) S.VOSB[i] = VOSB[i];
    S.EIG0 = EIG0;
}

void refdrv_set_qmr(int ntype, int qmrmax, double qmrtol)
{
    S.qmr_type = ntype;
    S.qmr_max = qmrmax;
    S.qmr_tol = qmrtol;
}

static void myhx(double c, const double *x, double *y)
{
    int n = S.myLen;
    for (int k = 0; k < n; k++) y[k] = (S.RES[k] - c) * x[k];
    for (int l = 0; l < S.sF; l++) {
        if (S.OUTH && l == S.sF - 1) h3x_(&S.myDim[l], &S.sN[l], S.OUTH, x, S.tmp);   /* osb.hx.h:93-94 */
        else if (S.sDEP[l]) h1x_dep_(&S.myDim[l], &S.sN[l], &S.myBlk[l], S.H0[l], S.DEP[l], x, S.tmp);
        else h1x_xyz_(&S.myDim[l], &S.sN[l], &S.myBlk[l], S.H0[l], x, S.tmp);
        for (int k = 0; k < n; k++) y[k] = y[k] + S.tmp[k];
    }
}

/* HX0 (TA0): Y = H X */
void refdrv_hx0(const int *n, const double *x, double *y) { (void)n; myhx(0.0, x, y); }
/* EHX / HX with sHC=TA1: Y = (H - E0) X.  TA0 subtracts nothing, TA1 subtracts E0;
 * y = (RES - E0) x computed as in osb.hx.h:86. */
void refdrv_hx(const int *n, const double *x, double *y)
{
    (void)n;
    int nn = S.myLen;
    for (int k = 0; k < nn; k++) y[k] = (S.RES[k] - S.E0) * x[k];
    for (int l = 0; l < S.sF; l++) {
        if (S.OUTH && l == S.sF - 1) h3x_(&S.myDim[l], &S.sN[l], S.OUTH, x, S.tmp);   /* osb.hx.h:93-94 */
        else if (S.sDEP[l]) h1x_dep_(&S.myDim[l], &S.sN[l], &S.myBlk[l], S.H0[l], S.DEP[l], x, S.tmp);
        else h1x_xyz_(&S.myDim[l], &S.sN[l], &S.myBlk[l], S.H0[l], x, S.tmp);
        for (int k = 0; k < nn; k++) y[k] = y[k] + S.tmp[k];
    }
}

/* PX with sOSB=TOSB, sPC=TA1 */
void refdrv_px(const int *n, const double *x, double *y)
{
    (void)n;
    int nn = S.myLen;
    memcpy(y, x, sizeof(double) * nn);
    for (int l = S.sF - 1; l >= 0; l--) vtx_(&S.myDim[l], &S.sN[l], &S.myBlk[l], S.VOSB[l], y);
    for (int k = 0; k < nn; k++) y[k] = y[k] / (S.EIG0[k] - S.E0);
    for (int l = 0; l < S.sF; l++) vx_(&S.myDim[l], &S.sN[l], &S.myBlk[l], S.VOSB[l], y);
}

/* OSB_QMR: the LinSolv callback of PIST */
int refdrv_linsolv(const int *n, const double *b, double *x, double *res)
{
    return qmr_(&S.qmr_type, &S.qmr_max, &S.qmr_tol, n, b, refdrv_hx, refdrv_px, x, res);
}

/* myDiag for the real case (osb.diag.h:111-137 with myDiagInit :18-31); EIG0
 * starts as RES.  VOSB[l] and EIG0 are caller-allocated. */
int refdrv_diag(int bjmax, double bjtol, double **VOSB, double *EIG0)
{
    int n = S.myLen;
    memcpy(EIG0, S.RES, sizeof(double) * n);
    for (int l = S.sF - 1; l >= 0; l--) {
        double *hosb = (double *)malloc(sizeof(double) * (size_t)S.sN[l] * n);
        if (!hosb) return 0;
        if (S.OUTH && l == S.sF - 1) memcpy(hosb, S.OUTH, sizeof(double) * (size_t)S.sN[l] * n);   /* osb.diag.h:22-23 */
        else if (S.sDEP[l]) hosbinit_dep_(&S.myBlk[l], &S.sN[l], &S.myDim[l], S.H0[l], S.DEP[l], EIG0, hosb);
        else hosbinit_xyz_(&S.myBlk[l], &S.sN[l], &S.myDim[l], S.H0[l], EIG0, hosb);
        hosbdiag_(&bjmax, &bjtol, &S.myBlk[l], &S.sN[l], &S.myDim[l], hosb, VOSB[l], EIG0);
        free(hosb);
    }
    return 1;
}
