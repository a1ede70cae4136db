/*
 * TEST INFRASTRUCTURE ONLY.  Glue for the reference checker library
 * (_ref/libscalit_ref.so).  H*X and P*X are the reference's own module
 * routines MYHX / MYPX (src/iterate/module/osb.hx.h:71-110, osb.px.h:125-219),
 * called through ref_osbop.f90 on the state this file hands over; the
 * linear-solver callback (osbw.qmr.h:13-23) and the OSB set-up loop
 * (osb.diag.h:111-137) are restated here around the reference's compiled
 * QMR, HOSBINIT_* and HosbDiag, which do all the arithmetic.
 */
#include <stdlib.h>
#include <string.h>

#define FMAX 8

extern void hosbinit_xyz_(const int *blk, const int *mak, const int *ndk, const double *h1, const double *res,
                          double *hosb);
extern void hosbinit_dep_(const int *blk, const int *mak, const int *ndk, const double *h1, const double *dep,
                          const double *res, double *hosb);
extern void hosbdiag_(const int *bjmax, const double *bjtol, const int *blk, const int *mak, const int *ndk,
                      double *hosb, double *vosb, double *eigval);
typedef void (*fop)(const int *, const double *, double *);
/* ref_osbop.f90: module osb's state and its MYHX / MYPX */
extern void refop_set(int f, const int *n, const int *dep, int ndvr, const double *h0, const double *dep_v,
                      const double *res, const double *outh);
extern void refop_set_shift(double e0, double de, double beta);
extern void refop_set_precond(const double *vosb, const double *eig0);
extern void refop_hx(int ntype, const double *x, double *y);
extern void refop_px(int nosb, int npc, const double *x, double *y);
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
    /* module osb's concatenated H0 (sN(l)^2 per level, none for the sNDVR
     * layer) and DEP (myDim(l) per dependent level), osb.init.h:22-46 */
    long nh = 0, nd = 0;
    for (int i = 0; i < sF; i++) {
        if (!(OUTH && i == sF - 1)) nh += (long)sN[i] * sN[i];
        if (sDEP[i] && !(OUTH && i == sF - 1)) nd += S.myDim[i];
    }
    double *h = (double *)malloc(sizeof(double) * (nh > 0 ? nh : 1));
    double *d = (double *)malloc(sizeof(double) * (nd > 0 ? nd : 1));
    long oh = 0, od = 0;
    int dep_eff[FMAX];
    for (int i = 0; i < sF; i++) {
        const int layer = OUTH && i == sF - 1;
        dep_eff[i] = layer ? 0 : sDEP[i];
        if (!layer) {
            memcpy(h + oh, H0[i], sizeof(double) * sN[i] * sN[i]);
            oh += (long)sN[i] * sN[i];
        }
        if (dep_eff[i]) {
            memcpy(d + od, DEP[i], sizeof(double) * S.myDim[i]);
            od += S.myDim[i];
        }
    }
    refop_set(sF, sN, dep_eff, OUTH ? 1 : 0, h, d, RES, OUTH);
    refop_set_shift(E0, 1.0, 0.0);
    free(h);
    free(d);
}

void refdrv_set_shift(double e0, double de, double beta)
{
    S.E0 = e0;
    refop_set_shift(e0, de, beta);
}

void refdrv_set_precond(const double **VOSB, const double *EIG0)
{
    long nv = 0;
    for (int i = 0; i < S.sF; i++) {
        S.VOSB[i] = VOSB[i];
        nv += (long)S.myBlk[i] * S.sN[i] * S.sN[i];
    }
    S.EIG0 = EIG0;
    double *v = (double *)malloc(sizeof(double) * nv);
    long o = 0;
    for (int i = 0; i < S.sF; i++) {
        const long len = (long)S.myBlk[i] * S.sN[i] * S.sN[i];
        memcpy(v + o, VOSB[i], sizeof(double) * len);
        o += len;
    }
    refop_set_precond(v, EIG0);
    free(v);
}

void refdrv_set_qmr(int ntype, int qmrmax, double qmrtol)
{
    S.qmr_type = ntype;
    S.qmr_max = qmrmax;
    S.qmr_tol = qmrtol;
}

/* HX0 (TA0): Y = H X, module osb's MYHX */
void refdrv_hx0(const int *n, const double *x, double *y)
{
    (void)n;
    refop_hx(0, x, y);
}
/* EHX / HX with sHC=TA1: Y = (H - E0) X, module osb's MYHX */
void refdrv_hx(const int *n, const double *x, double *y)
{
    (void)n;
    refop_hx(1, x, y);
}

/* PX with sOSB=TOSB, sPC=TA1: module osb's MYPX */
void refdrv_px(const int *n, const double *x, double *y)
{
    (void)n;
    refop_px(1, 1, x, y);
}
/* MYPX with kernel nosb (0 TOSB, 1 TOSBD1, 2 TOSBD2, this library's numbering)
 * and npc (0 TA0, 1 TA1) */
void refdrv_px_kind(int nosb, int npc, const double *x, double *y) { refop_px(nosb + 1, npc, x, y); }

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
