/*
 * CPU oracle for the ScalIT H*X / OSB / QMR / PIST / Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's Fortran,
 * loop by loop and in the same floating-point evaluation order, so that it can
 * be checked against the reference routines compiled from their own sources.
 * Only tests, the smoke check and the benchmark's CPU-baseline leg use it; the
 * product library never does.
 *
 * Every function cites the reference file:line it follows.  The two small
 * LAPACK eigensolvers the reference calls (DSYEV('N'), DSTEV('N')) are not part
 * of the reference sources; they are replaced by a cyclic Jacobi eigensolver
 * and checked against scipy/numpy in tests/test_oracle_eig.py.
 */
#include "scalit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX2(i, j, ld) ((long)(i) + (long)(ld) * (long)(j))

/* src/iterate/module/osb.init.h:2-16 */
void orc_cal_osb(orc_osb *o)
{
    int i;
    o->myDim[0] = 1;
    for (i = 0; i < o->sF - 1; i++) o->myDim[i + 1] = o->myDim[i] * o->sN[i];
    o->myBlk[o->sF - 1] = 1;
    for (i = o->sF - 1; i >= 1; i--) o->myBlk[i - 1] = o->myBlk[i] * o->sN[i];
    o->myLen = (long)o->myBlk[0] * o->sN[0];
}

/* src/libutil/serial/base/matrix.f90:91-107  XOUT(:,i) = sum_j H1(i,j) X(:,j) */
void orc_h1x(int nin, int nout, const double *h1, const double *x, double *xout)
{
    for (int i = 0; i < nout; i++) {
        double *xo = xout + (long)nin * i;
        double h = h1[IDX2(i, 0, nout)];
        for (int k = 0; k < nin; k++) xo[k] = h * x[k];
        for (int j = 1; j < nout; j++) {
            h = h1[IDX2(i, j, nout)];
            const double *xj = x + (long)nin * j;
            for (int k = 0; k < nin; k++) xo[k] = xo[k] + h * xj[k];
        }
    }
}

/* src/libutil/serial/base/matrix.f90:116-130  XOUT(:,i) = sum_j H1(j,i) X(:,j) */
void orc_h1tx(int nin, int nout, const double *h1, const double *x, double *xout)
{
    for (int i = 0; i < nout; i++) {
        double *xo = xout + (long)nin * i;
        double h = h1[IDX2(0, i, nout)];
        for (int k = 0; k < nin; k++) xo[k] = h * x[k];
        for (int j = 1; j < nout; j++) {
            h = h1[IDX2(j, i, nout)];
            const double *xj = x + (long)nin * j;
            for (int k = 0; k < nin; k++) xo[k] = xo[k] + h * xj[k];
        }
    }
}

/* src/libutil/serial/base/matrix.f90:139-155  (H1(i,j)*CORX)*X, evaluated left to right */
void orc_h2x(int nin, int nout, const double *h1, const double *corx, const double *x, double *xout)
{
    for (int i = 0; i < nout; i++) {
        double *xo = xout + (long)nin * i;
        double h = h1[IDX2(i, 0, nout)];
        for (int k = 0; k < nin; k++) xo[k] = h * corx[k] * x[k];
        for (int j = 1; j < nout; j++) {
            h = h1[IDX2(i, j, nout)];
            const double *xj = x + (long)nin * j;
            for (int k = 0; k < nin; k++) xo[k] = xo[k] + h * corx[k] * xj[k];
        }
    }
}

/* src/libutil/serial/base/matrix.f90:188-203  XOUT(:,i) = sum_j A(:,i,j) X(:,j) */
void orc_h3x(int nin, int nout, const double *a, const double *x, double *xout)
{
    for (int i = 0; i < nout; i++) {
        double *xo = xout + (long)nin * i;
        const double *ai = a + (long)nin * i;
        for (int k = 0; k < nin; k++) xo[k] = ai[k] * x[k];
        for (int j = 1; j < nout; j++) {
            const double *aij = a + (long)nin * (i + (long)nout * j);
            const double *xj = x + (long)nin * j;
            for (int k = 0; k < nin; k++) xo[k] = xo[k] + aij[k] * xj[k];
        }
    }
}

void orc_h3tx(int nin, int nout, const double *a, const double *x, double *xout)
{
    for (int i = 0; i < nout; i++) {
        double *xo = xout + (long)nin * i;
        for (int k = 0; k < nin; k++) xo[k] = 0.0;
        for (int j = 0; j < nout; j++) {
            const double *aji = a + (long)nin * (j + (long)nout * i);
            const double *xj = x + (long)nin * j;
            for (int k = 0; k < nin; k++) xo[k] = xo[k] + aji[k] * xj[k];
        }
    }
}

/* src/iterate/libiterate/hpx/hpx.f90:5-18 */
void orc_h1x_xyz(int nin, int nout, int blk, const double *h1, const double *x, double *y)
{
    long bs = (long)nin * nout;
    for (int i = 0; i < blk; i++) orc_h1x(nin, nout, h1, x + bs * i, y + bs * i);
}

/* src/iterate/libiterate/hpx/hpx.f90:21-35 */
void orc_h1x_dep(int nin, int nout, int blk, const double *h1, const double *corx, const double *x, double *y)
{
    long bs = (long)nin * nout;
    for (int i = 0; i < blk; i++) orc_h2x(nin, nout, h1, corx, x + bs * i, y + bs * i);
}

/* src/iterate/libiterate/hpx/hpx.f90:40-54  X(:,:,k) <- V_k X(:,:,k) */
void orc_vx(int nin, int nout, int blk, const double *vout, double *x)
{
    long bs = (long)nin * nout;
    double *tmp = (double *)malloc(sizeof(double) * bs);
    for (int i = 0; i < blk; i++) {
        orc_h1x(nin, nout, vout + (long)nout * nout * i, x + bs * i, tmp);
        memcpy(x + bs * i, tmp, sizeof(double) * bs);
    }
    free(tmp);
}

/* src/iterate/libiterate/hpx/hpx.f90:59-73  X(:,:,k) <- V_k^T X(:,:,k) */
void orc_vtx(int nin, int nout, int blk, const double *vout, double *x)
{
    long bs = (long)nin * nout;
    double *tmp = (double *)malloc(sizeof(double) * bs);
    for (int i = 0; i < blk; i++) {
        orc_h1tx(nin, nout, vout + (long)nout * nout * i, x + bs * i, tmp);
        memcpy(x + bs * i, tmp, sizeof(double) * bs);
    }
    free(tmp);
}

/* src/iterate/module/osb.hx.h:71-110 (sNDVR=F form; levels 1..sF, Y += tmp per level) */
void orc_myhx(const orc_osb *o, int ntype, const double *x, double *y)
{
    long n = o->myLen;
    double *tmp = (double *)malloc(sizeof(double) * n);
    if (ntype == ORC_TA0) {
        for (long k = 0; k < n; k++) y[k] = o->RES[k] * x[k];
    } else {
        for (long k = 0; k < n; k++) y[k] = (o->RES[k] - o->E0) * x[k];
    }
    for (int l = 0; l < o->sF; l++) {
        if (o->ndvr && l == o->sF - 1)
            orc_h3x(o->myDim[l], o->sN[l], o->OUTH, x, tmp);
        else if (o->sDEP[l])
            orc_h1x_dep(o->myDim[l], o->sN[l], o->myBlk[l], o->H0[l], o->DEP[l], x, tmp);
        else
            orc_h1x_xyz(o->myDim[l], o->sN[l], o->myBlk[l], o->H0[l], x, tmp);
        for (long k = 0; k < n; k++) y[k] = y[k] + tmp[k];
    }
    free(tmp);
}

/* Adjoint of orc_myhx: every level uses H0^T (H2X with the transposed matrix,
 * matrix.f90:163-179 H2TX); RES and DEP are diagonal so they are unchanged. */
void orc_myhx_t(const orc_osb *o, int ntype, const double *x, double *y)
{
    long n = o->myLen;
    double *tmp = (double *)malloc(sizeof(double) * n);
    if (ntype == ORC_TA0) {
        for (long k = 0; k < n; k++) y[k] = o->RES[k] * x[k];
    } else {
        for (long k = 0; k < n; k++) y[k] = (o->RES[k] - o->E0) * x[k];
    }
    for (int l = 0; l < o->sF; l++) {
        int nin = o->myDim[l], nout = o->sN[l], blk = o->myBlk[l];
        long bs = (long)nin * nout;
        if (o->ndvr && l == o->sF - 1) {
            orc_h3tx(nin, nout, o->OUTH, x, tmp);
            for (long k = 0; k < n; k++) y[k] = y[k] + tmp[k];
            continue;
        }
        for (int b = 0; b < blk; b++) {
            const double *xb = x + bs * b;
            double *tb = tmp + bs * b;
            for (int i = 0; i < nout; i++) {
                double *xo = tb + (long)nin * i;
                for (int k = 0; k < nin; k++) xo[k] = 0.0;
                for (int j = 0; j < nout; j++) {
                    double h = o->H0[l][IDX2(j, i, nout)];
                    for (int k = 0; k < nin; k++)
                        xo[k] += (o->sDEP[l] ? h * o->DEP[l][k] : h) * xb[(long)nin * j + k];
                }
            }
        }
        for (long k = 0; k < n; k++) y[k] = y[k] + tmp[k];
    }
    free(tmp);
}

/* src/iterate/module/osb.px.h:125-219 (TOSB, TOSBD1, TOSBD2 kernels) */
void orc_mypx(const orc_osb *o, int nosb, int npc, const double *x, double *y)
{
    long n = o->myLen;
    double de = fabs(o->DE), rde = 1.0 / de, e0;
    memcpy(y, x, sizeof(double) * n);
    for (int l = o->sF - 1; l >= 0; l--) orc_vtx(o->myDim[l], o->sN[l], o->myBlk[l], o->VOSB[l], y);
    e0 = (npc == ORC_TA0) ? 0.0 : o->E0;
    if (nosb == ORC_TOSBD1) {
        for (long i = 0; i < n; i++) {
            double ade = o->EIG0[i] - e0;
            if (fabs(ade) <= de) y[i] = (ade < 0.0) ? -y[i] * rde : y[i] * rde;
            else y[i] = y[i] / ade;
        }
    } else if (nosb == ORC_TOSBD2) {
        double beta = 1.0 + o->beta, alpha = -o->beta * de;
        for (long i = 0; i < n; i++) {
            double ade = o->EIG0[i] - e0, t;
            if (fabs(ade) <= de) t = (ade < 0.0) ? (-alpha + beta * ade) : (alpha + beta * ade);
            else t = ade;
            y[i] = y[i] / t;
        }
    } else {
        for (long i = 0; i < n; i++) y[i] = y[i] / (o->EIG0[i] - e0);
    }
    for (int l = 0; l < o->sF; l++) orc_vx(o->myDim[l], o->sN[l], o->myBlk[l], o->VOSB[l], y);
}

/* src/iterate/libiterate/diag/hinitxyz.f90:7-29 */
void orc_hosbinit_xyz(int blk, int mak, int ndk, const double *h1, const double *res, double *hosb)
{
    for (int i = 0; i < blk; i++)
        for (int j = 0; j < mak; j++)
            for (int k = 0; k < mak; k++) {
                double *h = hosb + (long)ndk * (k + (long)mak * (j + (long)mak * i));
                for (int d = 0; d < ndk; d++) h[d] = h1[IDX2(k, j, mak)];
            }
    for (int j = 0; j < blk; j++)
        for (int k = 0; k < mak; k++) {
            double *h = hosb + (long)ndk * (k + (long)mak * (k + (long)mak * j));
            const double *r = res + (long)ndk * (k + (long)mak * j);
            for (int d = 0; d < ndk; d++) h[d] = h[d] + r[d];
        }
}

/* src/iterate/libiterate/diag/hinitdep.f90:10-34 */
void orc_hosbinit_dep(int blk, int mak, int ndk, const double *h1, const double *dep, const double *res, double *hosb)
{
    for (int i = 0; i < blk; i++)
        for (int j = 0; j < mak; j++)
            for (int k = 0; k < mak; k++) {
                double *h = hosb + (long)ndk * (k + (long)mak * (j + (long)mak * i));
                for (int d = 0; d < ndk; d++) h[d] = h1[IDX2(k, j, mak)] * dep[d];
            }
    for (int j = 0; j < blk; j++)
        for (int k = 0; k < mak; k++) {
            double *h = hosb + (long)ndk * (k + (long)mak * (k + (long)mak * j));
            const double *r = res + (long)ndk * (k + (long)mak * j);
            for (int d = 0; d < ndk; d++) h[d] = h[d] + r[d];
        }
}

/* src/iterate/libiterate/diag/sumdiag.f90:6-22 */
double orc_sum_off_diag(int nin, int nout, const double *hout)
{
    double s = 0.0;
    for (int i = 0; i < nout; i++)
        for (int j = 0; j < i; j++) {
            const double *h = hout + (long)nin * (j + (long)nout * i);
            double d = 0.0;
            for (int k = 0; k < nin; k++) d += h[k] * h[k];
            s = s + d;
        }
    return 2.0 * s;
}

/* src/iterate/libiterate/diag/jacobi.f90:2-83  (one joint rotation of the NIN matrices) */
double orc_mjacobi(int p1, int q1, int nout, int nin, double *hout)
{
    const double PIE = 3.14159265358979323846264338328;
    int p = p1 - 1, q = q1 - 1;
#define HO(d, r, c) hout[(long)(d) + (long)nin * ((long)(r) + (long)nout * (long)(c))]
    double *odg = (double *)malloc(sizeof(double) * nin * 5);
    double *ddg = odg + nin, *newp = odg + 2 * (long)nin, *newq = odg + 3 * (long)nin, *newqp = odg + 4 * (long)nin;
    double A = 0.0, B = 0.0, C = 0.0, X, SN, CS, C2, S2, SC;
    for (int d = 0; d < nin; d++) {
        odg[d] = HO(d, p, q);
        ddg[d] = 0.5 * HO(d, p, p) - 0.5 * HO(d, q, q);
    }
    for (int d = 0; d < nin; d++) A += odg[d] * odg[d];
    for (int d = 0; d < nin; d++) B += odg[d] * ddg[d];
    B = 2.0 * B;
    for (int d = 0; d < nin; d++) C += ddg[d] * ddg[d];

    if (A == C) {
        if (B == 0.0) X = 0.0;
        else X = -copysign(PIE / 8.0, B);
    } else {
        X = atan(B / (A - C)) / 4.0;
        if (A > C) {
            if (X < 0.0) X = X + PIE / 4.0;
            else X = X - PIE / 4.0;
        }
    }
    SN = sin(X); CS = cos(X);
    C2 = CS * CS; S2 = SN * SN; SC = SN * CS;

    for (int d = 0; d < nin; d++) {
        newp[d] = C2 * HO(d, p, p) + S2 * HO(d, q, q) - 2 * SC * odg[d];
        newq[d] = C2 * HO(d, q, q) + S2 * HO(d, p, p) + 2 * SC * odg[d];
        newqp[d] = (C2 - S2) * odg[d] + 2 * SC * ddg[d];
    }
    for (int d = 0; d < nin; d++) {
        HO(d, p, p) = newp[d];
        HO(d, q, q) = newq[d];
        HO(d, p, q) = newqp[d];
        HO(d, q, p) = newqp[d];
    }
    for (int r = 0; r < nout; r++) {
        if (r != p && r != q) {
            for (int d = 0; d < nin; d++) {
                double np = CS * HO(d, r, p) - SN * HO(d, r, q);
                double nq = CS * HO(d, r, q) + SN * HO(d, r, p);
                HO(d, r, p) = np; HO(d, p, r) = np;
                HO(d, r, q) = nq; HO(d, q, r) = nq;
            }
        }
    }
#undef HO
    free(odg);
    return X;
}

/* src/iterate/libiterate/diag/vupdate.f90:4-18 */
void orc_vinit(int nout, double *vout)
{
    memset(vout, 0, sizeof(double) * nout * nout);
    for (int i = 0; i < nout; i++) vout[IDX2(i, i, nout)] = 1.0;
}

/* src/iterate/libiterate/diag/vupdate.f90:21-36 */
void orc_vupdate(int p1, int q1, double phi, int nout, double *vout)
{
    double c = cos(phi), s = sin(phi);
    double *vp = vout + (long)nout * (p1 - 1), *vq = vout + (long)nout * (q1 - 1);
    for (int i = 0; i < nout; i++) {
        double t = c * vp[i] - s * vq[i];
        vq[i] = c * vq[i] + s * vp[i];
        vp[i] = t;
    }
}

/* src/iterate/libiterate/diag/hosbdiag.f90:5-75  (stats5 = the reference's eff(1:5)) */
void orc_hosb_diag(int bjmax, double bjtol, int blk, int mak, int ndk, double *hosb, double *vosb, double *eigval,
                   double *stats5)
{
    double eff[5] = {0, 0, 0, 0, 0};
    long bs = (long)ndk * mak * mak;
    for (int it = 0; it < blk; it++) {
        double *h = hosb + bs * it;
        double *v = vosb + (long)mak * mak * it;
        double initye = orc_sum_off_diag(ndk, mak, h);
        double ye = initye, dye = initye, eff0;
        int k = 1;
        orc_vinit(mak, v);
        while (fabs(dye) > initye * bjtol && k <= bjmax) {
            for (int i = 1; i <= mak - 1; i++)
                for (int j = 1; j <= mak - i; j++) {
                    double phi = orc_mjacobi(j, j + i, mak, ndk, h);
                    orc_vupdate(j, j + i, phi, mak, v);
                }
            dye = orc_sum_off_diag(ndk, mak, h) - ye;
            ye = ye + dye;
            k = k + 1;
        }
        eff0 = fabs(1.0 - ye / initye);
        if (it == 0) { eff[0] = eff0; eff[1] = eff0; }
        else { if (eff0 < eff[0]) eff[0] = eff0; if (eff0 > eff[1]) eff[1] = eff0; }
        eff[2] += eff0;
        eff[4] += k;
        for (int i = 0; i < mak; i++) {
            double *hd = h + (long)ndk * (i + (long)mak * i);
            double *e = eigval + (long)ndk * (i + (long)mak * it);
            for (int d = 0; d < ndk; d++) e[d] = hd[d];
        }
        for (int i = 0; i < mak; i++) {
            double *hd = h + (long)ndk * (i + (long)mak * i);
            for (int d = 0; d < ndk; d++) hd[d] = 0.0;
        }
    }
    if (stats5) memcpy(stats5, eff, sizeof(eff));
}

/* src/iterate/module/osb.diag.h:53-151 with myDiagInit (osb.diag.h:19-32) and
 * EIG0 = RES before the first level (osb.io.h:79).  VOSB[l] and EIG0 must be
 * allocated by the caller. */
int orc_my_diag(orc_osb *o, int bjmax, double bjtol)
{
    long n = o->myLen;
    double *hosb;
    memcpy(o->EIG0, o->RES, sizeof(double) * n);
    for (int l = o->sF - 1; l >= 0; l--) {
        hosb = (double *)malloc(sizeof(double) * (long)o->sN[l] * n);
        if (!hosb) return 0;
        if (o->ndvr && l == o->sF - 1)   /* myDiagInit: HOSB(sF) = OUTH (osb.diag.h:22-23) */
            memcpy(hosb, o->OUTH, sizeof(double) * (long)o->sN[l] * n);
        else if (o->sDEP[l])
            orc_hosbinit_dep(o->myBlk[l], o->sN[l], o->myDim[l], o->H0[l], o->DEP[l], o->EIG0, hosb);
        else
            orc_hosbinit_xyz(o->myBlk[l], o->sN[l], o->myDim[l], o->H0[l], o->EIG0, hosb);
        orc_hosb_diag(bjmax, bjtol, o->myBlk[l], o->sN[l], o->myDim[l], hosb, o->VOSB[l], o->EIG0, NULL);
        free(hosb);
    }
    return 1;
}

/* ---- solver callbacks on an orc_osb (the reference's OSB_QMR glue,
 *      src/iterate/module/osbw.qmr.h:13-23, with sHC = sPC = TA1) ---- */
void orc_cb_hx_e(void *ctx, long n, const double *x, double *y)
{
    orc_solver_ctx *c = (orc_solver_ctx *)ctx;
    (void)n;
    c->hx_calls++;
    if (c->threaded) orc_myhx_omp(c->o, ORC_TA1, x, y, c->tmp);
    else orc_myhx(c->o, ORC_TA1, x, y);
}

void orc_cb_hx0(void *ctx, long n, const double *x, double *y)
{
    orc_solver_ctx *c = (orc_solver_ctx *)ctx;
    (void)n;
    c->hx_calls++;
    if (c->threaded) orc_myhx_omp(c->o, ORC_TA0, x, y, c->tmp);
    else orc_myhx(c->o, ORC_TA0, x, y);
}

void orc_cb_px(void *ctx, long n, const double *x, double *y)
{
    orc_solver_ctx *c = (orc_solver_ctx *)ctx;
    (void)n;
    c->px_calls++;
    if (c->threaded) orc_mypx_omp(c->o, c->o->osb_kind, ORC_TA1, x, y);
    else orc_mypx(c->o, c->o->osb_kind, ORC_TA1, x, y);
}

int orc_cb_linsolv(void *ctx, long n, const double *b, double *x, double *res)
{
    orc_solver_ctx *c = (orc_solver_ctx *)ctx;
    int r = orc_qmr(c->qmr_type, c->qmr_max, c->qmr_tol, n, b, orc_cb_hx_e, orc_cb_px, ctx, x, res);
    c->solves++;
    if (r > 0) c->qmr_iters += r;
    return r;
}

/* Fortran dot_product: a left-to-right sum */
double orc_dot(long n, const double *a, const double *b)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* src/iterate/libiterate/qmr/qmr.f90:5-175 */
int orc_qmr(int ntype, int qmrmax, double qmrtol, long n, const double *b,
            orc_op_fn hx, orc_op_fn px, void *ctx, double *x, double *res)
{
    double *R = (double *)malloc(sizeof(double) * n * 8);
    if (!R) return -1;
    double *UX = R + n, *UR = R + 2 * n, *Q = R + 3 * n, *S = R + 4 * n, *V = R + 5 * n, *Z = R + 6 * n,
           *HX0 = R + 7 * n;
    double BNRM, ZNRM, VNRM, VNRMPO = 0, ZVRAT, QSDOT = 0, VZDOT, BETA, ETA, GAMAMO, GAMA, THETAMO = 0, THETA;
    double RES1, RES2, RES3, XNRM, r = 0.0;
    int ret = -1, i;
    int at = abs(ntype);

    ZVRAT = 1.0;
    if (ntype <= 0) memset(x, 0, sizeof(double) * n);
    memcpy(R, b, sizeof(double) * n);
    memcpy(V, b, sizeof(double) * n);
    BNRM = sqrt(orc_dot(n, b, b));
    VNRM = BNRM;
    px(ctx, n, b, Z);
    ZNRM = sqrt(orc_dot(n, Z, Z));
    GAMAMO = 1.0; ETA = -1.0;

    ret = 0;
    for (i = 1; i <= qmrmax; i++) {
        if (VNRM == 0.0 || ZNRM == 0.0) break;
        ZVRAT = ZVRAT * ZNRM / VNRM;
        for (long k = 0; k < n; k++) { V[k] = V[k] / VNRM; Z[k] = Z[k] / ZNRM; }
        VZDOT = orc_dot(n, V, Z);
        if (VZDOT == 0.0) break;
        if (i == 1) memcpy(Q, Z, sizeof(double) * n);
        else {
            double c = VNRM * VZDOT / QSDOT;
            for (long k = 0; k < n; k++) Q[k] = Z[k] - c * Q[k];
        }
        hx(ctx, n, Q, S);
        for (long k = 0; k < n; k++) S[k] = ZVRAT * S[k];
        QSDOT = orc_dot(n, Q, S);
        BETA = QSDOT / VZDOT;
        if (QSDOT == 0.0 || BETA == 0.0) break;
        for (long k = 0; k < n; k++) V[k] = S[k] - BETA * V[k];
        VNRMPO = sqrt(orc_dot(n, V, V));
        px(ctx, n, V, Z);
        for (long k = 0; k < n; k++) Z[k] = Z[k] / ZVRAT;
        ZNRM = sqrt(orc_dot(n, Z, Z));
        THETA = VNRMPO / (GAMAMO * fabs(BETA));
        GAMA = 1.0 / sqrt(1.0 + THETA * THETA);
        ETA = -ETA * VNRM * GAMA * GAMA / (BETA * GAMAMO * GAMAMO);
        if (GAMA == 0.0) break;
        if (i == 1) {
            for (long k = 0; k < n; k++) { UX[k] = (ETA * ZVRAT) * Q[k]; UR[k] = ETA * S[k]; }
        } else {
            double tg = (THETAMO * GAMA) * (THETAMO * GAMA);
            for (long k = 0; k < n; k++) {
                UX[k] = (ETA * ZVRAT) * Q[k] + tg * UX[k];
                UR[k] = ETA * S[k] + tg * UR[k];
            }
        }
        for (long k = 0; k < n; k++) { x[k] = x[k] + UX[k]; R[k] = R[k] - UR[k]; }
        if (at == 1) {
            hx(ctx, n, x, HX0);
            XNRM = sqrt(orc_dot(n, HX0, HX0));
            for (long k = 0; k < n; k++) HX0[k] = HX0[k] - b[k];
            r = sqrt(orc_dot(n, HX0, HX0)) / XNRM;
        } else if (at == 2) {
            RES1 = sqrt(orc_dot(n, R, R)) / BNRM;
            RES2 = sqrt(orc_dot(n, UX, UX) / orc_dot(n, x, x));
            RES3 = sqrt(orc_dot(n, UR, UR) / orc_dot(n, R, R));
            r = RES1 > RES2 ? RES1 : RES2;
            r = r > RES3 ? r : RES3;
        } else {
            r = sqrt(orc_dot(n, R, R)) / BNRM;
        }
        *res = r;
        if (fabs(r) < qmrtol) { ret = i; break; }
        ret = -i;
        VNRM = VNRMPO; GAMAMO = GAMA; THETAMO = THETA;
    }
    free(R);
    return ret;
}

/* src/iterate/libiterate/orth/gsorth.f90:42-72 (VJ(:,M+1) from WJ, M old columns) */
int orc_mgs_orth(long n, const double *wj, int m, double *vj)
{
    double *v0 = vj + n * m, normw, d;
    memcpy(v0, wj, sizeof(double) * n);
    normw = sqrt(orc_dot(n, wj, wj));
    if (normw == 0.0) return 0;
    for (int i = 0; i < m; i++) {
        d = orc_dot(n, v0, vj + n * i);
        for (long k = 0; k < n; k++) v0[k] = v0[k] - d * vj[n * i + k];
    }
    d = sqrt(orc_dot(n, v0, v0));
    if (d != 0.0) {
        for (long k = 0; k < n; k++) v0[k] = v0[k] / d;
        return 1;
    }
    return 0;
}

/* src/iterate/libiterate/orth/gsorth.f90:122-152 */
int orc_lan_mgs_orth(long n, const double *wj, int m, double *vj, double *beta)
{
    double *v0 = vj + n * m, normw, d;
    memcpy(v0, wj, sizeof(double) * n);
    normw = sqrt(orc_dot(n, wj, wj));
    if (normw == 0.0) return 0;   /* the reference leaves its result undefined here */
    for (int i = 0; i < m; i++) {
        d = orc_dot(n, v0, vj + n * i);
        for (long k = 0; k < n; k++) v0[k] = v0[k] - d * vj[n * i + k];
    }
    d = sqrt(orc_dot(n, v0, v0));
    *beta = d;
    if (d == 0.0) return 0;
    for (long k = 0; k < n; k++) v0[k] = v0[k] / d;
    return 1;
}

/* src/iterate/libiterate/pist/pistHij.f90:328-383.  vj is N x m_high (column j
 * at vj + n*(j-1)), hmat is m_max x m_high.  qmr_iters (optional) accumulates
 * the linear-solver return values. */
int orc_pist_h0(long n, const double *x, int m_max, int m_low, int m_high,
                orc_op_fn h0x, orc_linsolv_fn linsolv, void *ctx, double *vj, double *hmat, int *qmr_iters)
{
    int m_start, num;
    double tmp, res;
    double *wj;
    if (m_low < 1 || m_low >= m_high || m_high > m_max) return 0;
    if (m_low == 1) {
        m_start = 1;
        tmp = sqrt(orc_dot(n, x, x));
        if (tmp == 0.0) return 0;
        for (long k = 0; k < n; k++) vj[k] = x[k] / tmp;
    } else {
        m_start = m_low - 1;
    }
    wj = (double *)malloc(sizeof(double) * n);
    for (int j = m_start; j <= m_high - 1; j++) {
        num = linsolv(ctx, n, vj + n * (j - 1), wj, &res);
        if (qmr_iters && num > 0) *qmr_iters += num;
        if (num <= 0) { free(wj); return num; }
        if (!orc_mgs_orth(n, wj, j, vj)) { free(wj); return -j; }
    }
    for (int j = m_low; j <= m_high; j++) {
        h0x(ctx, n, vj + n * (j - 1), wj);
        for (int i = 1; i <= m_high; i++) {
            double d = orc_dot(n, vj + n * (i - 1), wj);
            hmat[IDX2(i - 1, j - 1, m_max)] = d;
            hmat[IDX2(j - 1, i - 1, m_max)] = d;
        }
    }
    free(wj);
    return m_high;
}

/* src/iterate/libiterate/pist/pist.f90:7-98 */
int orc_pist_conv(double e0, double etol, int ntype, long n, const double *x, int startcnt, int stepcnt,
                  int stepeig, int m, int m_max, orc_op_fn h0x, orc_linsolv_fn linsolv, void *ctx,
                  double *eig, double *res, int *qmr_iters)
{
    int cnt_low, cnt_high, num, ret = 0;
    double *vj, *hmat_old, *hmat, *tmpeig, *oldeig;
    int *ind;
    if (stepeig >= m_max || startcnt > m_max || m > m_max) return 0;
    vj = (double *)malloc(sizeof(double) * n * m_max);
    hmat_old = (double *)calloc((size_t)m_max * m_max, sizeof(double));
    hmat = (double *)malloc(sizeof(double) * m_max * m_max);
    tmpeig = (double *)malloc(sizeof(double) * m_max);
    oldeig = (double *)malloc(sizeof(double) * m);
    ind = (int *)malloc(sizeof(int) * m);
    cnt_low = 1;
    cnt_high = startcnt;
    if (m > cnt_high) cnt_high = m;
    if (stepeig + 1 > cnt_high) cnt_high = stepeig + 1;
    while (cnt_high <= m_max) {
        ret = orc_pist_h0(n, x, m_max, cnt_low, cnt_high, h0x, linsolv, ctx, vj, hmat_old, qmr_iters);
        if (ret < 0) { ret = -cnt_high; goto done; }
        num = cnt_high - stepeig;
        if (orc_sym_eigvals(num, hmat_old, m_max, tmpeig) != 0) { ret = -num; goto done; }
        orc_get_window0(e0, num, tmpeig, m, ind, oldeig);
        num = cnt_high;
        if (orc_sym_eigvals(num, hmat_old, m_max, tmpeig) != 0) { ret = -num; goto done; }
        orc_get_window0(e0, num, tmpeig, m, ind, eig);
        if (ntype <= 0) *res = orc_get_diff_max(m, eig, oldeig);
        else *res = orc_get_sep_max(m, oldeig, cnt_high, tmpeig);
        if (*res <= etol) { ret = cnt_high; break; }
        ret = -cnt_high;
        cnt_low = cnt_high + 1;
        if (cnt_low >= m_max) { ret = -m_max; break; }
        cnt_high = cnt_high + stepcnt;
        if (cnt_high > m_max) cnt_high = m_max;
    }
    orc_reorder_asc(m, eig);
done:
    free(vj); free(hmat_old); free(hmat); free(tmpeig); free(oldeig); free(ind);
    return ret;
}

/* src/iterate/libiterate/lanczos/lanHij.f90:6-45.  vj is N x (m_high+1). */
int orc_lan_h0(long n, const double *x, double *vj, int m_low, int m_high, orc_op_fn hx, void *ctx,
               double *alpha, double *beta)
{
    double tmp, *wj;
    if (m_low < 1 || m_low >= m_high) return 0;
    if (m_low == 1) {
        tmp = sqrt(orc_dot(n, x, x));
        if (tmp == 0.0) return 0;
        for (long k = 0; k < n; k++) vj[k] = x[k] / tmp;
    }
    wj = (double *)malloc(sizeof(double) * n);
    for (int j = m_low; j <= m_high; j++) {
        const double *vjj = vj + n * (j - 1);
        hx(ctx, n, vjj, wj);
        if (j != 1) {
            const double *vjm = vj + n * (j - 2);
            for (long k = 0; k < n; k++) wj[k] = wj[k] - beta[j - 2] * vjm[k];
        }
        alpha[j - 1] = orc_dot(n, wj, vjj);
        for (long k = 0; k < n; k++) wj[k] = wj[k] - alpha[j - 1] * vjj[k];
        if (!orc_lan_mgs_orth(n, wj, j, vj, &beta[j - 1])) { free(wj); return -j; }
    }
    free(wj);
    return m_high;
}

/* src/iterate/libiterate/lanczos/lan.f90:7-105 */
int orc_lan_conv(double e0, double etol, int ntype, long n, const double *x, int startcnt, int stepcnt,
                 int stepeig, int m, int m_max, orc_op_fn hx, void *ctx, double *eig, double *res)
{
    int cnt_low, cnt_high, num, ret = 0;
    double *vj, *alpha, *beta, *w, *oldeig;
    int *ind;
    if (stepeig >= m_max || startcnt > m_max || m > m_max) return 0;
    vj = (double *)malloc(sizeof(double) * n * (m_max + 1));
    alpha = (double *)calloc(m_max, sizeof(double));
    beta = (double *)calloc(m_max, sizeof(double));
    w = (double *)malloc(sizeof(double) * m_max);
    oldeig = (double *)malloc(sizeof(double) * m);
    ind = (int *)malloc(sizeof(int) * m);
    cnt_low = 1;
    cnt_high = startcnt;
    if (m > cnt_high) cnt_high = m;
    if (stepeig + 1 > cnt_high) cnt_high = stepeig + 1;
    while (cnt_high <= m_max) {
        num = orc_lan_h0(n, x, vj, cnt_low, cnt_high, hx, ctx, alpha, beta);
        if (num <= 0) { ret = -(cnt_low - 1); goto done; }
        num = cnt_high - stepeig;
        if (orc_tridiag_eigvals(num, alpha, beta, w) != 0) { ret = -num; goto done; }
        orc_get_window0(e0, num, w, m, ind, oldeig);
        num = cnt_high;
        if (orc_tridiag_eigvals(num, alpha, beta, w) != 0) { ret = -num; goto done; }
        orc_get_window0(e0, num, w, m, ind, eig);
        if (ntype <= 0) *res = orc_get_diff_max(m, eig, oldeig);
        else *res = orc_get_sep_max(m, oldeig, cnt_high, w);
        if (*res < etol) { ret = cnt_high; break; }
        ret = -cnt_high;
        cnt_low = cnt_high + 1;
        if (cnt_low >= m_max) { ret = -m_max; break; }
        cnt_high = cnt_high + stepcnt;
        if (cnt_high > m_max) cnt_high = m_max;
    }
    orc_reorder_asc(m, eig);
done:
    free(vj); free(alpha); free(beta); free(w); free(oldeig); free(ind);
    return ret;
}

/* Cyclic Jacobi eigensolver for a symmetric matrix (upper triangle of a is
 * read, as DSYEV('U') does).  Rutishauser's rotation formulas; stops when the
 * off-diagonal norm is below 1e-17 of the Frobenius norm. */
static int sym_jacobi(int n, double *a, double *v, double *w)
{
    if (v) {
        memset(v, 0, sizeof(double) * n * n);
        for (int i = 0; i < n; i++) v[i + (long)n * i] = 1.0;
    }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                double t = a[i + (long)n * j] * a[i + (long)n * j];
                tot += t;
                if (i != j) off += t;
            }
        if (off <= 1e-34 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = a[p + (long)n * q];
                if (apq == 0.0) continue;
                double app = a[p + (long)n * p], aqq = a[q + (long)n * q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[k + (long)n * p], akq = a[k + (long)n * q];
                    a[k + (long)n * p] = c * akp - s * akq;
                    a[k + (long)n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[p + (long)n * k], aqk = a[q + (long)n * k];
                    a[p + (long)n * k] = c * apk - s * aqk;
                    a[q + (long)n * k] = s * apk + c * aqk;
                }
                if (v)
                    for (int k = 0; k < n; k++) {
                        double vkp = v[k + (long)n * p], vkq = v[k + (long)n * q];
                        v[k + (long)n * p] = c * vkp - s * vkq;
                        v[k + (long)n * q] = s * vkp + c * vkq;
                    }
            }
    }
    for (int i = 0; i < n; i++) w[i] = a[i + (long)n * i];
    return 0;
}

static void sort_pairs(int n, double *w, double *z, int ldz)
{
    for (int i = 1; i < n; i++) {
        for (int j = i; j > 0 && w[j - 1] > w[j]; j--) {
            double t = w[j]; w[j] = w[j - 1]; w[j - 1] = t;
            if (z)
                for (int k = 0; k < n; k++) {
                    double u = z[k + (long)ldz * j];
                    z[k + (long)ldz * j] = z[k + (long)ldz * (j - 1)];
                    z[k + (long)ldz * (j - 1)] = u;
                }
        }
    }
}

int orc_sym_eigvals(int n, const double *a, int lda, double *w)
{
    double *b = (double *)malloc(sizeof(double) * n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j; i++) {
            b[i + (long)n * j] = a[i + (long)lda * j];
            b[j + (long)n * i] = a[i + (long)lda * j];
        }
    sym_jacobi(n, b, NULL, w);
    sort_pairs(n, w, NULL, 0);
    free(b);
    return 0;
}

int orc_tridiag_eigvals(int n, const double *d, const double *e, double *w)
{
    double *b = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++) b[i + (long)n * i] = d[i];
    for (int i = 0; i + 1 < n; i++) { b[i + (long)n * (i + 1)] = e[i]; b[i + 1 + (long)n * i] = e[i]; }
    sym_jacobi(n, b, NULL, w);
    sort_pairs(n, w, NULL, 0);
    free(b);
    return 0;
}

/* src/libutil/serial/base/util.f90:669-717 (first index wins ties) */
void orc_get_window0(double e0, int n, const double *vec, int m, int *vecind, double *newvec)
{
    if (n < m) {
        for (int i = 0; i < n; i++) { newvec[i] = vec[i]; vecind[i] = i + 1; }
        for (int i = n - 1; i < m; i++) { newvec[i] = vec[n - 1]; vecind[i] = n; }
        return;
    }
    char *label = (char *)malloc(n);
    double *absv = (double *)malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) { label[j] = 1; absv[j] = fabs(vec[j] - e0); }
    for (int i = 0; i < m; i++) {
        int index = -1;
        double val0 = 0.0;
        for (int j = 0; j < n; j++) {
            if (!label[j]) continue;
            if (index < 0) { index = j; val0 = absv[j]; }
            else if (val0 > absv[j]) { val0 = absv[j]; index = j; }
        }
        vecind[i] = index + 1;
        newvec[i] = vec[index];
        label[index] = 0;
    }
    free(label); free(absv);
}

/* src/libutil/serial/base/util.f90:84-100 */
double orc_get_diff_max(int n, const double *v1, const double *v2)
{
    double m = fabs(v1[0] - v2[0]);
    for (int i = 1; i < n; i++) { double v = fabs(v1[i] - v2[i]); if (m < v) m = v; }
    return m;
}

/* src/libutil/serial/base/util.f90:124-149 */
double orc_get_sep_max(int n1, const double *v1, int n2, const double *v2)
{
    double g = 0.0;
    for (int i = 0; i < n1; i++) {
        double mm = fabs(v1[i] - v2[0]);
        for (int j = 1; j < n2; j++) { double v = fabs(v1[i] - v2[j]); if (mm > v) mm = v; }
        if (i == 0 || g < mm) g = mm;
    }
    return g;
}

/* src/libutil/serial/base/util.f90:415-428 Reorder('A') */
void orc_reorder_asc(int n, double *vec)
{
    sort_pairs(n, vec, NULL, 0);
}

/* src/ham/libham/dvr/sinc.f90:67-109 H1_SINC2 in hartree (isCMU=.false.) */
void orc_h1_sinc2(int n, double a0, double b0, double mass, double *h1)
{
    const double PI = 3.14159265358979323846;
    const double PI3 = PI * PI / 3.0;
    double dx = (b0 - a0) / n;
    int i0 = (int)(a0 / dx);
    double scale = 0.5 / (dx * dx * mass);
    for (int i = 1; i <= n; i++) {
        for (int j = 1; j <= i - 1; j++) {
            int ij1 = i - j, ij2 = i + j + i0 + i0;
            double h = 2.0 / (double)(ij1 * ij1) - 2.0 / (double)(ij2 * ij2);
            if ((ij1 / 2) * 2 != ij1) h = -h;
            h1[IDX2(i - 1, j - 1, n)] = h;
            h1[IDX2(j - 1, i - 1, n)] = h;
        }
    }
    for (int i = 1; i <= n; i++) h1[IDX2(i - 1, i - 1, n)] = PI3 - 0.5 / (double)((i + i0) * (i + i0));
    for (long k = 0; k < (long)n * n; k++) h1[k] = scale * h1[k];
}

/* src/ham/libham/dvr/sinc_xh.f90:50-67 Sinc2_DVR */
void orc_sinc2_dvr(int n, double a0, double b0, double mass, double *x0, double *h0)
{
    double xmin = a0 < b0 ? a0 : b0, xmax = a0 < b0 ? b0 : a0;
    double dx = (xmax - xmin) / n;
    int i0 = (int)(xmin / dx);
    for (int i = 1; i <= n; i++) x0[i - 1] = dx * (i0 + i);
    orc_h1_sinc2(n, xmin, xmax, mass, h0);
}

/* src/ham/libham/poly/gauLege.f90:63-103 Gau_legendre (Newton, EPSI=1e-14) */
void orc_gau_legendre(double x1, double x2, int n, double *x, double *w)
{
    const double EPSI = 1.0e-14, PIE = 3.14159265358979323846264338328;
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; i++) {
        double z = cos(PIE * (i - .25) / (n + .5)), z1, p1, p2, p3, pp;
        for (;;) {
            p1 = 1.0; p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z; z = z1 - p1 / pp;
            if (fabs(z - z1) <= EPSI) break;
        }
        x[i - 1] = xm - xl * z; x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* src/ham/libham/poly/yj.f90:16-40 YjPolys (normalised Legendre, with the
 * reference's 14-digit constants for sqrt(1/2) and sqrt(3/2)) */
void orc_yj_polys(int n0, int l, const double *x, double *yj)
{
    const double SQRT_1_2 = 0.70710678118655, SQRT_3_2 = 1.22474487139159;
    for (int k = 0; k < n0; k++) yj[k] = SQRT_1_2;
    if (l >= 1)
        for (int k = 0; k < n0; k++) yj[k + n0] = SQRT_3_2 * x[k];
    for (int i = 3; i <= l + 1; i++) {
        double re0 = 1.0 / (i - 1.0);
        double re1 = sqrt((-1.0 + i + i) * (-3.0 + i + i)) * re0;
        double re2 = sqrt((-1.0 + i + i) / (-5.0 + i + i)) * (1.0 - re0);
        for (int k = 0; k < n0; k++)
            yj[k + (long)n0 * (i - 1)] = re1 * x[k] * yj[k + (long)n0 * (i - 2)] - re2 * yj[k + (long)n0 * (i - 3)];
    }
}

/* Legendre DVR of L^2: T(i,j) = sqrt(w_i) P~_j(x_i) (j = 0..n-1) is the
 * DVR<-FBR transform; L2 = T diag(j(j+1)) T^T.  Nodes ascending in cos(theta). */
void orc_legendre_l2(int n, double *x, double *l2)
{
    double *w = (double *)malloc(sizeof(double) * n);
    double *yj = (double *)malloc(sizeof(double) * n * n);
    orc_gau_legendre(-1.0, 1.0, n, x, w);
    orc_yj_polys(n, n - 1, x, yj);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            double s = 0.0;
            for (int j = 0; j < n; j++)
                s += sqrt(w[i]) * yj[i + (long)n * j] * (double)j * (j + 1.0) * sqrt(w[k]) * yj[k + (long)n * j];
            l2[IDX2(i, k, n)] = s;
        }
    free(w); free(yj);
}

/* ---- threaded H*X: same per-block loop nests, blocks split across threads ---- */
int orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_myhx_omp(const orc_osb *o, int ntype, const double *x, double *y, double *tmp)
{
    long n = o->myLen;
    double c = (ntype == ORC_TA0) ? 0.0 : o->E0;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; k++) y[k] = (o->RES[k] - c) * x[k];
    for (int l = 0; l < o->sF; l++) {
        int nin = o->myDim[l], nout = o->sN[l], blk = o->myBlk[l];
        long bs = (long)nin * nout;
        if (o->ndvr && l == o->sF - 1) {
            /* the outer dense layer: H3X with the inner points split over the threads */
#pragma omp parallel for schedule(static)
            for (int kc = 0; kc < nin; kc += 64) {
                int ke = kc + 64 < nin ? kc + 64 : nin;
                for (int i = 0; i < nout; i++) {
                    double *xo = tmp + (long)nin * i;
                    for (int k = kc; k < ke; k++) xo[k] = o->OUTH[(long)nin * i + k] * x[k];
                    for (int j = 1; j < nout; j++) {
                        const double *aij = o->OUTH + (long)nin * (i + (long)nout * j);
                        const double *xj = x + (long)nin * j;
                        for (int k = kc; k < ke; k++) xo[k] = xo[k] + aij[k] * xj[k];
                    }
                }
            }
        } else if (blk >= 64 || nin < 64) {
#pragma omp parallel for schedule(static)
            for (int b = 0; b < blk; b++) {
                if (o->sDEP[l]) orc_h2x(nin, nout, o->H0[l], o->DEP[l], x + bs * b, tmp + bs * b);
                else orc_h1x(nin, nout, o->H0[l], x + bs * b, tmp + bs * b);
            }
        } else {
            /* few large blocks (outer levels): split the inner NIN range instead */
            for (int b = 0; b < blk; b++) {
                const double *xb = x + bs * b;
                double *tb = tmp + bs * b;
                const double *h1 = o->H0[l];
#pragma omp parallel for schedule(static)
                for (int kc = 0; kc < nin; kc += 64) {
                    int ke = kc + 64 < nin ? kc + 64 : nin;
                    for (int i = 0; i < nout; i++) {
                        double *xo = tb + (long)nin * i;
                        double h = h1[IDX2(i, 0, nout)];
                        if (o->sDEP[l])
                            for (int k = kc; k < ke; k++) xo[k] = h * o->DEP[l][k] * xb[k];
                        else
                            for (int k = kc; k < ke; k++) xo[k] = h * xb[k];
                        for (int j = 1; j < nout; j++) {
                            const double *xj = xb + (long)nin * j;
                            h = h1[IDX2(i, j, nout)];
                            if (o->sDEP[l])
                                for (int k = kc; k < ke; k++) xo[k] = xo[k] + h * o->DEP[l][k] * xj[k];
                            else
                                for (int k = kc; k < ke; k++) xo[k] = xo[k] + h * xj[k];
                        }
                    }
                }
            }
        }
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; k++) y[k] = y[k] + tmp[k];
    }
}

/* ---- threaded P*X: MYPX's per-block rotations with the blocks (or, for
 *      the few large blocks of outer levels, chunks of 64 inner points)
 *      split over the host threads, as MPI ranks split them; every value is
 *      summed in orc_h1x / orc_h1tx order, so the result is orc_mypx's ---- */
static void vx_omp(int nin, int nout, int blk, const double *v, double *x, int trans)
{
    const long nch = (nin + 63) / 64, bs = (long)nin * nout;
#pragma omp parallel
    {
        double *t = (double *)malloc(sizeof(double) * 64 * nout);
#pragma omp for schedule(static)
        for (long w = 0; w < (long)blk * nch; w++) {
            const long b = w / nch;
            const int p0 = (int)(w % nch) * 64, p1 = p0 + 64 < nin ? p0 + 64 : nin, np = p1 - p0;
            const double *vb = v + (long)nout * nout * b;
            double *xb = x + bs * b;
            for (int i = 0; i < nout; i++) {
                double *to = t + (long)64 * i;
                double h = trans ? vb[IDX2(0, i, nout)] : vb[IDX2(i, 0, nout)];
                for (int k = 0; k < np; k++) to[k] = h * xb[p0 + k];
                for (int j = 1; j < nout; j++) {
                    h = trans ? vb[IDX2(j, i, nout)] : vb[IDX2(i, j, nout)];
                    const double *xj = xb + (long)nin * j + p0;
                    for (int k = 0; k < np; k++) to[k] = to[k] + h * xj[k];
                }
            }
            for (int i = 0; i < nout; i++)
                for (int k = 0; k < np; k++) xb[(long)nin * i + p0 + k] = t[(long)64 * i + k];
        }
        free(t);
    }
}

void orc_mypx_omp(const orc_osb *o, int nosb, int npc, const double *x, double *y)
{
    long n = o->myLen;
    double de = fabs(o->DE), rde = 1.0 / de, e0 = (npc == ORC_TA0) ? 0.0 : o->E0;
    memcpy(y, x, sizeof(double) * n);
    for (int l = o->sF - 1; l >= 0; l--) vx_omp(o->myDim[l], o->sN[l], o->myBlk[l], o->VOSB[l], y, 1);
    const double beta = 1.0 + o->beta, alpha = -o->beta * de;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        const double ade = o->EIG0[i] - e0;
        if (nosb == ORC_TOSBD1) {
            if (fabs(ade) <= de) y[i] = (ade < 0.0) ? -y[i] * rde : y[i] * rde;
            else y[i] = y[i] / ade;
        } else if (nosb == ORC_TOSBD2) {
            double t;
            if (fabs(ade) <= de) t = (ade < 0.0) ? (-alpha + beta * ade) : (alpha + beta * ade);
            else t = ade;
            y[i] = y[i] / t;
        } else {
            y[i] = y[i] / ade;
        }
    }
    for (int l = 0; l < o->sF; l++) vx_omp(o->myDim[l], o->sN[l], o->myBlk[l], o->VOSB[l], y, 0);
}
