// Complex (absorbing-potential) jobs: the resonance path of the reference's
// OSB module.
//
//  * MYHX_DX (src/iterate/module/osb.hxdx.h:177-240): Y = (RES - E0 +- i A) X
//    plus the real level products, on complex X.  The real operator applies to
//    the real and imaginary planes as one batch of 2 nvec real vectors (the
//    fused H*X kernels unchanged); one kernel adds the +-i A X term.
//  * MYPX_DX (src/iterate/module/osb.mypxdx.h): kernels_px.hip, px_apply_cx.
//  * QMR_CX (src/iterate/libiterate/qmr/qmrcx.f90:4-179) as OSB_QMRDX
//    (src/iterate/module/osbw.qmrdx.h:16-27) calls it: HX_DX = MYHX_DX(sHC),
//    PX_DX = MYPX_DX(sOSB, sPC).  Vectors stay on the GPU; the recurrence's
//    scalars are formed on the host from deterministic two-stage reductions.
//  * PIST_CONV_SX (src/iterate/libiterate/pist/pist_sx.f90:9-110) and its
//    Hermitian twin PIST_CONV_CX (pist_cx.f90, pistHij_cx.f90, MGS_ORTH_CX
//    orth/gsorth_cx.f90:51-79; OSB_PISTCONVDX2, osbw.pistconv.h:204-244) with
//    pistH0_SX (pistHij_sx.f90:58-121) and MGS_ORTH_SX (orth/gsorth_sx.f90:
//    47-74), as OSB_PISTCONVDX1 (src/iterate/module/osbw.pistconv.h:98-140)
//    drives it: H0X = MYHX_DX(TMAP0), LinSolv = OSB_QMRDX with sHC = TMAP and
//    sPC = TMAP (sAP) or TA1.  The Gram-Schmidt projections chain on the
//    device (no host round trip per projection); the Krylov matrix's complex
//    eigenvalues (ZGEEV in the reference) come from a Hessenberg QR on the
//    host, over at most M_MAX values as the reference's root rank does.
//
// Complex vectors use the split layout: vector v's real part at x + 2 v N and
// its imaginary part at x + (2 v + 1) N.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "scalit_impl.hpp"

namespace scalit {

using cd = std::complex<double>;

constexpr int CX_G = 512;   // blocks of a reduction's first stage

struct CxState {
    DBuf<double> vec;    // QMR vectors: R, UX, UR, Q, S, V, Z, HX0, X, B (2 N each)
    DBuf<double> part;   // first-stage partial sums, CX_G x 2 per slot
    DBuf<double> scal;   // reduced values, 2 per slot
    std::vector<double> hscal;
    DBuf<double> qs;     // QMR_CX's scalars on the device
    // one QMR_CX iteration as a HIP graph, for the operator generation, the
    // shift E0 / DE / beta (kernel arguments) and the solver options it was
    // recorded with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long g_gen = -1;
    int g_key[4] = {0, 0, 0, 0};
    double g_dkey[4] = {0.0, 0.0, 0.0, 0.0};
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~CxState() { drop_graph(); }
    DBuf<double> vj;     // PIST_CONV_SX's Krylov basis
    DBuf<double> wj;     // its solve and H0X vectors
    int slots = 0;
};

static CxState &cxstate(scalit_ctx *c)
{
    if (!c->cx) c->cx = new CxState();
    return *c->cx;
}

void release_complex(scalit_ctx *c)
{
    delete c->cx;
    c->cx = nullptr;
}

const double *cx_imag_diag(const scalit_ctx *c, int ntype, double *sign, int *rc)
{
    int which = -1;
    double sg = 0.0;
    switch (ntype) {
    case SCALIT_TAAP0:
    case SCALIT_TAAP: which = SCALIT_AP; sg = 1.0; break;
    case SCALIT_TAAPP: which = SCALIT_APP; sg = 1.0; break;
    case SCALIT_TAAPR: which = SCALIT_APR; sg = 1.0; break;
    case SCALIT_TMAP0:
    case SCALIT_TMAP: which = SCALIT_AP; sg = -1.0; break;
    case SCALIT_TMAPP: which = SCALIT_APP; sg = -1.0; break;
    case SCALIT_TMAPR: which = SCALIT_APR; sg = -1.0; break;
    case SCALIT_TA0:
    case SCALIT_TA1: break;
    default: *rc = SCALIT_EINVAL; return nullptr;
    }
    *sign = sg;
    if (which < 0) return nullptr;
    if (!c->has_ap[which]) {
        *rc = SCALIT_ESTATE;
        return nullptr;
    }
    return c->ap[which].p;
}

// y += i sign a x on split complex vectors
__global__ void k_cx_imag(long N, int nvec, const double *__restrict__ a, double sign, const double *__restrict__ x,
                          double *__restrict__ y)
{
    const int v = blockIdx.y;
    const double *xr = x + 2L * v * N, *xi = xr + N;
    double *yr = y + 2L * v * N, *yi = yr + N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const double t = sign * a[e];
        const double r = xr[e], i = xi[e];
        yr[e] -= t * i;
        yi[e] += t * r;
    }
}

// first stage of sum_e a_e b_e (conj = 0, the bilinear DOT_CX) or
// sum_e conj(a_e) b_e (conj = 1): one (re, im) partial per block
__global__ void __launch_bounds__(256) k_cx_dot(long N, const double *__restrict__ a, const double *__restrict__ b,
                                                int conj, double *__restrict__ part)
{
    __shared__ double sr[256], si[256];
    const double *ar = a, *ai = a + N, *br = b, *bi = b + N;
    double r = 0.0, i = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double xr = ar[e], xi = conj ? -ai[e] : ai[e], yr = br[e], yi = bi[e];
        r += xr * yr - xi * yi;
        i += xr * yi + xi * yr;
    }
    sr[threadIdx.x] = r;
    si[threadIdx.x] = i;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            si[threadIdx.x] += si[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sr[0];
        part[2 * blockIdx.x + 1] = si[0];
    }
}

// second stage in a fixed order; op 1 takes the principal complex square root,
// op 2 the square root of the real part (DSQRT(DBLE(.)) of a Hermitian norm)
__global__ void __launch_bounds__(256) k_cx_fin(int G, const double *__restrict__ part, int op,
                                                double *__restrict__ out)
{
    __shared__ double sr[256], si[256];
    double r = 0.0, i = 0.0;
    for (int b = threadIdx.x; b < G; b += 256) {
        r += part[2 * b];
        i += part[2 * b + 1];
    }
    sr[threadIdx.x] = r;
    si[threadIdx.x] = i;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            si[threadIdx.x] += si[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double a = sr[0], b = si[0];
        if (op == 1) {
            const double m = hypot(a, b);
            const double re = sqrt(0.5 * (m + a)), im = copysign(sqrt(0.5 * (m - a)), b);
            a = re;
            b = im;
        } else if (op == 2) {
            a = sqrt(a);
            b = 0.0;
        }
        out[0] = a;
        out[1] = b;
    }
}

// out = a x + b y (x or y may be null for a zero term; out may alias either)
__global__ void k_cx_axpby(long N, double a_r, double a_i, const double *x, double b_r, double b_i, const double *y,
                           double *out)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        double r = 0.0, i = 0.0;
        if (x) {
            const double xr = x[e], xi = x[e + N];
            r += a_r * xr - a_i * xi;
            i += a_r * xi + a_i * xr;
        }
        if (y) {
            const double yr = y[e], yi = y[e + N];
            r += b_r * yr - b_i * yi;
            i += b_r * yi + b_i * yr;
        }
        out[e] = r;
        out[e + N] = i;
    }
}

// y -= s x, s a reduced device value
__global__ void k_cx_axpy_dev(long N, const double *__restrict__ s, const double *__restrict__ x, double *__restrict__ y)
{
    const double sr = s[0], si = s[1];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const double xr = x[e], xi = x[e + N];
        y[e] -= sr * xr - si * xi;
        y[e + N] -= sr * xi + si * xr;
    }
}

// y /= s, s a reduced device value (left unchanged when s = 0)
__global__ void k_cx_div_dev(long N, const double *__restrict__ s, double *__restrict__ y)
{
    const double sr = s[0], si = s[1];
    const double den = sr * sr + si * si;
    if (den == 0.0) return;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const double yr = y[e], yi = y[e + N];
        y[e] = (yr * sr + yi * si) / den;
        y[e + N] = (yi * sr - yr * si) / den;
    }
}

static unsigned cx_grid(long N) { return (unsigned)std::min<long>((N + 255) / 256, 2048); }
static int cx_blocks(long N) { return (int)std::min<long>((N + 255) / 256, CX_G); }

static int cx_slots(CxState &X, int n)
{
    if (X.slots >= n) return SCALIT_OK;
    X.drop_graph();   // the recorded QMR iteration addresses the partial sums
    SC_TRY(X.part.alloc((size_t)n * CX_G * 2));
    SC_TRY(X.scal.alloc((size_t)n * 2));
    X.hscal.assign((size_t)n * 2, 0.0);
    X.slots = n;
    return SCALIT_OK;
}

// slot <- sum a b (or conj(a) b); op 1: its square root
static void cx_dot(CxState &X, long N, const double *a, const double *b, int conj, int slot, int op, hipStream_t s)
{
    const int G = cx_blocks(N);
    double *part = X.part.p + (size_t)slot * CX_G * 2;
    hipLaunchKernelGGL(k_cx_dot, dim3(G), dim3(256), 0, s, N, a, b, conj, part);
    hipLaunchKernelGGL(k_cx_fin, dim3(1), dim3(256), 0, s, G, (const double *)part, op, X.scal.p + 2 * slot);
}

// the first n slots to the host
static int cx_read(CxState &X, int n, hipStream_t s)
{
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(X.hscal.data(), X.scal.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

static cd slot_val(const CxState &X, int k) { return cd(X.hscal[2 * k], X.hscal[2 * k + 1]); }

static void cx_axpby(long N, cd a, const double *x, cd b, const double *y, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_cx_axpby, dim3(cx_grid(N)), dim3(256), 0, s, N, a.real(), a.imag(), x, b.real(), b.imag(),
                       y, out);
}

// MYHX_DX on nvec split complex vectors (MYHX_CX when sCX is on)
static int hx_apply_cx(scalit_ctx *c, int ntype, int nvec, const double *x, double *y, hipStream_t s)
{
    if (c->scx) return hx_apply_cxh(c, ntype, nvec, x, y, s);   // complex level matrices: MYHX_CX
    if (!real_levels_ready(c)) return SCALIT_ESTATE;
    int rc = SCALIT_OK;
    double sign = 0.0;
    const double *a = cx_imag_diag(c, ntype, &sign, &rc);
    SC_TRY(rc);
    const long N = c->N;
    const double sh = (ntype == SCALIT_TA0 || ntype == SCALIT_TAAP0 || ntype == SCALIT_TMAP0) ? 0.0 : c->e0;
    SC_TRY(hx_apply_f64(c, nullptr, sh, 2 * nvec, x, N, y, N, false, nullptr, s));
    if (a)
        hipLaunchKernelGGL(k_cx_imag, dim3(cx_grid(N), nvec), dim3(256), 0, s, N, nvec, a, sign, x, y);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// ---------------------------------------------------------------- QMR_CX
// The recurrence's scalars live on the device (QS below): each step's vector
// kernel writes one partial sum per block for the reductions it feeds, and a
// one-workgroup step kernel sums them in a fixed order and advances the
// scalars.  A stop flag set by a breakdown, by convergence or by a zero norm
// makes every later kernel of the solve return at once, so the host launches
// iterations in groups and reads the flag once per group.
enum QsIdx {
    QS_ZVRAT, QS_VNRM, QS_ZNRM, QS_VZ, QS_VZI, QS_QS, QS_QSI, QS_BETA, QS_BETAI, QS_ETA, QS_ETAI, QS_GAMAMO,
    QS_THETAMO, QS_THETA, QS_GAMA, QS_VNRMPO, QS_BNRM, QS_RES, QS_STOP, QS_CODE, QS_ITER, QS_CQ, QS_CQI, QS_TG,
    QS_INVV, QS_INVZ, QS_COUNT
};
constexpr int QS_LEN = 32;

__device__ __forceinline__ bool qs_stopped(const double *qs) { return qs[QS_STOP] != 0.0; }

// block partial sums of up to four quantities into part[slot][block][2]
template <int K>
__device__ __forceinline__ void cx_block_sums(double (&v)[K][2], double *part)
{
    __shared__ double sh[K][2][256];
#pragma unroll
    for (int k = 0; k < K; k++) {
        sh[k][0][threadIdx.x] = v[k][0];
        sh[k][1][threadIdx.x] = v[k][1];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < K; k++) {
                sh[k][0][threadIdx.x] += sh[k][0][threadIdx.x + o];
                sh[k][1][threadIdx.x] += sh[k][1][threadIdx.x + o];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) {
            part[(size_t)k * CX_G * 2 + 2 * blockIdx.x] = sh[k][0][0];
            part[(size_t)k * CX_G * 2 + 2 * blockIdx.x + 1] = sh[k][1][0];
        }
}

// V *= 1/VNRM, Z *= 1/ZNRM; partial V^T Z
__global__ void __launch_bounds__(256) k_cxq_vz(long N, const double *__restrict__ qs, double *__restrict__ V,
                                                double *__restrict__ Z, double *__restrict__ part)
{
    if (qs_stopped(qs)) return;
    const double iv = qs[QS_INVV], iz = qs[QS_INVZ];
    double acc[1][2] = {{0.0, 0.0}};
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double vr = V[e] * iv, vi = V[e + N] * iv, zr = Z[e] * iz, zi = Z[e + N] * iz;
        V[e] = vr;
        V[e + N] = vi;
        Z[e] = zr;
        Z[e + N] = zi;
        acc[0][0] += vr * zr - vi * zi;
        acc[0][1] += vr * zi + vi * zr;
    }
    cx_block_sums<1>(acc, part);
}

// Q = Z (first iteration) or Z + CQ Q
__global__ void k_cxq_q(long N, const double *__restrict__ qs, int first, const double *__restrict__ Z,
                        double *__restrict__ Q)
{
    if (qs_stopped(qs)) return;
    const double cr = qs[QS_CQ], ci = qs[QS_CQI];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        double r = Z[e], i = Z[e + N];
        if (!first) {
            const double qr = Q[e], qi = Q[e + N];
            r += cr * qr - ci * qi;
            i += cr * qi + ci * qr;
        }
        Q[e] = r;
        Q[e + N] = i;
    }
}

// S *= ZVRAT; partial Q^T S
__global__ void __launch_bounds__(256) k_cxq_qs(long N, const double *__restrict__ qs, const double *__restrict__ Q,
                                                double *__restrict__ S, double *__restrict__ part)
{
    if (qs_stopped(qs)) return;
    const double zv = qs[QS_ZVRAT];
    double acc[1][2] = {{0.0, 0.0}};
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double sr = S[e] * zv, si = S[e + N] * zv, qr = Q[e], qi = Q[e + N];
        S[e] = sr;
        S[e + N] = si;
        acc[0][0] += qr * sr - qi * si;
        acc[0][1] += qr * si + qi * sr;
    }
    cx_block_sums<1>(acc, part);
}

// V = S - BETA V; partial |V|^2
__global__ void __launch_bounds__(256) k_cxq_v(long N, const double *__restrict__ qs, const double *__restrict__ S,
                                               double *__restrict__ V, double *__restrict__ part)
{
    if (qs_stopped(qs)) return;
    const double br = qs[QS_BETA], bi = qs[QS_BETAI];
    double acc[1][2] = {{0.0, 0.0}};
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double vr0 = V[e], vi0 = V[e + N];
        const double vr = S[e] - (br * vr0 - bi * vi0), vi = S[e + N] - (br * vi0 + bi * vr0);
        V[e] = vr;
        V[e + N] = vi;
        acc[0][0] += vr * vr + vi * vi;
    }
    cx_block_sums<1>(acc, part);
}

// Z /= ZVRAT; partial |Z|^2
__global__ void __launch_bounds__(256) k_cxq_z(long N, const double *__restrict__ qs, double *__restrict__ Z,
                                               double *__restrict__ part)
{
    if (qs_stopped(qs)) return;
    const double iz = 1.0 / qs[QS_ZVRAT];
    double acc[1][2] = {{0.0, 0.0}};
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double zr = Z[e] * iz, zi = Z[e + N] * iz;
        Z[e] = zr;
        Z[e + N] = zi;
        acc[0][0] += zr * zr + zi * zi;
    }
    cx_block_sums<1>(acc, part);
}

// UX = CQ Q + TG UX, UR = ETA S + TG UR (TG = 0 and no read on the first
// iteration), X += UX, R -= UR; partials |R|^2, |UX|^2, |X|^2, |UR|^2
__global__ void __launch_bounds__(256) k_cxq_x(long N, const double *__restrict__ qs, int first,
                                               const double *__restrict__ Q, const double *__restrict__ S,
                                               double *__restrict__ UX, double *__restrict__ UR,
                                               double *__restrict__ X, double *__restrict__ R,
                                               double *__restrict__ part)
{
    if (qs_stopped(qs)) return;
    const double cr = qs[QS_CQ], ci = qs[QS_CQI], er = qs[QS_ETA], ei = qs[QS_ETAI], tg = qs[QS_TG];
    double acc[4][2] = {};
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double qr = Q[e], qi = Q[e + N], sr = S[e], si = S[e + N];
        double uxr = cr * qr - ci * qi, uxi = cr * qi + ci * qr;
        double urr = er * sr - ei * si, uri = er * si + ei * sr;
        if (!first) {
            uxr += tg * UX[e];
            uxi += tg * UX[e + N];
            urr += tg * UR[e];
            uri += tg * UR[e + N];
        }
        UX[e] = uxr;
        UX[e + N] = uxi;
        UR[e] = urr;
        UR[e + N] = uri;
        const double xr = X[e] + uxr, xi = X[e + N] + uxi, rr = R[e] - urr, ri = R[e + N] - uri;
        X[e] = xr;
        X[e + N] = xi;
        R[e] = rr;
        R[e + N] = ri;
        acc[0][0] += rr * rr + ri * ri;
        acc[1][0] += uxr * uxr + uxi * uxi;
        acc[2][0] += xr * xr + xi * xi;
        acc[3][0] += urr * urr + uri * uri;
    }
    cx_block_sums<4>(acc, part);
}

// partials |HX0|^2 and |HX0 - B|^2 (criterion |HX - B| / |HX|)
__global__ void __launch_bounds__(256) k_cxq_hx0(long N, const double *__restrict__ qs, const double *__restrict__ H,
                                                 const double *__restrict__ B, double *__restrict__ part)
{
    if (qs_stopped(qs)) return;
    double acc[2][2] = {};
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long)gridDim.x * 256) {
        const double hr = H[e], hi = H[e + N], dr = hr - B[e], di = hi - B[e + N];
        acc[0][0] += hr * hr + hi * hi;
        acc[1][0] += dr * dr + di * di;
    }
    cx_block_sums<2>(acc, part);
}

enum CxqStep { CXQ_BEGIN, CXQ_AFTER_VZ, CXQ_AFTER_QS, CXQ_THETA, CXQ_END };

// the recurrence's scalar steps (qmrcx.f90:56-175), one workgroup; part holds
// the partial sums of slots 0..(nsum-1), G blocks each
__global__ void __launch_bounds__(256) k_cxq_step(int step, int G, int nsum, const double *__restrict__ part,
                                                  double *__restrict__ qs, int ntype, double tol, int first)
{
    __shared__ double sr[6][256], si[6][256];
    if (qs_stopped(qs)) return;
    for (int k = 0; k < nsum; k++) {
        double r = 0.0, i = 0.0;
        for (int b = threadIdx.x; b < G; b += 256) {
            r += part[(size_t)k * CX_G * 2 + 2 * b];
            i += part[(size_t)k * CX_G * 2 + 2 * b + 1];
        }
        sr[k][threadIdx.x] = r;
        si[k][threadIdx.x] = i;
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < nsum; k++) {
                sr[k][threadIdx.x] += sr[k][threadIdx.x + o];
                si[k][threadIdx.x] += si[k][threadIdx.x + o];
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    using C2 = double2;
    auto cmul = [](C2 a, C2 b) { return C2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; };
    auto cdiv = [](C2 a, C2 b) {
        const double d = b.x * b.x + b.y * b.y;
        return C2{(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
    };
    auto cabs = [](C2 a) { return hypot(a.x, a.y); };
    switch (step) {
    case CXQ_BEGIN: {
        const double vn = qs[QS_VNRM], zn = qs[QS_ZNRM];
        if (vn == 0.0 || zn == 0.0) {
            qs[QS_STOP] = 1.0;
            return;
        }
        qs[QS_ZVRAT] = qs[QS_ZVRAT] * zn / vn;
        qs[QS_INVV] = 1.0 / vn;
        qs[QS_INVZ] = 1.0 / zn;
        qs[QS_ITER] += 1.0;
        break;
    }
    case CXQ_AFTER_VZ: {
        const C2 vz{sr[0][0], si[0][0]};
        qs[QS_VZ] = vz.x;
        qs[QS_VZI] = vz.y;
        if (cabs(vz) == 0.0) {
            qs[QS_STOP] = 1.0;
            return;
        }
        if (!first) {
            const C2 c = cdiv(C2{qs[QS_VNRM] * vz.x, qs[QS_VNRM] * vz.y}, C2{qs[QS_QS], qs[QS_QSI]});
            qs[QS_CQ] = -c.x;
            qs[QS_CQI] = -c.y;
        }
        break;
    }
    case CXQ_AFTER_QS: {
        const C2 q{sr[0][0], si[0][0]};
        qs[QS_QS] = q.x;
        qs[QS_QSI] = q.y;
        const C2 beta = cdiv(q, C2{qs[QS_VZ], qs[QS_VZI]});
        qs[QS_BETA] = beta.x;
        qs[QS_BETAI] = beta.y;
        if (cabs(q) == 0.0 || cabs(beta) == 0.0) qs[QS_STOP] = 1.0;
        break;
    }
    case CXQ_THETA: {
        const double vnrmpo = sqrt(fabs(sr[0][0])), znrm = sqrt(fabs(sr[1][0]));
        const C2 beta{qs[QS_BETA], qs[QS_BETAI]};
        const double gamamo = qs[QS_GAMAMO], vnrm = qs[QS_VNRM];
        const double theta = vnrmpo / (gamamo * cabs(beta));
        const double gama = 1.0 / sqrt(1.0 + theta * theta);
        // ETA = -ETA VNRM GAMA^2 / (BETA GAMAMO^2)
        C2 eta = cdiv(C2{-qs[QS_ETA] * vnrm * gama * gama, -qs[QS_ETAI] * vnrm * gama * gama},
                      C2{beta.x * gamamo * gamamo, beta.y * gamamo * gamamo});
        qs[QS_VNRMPO] = vnrmpo;
        qs[QS_ZNRM] = znrm;
        qs[QS_THETA] = theta;
        qs[QS_GAMA] = gama;
        qs[QS_ETA] = eta.x;
        qs[QS_ETAI] = eta.y;
        if (gama == 0.0) {
            qs[QS_STOP] = 1.0;
            return;
        }
        const C2 cq = cmul(eta, C2{qs[QS_ZVRAT], 0.0});
        qs[QS_CQ] = cq.x;
        qs[QS_CQI] = cq.y;
        const double tg = qs[QS_THETAMO] * gama;
        qs[QS_TG] = first ? 0.0 : tg * tg;
        break;
    }
    case CXQ_END: {
        const int t = ntype < 0 ? -ntype : ntype;
        double res;
        if (t == 1) {
            res = sqrt(fabs(sr[1][0])) / sqrt(fabs(sr[0][0]));
        } else if (t == 2) {
            const double r1 = sqrt(fabs(sr[0][0])) / qs[QS_BNRM];
            const double r2 = sqrt(fabs(sr[1][0] / sr[2][0]));
            const double r3 = sqrt(fabs(sr[3][0] / sr[0][0]));
            res = fmax(r1, fmax(r2, r3));
        } else {
            res = sqrt(fabs(sr[0][0])) / qs[QS_BNRM];
        }
        qs[QS_RES] = res;
        const double it = qs[QS_ITER];
        if (fabs(res) < tol) {
            qs[QS_CODE] = it;
            qs[QS_STOP] = 1.0;
            return;
        }
        qs[QS_CODE] = -it;
        qs[QS_VNRM] = qs[QS_VNRMPO];
        qs[QS_GAMAMO] = qs[QS_GAMA];
        qs[QS_THETAMO] = qs[QS_THETA];
        break;
    }
    }
}

// QMR_CX (qmrcx.f90:4-179).  x (device, split) holds the start vector when
// ntype > 0; *ret gets QMR_CX's return value.
static int qmr_cx_run(scalit_ctx *c, int nosb, int hc, int pc, int ntype, int qmrmax, double tol, const double *b_in,
                      double *x_out, int *ret, double *res_out, hipStream_t s)
{
    CxState &X = cxstate(c);
    const long N = c->N, L = 2 * N;
    {
        double *old = X.vec.p;
        double *old_qs = X.qs.p;
        SC_TRY(X.vec.alloc((size_t)10 * L));
        SC_TRY(cx_slots(X, 4));
        SC_TRY(X.qs.alloc(QS_LEN));
        if (X.vec.p != old || X.qs.p != old_qs) X.drop_graph();
    }
    double *R = X.vec.p, *UX = R + L, *UR = UX + L, *Q = UR + L, *S = Q + L, *V = S + L, *Z = V + L, *HX0 = Z + L;
    // the solve works on its own copies of x and b, so that the recorded
    // iteration serves every right-hand side
    double *x = HX0 + L, *b = x + L;
    double *qs = X.qs.p, *part = X.part.p;
    const int G = cx_blocks(N), t = std::abs(ntype);
    const unsigned g = cx_grid(N);
    auto nrm = [&](const double *v, int slot) { cx_dot(X, N, v, v, 1, slot, 0, s); };
    HIP_TRY(hipMemcpyAsync(b, b_in, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
    if (ntype <= 0) HIP_TRY(hipMemsetAsync(x, 0, sizeof(double) * L, s));
    else HIP_TRY(hipMemcpyAsync(x, x_out, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(R, b, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(V, b, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
    nrm(b, 0);
    SC_TRY(px_apply_cx(c, nosb, pc, 1, b, Z, s));
    nrm(Z, 1);
    SC_TRY(cx_read(X, 2, s));
    // the initial state: ZVRAT = 1, VNRM = BNRM, GAMAMO = 1, ETA = -1, code 0
    std::vector<double> q0(QS_LEN, 0.0);
    q0[QS_ZVRAT] = 1.0;
    q0[QS_BNRM] = std::sqrt(std::abs(slot_val(X, 0).real()));
    q0[QS_VNRM] = q0[QS_BNRM];
    q0[QS_ZNRM] = std::sqrt(std::abs(slot_val(X, 1).real()));
    q0[QS_GAMAMO] = 1.0;
    q0[QS_ETA] = -1.0;
    HIP_TRY(hipMemcpyAsync(qs, q0.data(), sizeof(double) * QS_LEN, hipMemcpyHostToDevice, s));
    auto iteration = [&](int first) -> int {
        hipLaunchKernelGGL(k_cxq_step, dim3(1), dim3(256), 0, s, (int)CXQ_BEGIN, G, 0, (const double *)part, qs,
                           ntype, tol, first);
        hipLaunchKernelGGL(k_cxq_vz, dim3(G), dim3(256), 0, s, N, (const double *)qs, V, Z, part);
        hipLaunchKernelGGL(k_cxq_step, dim3(1), dim3(256), 0, s, (int)CXQ_AFTER_VZ, G, 1, (const double *)part, qs,
                           ntype, tol, first);
        hipLaunchKernelGGL(k_cxq_q, dim3(g), dim3(256), 0, s, N, (const double *)qs, first, (const double *)Z, Q);
        SC_TRY(hx_apply_cx(c, hc, 1, Q, S, s));
        hipLaunchKernelGGL(k_cxq_qs, dim3(G), dim3(256), 0, s, N, (const double *)qs, (const double *)Q, S, part);
        hipLaunchKernelGGL(k_cxq_step, dim3(1), dim3(256), 0, s, (int)CXQ_AFTER_QS, G, 1, (const double *)part, qs,
                           ntype, tol, first);
        hipLaunchKernelGGL(k_cxq_v, dim3(G), dim3(256), 0, s, N, (const double *)qs, (const double *)S, V, part);
        SC_TRY(px_apply_cx(c, nosb, pc, 1, V, Z, s));
        hipLaunchKernelGGL(k_cxq_z, dim3(G), dim3(256), 0, s, N, (const double *)qs, Z, part + (size_t)CX_G * 2);
        hipLaunchKernelGGL(k_cxq_step, dim3(1), dim3(256), 0, s, (int)CXQ_THETA, G, 2, (const double *)part, qs,
                           ntype, tol, first);
        hipLaunchKernelGGL(k_cxq_x, dim3(G), dim3(256), 0, s, N, (const double *)qs, first, (const double *)Q,
                           (const double *)S, UX, UR, x, R, part);
        if (t == 1) {
            SC_TRY(hx_apply_cx(c, hc, 1, x, HX0, s));
            hipLaunchKernelGGL(k_cxq_hx0, dim3(G), dim3(256), 0, s, N, (const double *)qs, (const double *)HX0, b,
                               part);
        }
        hipLaunchKernelGGL(k_cxq_step, dim3(1), dim3(256), 0, s, (int)CXQ_END, G, t == 2 ? 4 : (t == 1 ? 2 : 1),
                           (const double *)part, qs, ntype, tol, first);
        HIP_TRY(hipGetLastError());
        return SCALIT_OK;
    };
    double hq[QS_LEN];
    hq[QS_CODE] = 0.0;
    hq[QS_RES] = 0.0;
    int done = 0;
    bool stop = false;
    while (done < qmrmax && !stop) {
        const int k = std::min(done == 0 ? 1 : 8, qmrmax - done);
        for (int i = 0; i < k; i++) {
            if (done + i == 0) {
                SC_TRY(iteration(1));   // the first (and any allocation or tuning it needs) runs as launched
                continue;
            }
            const int key[4] = {nosb, hc, pc, ntype};
            const double dkey[4] = {tol, c->e0, c->de, c->beta};
            if (!X.exec || X.g_gen != c->gen || !std::equal(key, key + 4, X.g_key) ||
                !std::equal(dkey, dkey + 4, X.g_dkey)) {
                X.drop_graph();
                HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const int rc = iteration(0);
                hipGraph_t gr = nullptr;
                const hipError_t e = hipStreamEndCapture(s, &gr);
                if (rc != SCALIT_OK) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return rc;
                }
                HIP_TRY(e);
                X.graph = gr;
                HIP_TRY(hipGraphInstantiate(&X.exec, gr, nullptr, nullptr, 0));
                X.g_gen = c->gen;
                std::copy(key, key + 4, X.g_key);
                std::copy(dkey, dkey + 4, X.g_dkey);
            }
            HIP_TRY(hipGraphLaunch(X.exec, s));
        }
        done += k;
        HIP_TRY(hipMemcpyAsync(hq, qs, sizeof(double) * QS_LEN, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        stop = hq[QS_STOP] != 0.0;
    }
    HIP_TRY(hipMemcpyAsync(x_out, x, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    *ret = (int)hq[QS_CODE];
    if (res_out) *res_out = hq[QS_RES];
    return SCALIT_OK;
}

// ---------------------------------------------------------------- host eigenvalues
// Eigenvalues of a general complex matrix (ZGEEV with JOBVL = JOBVR = 'N'):
// Householder reduction to upper Hessenberg form, then implicit single-shift
// QR sweeps (Wilkinson shifts, exceptional shifts at 10 and 20 sweeps without
// deflation) on the active block, deflating on negligible subdiagonals.
static int cx_eigvals(int n, std::vector<cd> &A, std::vector<cd> &w)
{
    auto a = [&](int i, int j) -> cd & { return A[i + (size_t)n * j]; };
    w.assign(n, cd(0.0, 0.0));
    for (int k = 0; k + 2 < n; k++) {
        double xn = 0.0;
        for (int i = k + 1; i < n; i++) xn += std::norm(a(i, k));
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const cd x0 = a(k + 1, k);
        const cd ph = std::abs(x0) == 0.0 ? cd(1.0, 0.0) : x0 / std::abs(x0);
        const cd alpha = -ph * xn;
        std::vector<cd> v(n, cd(0.0, 0.0));
        for (int i = k + 1; i < n; i++) v[i] = a(i, k);
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int i = k + 1; i < n; i++) vn += std::norm(v[i]);
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = k + 1; i < n; i++) v[i] /= vn;
        // A <- (I - 2 v v^H) A (I - 2 v v^H)
        for (int j = 0; j < n; j++) {
            cd t(0.0, 0.0);
            for (int i = k + 1; i < n; i++) t += std::conj(v[i]) * a(i, j);
            for (int i = k + 1; i < n; i++) a(i, j) -= 2.0 * v[i] * t;
        }
        for (int i = 0; i < n; i++) {
            cd t(0.0, 0.0);
            for (int j = k + 1; j < n; j++) t += a(i, j) * v[j];
            for (int j = k + 1; j < n; j++) a(i, j) -= 2.0 * t * std::conj(v[j]);
        }
        for (int i = k + 2; i < n; i++) a(i, k) = cd(0.0, 0.0);
    }
    const double eps = 2.220446049250313e-16;
    int ihi = n - 1, iter = 0, total = 0;
    while (ihi >= 0) {
        int l = ihi;
        for (; l > 0; l--) {
            const double h = std::abs(a(l - 1, l - 1)) + std::abs(a(l, l));
            if (std::abs(a(l, l - 1)) <= eps * (h == 0.0 ? 1.0 : h)) {
                a(l, l - 1) = cd(0.0, 0.0);
                break;
            }
        }
        if (l == ihi) {
            w[ihi] = a(ihi, ihi);
            ihi--;
            iter = 0;
            continue;
        }
        if (++total > 60 * n + 60) return SCALIT_ESTATE;
        iter++;
        cd mu;
        if (iter == 10 || iter == 20) {
            mu = a(ihi, ihi) + 0.75 * std::abs(a(ihi, ihi - 1).real());
        } else {
            const cd p = a(ihi - 1, ihi - 1), q = a(ihi - 1, ihi), r = a(ihi, ihi - 1), d = a(ihi, ihi);
            const cd half = 0.5 * (p - d);
            const cd disc = std::sqrt(half * half + q * r);
            const cd m1 = d - q * r / (half + disc), m2 = d - q * r / (half - disc);
            const bool ok1 = std::abs(half + disc) > 0.0, ok2 = std::abs(half - disc) > 0.0;
            if (ok1 && ok2) mu = std::abs(m1 - d) <= std::abs(m2 - d) ? m1 : m2;
            else mu = ok1 ? m1 : (ok2 ? m2 : d);
        }
        // implicit QR sweep over rows/columns l..ihi
        for (int k = l; k < ihi; k++) {
            cd x, y;
            if (k == l) {
                x = a(l, l) - mu;
                y = a(l + 1, l);
            } else {
                x = a(k, k - 1);
                y = a(k + 1, k - 1);
            }
            const double ax = std::abs(x), ay = std::abs(y);
            if (ay == 0.0) continue;
            const double rr = std::hypot(ax, ay);
            double cs;
            cd sn;
            if (ax == 0.0) {
                cs = 0.0;
                sn = std::conj(y) / ay;
            } else {
                cs = ax / rr;
                sn = (x / ax) * std::conj(y) / rr;
            }
            // rows k, k+1 from the left
            for (int j = std::max(l, k - 1); j <= ihi; j++) {
                const cd u = a(k, j), v = a(k + 1, j);
                a(k, j) = cs * u + sn * v;
                a(k + 1, j) = -std::conj(sn) * u + cs * v;
            }
            if (k > l) a(k + 1, k - 1) = cd(0.0, 0.0);
            // columns k, k+1 from the right
            for (int i = l; i <= std::min(k + 2, ihi); i++) {
                const cd u = a(i, k), v = a(i, k + 1);
                a(i, k) = cs * u + std::conj(sn) * v;
                a(i, k + 1) = -sn * u + cs * v;
            }
        }
    }
    return SCALIT_OK;
}

// util_cx.f90 getOpCX: 1 = real part, 6 = |z|^2
static double op_cx(cd z, int t) { return t == 1 ? z.real() : z.real() * z.real() + z.imag() * z.imag(); }

// getWindow0 (util.f90:669-717) on op_cx(., 1): the m values whose real part
// is nearest e0, first index winning ties
static void window_cx(double e0, const std::vector<cd> &vec, int m, std::vector<cd> &out)
{
    const int n = (int)vec.size();
    out.assign(m, cd(0.0, 0.0));
    if (n < m) {
        for (int i = 0; i < n; i++) out[i] = vec[i];
        for (int i = std::max(n - 1, 0); i < m; i++) out[i] = vec[n - 1];
        return;
    }
    std::vector<char> used(n, 0);
    for (int i = 0; i < m; i++) {
        int idx = -1;
        double best = 0.0;
        for (int j = 0; j < n; j++) {
            if (used[j]) continue;
            const double d = std::fabs(op_cx(vec[j], 1) - e0);
            if (idx < 0 || best > d) {
                idx = j;
                best = d;
            }
        }
        out[i] = vec[idx];
        used[idx] = 1;
    }
}

struct PistSx {
    double e0, etol;
    int ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, pc;
    double qmrtol;
    // 0: PIST_CONV_SX (x^T y, complex square-root norms, MGS_ORTH_SX);
    // 1: PIST_CONV_CX (x^H y, real norms, MGS_ORTH_CX, pistHij_cx.f90)
    int herm;
};

// PIST_CONV_SX.  x0: device split start vector; eig: host, 2 m doubles
// (re, im pairs, ascending real part); *ret the reference's return value.
static int pist_sx_run(scalit_ctx *c, const PistSx &p, const double *x0, double *eig, int *ret, double *res_out,
                       int *stats, hipStream_t s)
{
    CxState &X = cxstate(c);
    const long N = c->N, L = 2 * N;
    const int m = p.m, mm = p.m_max;
    *ret = 0;
    if (stats)
        for (int k = 0; k < 4; k++) stats[k] = 0;
    if (p.stepeig >= mm || p.startcnt > mm || m > mm) return SCALIT_OK;   // PIST_CONV_SX = 0
    SC_TRY(X.vj.alloc((size_t)mm * L));
    SC_TRY(X.wj.alloc((size_t)2 * L));
    SC_TRY(cx_slots(X, std::max(mm, 4)));
    double *W = X.wj.p, *HW = W + L;
    std::vector<cd> hmat((size_t)mm * mm, cd(0.0, 0.0));
    auto H = [&](int i, int j) -> cd & { return hmat[(i - 1) + (size_t)mm * (j - 1)]; };
    auto VJ = [&](int j) { return X.vj.p + (size_t)(j - 1) * L; };
    int cnt_low = 1, cnt_high = std::max(std::max(p.startcnt, m), p.stepeig + 1);
    std::vector<cd> tmp, old(m), ev(m);
    double res = 0.0;
    int code = 0;
    auto eigs = [&](int num, std::vector<cd> &w) {
        std::vector<cd> A((size_t)num * num);
        for (int j = 1; j <= num; j++)
            for (int i = 1; i <= num; i++) A[(i - 1) + (size_t)num * (j - 1)] = H(i, j);
        return cx_eigvals(num, A, w);
    };
    while (cnt_high <= mm) {
        // pistH0_SX (pistHij_sx.f90:58-121)
        int m_start;
        const int cj = p.herm, nop = p.herm ? 2 : 1;
        if (cnt_low == 1) {
            m_start = 1;
            cx_dot(X, N, x0, x0, cj, 0, nop, s);
            SC_TRY(cx_read(X, 1, s));
            const cd t = slot_val(X, 0);
            if (t == cd(0.0, 0.0)) {
                code = -cnt_high;
                break;
            }
            cx_axpby(N, 1.0 / t, x0, 0.0, nullptr, VJ(1), s);
        } else {
            m_start = cnt_low - 1;
        }
        bool fail = false;
        for (int j = m_start; j <= cnt_high - 1 && !fail; j++) {
            int num = 0;
            double qres = 0.0;
            SC_TRY(qmr_cx_run(c, c->osb, SCALIT_TMAP, p.pc, p.qmr_type, p.qmrmax, p.qmrtol, VJ(j), W, &num, &qres, s));
            if (stats) {
                stats[0] += std::abs(num);
                stats[1] += 1;
            }
            if (num <= 0) {
                fail = true;
                break;
            }
            // MGS_ORTH_SX: the projections chained on the device
            double *vn = VJ(j + 1);
            HIP_TRY(hipMemcpyAsync(vn, W, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
            for (int i = 1; i <= j; i++) {
                cx_dot(X, N, VJ(i), vn, cj, 0, 0, s);
                hipLaunchKernelGGL(k_cx_axpy_dev, dim3(cx_grid(N)), dim3(256), 0, s, N, (const double *)X.scal.p,
                                   (const double *)VJ(i), vn);
            }
            cx_dot(X, N, vn, vn, cj, 1, nop, s);
            SC_TRY(cx_read(X, 2, s));
            if (slot_val(X, 1) == cd(0.0, 0.0)) {
                fail = true;
                break;
            }
            hipLaunchKernelGGL(k_cx_div_dev, dim3(cx_grid(N)), dim3(256), 0, s, N, (const double *)(X.scal.p + 2), vn);
        }
        if (fail) {
            code = -cnt_high;
            break;
        }
        // HMAT(i, j) = v_i^T H0 v_j
        for (int j = 1; j <= cnt_high; j++) {
            const int i0 = j >= cnt_low ? 1 : cnt_low;
            SC_TRY(hx_apply_cx(c, SCALIT_TMAP0, 1, VJ(j), HW, s));
            if (stats) stats[3] += 1;
            for (int i = i0; i <= cnt_high; i++) cx_dot(X, N, VJ(i), HW, cj, i - i0, 0, s);
            SC_TRY(cx_read(X, cnt_high - i0 + 1, s));
            for (int i = i0; i <= cnt_high; i++) H(i, j) = slot_val(X, i - i0);
        }
        int num = cnt_high - p.stepeig;
        if (eigs(num, tmp) != SCALIT_OK) {
            code = -num;
            break;
        }
        window_cx(p.e0, tmp, m, old);
        if (eigs(cnt_high, tmp) != SCALIT_OK) {
            code = -cnt_high;
            break;
        }
        window_cx(p.e0, tmp, m, ev);
        // getDiffMax_CX with nConvType 6 (util.f90:84-100 on |z|^2)
        res = std::fabs(op_cx(ev[0], 6) - op_cx(old[0], 6));
        for (int k = 1; k < m; k++) res = std::max(res, std::fabs(op_cx(ev[k], 6) - op_cx(old[k], 6)));
        if (stats) stats[2] = cnt_high;
        if (res <= p.etol) {
            code = cnt_high;
            break;
        }
        code = -cnt_high;
        cnt_low = cnt_high + 1;
        if (cnt_low >= mm) {
            code = -mm;
            break;
        }
        cnt_high = std::min(cnt_high + p.stepcnt, mm);
    }
    // Reorder_CX('A', 1): ascending real part, first index first among ties
    std::vector<int> ord(m);
    for (int k = 0; k < m; k++) ord[k] = k;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ev[a].real() < ev[b].real(); });
    for (int k = 0; k < m; k++) {
        eig[2 * k] = ev[ord[k]].real();
        eig[2 * k + 1] = ev[ord[k]].imag();
    }
    *ret = code;
    if (res_out) *res_out = res;
    return SCALIT_OK;
}

// ---------------------------------------------------------------- CRP
// AP = APP + APR and SQ_APR = sqrt(APR), as osbCRP sets them
__global__ void k_crp_prep(long N, const double *__restrict__ app, const double *__restrict__ apr,
                           double *__restrict__ ap, double *__restrict__ sq)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        ap[e] = app[e] + apr[e];
        sq[e] = sqrt(apr[e]);
    }
}

// y = a o x on one split complex vector
__global__ void k_cx_diag(long N, const double *__restrict__ a, const double *__restrict__ x, double *__restrict__ y)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        y[e] = a[e] * x[e];
        y[e + N] = a[e] * x[e + N];
    }
}

struct CrpArgs {
    int nm, qmr_type, qmrmax, sap;
    double qmrtol;
};

// CRP_HXDX (src/iterate/module/osbw.crp.h:79-131): out = APR^1/2 (H - i AP - E)^-1
// APP (H + i AP - E)^-1 APR^1/2 in, the two solves by OSB_QMRDX
static int crp_apply(scalit_ctx *c, const CrpArgs &a, const double *sq, const double *in, double *out, double *t1,
                     double *t2, int *its, hipStream_t s)
{
    const long N = c->N;
    const unsigned g = cx_grid(N);
    int n1 = 0, n2 = 0;
    double r = 0.0;
    hipLaunchKernelGGL(k_cx_diag, dim3(g), dim3(256), 0, s, N, sq, in, t1);
    SC_TRY(qmr_cx_run(c, c->osb, SCALIT_TMAP, a.sap ? SCALIT_TMAP : SCALIT_TA1, a.qmr_type, a.qmrmax, a.qmrtol, t1, t2,
                      &n1, &r, s));
    hipLaunchKernelGGL(k_cx_diag, dim3(g), dim3(256), 0, s, N, (const double *)c->ap[SCALIT_APP].p,
                       (const double *)t2, t1);
    SC_TRY(qmr_cx_run(c, c->osb, SCALIT_TAAP, a.sap ? SCALIT_TAAP : SCALIT_TA1, a.qmr_type, a.qmrmax, a.qmrtol, t1, t2,
                      &n2, &r, s));
    hipLaunchKernelGGL(k_cx_diag, dim3(g), dim3(256), 0, s, N, sq, (const double *)t2, out);
    HIP_TRY(hipGetLastError());
    *its += std::abs(n1) + std::abs(n2);
    return SCALIT_OK;
}

// calOSBCRP (osbw.crp.h:54-71) with LAN_DX (src/iterate/libiterate/lanczos/
// lan_dx.f90:299-325) and lanHij_dx (lanHij_dx.f90:59-97): nm steps of
// Lanczos with full re-orthogonalisation (LAN_MGS_ORTH_CX, orth/gsorth_cx.f90:
// 132-162) on the Hermitian CRP operator, the tridiagonal's eigenvalues
// (DSTEV), 4 times their sum; 0 when the Lanczos or the eigensolve fails.
static int crp_one(scalit_ctx *c, const CrpArgs &a, const double *sq, const double *x0, double *crp, int *its,
                   hipStream_t s)
{
    CxState &X = cxstate(c);
    const long N = c->N, L = 2 * N;
    const int M = a.nm;
    double *V = X.vj.p, *W = X.wj.p, *T1 = W + L, *T2 = T1 + L;
    auto VJ = [&](int j) { return V + (size_t)(j - 1) * L; };
    std::vector<double> alpha(M), beta(M);
    *crp = 0.0;
    cx_dot(X, N, x0, x0, 1, 0, 0, s);
    SC_TRY(cx_read(X, 1, s));
    const double t = std::sqrt(slot_val(X, 0).real());
    if (t == 0.0) return SCALIT_OK;
    cx_axpby(N, 1.0 / t, x0, 0.0, nullptr, VJ(1), s);
    for (int j = 1; j <= M; j++) {
        SC_TRY(crp_apply(c, a, sq, VJ(j), W, T1, T2, its, s));
        if (j != 1) cx_axpby(N, 1.0, W, -beta[j - 2], VJ(j - 1), W, s);
        cx_dot(X, N, W, VJ(j), 1, 0, 0, s);
        SC_TRY(cx_read(X, 1, s));
        alpha[j - 1] = slot_val(X, 0).real();
        double *vn = VJ(j + 1);
        HIP_TRY(hipMemcpyAsync(vn, W, sizeof(double) * L, hipMemcpyDeviceToDevice, s));
        for (int i = 1; i <= j; i++) {
            cx_dot(X, N, VJ(i), vn, 1, 0, 0, s);
            hipLaunchKernelGGL(k_cx_axpy_dev, dim3(cx_grid(N)), dim3(256), 0, s, N, (const double *)X.scal.p,
                               (const double *)VJ(i), vn);
        }
        cx_dot(X, N, vn, vn, 1, 1, 0, s);
        SC_TRY(cx_read(X, 2, s));
        const double nw = std::sqrt(slot_val(X, 1).real());
        beta[j - 1] = nw;
        if (nw == 0.0) return SCALIT_OK;   // lanHij_dx = -j: LAN_DX fails, the CRP is 0
        cx_axpby(N, 1.0 / nw, vn, 0.0, nullptr, vn, s);
    }
    std::vector<cd> T((size_t)M * M, cd(0.0, 0.0)), ev;
    for (int k = 0; k < M; k++) {
        T[k + (size_t)M * k] = alpha[k];
        if (k + 1 < M) {
            T[k + 1 + (size_t)M * k] = beta[k];
            T[k + (size_t)M * (k + 1)] = beta[k];
        }
    }
    if (cx_eigvals(M, T, ev) != SCALIT_OK) return SCALIT_OK;
    double sum = 0.0;
    std::vector<double> w(M);
    for (int k = 0; k < M; k++) w[k] = ev[k].real();
    std::sort(w.begin(), w.end());
    for (int k = 0; k < M; k++) sum += w[k];
    *crp = 4.0 * sum;
    return SCALIT_OK;
}

}  // namespace scalit

using namespace scalit;

extern "C" {

int scalit_set_ap(scalit_ctx *c, int which, const double *ap)
{
    if (!c || !ap || which < SCALIT_AP || which > SCALIT_APR) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    SC_TRY(c->ap[which].alloc(c->N));
    HIP_TRY(hipMemcpy(c->ap[which].p, ap, sizeof(double) * c->N, hipMemcpyHostToDevice));
    c->has_ap[which] = true;
    c->ap_ver++;
    c->gen++;
    return SCALIT_OK;
}

int scalit_hx_cx(scalit_ctx *c, int ntype, int nvec, const double *x, double *y, scalit_stream stream)
{
    if (!c || nvec < 1 || !x || !y || x == y) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    if (!c->has_res) return SCALIT_ESTATE;
    return hx_apply_cx(c, ntype, nvec, x, y, (hipStream_t)stream);
}

int scalit_px_cx(scalit_ctx *c, int nosb, int npc, int nvec, const double *x, double *y, scalit_stream stream)
{
    if (!c || nvec < 1 || !x || !y || x == y) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    return px_apply_cx(c, nosb, npc, nvec, x, y, (hipStream_t)stream);
}

int scalit_qmr_cx(scalit_ctx *c, int qmr_type, int hc, int pc, int qmrmax, double qmrtol, const double *b, double *x,
                  int *ret, double *res)
{
    if (!c || !b || !x || !ret || b == x || qmrmax < 0) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    if (!c->has_res || !c->has_precond) return SCALIT_ESTATE;
    if (c->osb == SCALIT_TOSBW && c->osbw_n <= 0) return SCALIT_ESTATE;   // TOSBW needs its window
    HIP_TRY(hipDeviceSynchronize());
    return qmr_cx_run(c, c->osb, hc, pc, qmr_type, qmrmax, qmrtol, b, x, ret, res, c->stream);
}

int scalit_pist_conv_sx(scalit_ctx *c, double e0, double etol, int ntype, const double *x0, int startcnt, int stepcnt,
                        int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int pc, double *eig,
                        int *ret, double *res, int *stats)
{
    if (!c || !x0 || !eig || !ret || m < 1 || m_max < 1 || ntype > 0) return SCALIT_EINVAL;
    if (pc != SCALIT_TMAP && pc != SCALIT_TA1) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    if (!c->has_res || !c->has_precond || !c->has_ap[SCALIT_AP]) return SCALIT_ESTATE;
    if (c->osb == SCALIT_TOSBW && c->osbw_n <= 0) return SCALIT_ESTATE;   // TOSBW needs its window
    HIP_TRY(hipDeviceSynchronize());
    PistSx p{e0, etol, ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, pc, qmrtol, 0};
    return pist_sx_run(c, p, x0, eig, ret, res, stats, c->stream);
}

int scalit_pist_conv_cx(scalit_ctx *c, double e0, double etol, int ntype, const double *x0, int startcnt, int stepcnt,
                        int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int pc, double *eig,
                        int *ret, double *res, int *stats)
{
    if (!c || !x0 || !eig || !ret || m < 1 || m_max < 1 || ntype > 0) return SCALIT_EINVAL;
    if (pc != SCALIT_TMAP && pc != SCALIT_TA1) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    if (!c->has_res || !c->has_precond || !c->has_ap[SCALIT_AP]) return SCALIT_ESTATE;
    if (c->osb == SCALIT_TOSBW && c->osbw_n <= 0) return SCALIT_ESTATE;   // TOSBW needs its window
    HIP_TRY(hipDeviceSynchronize());
    PistSx p{e0, etol, ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, pc, qmrtol, 1};
    return pist_sx_run(c, p, x0, eig, ret, res, stats, c->stream);
}

int scalit_crp(scalit_ctx *c, int ncrp, const double *energies, int nm, const double *x0, int qmr_type, int qmrmax,
               double qmrtol, int sap, double *crp, int *stats)
{
    if (!c || ncrp < 1 || !energies || nm < 1 || !x0 || !crp) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    if (!c->has_res || !c->has_precond || !c->has_ap[SCALIT_APP] || !c->has_ap[SCALIT_APR]) return SCALIT_ESTATE;
    if (c->osb == SCALIT_TOSBW && c->osbw_n <= 0) return SCALIT_ESTATE;   // TOSBW needs its window
    HIP_TRY(hipDeviceSynchronize());
    hipStream_t s = c->stream;
    CxState &X = cxstate(c);
    const long N = c->N, L = 2 * N;
    SC_TRY(c->ap[SCALIT_AP].alloc(N));
    SC_TRY(X.vj.alloc((size_t)(nm + 1) * L));
    SC_TRY(X.wj.alloc((size_t)3 * L));
    SC_TRY(cx_slots(X, 4));
    DBuf<double> sq;
    SC_TRY(sq.alloc(N));
    hipLaunchKernelGGL(k_crp_prep, dim3(cx_grid(N)), dim3(256), 0, s, N, (const double *)c->ap[SCALIT_APP].p,
                       (const double *)c->ap[SCALIT_APR].p, c->ap[SCALIT_AP].p, sq.p);
    HIP_TRY(hipGetLastError());
    c->has_ap[SCALIT_AP] = true;
    c->ap_ver++;
    const double e_keep = c->e0;
    CrpArgs a{nm, qmr_type, qmrmax, sap ? 1 : 0, qmrtol};
    int its = 0;
    for (int k = 0; k < ncrp; k++) {
        c->e0 = energies[k];   // sOSBW%mE0 = engy(i)
        const int rc = crp_one(c, a, sq.p, x0 + (size_t)k * L, crp + k, &its, s);
        if (rc != SCALIT_OK) {
            c->e0 = e_keep;
            return rc;
        }
    }
    c->e0 = e_keep;
    HIP_TRY(hipStreamSynchronize(s));
    if (stats) {
        stats[0] = its;
        stats[1] = ncrp * nm * 2;
    }
    return SCALIT_OK;
}

int scalit_small_eig_cx(int n, const double *a, int lda, double *w)
{
    if (n < 1 || !a || !w || lda < n) return SCALIT_EINVAL;
    std::vector<cd> A((size_t)n * n), ev;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            A[i + (size_t)n * j] = cd(a[2 * (i + (size_t)lda * j)], a[2 * (i + (size_t)lda * j) + 1]);
    SC_TRY(cx_eigvals(n, A, ev));
    for (int i = 0; i < n; i++) {
        w[2 * i] = ev[i].real();
        w[2 * i + 1] = ev[i].imag();
    }
    return SCALIT_OK;
}

}  // extern "C"
