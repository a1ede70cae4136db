// Complex level matrices (the reference's sCX = T): MYHX_CX on split complex
// vectors, and the C-ABI that loads H0CX / DEPCX / OUTHCX and switches them
// on.  (The sCX set-up of the OSB preconditioner is block_diag's complex
// branch in kernels_diag.hip; P*X keeps the real MYPX_DX.)
//
// MYHX_CX (src/iterate/module/osb.hxcx.h:191-245):
//   Y = (RES - E0 +- i A) X + sum_l level_l(X), with
//   level_l(X)(d, i, b) = [DEPCX(d)] sum_j H0CX(i, j) X(d, j, b)
//     (H1X_XYZ_CX / H1X_DEP_CX, src/iterate/libiterate/hpx/hpx_cx.f90:6-39,
//      over H1X_CX / H2X_CX, src/libutil/serial/base/matrix_cx.f90:97-160)
//   and, for the sNDVR layer, Y(p, i) = sum_j OUTHCX(p, i, j) X(p, j)
//     (H3X_CX, matrix_cx.f90:194-208).
// TMAPR takes -APR, as the parallel MYHX_CX (mosb3.hxcx.h:92-94) does; the
// serial one (osb.hxcx.h:234-235) takes -APP there.
//
// Vectors are split (N real parts, then N imaginary parts, per vector); the
// level matrices keep the reference's (re, im) pairs.  Each level's product
// is a batched complex matrix-vector product: one thread per output point
// (d, i, b), its row of H0CX read by every thread of the same i (one load
// for a wave walking d), its X column walking d with unit stride.  Every
// level reads and writes Y once; with the level's X columns staged in LDS
// (levels of up to 256 points) X is read once too: 48 bytes and 8 n_l FLOP
// per point and vector.
#include "scalit_impl.hpp"
#include "solver_state.hpp"

#include <algorithm>
#include <vector>

namespace scalit {

static unsigned cxh_grid(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 4096)); }

// y = (RES - sh) x + i sign a x (a null: no imaginary diagonal)
__global__ void __launch_bounds__(256) k_cxh_diag(long N, const double *__restrict__ res, double sh,
                                                  const double *__restrict__ a, double sign,
                                                  const double *__restrict__ x, double *__restrict__ y)
{
    const int v = blockIdx.y;
    const double *xr = x + 2L * v * N, *xi = xr + N;
    double *yr = y + 2L * v * N, *yi = yr + N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        const double d = res[e] - sh, t = a ? sign * a[e] : 0.0;
        const double r = xr[e], i = xi[e];
        yr[e] = d * r - t * i;
        yi[e] = d * i + t * r;
    }
}

// y(d, i, b) += [DEP(d)] sum_j H(i, j) x(d, j, b) on every vector
__global__ void __launch_bounds__(256) k_cxh_level(int nin, int n, int blk, const dcx *__restrict__ h,
                                                   const dcx *__restrict__ dep, long N,
                                                   const double *__restrict__ x, double *__restrict__ y)
{
    const int v = blockIdx.y;
    const double *xr = x + 2L * v * N, *xi = xr + N;
    double *yr = y + 2L * v * N, *yi = yr + N;
    const long total = (long)nin * n * blk;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int d = (int)(idx % nin);
        const long t = idx / nin;
        const int i = (int)(t % n);
        const long base = (t / n) * (long)nin * n + d;
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < n; j++) {
            const dcx hv = h[i + (long)n * j];
            const double pr = xr[base + (long)nin * j], pi = xi[base + (long)nin * j];
            ar += hv.re * pr - hv.im * pi;
            ai += hv.re * pi + hv.im * pr;
        }
        if (dep) {
            const dcx dv = dep[d];
            const double r = dv.re * ar - dv.im * ai;
            ai = dv.re * ai + dv.im * ar;
            ar = r;
        }
        yr[idx] += ar;
        yi[idx] += ai;
    }
}

// The same with the X columns staged in LDS.  Column c = (d, b) of a level
// is the n-vector x(d, ., b) (stride nin); a workgroup stages XC columns
// (XC consecutive d of one b, or XC consecutive b when nin = 1) once, and its
// threads, XC apart in the column and 256 / XC apart in the row, form every
// output of those columns from LDS: each X value is read from HBM once per
// level instead of once per output row.
constexpr int XC = 16;
__global__ void __launch_bounds__(256) k_cxh_level_lds(int nin, int n, int blk, const dcx *__restrict__ h,
                                                       const dcx *__restrict__ dep, long N,
                                                       const double *__restrict__ x, double *__restrict__ y)
{
    extern __shared__ char smem_raw[];
    dcx *xt = (dcx *)smem_raw;   // xt[j XC + cc]
    const int v = blockIdx.y;
    const double *xr = x + 2L * v * N, *xi = xr + N;
    double *yr = y + 2L * v * N, *yi = yr + N;
    const long ncol = (long)nin * blk, c0 = (long)blockIdx.x * XC;
    for (int e = threadIdx.x; e < n * XC; e += 256) {
        // nin = 1: a column is contiguous, so walk it first
        const int cc = nin == 1 ? e / n : e % XC, j = nin == 1 ? e % n : e / XC;
        const long c = c0 + cc;
        dcx val{0.0, 0.0};
        if (c < ncol) {
            const long off = (c % nin) + (long)nin * (j + (long)n * (c / nin));
            val = {xr[off], xi[off]};
        }
        xt[j * XC + cc] = val;
    }
    __syncthreads();
    constexpr int NG = 256 / XC;
    const int cc = threadIdx.x % XC, ig = threadIdx.x / XC;
    const long c = c0 + cc;
    if (c >= ncol) return;   // no barrier follows
    const long d = c % nin, b = c / nin;
    for (int i = ig; i < n; i += NG) {
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < n; j++) {
            const dcx hv = h[i + (long)n * j], xv = xt[j * XC + cc];
            ar += hv.re * xv.re - hv.im * xv.im;
            ai += hv.re * xv.im + hv.im * xv.re;
        }
        if (dep) {
            const dcx dv = dep[d];
            const double r = dv.re * ar - dv.im * ai;
            ai = dv.re * ai + dv.im * ar;
            ar = r;
        }
        const long off = d + (long)nin * (i + (long)n * b);
        yr[off] += ar;
        yi[off] += ai;
    }
}

// the sNDVR layer: y(p, i) += sum_j A(p, i, j) x(p, j), A(p, i, j) at
// p + nin (i + n j)
__global__ void __launch_bounds__(256) k_cxh_outh(int nin, int n, const dcx *__restrict__ a, long N,
                                                  const double *__restrict__ x, double *__restrict__ y)
{
    const int v = blockIdx.y;
    const double *xr = x + 2L * v * N, *xi = xr + N;
    double *yr = y + 2L * v * N, *yi = yr + N;
    const long total = (long)nin * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int p = (int)(idx % nin), i = (int)(idx / nin);
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < n; j++) {
            const dcx av = a[p + (long)nin * (i + (long)n * j)];
            const double pr = xr[p + (long)nin * j], pi = xi[p + (long)nin * j];
            ar += av.re * pr - av.im * pi;
            ai += av.re * pi + av.im * pr;
        }
        yr[idx] += ar;
        yi[idx] += ai;
    }
}

static bool scx_complete(const scalit_ctx *c)
{
    for (int l = 0; l < c->sF; l++) {
        const Level &L = c->lv[l];
        if (c->ndvr && l == c->sF - 1) {
            if (!c->has_outhcx) return false;
        } else if (!L.has_h0cx || (L.dep && !L.has_depcx)) {
            return false;
        }
    }
    return true;
}

// the real level matrices that MYHX_DX applies are all there
bool real_levels_ready(const scalit_ctx *c)
{
    for (int l = 0; l < c->sF; l++) {
        const Level &L = c->lv[l];
        if (c->ndvr && l == c->sF - 1) continue;   // the OUTH layer
        if (!L.has_h0 || (L.dep && !L.has_dep)) return false;
    }
    return true;
}

int hx_apply_cxh(scalit_ctx *c, int ntype, int nvec, const double *x, double *y, hipStream_t s)
{
    int rc = SCALIT_OK;
    double sign = 0.0;
    const double *a = cx_imag_diag(c, ntype, &sign, &rc);
    SC_TRY(rc);
    if (!c->has_res || !scx_complete(c)) return SCALIT_ESTATE;
    const long N = c->N;
    const double sh = (ntype == SCALIT_TA0 || ntype == SCALIT_TAAP0 || ntype == SCALIT_TMAP0) ? 0.0 : c->e0;
    hipLaunchKernelGGL(k_cxh_diag, dim3(cxh_grid(N), nvec), dim3(256), 0, s, N, c->res.p, sh, a, sign, x, y);
    for (int l = 0; l < c->sF; l++) {
        const Level &L = c->lv[l];
        if (c->ndvr && l == c->sF - 1)
            hipLaunchKernelGGL(k_cxh_outh, dim3(cxh_grid((long)L.dim * L.n), nvec), dim3(256), 0, s, L.dim, L.n,
                               (const dcx *)c->outhcx.p, N, x, y);
        else if (L.n <= 256)
            hipLaunchKernelGGL(k_cxh_level_lds,
                               dim3((unsigned)(((long)L.dim * L.blk + XC - 1) / XC), nvec), dim3(256),
                               sizeof(dcx) * (size_t)XC * L.n, s, L.dim, L.n, L.blk, (const dcx *)L.h0cx.p,
                               L.dep ? (const dcx *)L.depcx.p : nullptr, N, x, y);
        else
            hipLaunchKernelGGL(k_cxh_level, dim3(cxh_grid(N), nvec), dim3(256), 0, s, L.dim, L.n, L.blk,
                               (const dcx *)L.h0cx.p, L.dep ? (const dcx *)L.depcx.p : nullptr, N, x, y);
    }
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// upload n complex values given as separate real and imaginary parts into
// buf as (re, im) pairs; a new allocation bumps the operator generation
static int upload_cx(scalit_ctx *c, DBuf<double> &buf, long n, const double *re, const double *im)
{
    std::vector<double> t(2 * (size_t)n);
    for (long k = 0; k < n; k++) {
        t[2 * k] = re[k];
        t[2 * k + 1] = im[k];
    }
    const double *old = buf.p;
    SC_TRY(buf.alloc(2 * (size_t)n));
    if (buf.p != old) c->gen++;
    HIP_TRY(hipMemcpy(buf.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    return SCALIT_OK;
}

}  // namespace scalit

using namespace scalit;

extern "C" {

int scalit_set_h0_cx(scalit_ctx *c, int level, const double *h0_re, const double *h0_im)
{
    if (!c || level < 1 || level > c->sF || !h0_re || !h0_im) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    Level &L = c->lv[level - 1];
    SC_TRY(upload_cx(c, L.h0cx, (long)L.n * L.n, h0_re, h0_im));
    L.has_h0cx = true;
    return SCALIT_OK;
}

int scalit_set_dep_cx(scalit_ctx *c, int level, const double *dep_re, const double *dep_im)
{
    if (!c || level < 1 || level > c->sF || !dep_re || !dep_im) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    Level &L = c->lv[level - 1];
    if (!L.dep) return SCALIT_EINVAL;
    SC_TRY(upload_cx(c, L.depcx, L.dim, dep_re, dep_im));
    L.has_depcx = true;
    return SCALIT_OK;
}

int scalit_set_outh_cx(scalit_ctx *c, const double *outh_re, const double *outh_im)
{
    if (!c || !outh_re || !outh_im) return SCALIT_EINVAL;
    if (c->sF < 2) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    SC_TRY(upload_cx(c, c->outhcx, c->N * (long)c->lv[c->sF - 1].n, outh_re, outh_im));
    c->has_outhcx = true;
    // sNDVR, as scalit_set_outh sets it: the outermost level is the layer
    // and has no DEP (osb.print.h:204)
    c->ndvr = true;
    c->lv[c->sF - 1].dep = false;
    c->fast3 = fast3_ok(c);
    return SCALIT_OK;
}

int scalit_set_scx(scalit_ctx *c, int on)
{
    if (!c) return SCALIT_EINVAL;
    if (on && !scx_complete(c)) return SCALIT_ESTATE;
    if (c->scx != (on != 0)) {
        c->scx = on != 0;
        c->gen++;
    }
    return SCALIT_OK;
}

}  // extern "C"
