// The windowed OSB preconditioner TOSBW (src/iterate/module/osb.px.h:191-206,
// set up by initOSBW, src/iterate/module/osb.osbw.h:2-50): the OSB states
// whose EIG0 lies within DE of E0 (DE adjusted until exactly mCnt do,
// adjustWinSize, osb.osbw.h:117-145) are treated exactly: P*X divides the
// other rotated components by EIG0 - E0 as TOSB does, and replaces the window
// components by the solution of (H_W - E0) y = x, with H_W the Hamiltonian in
// the window's OSB states (getVOSBH0, src/iterate/module/osb.hij.h:133-160:
// H_W(i,j) = v_i . H v_j off the diagonal, EIG0 on it).
//
// Here the window is set up on the GPU: every window state's OSB vector is a
// tensor product of one column per level (getLevelVi,
// src/iterate/module/osb.vi.h), written by one kernel for all states; H
// applies to all of them in one batched H*X; one kernel forms the upper
// triangle of H_W; and (H_W - E0)^-1 is formed once from its eigenvalues and
// eigenvectors (a symmetric solve equal to the reference's DSYTRF/DSYTRS up to
// rounding).  P*X keeps the fused level-1 pass of TOSB: the window's
// components of V1^T x are gathered and solved before it (one matrix-vector
// product of the window's size per vector), its divide reads EIG0 with +inf
// at the window (so those components become 0 rather than being divided),
// and V1 times the solved components is added to its output.
#include <cmath>
#include <vector>

#include "scalit_impl.hpp"

#include <complex>

namespace scalit {

struct OsbwLevels {
    int sF;
    int n[SC_FMAX];
    const double *v[SC_FMAX];
};

// X(:, k) = the OSB vector of grid index ind[k]: prod_l V_l[b_l](i_l, c_l)
__global__ void k_osbw_basis(long N, int cnt, OsbwLevels L, const int *__restrict__ ind, double *__restrict__ x,
                             long ldx)
{
    const int k = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N || k >= cnt) return;
    long ci = ind[k], ei = e;
    double val = 1.0;
    // level l's block index is the product index of the higher levels' columns
    long cols[SC_FMAX];
    for (int l = 0; l < L.sF; l++) {
        cols[l] = ci % L.n[l];
        ci /= L.n[l];
    }
    for (int l = 0; l < L.sF; l++) {
        const int n = L.n[l];
        const long il = ei % n;
        ei /= n;
        long b = 0;
        for (int m = L.sF - 1; m > l; m--) b = b * L.n[m] + cols[m];
        val *= L.v[l][b * n * n + il + (long)n * cols[l]];
    }
    x[(long)k * ldx + e] = val;
}

// H_W(i, j) = X(:, i) . HX(:, j) for i < j, both triangles, and
// H_W(i, i) = EIG0(ind(i)) - shift; one workgroup per pair
__global__ void __launch_bounds__(256) k_osbw_gram(long N, int cnt, const double *__restrict__ x, long ldx,
                                                   const double *__restrict__ hx, long ldh,
                                                   const double *__restrict__ eig0, const int *__restrict__ ind,
                                                   double shift, double *__restrict__ hw)
{
    __shared__ double red[4];
    const int i = blockIdx.x, j = blockIdx.y;
    if (i > j) return;
    if (i == j) {
        if (threadIdx.x == 0) hw[i + (long)cnt * i] = eig0[ind[i]] - shift;
        return;
    }
    const double *xi = x + (long)i * ldx, *hj = hx + (long)j * ldh;
    double acc = 0.0;
    for (long e = threadIdx.x; e < N; e += 256) acc += xi[e] * hj[e];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = red[0] + red[1] + red[2] + red[3];
        hw[i + (long)cnt * j] = v;
        hw[j + (long)cnt * i] = v;
    }
}

// M = Z diag(1 / w) Z^T (the inverse of a symmetric matrix from its eigenpairs)
__global__ void k_osbw_inverse(int n, const double *__restrict__ z, const double *__restrict__ w,
                               double *__restrict__ m)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    double acc = 0.0;
    for (int k = 0; k < n; k++) acc += z[i + (long)n * k] * z[j + (long)n * k] / w[k];
    m[e] = acc;
}

// EIG0 with +inf at the window's states
__global__ void k_osbw_mark(int n, const int *__restrict__ ind, double *__restrict__ e)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) e[ind[k]] = INFINITY;
}

// d = M z, z(k) = V1[b_k](:, c_k) . x(b_k block): the window's components of
// V1^T x solved; one workgroup of 16 waves per vector, one wave per window
// state for the dots, then each row of M z split over GATHER_KS partial sums
// (reduced through LDS in a fixed order)
constexpr int GATHER_NT = 1024, GATHER_KS = 4;
__global__ void __launch_bounds__(GATHER_NT) k_osbw_gather(int n1, int n, const int *__restrict__ ind,
                                                           const double *__restrict__ v1, const double *__restrict__ m,
                                                           const double *__restrict__ x, long ldx,
                                                           double *__restrict__ d, const int *__restrict__ active)
{
    extern __shared__ double zw[];   // n values of z, then GATHER_KS x n partial sums
    const int v = blockIdx.x;
    if (active && !active[v]) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = wv; k < n; k += GATHER_NT / 64) {
        const long g = ind[k], b = g / n1;
        const double *col = v1 + b * n1 * n1 + (long)n1 * (g - b * n1);
        const double *xb = x + (long)v * ldx + b * n1;
        double acc = 0.0;
        for (int i = lane; i < n1; i += 64) acc += col[i] * xb[i];
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (lane == 0) zw[k] = acc;
    }
    __syncthreads();
    double *part = zw + n;
    const int rows = GATHER_NT / GATHER_KS, q = threadIdx.x / rows, kq = (n + GATHER_KS - 1) / GATHER_KS;
    const int k0 = q * kq, k1 = min(n, k0 + kq);
    for (int i = threadIdx.x % rows; i < n; i += rows) {
        double acc = 0.0;
        for (int k = k0; k < k1; k++) acc += m[i + (long)n * k] * zw[k];
        part[q * n + i] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += GATHER_NT) {
        double acc = part[i];
        for (int j = 1; j < GATHER_KS; j++) acc += part[j * n + i];
        d[(long)v * n + i] = acc;
    }
}

// y(b block) += sum over the run's states k of V1[b](:, c_k) d(k): one
// workgroup per run of window states in one level-1 block and vector (runs
// touch disjoint blocks)
__global__ void __launch_bounds__(256) k_osbw_scatter(int n1, int n, const int *__restrict__ ind,
                                                      const int *__restrict__ runs, const double *__restrict__ v1,
                                                      const double *__restrict__ d, double *__restrict__ y, long ldy,
                                                      const int *__restrict__ active)
{
    const int r = blockIdx.x, v = blockIdx.y;
    if (active && !active[v]) return;
    const int k0 = runs[r], k1 = runs[r + 1];
    const long b = ind[k0] / n1;
    const double *vb = v1 + b * n1 * n1;
    double *yb = y + (long)v * ldy + b * n1;
    for (int i = threadIdx.x; i < n1; i += 256) {
        double acc = 0.0;
        for (int k = k0; k < k1; k++) acc += vb[i + (long)n1 * (ind[k] - b * n1)] * d[(long)v * n + k];
        yb[i] += acc;
    }
}

void osbw_gather(scalit_ctx *c, int nvec, const double *x, long ldx, const int *active, hipStream_t s)
{
    const int n = c->osbw_n;
    hipLaunchKernelGGL(k_osbw_gather, dim3(nvec), dim3(GATHER_NT), sizeof(double) * (size_t)n * (1 + GATHER_KS), s,
                       c->lv[0].n, n, c->osbw_ind.p, c->lv[0].vosb.p, c->osbw_m.p, x, ldx, c->osbw_d.p, active);
}

void osbw_scatter(scalit_ctx *c, int nvec, double *y, long ldy, const int *active, hipStream_t s)
{
    hipLaunchKernelGGL(k_osbw_scatter, dim3(c->osbw_nruns, nvec), dim3(256), 0, s, c->lv[0].n, c->osbw_n,
                       c->osbw_ind.p, c->osbw_runs.p, c->lv[0].vosb.p, c->osbw_d.p, y, ldy, active);
}

}  // namespace scalit

using namespace scalit;

namespace scalit {

// a[ind[k]] for the window's states
__global__ void k_osbw_take(int n, const int *__restrict__ ind, const double *__restrict__ a, double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = a[ind[k]];
}

// The window matrix of a complex job (createWH_DXCX, src/iterate/module/
// osb.pij.h:65-166; H_W complex for complex level matrices, HWCX):
// HWR = H_W + diag(-e0 + i sign A(WIND)),
// factorised there by ZGETRF (ZHETRF without an absorbing term) and applied
// by ZGETRS / ZHETRS in MYPX_DX's TOSBW case (osb.mypxdx.h:127-170).  Here
// its inverse is formed once per (nPC, e0, absorbing potentials, window) by
// Gauss-Jordan elimination with partial pivoting on the host (n <= 1024, as
// LAPACK runs on the host in the reference) and kept on the device, so that
// P*X applies it as one matrix-vector product.  A cache miss cannot be served
// under stream capture: the complex solvers' first P*X (eager) fills it.
int osbw_cx_matrix(scalit_ctx *c, int npc, double e0, const double *a, double sign, const dcx **m, hipStream_t s)
{
    const int n = c->osbw_n;
    if (n <= 0 || (int)c->osbw_ind_host.size() != n) return SCALIT_ESTATE;
    for (auto &S : c->osbw_cx)
        if (S.m.p && S.npc == npc && S.e0 == e0 && S.apv == c->ap_ver && S.wv == c->osbw_ver) {
            *m = (const dcx *)S.m.p;
            return SCALIT_OK;
        }
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(s, &st));
    if (st != hipStreamCaptureStatusNone) return SCALIT_ESTATE;
    int slot = -1;
    for (int k = 0; k < 4 && slot < 0; k++)
        if (c->osbw_cx[k].npc == npc) slot = k;
    if (slot < 0) slot = (c->osbw_cx_next++) & 3;
    auto &S = c->osbw_cx[slot];
    std::vector<double> aw(n, 0.0);
    if (a) {
        DBuf<double> t;
        SC_TRY(t.alloc(n));
        hipLaunchKernelGGL(k_osbw_take, dim3((n + 255) / 256), dim3(256), 0, s, n, c->osbw_ind.p, a, t.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(aw.data(), t.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    using cd = std::complex<double>;
    std::vector<cd> A((size_t)n * n), M((size_t)n * n, cd(0.0, 0.0));
    const bool cxw = !c->osbw_hwi_host.empty();   // complex level matrices: a complex H_W
    for (size_t e = 0; e < A.size(); e++) A[e] = cd(c->osbw_hw_host[e], cxw ? c->osbw_hwi_host[e] : 0.0);
    for (int i = 0; i < n; i++) {
        A[i + (size_t)n * i] += cd(-e0, sign * aw[i]);
        M[i + (size_t)n * i] = 1.0;
    }
    // Gauss-Jordan: rows of [A | M] reduced to [I | A^-1]
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int r = k + 1; r < n; r++)
            if (std::abs(A[r + (size_t)n * k]) > std::abs(A[p + (size_t)n * k])) p = r;
        if (A[p + (size_t)n * k] == cd(0.0, 0.0)) return SCALIT_ESTATE;   // singular (ZGETRF's INFO > 0)
        if (p != k)
            for (int j = 0; j < n; j++) {
                std::swap(A[k + (size_t)n * j], A[p + (size_t)n * j]);
                std::swap(M[k + (size_t)n * j], M[p + (size_t)n * j]);
            }
        const cd inv = 1.0 / A[k + (size_t)n * k];
        for (int j = 0; j < n; j++) {
            A[k + (size_t)n * j] *= inv;
            M[k + (size_t)n * j] *= inv;
        }
        for (int r = 0; r < n; r++) {
            if (r == k) continue;
            const cd f = A[r + (size_t)n * k];
            if (f == cd(0.0, 0.0)) continue;
            for (int j = 0; j < n; j++) {
                A[r + (size_t)n * j] -= f * A[k + (size_t)n * j];
                M[r + (size_t)n * j] -= f * M[k + (size_t)n * j];
            }
        }
    }
    const double *old = S.m.p;
    SC_TRY(S.m.alloc(2 * (size_t)n * n));
    if (S.m.p != old) c->gen++;
    HIP_TRY(hipMemcpy(S.m.p, M.data(), sizeof(cd) * M.size(), hipMemcpyHostToDevice));
    S.npc = npc;
    S.e0 = e0;
    S.apv = c->ap_ver;
    S.wv = c->osbw_ver;
    *m = (const dcx *)S.m.p;
    return SCALIT_OK;
}

}  // namespace scalit

extern "C" int scalit_set_osbw(scalit_ctx *c, int npc, double e0, double de, int cnt, int *hwlen, double *de_out)
{
    if (!c || cnt < 1 || cnt > 1024 || !(de > 0.0) || (npc != SCALIT_TA0 && npc != SCALIT_TA1)) return SCALIT_EINVAL;
    if (!c->has_precond || !c->has_res) return SCALIT_ESTATE;
    if (c->ndvr) return SCALIT_EINVAL;   // the level-1 rotation of an sNDVR operator is the outer layer
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const long N = c->N;
    c->osbw_n = 0;
    std::vector<double> eig(N);
    HIP_TRY(hipMemcpyAsync(eig.data(), c->eig0.p, 8 * (size_t)N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // adjustWinSize: the reference's search for a width holding exactly cnt states
    auto count = [&](double w) {
        long k = 0;
        for (long i = 0; i < N; i++)
            if (std::fabs(eig[i] - e0) <= w) k++;
        return k;
    };
    double de1 = 0.0;
    long cnt0 = count(de);
    int it = 0;
    while (cnt0 != cnt) {
        if (++it > 10000) return SCALIT_ESTATE;   // the search does not settle (ties in EIG0)
        if (cnt0 < cnt) {
            de1 = de;
            de = 2.0 * de;
        } else {
            de = 0.5 * (de1 + de);
        }
        cnt0 = count(de);
    }
    // getOSBWIndex: the window's states in grid order
    std::vector<int> ind;
    for (long i = 0; i < N && (int)ind.size() < cnt; i++)
        if (std::fabs(eig[i] - e0) <= de) ind.push_back((int)i);
    const int n = (int)ind.size();
    SC_TRY(c->osbw_ind.alloc(n));
    HIP_TRY(hipMemcpy(c->osbw_ind.p, ind.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    // the window's OSB vectors and H applied to them
    const long ld = (N + 63) & ~63L;
    DBuf<double> xv, hv, hw, w, z, work;
    SC_TRY(xv.alloc((size_t)n * ld));
    SC_TRY(hv.alloc((size_t)n * ld));
    OsbwLevels L;
    L.sF = c->sF;
    for (int l = 0; l < c->sF; l++) {
        L.n[l] = c->lv[l].n;
        L.v[l] = c->lv[l].vosb.p;
    }
    hipLaunchKernelGGL(k_osbw_basis, dim3((unsigned)((N + 255) / 256), n), dim3(256), 0, s, N, n, L,
                       c->osbw_ind.p, xv.p, ld);
    HIP_TRY(hipGetLastError());
    SC_TRY(hw.alloc((size_t)n * n));
    // the diagonal is EIG0 (getVOSBH0), less the shift (createWH: TA1)
    const double shift = npc == SCALIT_TA0 ? 0.0 : e0;
    c->osbw_hwi_host.clear();
    if (c->scx) {
        // getVOSBH0_CX (src/iterate/module/osb.hijcx.h:135-164): H_W(i,j) =
        // v_i . MYHX_CX(TA0) v_j above the diagonal, its conjugate below,
        // EIG0 on it: the real and imaginary planes' products, the lower
        // imaginary half negated on the host
        DBuf<double> xc, hc, hwi;
        SC_TRY(xc.alloc(2 * (size_t)n * N));
        SC_TRY(hc.alloc(2 * (size_t)n * N));
        SC_TRY(hwi.alloc((size_t)n * n));
        HIP_TRY(hipMemsetAsync(xc.p, 0, sizeof(double) * 2 * (size_t)n * N, s));
        HIP_TRY(hipMemcpy2DAsync(xc.p, sizeof(double) * 2 * N, xv.p, sizeof(double) * ld, sizeof(double) * N, n,
                                 hipMemcpyDeviceToDevice, s));
        SC_TRY(hx_apply_cxh(c, SCALIT_TA0, n, xc.p, hc.p, s));
        hipLaunchKernelGGL(k_osbw_gram, dim3(n, n), dim3(256), 0, s, N, n, xv.p, ld, (const double *)hc.p, 2 * N,
                           c->eig0.p, c->osbw_ind.p, shift, hw.p);
        hipLaunchKernelGGL(k_osbw_gram, dim3(n, n), dim3(256), 0, s, N, n, xv.p, ld, (const double *)hc.p + N,
                           2 * N, c->eig0.p, c->osbw_ind.p, 0.0, hwi.p);
        HIP_TRY(hipGetLastError());
        c->osbw_hwi_host.assign((size_t)n * n, 0.0);
        HIP_TRY(hipMemcpyAsync(c->osbw_hwi_host.data(), hwi.p, sizeof(double) * (size_t)n * n,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int j = 0; j < n; j++)
            for (int i = j; i < n; i++)
                c->osbw_hwi_host[i + (size_t)n * j] = i == j ? 0.0 : -c->osbw_hwi_host[j + (size_t)n * i];
    } else {
        SC_TRY(hx_apply_f64(c, nullptr, 0.0, n, xv.p, ld, hv.p, ld, false, nullptr, s));   // HijX: MYHX(TA0)
        hipLaunchKernelGGL(k_osbw_gram, dim3(n, n), dim3(256), 0, s, N, n, xv.p, ld, hv.p, ld, c->eig0.p,
                           c->osbw_ind.p, shift, hw.p);
        HIP_TRY(hipGetLastError());
    }
    SC_TRY(w.alloc(n));
    SC_TRY(z.alloc((size_t)n * n));
    SC_TRY(work.alloc(small_sym_eig_work(n)));
    SC_TRY(small_sym_eig(n, hw.p, n, w.p, z.p, n, work.p, s));
    // the unshifted H_W and the indices on the host, for the complex jobs'
    // window matrices (osbw_cx_matrix)
    c->osbw_hw_host.assign((size_t)n * n, 0.0);
    HIP_TRY(hipMemcpyAsync(c->osbw_hw_host.data(), hw.p, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) c->osbw_hw_host[i + (size_t)n * i] += shift;
    c->osbw_ind_host = ind;
    SC_TRY(c->osbw_m.alloc((size_t)n * n));
    hipLaunchKernelGGL(k_osbw_inverse, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0, s, n, z.p, w.p,
                       c->osbw_m.p);
    HIP_TRY(hipGetLastError());
    // runs of window states in one level-1 block (the indices ascend)
    const int n1 = c->lv[0].n;
    std::vector<int> runs{0};
    for (int k = 1; k < n; k++)
        if (ind[k] / n1 != ind[k - 1] / n1) runs.push_back(k);
    runs.push_back(n);
    SC_TRY(c->osbw_runs.alloc(runs.size()));
    HIP_TRY(hipMemcpy(c->osbw_runs.p, runs.data(), sizeof(int) * runs.size(), hipMemcpyHostToDevice));
    SC_TRY(c->osbw_eig0.alloc(N));
    HIP_TRY(hipMemcpyAsync(c->osbw_eig0.p, c->eig0.p, 8 * (size_t)N, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_osbw_mark, dim3((n + 255) / 256), dim3(256), 0, s, n, c->osbw_ind.p, c->osbw_eig0.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    c->osbw_n = n;
    c->osbw_nruns = (int)runs.size() - 1;
    c->osbw_shift = shift;
    c->osbw_ver++;
    c->gen++;
    if (hwlen) *hwlen = n;
    if (de_out) *de_out = de;
    return SCALIT_OK;
}
