// GPU set-up of the OSB preconditioner and the small dense eigensolves.
//
// OSB set-up: myDiag (src/iterate/module/osb.diag.h:53-151) for levels sF..1:
//   HOSB(d,k,j,b) = H0(k,j) [* DEP(d)] + delta_kj EIG0(d,k,b)
//     (src/iterate/libiterate/diag/hinitxyz.f90:7-29, hinitdep.f90:10-34)
//   joint Jacobi sweeps per block b with one rotation angle per pair (p,q)
//   shared by all inner points d (MJACOBI, jacobi.f90:2-83), pairs in the
//   reference's order (i = 1..MAK-1, j = 1..MAK-i, pair (j, j+i)), until
//   |d offdiag| <= tol * initial or BJMAX sweeps (hosbdiag.f90:5-75);
//   VOSB accumulates the rotations (vupdate.f90:21-36), EIG0 <- diag(HOSB).
//
// Level 1 (NDK = 1): one wave per block runs the whole Jacobi in LDS.
// Levels with NDK > 1: one launch per rotation; each workgroup owns a slice of
// 64 inner points d (its 4 waves split the rotated rows and columns), reduces
// the A/B/C sums of its pair from the per-slice partials written by the
// previous rotation (every workgroup in the same fixed order, so all agree on
// the angle), applies the rotation to its slice, and writes the partial sums
// of the next pair.  A sweep of launches is recorded once in a HIP graph and
// replayed.
#include "scalit_impl.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace scalit {

__device__ __forceinline__ double jacobi_angle(double A, double B, double C)
{
    const double PIE = 3.14159265358979323846264338328;
    double X;
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
    return X;
}

// Element arithmetic of the set-up: double for a real H0, dcx for complex
// level matrices (HosbDiagDX, hosbdiag_dx.f90:6-73, with MJACOBI_CX,
// jacobi_cx.f90:5-96: the same real rotations of a complex symmetric HOSB,
// the angle from A = sum |H(p,q)|^2, B = 2 Re sum conj(H(p,q)) DDG,
// C = sum |DDG|^2).  rdot(a, b) is Re(conj(a) b), a term of dot_product.
__device__ __forceinline__ dcx operator+(dcx a, dcx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ dcx operator-(dcx a, dcx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ dcx operator*(double s, dcx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ double cmul(double a, double b) { return a * b; }
__device__ __forceinline__ dcx cmul(dcx a, dcx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ double rdot(double a, double b) { return a * b; }
__device__ __forceinline__ double rdot(dcx a, dcx b) { return a.re * b.re + a.im * b.im; }
__device__ __forceinline__ double real_of(double a) { return a; }
__device__ __forceinline__ double real_of(dcx a) { return a.re; }
__device__ __forceinline__ double add_real(double a, double r) { return a + r; }
__device__ __forceinline__ dcx add_real(dcx a, double r) { return {a.re + r, a.im}; }

// ---------------------------------------------------------------- level 1
// One wave per block b.  H (n x n) built in LDS from H0 and EIG0; V written
// to VOSB (global); EIG0 overwritten with the (real part of the) diagonal at
// the end.  The off-diagonal measure reads the block as doubles at
// j + n i (j < i): for a real H that is H(j,i); for a complex one it is what
// sumOffDiagCX (sumdiag.f90:49-66), whose dummy is declared double precision,
// reads from the complex block's (re, im) storage.
template <typename E>
__global__ void __launch_bounds__(64) k_osb_jacobi_ndk1(int n, int nblk, const E *__restrict__ h0,
                                                      double *__restrict__ eig0, double *__restrict__ vosb, int bjmax,
                                                      double bjtol, int *sweeps_out)
{
    extern __shared__ char smem_raw[];
    E *H = (E *)smem_raw;   // H(r,c) at r + n c
    const double *Hd = (const double *)smem_raw;
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    double *eb = eig0 + (long)b * n;
    double *V = vosb + (long)b * n * n;
    for (int e = lane; e < n * n; e += 64) {
        int r = e % n, c = e / n;
        E v = h0[e];
        if (r == c) v = add_real(v, eb[r]);
        H[e] = v;
        V[e] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    auto offdiag = [&]() {
        double s = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < i; j++) {
                double h = Hd[j + n * i];
                s = s + (0.0 + h * h);
            }
        return 2.0 * s;
    };
    double initye = 0.0;
    if (lane == 0) initye = offdiag();
    initye = __shfl(initye, 0);
    double ye = initye, dye = initye;
    int k = 1;
    while (fabs(dye) > initye * bjtol && k <= bjmax) {
        for (int i = 1; i <= n - 1; i++) {
            for (int j = 1; j <= n - i; j++) {
                const int p = j - 1, q = j + i - 1;
                const E hpq = H[p + n * q], hpp = H[p + n * p], hqq = H[q + n * q];
                const E odg = hpq;
                const E ddg = 0.5 * hpp - 0.5 * hqq;
                const double A = 0.0 + rdot(odg, odg);
                const double B = 2.0 * (0.0 + rdot(odg, ddg));
                const double C = 0.0 + rdot(ddg, ddg);
                const double X = jacobi_angle(A, B, C);
                const double SN = sin(X), CS = cos(X);
                const double C2 = CS * CS, S2 = SN * SN, SC = SN * CS;
                __syncthreads();
                if (lane == 0) {
                    H[p + n * p] = C2 * hpp + S2 * hqq - 2 * SC * odg;
                    H[q + n * q] = C2 * hqq + S2 * hpp + 2 * SC * odg;
                    const E nqp = (C2 - S2) * odg + 2 * SC * ddg;
                    H[p + n * q] = nqp;
                    H[q + n * p] = nqp;
                }
                for (int r = lane; r < n; r += 64) {
                    if (r != p && r != q) {
                        const E hrp = H[r + n * p], hrq = H[r + n * q];
                        const E np = CS * hrp - SN * hrq;
                        const E nq = CS * hrq + SN * hrp;
                        H[r + n * p] = np;
                        H[p + n * r] = np;
                        H[r + n * q] = nq;
                        H[q + n * r] = nq;
                    }
                    // VUPDATE(p, q, X): columns p and q of V
                    const double vp = V[r + n * p], vq = V[r + n * q];
                    V[r + n * p] = CS * vp - SN * vq;
                    V[r + n * q] = CS * vq + SN * vp;
                }
                __syncthreads();
            }
        }
        double s = 0.0;
        if (lane == 0) s = offdiag();
        s = __shfl(s, 0);
        dye = s - ye;
        ye = ye + dye;
        k = k + 1;
    }
    for (int r = lane; r < n; r += 64) eb[r] = real_of(H[r + n * r]);
    if (lane == 0 && sweeps_out) atomicAdd(sweeps_out, k - 1);
}

// ---------------------------------------------------------------- NDK > 1
// HOSB(d,k,j,b) = H0(k,j) * [DEP(d)] + delta_kj EIG0(d,k,b)
template <typename E>
__global__ void k_hosb_init(int ndk, int mak, int nblk, const E *__restrict__ h0, const E *__restrict__ dep,
                            const double *__restrict__ eig0, E *__restrict__ hosb)
{
    long total = (long)ndk * mak * mak * nblk;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int d = (int)(e % ndk);
        long t = e / ndk;
        int kk = (int)(t % mak);
        t /= mak;
        int jj = (int)(t % mak);
        long b = t / mak;
        E v = h0[kk + (long)mak * jj];
        if (dep) v = cmul(v, dep[d]);
        if (kk == jj) v = add_real(v, eig0[d + (long)ndk * (kk + (long)mak * b)]);
        hosb[e] = v;
    }
}

struct JPart {
    double a, b, c;
};

__device__ __forceinline__ double block_sum_256(double v, double *red)
{
    // deterministic: fixed tree over the 256 threads
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] = red[tid] + red[tid + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// The rotation kernels give a workgroup 64 values of d (JD) and split the
// other index over its 4 waves; a pair's A/B/C partials are one per 64-point
// slice of d (nrc = ceil(ndk / 64) per block).
constexpr int JD = 64;

// sum over the workgroup of three values, a fixed tree (deterministic)
__device__ __forceinline__ void block_sum3_256(double &a, double &b, double &c, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = a;
    red[256 + tid] = b;
    red[512 + tid] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[tid] += red[tid + s];
            red[256 + tid] += red[256 + tid + s];
            red[512 + tid] += red[512 + tid + s];
        }
        __syncthreads();
    }
    a = red[0];
    b = red[256];
    c = red[512];
    __syncthreads();
}

// Rotation (p,q) of the 64-point slice of d at d0 of block hb (mak x mak per
// d) and of the block's vectors V (do_v).  Wave w rotates rows/columns
// r = w, w+4, ... in groups of RG whose loads are all issued before any of
// their stores (the row and column writes cannot alias the column reads of the
// same group, but the compiler cannot know that, and would otherwise make
// every element a full memory round trip).
template <typename E>
__device__ __forceinline__ void jrot_slice(E *hb, double *V, int ndk, int mak, int d0, int p, int q, double CS,
                                           double SN, bool do_v)
{
    constexpr int RG = 8;
    const double C2 = CS * CS, S2 = SN * SN, SC = SN * CS;
    const int lane = threadIdx.x & (JD - 1), w = threadIdx.x / JD;
    const int d = d0 + lane;
    const bool din = d < ndk;
#define HO(r, c) hb[d + (long)ndk * ((r) + (long)mak * (c))]
    if (din && w == 0) {
        const E hpp = HO(p, p), hqq = HO(q, q), odg = HO(p, q);
        const E ddg = 0.5 * hpp - 0.5 * hqq;
        HO(p, p) = C2 * hpp + S2 * hqq - 2 * SC * odg;
        HO(q, q) = C2 * hqq + S2 * hpp + 2 * SC * odg;
        const E nqp = (C2 - S2) * odg + 2 * SC * ddg;
        HO(p, q) = nqp;
        HO(q, p) = nqp;
    }
    if (din)
        for (int r0 = w; r0 < mak; r0 += 4 * RG) {
            E hp[RG], hq[RG];
#pragma unroll
            for (int k = 0; k < RG; k++) {
                const int r = r0 + 4 * k;
                const int rc = r < mak ? r : p;
                hp[k] = HO(rc, p);
                hq[k] = HO(rc, q);
            }
#pragma unroll
            for (int k = 0; k < RG; k++) {
                const int r = r0 + 4 * k;
                if (r >= mak || r == p || r == q) continue;
                const E np = CS * hp[k] - SN * hq[k];
                const E nq = CS * hq[k] + SN * hp[k];
                HO(r, p) = np;
                HO(p, r) = np;
                HO(r, q) = nq;
                HO(q, r) = nq;
            }
        }
#undef HO
    if (do_v)
        for (int r = threadIdx.x; r < mak; r += 256) {
            const double vp = V[r + (long)mak * p], vq = V[r + (long)mak * q];
            V[r + (long)mak * p] = CS * vp - SN * vq;
            V[r + (long)mak * q] = CS * vq + SN * vp;
        }
}

// A/B/C sums of pair (p,q) over the slice of d at d0 (wave 0 reads, the whole
// workgroup reduces)
template <typename E>
__device__ __forceinline__ void jpair_sums(const E *hb, int ndk, int mak, int d0, int p, int q, double &a,
                                           double &bb, double &c, double *red)
{
    const int d = d0 + threadIdx.x;
    E odg{}, ddg{};
    if (threadIdx.x < JD && d < ndk) {
        odg = hb[d + (long)ndk * (p + (long)mak * q)];
        ddg = 0.5 * hb[d + (long)ndk * (p + (long)mak * p)] - 0.5 * hb[d + (long)ndk * (q + (long)mak * q)];
    }
    a = rdot(odg, odg);
    bb = rdot(odg, ddg);
    c = rdot(ddg, ddg);
    block_sum3_256(a, bb, c, red);
}

// Partial A/B/C sums of pair (p,q) over each 64-point slice of d.
template <typename E>
__global__ void __launch_bounds__(256) k_jpart(int ndk, int mak, int p, int q, const E *__restrict__ hosb,
                                               JPart *__restrict__ part, const int *active)
{
    __shared__ double red[3 * 256];
    const int b = blockIdx.y;
    if (!active[b]) return;
    double a, bb, c;
    jpair_sums(hosb + (long)ndk * mak * mak * b, ndk, mak, blockIdx.x * JD, p, q, a, bb, c, red);
    if (threadIdx.x == 0) part[(long)b * gridDim.x + blockIdx.x] = {a, bb, c};
}

// One rotation (p,q) of every active block, then the partials of (pn,qn).
template <typename E>
__global__ void __launch_bounds__(256) k_jrot(int ndk, int mak, int nrc, int p, int q, int pn, int qn,
                                              E *__restrict__ hosb, double *__restrict__ vosb,
                                              const JPart *__restrict__ part_in, JPart *__restrict__ part_out,
                                              const int *active)
{
    __shared__ double red[3 * 256];
    const int b = blockIdx.y;
    if (!active[b]) return;
    double A = 0, B = 0, C = 0;
    const JPart *pb = part_in + (long)b * nrc;
    for (int i = 0; i < nrc; i++) {
        A += pb[i].a;
        B += pb[i].b;
        C += pb[i].c;
    }
    B = 2.0 * B;
    const double X = jacobi_angle(A, B, C);
    const long m2 = (long)mak * mak;
    E *hb = hosb + (long)ndk * m2 * b;
    jrot_slice(hb, vosb + m2 * b, ndk, mak, blockIdx.x * JD, p, q, cos(X), sin(X), blockIdx.x == 0);
    if (pn < 0) return;
    __syncthreads();   // every wave's updates of this slice are in memory
    double a, bb, c;
    jpair_sums(hb, ndk, mak, blockIdx.x * JD, pn, qn, a, bb, c, red);
    if (threadIdx.x == 0) part_out[(long)b * nrc + blockIdx.x] = {a, bb, c};
}

// A whole sweep for levels with ndk <= 64 (one slice per block): one
// workgroup per active block applies every rotation of the pair sequence in
// turn, the sums of each pair reduced in the workgroup.
template <typename E>
__global__ void __launch_bounds__(256) k_jsweep(int ndk, int mak, int npair, const int *__restrict__ pq,
                                                E *__restrict__ hosb, double *__restrict__ vosb,
                                                const int *active)
{
    __shared__ double red[3 * 256];
    const int b = blockIdx.x;
    if (!active[b]) return;
    const long m2 = (long)mak * mak;
    E *hb = hosb + (long)ndk * m2 * b;
    double *V = vosb + m2 * b;
    for (int r = 0; r < npair; r++) {
        const int p = pq[2 * r], q = pq[2 * r + 1];
        double a, bb, c;
        jpair_sums(hb, ndk, mak, 0, p, q, a, bb, c, red);
        const double X = jacobi_angle(a, 2.0 * bb, c);
        jrot_slice(hb, V, ndk, mak, 0, p, q, cos(X), sin(X), true);
        __syncthreads();   // this rotation is in memory before the next pair's sums
    }
}

// per-slice partial of sum_{i>j} D(d + ndk (j + mak i))^2, D the block's
// storage read as doubles (bstride doubles per block): H(d,j,i)^2 for a real
// HOSB, and for a complex one what sumOffDiagCX reads (see k_osb_jacobi_ndk1)
__global__ void __launch_bounds__(256) k_offdiag_part(int ndk, int mak, long bstride, const double *__restrict__ hosb,
                                                      double *__restrict__ part)
{
    __shared__ double red[256];
    const int b = blockIdx.y;
    const int d = blockIdx.x * 256 + threadIdx.x;
    const double *hb = hosb + bstride * b;
    double s = 0;
    if (d < ndk)
        for (int i = 0; i < mak; i++)
            for (int j = 0; j < i; j++) {
                double h = hb[d + (long)ndk * (j + (long)mak * i)];
                s += h * h;
            }
    double t = block_sum_256(s, red);
    if (threadIdx.x == 0) part[(long)b * gridDim.x + blockIdx.x] = t;
}

// convergence bookkeeping per block b (hosbdiag.f90:23-43)
__global__ void k_jconv(int nblk, int nch, int first, const double *__restrict__ part, double *__restrict__ ye,
                        double *__restrict__ initye, int *__restrict__ kcount, int *__restrict__ active, int bjmax,
                        double bjtol)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    double s = 0;
    for (int i = 0; i < nch; i++) s += part[(long)b * nch + i];
    s = 2.0 * s;
    if (first) {
        initye[b] = s;
        ye[b] = s;
        kcount[b] = 1;
        active[b] = (fabs(s) > s * bjtol && 1 <= bjmax) ? 1 : 0;
        return;
    }
    if (!active[b]) return;
    double dye = s - ye[b];
    ye[b] = ye[b] + dye;
    kcount[b] += 1;
    active[b] = (fabs(dye) > initye[b] * bjtol && kcount[b] <= bjmax) ? 1 : 0;
}

__global__ void k_vinit(int mak, int nblk, double *vosb)
{
    long total = (long)mak * mak * nblk;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        long m = e % ((long)mak * mak);
        vosb[e] = (m % mak == m / mak) ? 1.0 : 0.0;
    }
}

template <typename E>
__global__ void k_eig_extract(int ndk, int mak, int nblk, const E *__restrict__ hosb, double *__restrict__ eig0)
{
    long total = (long)ndk * mak * nblk;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int d = (int)(e % ndk);
        long t = e / ndk;
        int i = (int)(t % mak);
        long b = t / mak;
        eig0[e] = real_of(hosb[d + (long)ndk * (i + (long)mak * (i + (long)mak * b))]);
    }
}

// hosb_src: the level's HOSB given as is (the sNDVR = T layer, OUTH,
// osb.diag.h:22-23) instead of H0 [x DEP] + diag(EIG0); h0 / dep the
// level's matrix and DEP (dep null for an independent level).
template <typename E>
static int diag_level_ndk(scalit_ctx *c, int l, int bjmax, double bjtol, const E *hosb_src, const E *h0,
                          const E *dep, hipStream_t s)
{
    Level &L = c->lv[l];
    const int ndk = L.dim, mak = L.n, nblk = L.blk;
    const int nch = (ndk + 255) / 256;    // off-diagonal partials: 256-point slices of d
    const int nrc = (ndk + JD - 1) / JD;  // rotation partials: 64-point slices of d
    const long bstride = (long)(sizeof(E) / sizeof(double)) * ndk * mak * mak;   // one block, in doubles
    DBuf<E> hosb;
    DBuf<double> part_od, ye, initye;
    DBuf<JPart> pa, pb;
    DBuf<int> active, kcount, pq;
    SC_TRY(hosb.alloc((size_t)ndk * mak * mak * nblk));
    SC_TRY(part_od.alloc((size_t)nch * nblk));
    SC_TRY(ye.alloc(nblk));
    SC_TRY(initye.alloc(nblk));
    SC_TRY(pa.alloc((size_t)nrc * nblk));
    SC_TRY(pb.alloc((size_t)nrc * nblk));
    SC_TRY(active.alloc(nblk));
    SC_TRY(kcount.alloc(nblk));
    int rc = SCALIT_OK;
    {
        if (hosb_src) {
            if (hipMemcpyAsync(hosb.p, hosb_src, sizeof(E) * (size_t)ndk * mak * mak * nblk,
                               hipMemcpyDeviceToDevice, s) != hipSuccess)
                rc = SCALIT_EHIP;
        } else
            hipLaunchKernelGGL(k_hosb_init<E>, dim3(4096), dim3(256), 0, s, ndk, mak, nblk, h0, dep, c->eig0.p,
                               hosb.p);
        hipLaunchKernelGGL(k_vinit, dim3(1024), dim3(256), 0, s, mak, nblk, L.vosb.p);
        hipLaunchKernelGGL(k_offdiag_part, dim3(nch, nblk), dim3(256), 0, s, ndk, mak, bstride,
                           (const double *)hosb.p, part_od.p);
        hipLaunchKernelGGL(k_jconv, dim3((nblk + 255) / 256), dim3(256), 0, s, nblk, nch, 1, part_od.p, ye.p, initye.p,
                           kcount.p, active.p, bjmax, bjtol);
        // the pair sequence of one sweep
        std::vector<int> pp, qq;
        for (int i = 1; i <= mak - 1; i++)
            for (int j = 1; j <= mak - i; j++) {
                pp.push_back(j - 1);
                qq.push_back(j + i - 1);
            }
        const int npair = (int)pp.size();
        // ndk <= 64: one workgroup per block runs a whole sweep
        if (nrc == 1 && npair > 0) {
            std::vector<int> h(2 * (size_t)npair);
            for (int r = 0; r < npair; r++) {
                h[2 * r] = pp[r];
                h[2 * r + 1] = qq[r];
            }
            if (pq.alloc(h.size()) != SCALIT_OK) rc = SCALIT_ENOMEM;
            else if (hipMemcpyAsync(pq.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
                     hipStreamSynchronize(s) != hipSuccess)
                rc = SCALIT_EHIP;
        }
        std::vector<int> hact(nblk);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        bool captured = false;
        for (int sweep = 0; rc == SCALIT_OK && sweep < bjmax && npair > 0; sweep++) {
            if (hipMemcpyAsync(hact.data(), active.p, sizeof(int) * nblk, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = SCALIT_EHIP;
                break;
            }
            bool any = false;
            for (int b = 0; b < nblk; b++) any = any || hact[b];
            if (!any) break;
            if (nrc == 1) {
                hipLaunchKernelGGL(k_jsweep<E>, dim3(nblk), dim3(256), 0, s, ndk, mak, npair, pq.p, hosb.p, L.vosb.p,
                                   active.p);
                hipLaunchKernelGGL(k_offdiag_part, dim3(nch, nblk), dim3(256), 0, s, ndk, mak, bstride,
                                   (const double *)hosb.p, part_od.p);
                hipLaunchKernelGGL(k_jconv, dim3((nblk + 255) / 256), dim3(256), 0, s, nblk, nch, 0, part_od.p, ye.p,
                                   initye.p, kcount.p, active.p, bjmax, bjtol);
                if (hipGetLastError() != hipSuccess) rc = SCALIT_EHIP;
                continue;
            }
            if (!captured) {
                // record one sweep: partials of the first pair, every rotation, convergence
                if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                    rc = SCALIT_EHIP;
                    break;
                }
                hipLaunchKernelGGL(k_jpart<E>, dim3(nrc, nblk), dim3(256), 0, s, ndk, mak, pp[0], qq[0], hosb.p,
                                   pa.p, active.p);
                JPart *pin = pa.p, *pout = pb.p;
                for (int r = 0; r < npair; r++) {
                    int pn = (r + 1 < npair) ? pp[r + 1] : -1;
                    int qn = (r + 1 < npair) ? qq[r + 1] : -1;
                    hipLaunchKernelGGL(k_jrot<E>, dim3(nrc, nblk), dim3(256), 0, s, ndk, mak, nrc, pp[r], qq[r], pn,
                                       qn, hosb.p, L.vosb.p, pin, pout, active.p);
                    JPart *t = pin;
                    pin = pout;
                    pout = t;
                }
                hipLaunchKernelGGL(k_offdiag_part, dim3(nch, nblk), dim3(256), 0, s, ndk, mak, bstride,
                                   (const double *)hosb.p, part_od.p);
                hipLaunchKernelGGL(k_jconv, dim3((nblk + 255) / 256), dim3(256), 0, s, nblk, nch, 0, part_od.p, ye.p,
                                   initye.p, kcount.p, active.p, bjmax, bjtol);
                if (hipStreamEndCapture(s, &graph) != hipSuccess ||
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                    rc = SCALIT_EHIP;
                    break;
                }
                captured = true;
            }
            if (hipGraphLaunch(exec, s) != hipSuccess) {
                rc = SCALIT_EHIP;
                break;
            }
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (rc == SCALIT_OK) {
            hipLaunchKernelGGL(k_eig_extract<E>, dim3(4096), dim3(256), 0, s, ndk, mak, nblk, hosb.p, c->eig0.p);
            if (hipStreamSynchronize(s) != hipSuccess) rc = SCALIT_EHIP;
        }
    }
    pq.release();
    hosb.release();
    part_od.release();
    ye.release();
    initye.release();
    pa.release();
    pb.release();
    active.release();
    kcount.release();
    return rc;
}

// Level 1 with blocks too large for the one-wave LDS Jacobi (n1 > 143 does
// not fit 160 KB: the combined (r, R) layer of an energy-cut basis): every
// block's HOSB = H0 + diag(EIG0) (hosbdiag.f90 with NDK = 1) is
// diagonalised by the batched cyclic Jacobi eigensolver, one workgroup per
// block, to an off-diagonal norm of 1e-15 of the whole.  Its pair order is
// not MJACOBI's, so VOSB and EIG0 equal the reference's up to the order and
// signs of the eigenpairs: the same preconditioner.
static int diag_level_dense(scalit_ctx *c, int l, hipStream_t s)
{
    Level &L = c->lv[l];
    const int n = L.n, nb = L.blk;
    if (n > 1024) return SCALIT_EINVAL;
    DBuf<double> hosb, work;
    DBuf<EigvecJob> djobs;
    SC_TRY(hosb.alloc((size_t)nb * n * n));
    hipLaunchKernelGGL(k_hosb_init<double>, dim3(4096), dim3(256), 0, s, 1, n, nb, L.h0.p, (const double *)nullptr,
                       c->eig0.p, hosb.p);
    std::vector<EigvecJob> jobs(nb);
    for (int b = 0; b < nb; b++)
        jobs[b] = {hosb.p + (long)b * n * n, c->eig0.p + (long)b * n, L.vosb.p + (long)b * n * n, n, n, n};
    SC_TRY(small_sym_eig_batch(nb, jobs.data(), djobs, work, s));
    HIP_TRY(hipStreamSynchronize(s));   // the host job array and the scoped buffers end here
    return SCALIT_OK;
}

// The sCX = T set-up (the complex branch of myDiag, osb.diag.h:76-101): the
// same levels and rotations on the complex HOSB of H0CX [x DEPCX] (or
// OUTHCX) + diag(EIG0); VOSB stays real and EIG0 takes the real parts.
// Level-1 blocks too large for LDS take the global-memory sweep.
static int block_diag_cx(scalit_ctx *c, int bjmax, double bjtol, hipStream_t s)
{
    for (int l = c->sF - 1; l >= 0; l--) {
        Level &L = c->lv[l];
        const bool layer = c->ndvr && l == c->sF - 1;
        const dcx *h0 = (const dcx *)L.h0cx.p;
        const dcx *dep = L.dep ? (const dcx *)L.depcx.p : nullptr;
        if (!layer && L.dim == 1 && sizeof(dcx) * (size_t)L.n * L.n <= 160 * 1024) {
            size_t lds = sizeof(dcx) * (size_t)L.n * L.n;
            allow_full_lds((const void *)k_osb_jacobi_ndk1<dcx>);
            hipLaunchKernelGGL(k_osb_jacobi_ndk1<dcx>, dim3(L.blk), dim3(64), lds, s, L.n, L.blk, h0, c->eig0.p,
                               L.vosb.p, bjmax, bjtol, (int *)nullptr);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s));
        } else {
            SC_TRY(diag_level_ndk<dcx>(c, l, bjmax, bjtol, layer ? (const dcx *)c->outhcx.p : nullptr, h0, dep, s));
        }
    }
    return SCALIT_OK;
}

int block_diag(scalit_ctx *c, int bjmax, double bjtol)
{
    hipStream_t s = c->stream;
    for (int l = 0; l < c->sF; l++) {
        const bool layer = c->ndvr && l == c->sF - 1;
        if (c->scx) {
            if (layer ? !c->has_outhcx : (!c->lv[l].has_h0cx || (c->lv[l].dep && !c->lv[l].has_depcx)))
                return SCALIT_ESTATE;
        } else {
            if (!c->lv[l].has_h0 && !layer) return SCALIT_ESTATE;
            if (c->lv[l].dep && !c->lv[l].has_dep) return SCALIT_ESTATE;
        }
        SC_TRY(c->lv[l].vosb.alloc((size_t)c->lv[l].blk * c->lv[l].n * c->lv[l].n));
    }
    if (!c->has_res) return SCALIT_ESTATE;
    SC_TRY(c->eig0.alloc(c->N));
    HIP_TRY(hipMemcpyAsync(c->eig0.p, c->res.p, sizeof(double) * c->N, hipMemcpyDeviceToDevice, s));
    for (int l = c->sF - 1; !c->scx && l >= 0; l--) {
        Level &L = c->lv[l];
        if (L.dim == 1 && sizeof(double) * (size_t)L.n * L.n > 160 * 1024) {
            SC_TRY(diag_level_dense(c, l, s));
        } else if (L.dim == 1) {
            size_t lds = sizeof(double) * (size_t)L.n * L.n;
            allow_full_lds((const void *)k_osb_jacobi_ndk1<double>);
            hipLaunchKernelGGL(k_osb_jacobi_ndk1<double>, dim3(L.blk), dim3(64), lds, s, L.n, L.blk, L.h0.p,
                               c->eig0.p, L.vosb.p, bjmax, bjtol, (int *)nullptr);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(s));
        } else {
            SC_TRY(diag_level_ndk<double>(c, l, bjmax, bjtol, (c->ndvr && l == c->sF - 1) ? c->outh.p : nullptr,
                                          L.h0.p, L.dep ? L.depv.p : nullptr, s));
        }
    }
    if (c->scx) SC_TRY(block_diag_cx(c, bjmax, bjtol, s));
    c->has_precond = true;
    c->osbw_n = 0;
    // fp32 copies are rebuilt on demand
    for (int l = 0; l < c->sF; l++) {
        c->lv[l].vosb_f.release();
        c->lv[l].vosbt.release();
        c->lv[l].vosbt_f.release();
    }
    c->eig0_f.release();
    return SCALIT_OK;
}

// ---------------------------------------------------------------------------
// small dense symmetric eigensolver (one workgroup): cyclic Jacobi in the
// round-robin (tournament) ordering, n/2 disjoint rotations per round.  It
// stands in for LAPACK DSYEV('N'/'V') of the Rayleigh-Ritz matrices.
// work layout: A (n x n, ld n) | Z (n x n) | cs (2 * n) | pairs (2 * n ints)
// ---------------------------------------------------------------------------
size_t small_sym_eig_work(int n) { return (size_t)n * n * 2 + 4 * (size_t)n + 8; }

__device__ void small_sym_eig_body(int n, const double *__restrict__ a_in, int lda, double *__restrict__ w,
                                   double *__restrict__ z_out, int ldz, double *__restrict__ work)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    double *A = work;
    double *Z = A + (long)n * n;
    double *cs = Z + (long)n * n;   // c, s per pair
    __shared__ int pairs[2 * 512];
    __shared__ double red[1024];
    __shared__ int flag;
    const int m = (n + 1) / 2 * 2;   // even count with a dummy index n when n is odd
    for (int e = tid; e < n * n; e += nt) {
        int i = e % n, j = e / n;
        // symmetric from the upper triangle (DSYEV 'U')
        A[e] = (i <= j) ? a_in[i + (long)lda * j] : a_in[j + (long)lda * i];
        Z[e] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 60; sweep++) {
        // off-diagonal and total Frobenius norms
        double off = 0, tot = 0;
        for (int e = tid; e < n * n; e += nt) {
            double v = A[e] * A[e];
            tot += v;
            if (e % n != e / n) off += v;
        }
        red[tid] = off;
        __syncthreads();
        for (int s2 = nt / 2; s2 > 0; s2 >>= 1) {
            if (tid < s2) red[tid] += red[tid + s2];
            __syncthreads();
        }
        off = red[0];
        __syncthreads();
        red[tid] = tot;
        __syncthreads();
        for (int s2 = nt / 2; s2 > 0; s2 >>= 1) {
            if (tid < s2) red[tid] += red[tid + s2];
            __syncthreads();
        }
        tot = red[0];
        __syncthreads();
        // stop once the off-diagonal part is below 1e-15 of the whole (in
        // Frobenius norm; off and tot are sums of squares): a tighter test
        // sits under the rounding floor and would run every sweep
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int round = 0; round < m - 1; round++) {
            // tournament pairing: position 0 fixed, others rotate
            for (int k = tid; k < m / 2; k += nt) {
                int a = (k == 0) ? 0 : 1 + (k - 1 + round) % (m - 1);
                int b = 1 + (m - 2 - k + round) % (m - 1);
                if (k == 0) b = 1 + (m - 2 + round) % (m - 1);
                int p = a < b ? a : b, q = a < b ? b : a;
                pairs[2 * k] = p;
                pairs[2 * k + 1] = q;
                double c = 1.0, s = 0.0;
                if (q < n) {
                    double apq = A[p + (long)n * q];
                    if (apq != 0.0) {
                        double app = A[p + (long)n * p], aqq = A[q + (long)n * q];
                        double theta = (aqq - app) / (2.0 * apq);
                        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                    }
                }
                cs[2 * k] = c;
                cs[2 * k + 1] = s;
            }
            __syncthreads();
            // rows: A <- J^T A  (each pair touches rows p and q)
            for (int e = tid; e < (m / 2) * n; e += nt) {
                int k = e / n, col = e % n;
                int p = pairs[2 * k], q = pairs[2 * k + 1];
                if (q >= n) continue;
                double c = cs[2 * k], s = cs[2 * k + 1];
                if (s == 0.0) continue;
                double ap = A[p + (long)n * col], aq = A[q + (long)n * col];
                A[p + (long)n * col] = c * ap - s * aq;
                A[q + (long)n * col] = s * ap + c * aq;
            }
            __syncthreads();
            // columns: A <- A J ; Z <- Z J
            for (int e = tid; e < (m / 2) * n; e += nt) {
                int k = e / n, row = e % n;
                int p = pairs[2 * k], q = pairs[2 * k + 1];
                if (q >= n) continue;
                double c = cs[2 * k], s = cs[2 * k + 1];
                if (s == 0.0) continue;
                double ap = A[row + (long)n * p], aq = A[row + (long)n * q];
                A[row + (long)n * p] = c * ap - s * aq;
                A[row + (long)n * q] = s * ap + c * aq;
                double zp = Z[row + (long)n * p], zq = Z[row + (long)n * q];
                Z[row + (long)n * p] = c * zp - s * zq;
                Z[row + (long)n * q] = s * zp + c * zq;
            }
            __syncthreads();
        }
    }
    (void)flag;
    // sort ascending (selection by rank: stable on ties by index)
    for (int i = tid; i < n; i += nt) {
        double di = A[i + (long)n * i];
        int rank = 0;
        for (int j = 0; j < n; j++) {
            double dj = A[j + (long)n * j];
            if (dj < di || (dj == di && j < i)) rank++;
        }
        w[rank] = di;
        if (z_out)
            for (int r = 0; r < n; r++) z_out[r + (long)ldz * rank] = Z[r + (long)n * i];
    }
}

__global__ void __launch_bounds__(1024) k_small_sym_eig(int n, const double *__restrict__ a_in, int lda,
                                                        double *__restrict__ w, double *__restrict__ z_out, int ldz,
                                                        double *__restrict__ work)
{
    small_sym_eig_body(n, a_in, lda, w, z_out, ldz, work);
}

// one problem per workgroup (blockIdx.x picks the job; work: wstride doubles each)
__global__ void __launch_bounds__(1024) k_small_sym_eig_batch(const EigvecJob *__restrict__ jobs,
                                                              double *__restrict__ work, long wstride)
{
    const EigvecJob jb = jobs[blockIdx.x];
    small_sym_eig_body(jb.n, jb.a, jb.lda, jb.w, jb.z, jb.ldz, work + (long)blockIdx.x * wstride);
}

// ---------------------------------------------------------------------------
// Eigenvalues only (DSYEV 'N'): Householder reduction to tridiagonal form
// (Golub & Van Loan, Matrix Computations, alg. 8.3.1), then every eigenvalue
// by multisection on the Sturm sequence of the tridiagonal matrix (Demmel,
// Applied Numerical Linear Algebra, sec. 5.3.4: the count of negative pivots
// of T - xI is the number of eigenvalues below x).  T threads share one
// eigenvalue index and each round count at T points of its interval, which
// shrinks the interval (T+1)-fold per round instead of 2-fold; the index also
// gives the ascending order.  One workgroup of 1024 threads per problem
// (blockIdx.x picks the job); the matrix lives in LDS when it fits (LDS_A) and
// in the work buffer (wstride doubles per job) otherwise.
// ---------------------------------------------------------------------------
constexpr int EIGV_NT = 1024;
constexpr int EIGV_LDS_N = 112;   // largest n whose n x (n+1) matrix sits in LDS with the vectors

__device__ double block_sum_1024(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();   // red may still be read by the previous reduction
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0;
#pragma unroll
    for (int k = 0; k < EIGV_NT / 64; k++) t += red[k];
    return t;
}

// Eigenvalues, ascending, of the symmetric tridiagonal matrix with diagonal
// d[0..n) and off-diagonal e[0..n-1) into w, by the whole workgroup
// (EIGV_NT threads): multisection on the Sturm count, T threads per
// eigenvalue (a power of two <= 64, so a group never spans two wavefronts).
__device__ void tridiag_bisect(int n, const double *d, const double *e, double *__restrict__ w)
{
    const int tid = threadIdx.x;
    // Gershgorin interval of the tridiagonal matrix
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < n; i++) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
    }
    const double span = fmax(hi - lo, 1e-300);
    lo -= 1e-12 * span;
    hi += 1e-12 * span;
    const double pivmin = 1e-290;
    // T = threads per eigenvalue: a power of two <= 64 (a group never spans
    // two wavefronts, so the group reductions are lane shuffles)
    int T = 1;
    while (T < 64 && 2 * T * n <= EIGV_NT) T *= 2;   // every eigenvalue in one pass of the workgroup
    const int sub = tid & (T - 1);
    const double frac = (double)(sub + 1) / (double)(T + 1);
    for (int base = 0; base < n; base += EIGV_NT / T) {
        // the index this group brackets: the smallest x with count(x) > idx
        const int idx = base + tid / T;
        const int ie = idx < n ? idx : n - 1;
        double a = lo, b = hi;   // count(a) <= ie < count(b), identical across the group
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (a + b);
            if (mid <= a || mid >= b) break;   // interval at one ulp
            const double x = T == 1 ? mid : a + (b - a) * frac;
            int cnt = 0;
            double q = d[0] - x;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0;
            for (int i = 1; i < n; i++) {
                q = d[i] - x - e[i - 1] * e[i - 1] / q;
                if (fabs(q) < pivmin) q = -pivmin;
                cnt += q < 0;
            }
            double bn = cnt > ie ? x : b;
            for (int o = 1; o < T; o <<= 1) bn = fmin(bn, __shfl_xor(bn, o));
            double an = (cnt <= ie && x < bn) ? x : a;
            for (int o = 1; o < T; o <<= 1) an = fmax(an, __shfl_xor(an, o));
            if (an == a && bn == b) break;   // every point fell on an end: a few ulps apart
            a = an;
            b = bn;
        }
        if (idx < n && sub == 0) w[idx] = 0.5 * (a + b);
    }
}

template <bool LDS_A>
__global__ void __launch_bounds__(EIGV_NT) k_sym_eigvals(const EigvalJob *__restrict__ jobs, EigvalJob one,
                                                         double *__restrict__ gwork, long wstride)
{
    extern __shared__ double sm_eig[];
    const EigvalJob jb = jobs ? jobs[blockIdx.x] : one;
    const int n = jb.n, lda = jb.lda;
    const double *__restrict__ a_in = jb.a;
    double *__restrict__ w = jb.w;
    const int tid = threadIdx.x;
    const int ld = n + 1;
    double *A = LDS_A ? sm_eig : gwork + (long)blockIdx.x * wstride;   // A(i,j) at i + ld j
    double *vec = LDS_A ? sm_eig + (long)n * ld : sm_eig;  // v | p | d | e (n each) | red (16)
    double *v = vec, *pw = vec + n, *d = vec + 2 * n, *e = vec + 3 * n, *red = vec + 4 * n;
    for (int x = tid; x < n * n; x += EIGV_NT) {
        const int i = x % n, j = x / n;
        A[i + ld * j] = (i <= j) ? a_in[i + (long)lda * j] : a_in[j + (long)lda * i];   // from the upper triangle
    }
    __syncthreads();
    // One Householder step per column, three barriers each: wave 0 forms the
    // reflector (v, tau) alone; every thread then takes part in p = tau A22 v;
    // each wave forms p.v by itself (the same sum in every wave) and the
    // rank-2 update A22 -= v w^T + w v^T uses w = p - (tau p.v / 2) v on the
    // fly.  tau goes through red[k & 1], so wave 0 may start the next column
    // while the other waves still read this one's.
    const int lane = tid & 63;
    for (int k = 0; k + 2 < n; k++) {
        const int m = n - k - 1;   // trailing size; x = A(k+1.., k)
        if (tid < 64) {
            double ss = 0;
            for (int i = 1 + lane; i < m; i += 64) {
                const double xi = A[(k + 1 + i) + ld * k];
                ss += xi * xi;
            }
            for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);   // below the subdiagonal
            const double x0 = A[(k + 1) + ld * k];
            if (lane == 0) d[k] = A[k + ld * k];
            if (ss == 0.0) {   // column already reduced
                if (lane == 0) {
                    e[k] = x0;
                    red[k & 1] = 0.0;
                }
            } else {
                const double nrm = sqrt(x0 * x0 + ss);
                const double alpha = x0 <= 0 ? nrm : -nrm;   // new subdiagonal value, no cancellation in v0
                if (lane == 0) {
                    e[k] = alpha;
                    red[k & 1] = 2.0 / ((x0 - alpha) * (x0 - alpha) + ss);
                }
                for (int i = lane; i < m; i += 64) v[i] = (i == 0) ? x0 - alpha : A[(k + 1 + i) + ld * k];
            }
        }
        __syncthreads();
        const double tau = red[k & 1];
        if (tau == 0.0) continue;
        // p = tau A22 v: eight threads per row (neighbouring lanes) while the
        // rows fit, their partial sums joined by lane shuffles
        if (m <= EIGV_NT / 8) {
            const int i = tid >> 3, sub = tid & 7;
            double acc = 0;
            if (i < m)
                for (int j = sub; j < m; j += 8) acc += A[(k + 1 + i) + ld * (k + 1 + j)] * v[j];
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            acc += __shfl_xor(acc, 4);
            if (i < m && sub == 0) pw[i] = tau * acc;
        } else {
            for (int i = tid; i < m; i += EIGV_NT) {
                double acc = 0;
                for (int j = 0; j < m; j++) acc += A[(k + 1 + i) + ld * (k + 1 + j)] * v[j];
                pw[i] = tau * acc;
            }
        }
        __syncthreads();
        double pv = 0;
        for (int i = lane; i < m; i += 64) pv += pw[i] * v[i];
        for (int o = 32; o > 0; o >>= 1) pv += __shfl_xor(pv, o);
        const double K = 0.5 * tau * pv;
        // A22 -= v w^T + w v^T, w = p - K v
        for (int x = tid; x < m * m; x += EIGV_NT) {
            const int i = x % m, j = x / m;
            const double wi = pw[i] - K * v[i], wj = pw[j] - K * v[j];
            A[(k + 1 + i) + ld * (k + 1 + j)] -= v[i] * wj + wi * v[j];
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (n >= 2) {
            d[n - 2] = A[(n - 2) + ld * (n - 2)];
            e[n - 2] = A[(n - 1) + ld * (n - 2)];
        }
        d[n - 1] = A[(n - 1) + ld * (n - 1)];
    }
    __syncthreads();
    tridiag_bisect(n, d, e, w);
}

int small_sym_eig(int n, const double *a, int lda, double *w, double *z, int ldz, double *work, hipStream_t s)
{
    if (z == nullptr && n > 0) {
        allow_full_lds((const void *)k_sym_eigvals<true>);
        const size_t vec_bytes = (4 * (size_t)n + 16) * sizeof(double);
        const EigvalJob one{a, w, n, lda};
        if (n <= EIGV_LDS_N)
            hipLaunchKernelGGL(k_sym_eigvals<true>, dim3(1), dim3(EIGV_NT), (size_t)n * (n + 1) * 8 + vec_bytes, s,
                               nullptr, one, work, 0L);
        else
            hipLaunchKernelGGL(k_sym_eigvals<false>, dim3(1), dim3(EIGV_NT), vec_bytes, s, nullptr, one, work, 0L);
        HIP_TRY(hipGetLastError());
        return SCALIT_OK;
    }
    if (n < 1 || n > 1024) return SCALIT_EINVAL;
    // a sweep is a few barriers per rotation round, so small matrices run on
    // fewer threads, where a barrier costs less (the loops and the tree
    // reduction take any power of two)
    const int nt = n <= 32 ? 64 : (n <= 96 ? 256 : 1024);
    hipLaunchKernelGGL(k_small_sym_eig, dim3(1), dim3(nt), 0, s, n, a, lda, w, z, ldz, work);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

int small_sym_eig_batch(int njobs, const EigvecJob *jobs, DBuf<EigvecJob> &djobs, DBuf<double> &work,
                        hipStream_t s)
{
    if (njobs < 0) return SCALIT_EINVAL;
    if (njobs == 0) return SCALIT_OK;
    int nmax = 0;
    for (int j = 0; j < njobs; j++) {
        if (jobs[j].n < 1 || jobs[j].n > 1024 || jobs[j].lda < jobs[j].n || jobs[j].ldz < jobs[j].n || !jobs[j].a ||
            !jobs[j].w || !jobs[j].z)
            return SCALIT_EINVAL;
        nmax = std::max(nmax, jobs[j].n);
    }
    const long stride = (long)small_sym_eig_work(nmax);
    SC_TRY(djobs.alloc(njobs));
    SC_TRY(work.alloc((size_t)njobs * stride));
    HIP_TRY(hipMemcpyAsync(djobs.p, jobs, sizeof(EigvecJob) * njobs, hipMemcpyHostToDevice, s));
    const int nt = nmax <= 32 ? 64 : (nmax <= 96 ? 256 : 1024);
    hipLaunchKernelGGL(k_small_sym_eig_batch, dim3(njobs), dim3(nt), 0, s, djobs.p, work.p, stride);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

int small_sym_eigvals_batch(int njobs, const EigvalJob *jobs, DBuf<EigvalJob> &djobs, DBuf<double> &work,
                            hipStream_t s)
{
    if (njobs < 0) return SCALIT_EINVAL;
    if (njobs == 0) return SCALIT_OK;
    int nmax = 0;
    for (int j = 0; j < njobs; j++) {
        if (jobs[j].n < 1 || jobs[j].lda < jobs[j].n || !jobs[j].a || !jobs[j].w) return SCALIT_EINVAL;
        nmax = std::max(nmax, jobs[j].n);
    }
    allow_full_lds((const void *)k_sym_eigvals<true>);
    SC_TRY(djobs.alloc(njobs));
    HIP_TRY(hipMemcpyAsync(djobs.p, jobs, sizeof(EigvalJob) * njobs, hipMemcpyHostToDevice, s));
    const size_t vec_bytes = (4 * (size_t)nmax + 16) * sizeof(double);
    const EigvalJob none{nullptr, nullptr, 0, 0};
    if (nmax <= EIGV_LDS_N) {
        hipLaunchKernelGGL(k_sym_eigvals<true>, dim3(njobs), dim3(EIGV_NT), (size_t)nmax * (nmax + 1) * 8 + vec_bytes,
                           s, djobs.p, none, nullptr, 0L);
    } else {
        const long stride = (long)nmax * (nmax + 1);
        SC_TRY(work.alloc((size_t)njobs * stride));
        hipLaunchKernelGGL(k_sym_eigvals<false>, dim3(njobs), dim3(EIGV_NT), vec_bytes, s, djobs.p, none, work.p,
                           stride);
    }
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// Eigenvalues of symmetric tridiagonal matrices (the Lanczos DSTEV steps),
// one workgroup per job: the diagonal and off-diagonal are staged in LDS
// when they fit, then the same Sturm multisection as k_sym_eigvals.
constexpr int TRI_LDS_N = 9000;
__global__ void __launch_bounds__(EIGV_NT) k_tri_eigvals(const TriEigJob *__restrict__ jobs)
{
    extern __shared__ double sm_tri[];
    const TriEigJob jb = jobs[blockIdx.x];
    const int n = jb.n;
    if (n <= TRI_LDS_N) {
        double *d = sm_tri, *e = sm_tri + n;
        for (int i = threadIdx.x; i < n; i += EIGV_NT) {
            d[i] = jb.d[i];
            e[i] = i + 1 < n ? jb.e[i] : 0.0;
        }
        __syncthreads();
        tridiag_bisect(n, d, e, jb.w);
    } else {
        tridiag_bisect(n, jb.d, jb.e, jb.w);
    }
}

int tridiag_eigvals_batch(int njobs, const TriEigJob *jobs, DBuf<TriEigJob> &djobs, hipStream_t s)
{
    if (njobs < 0) return SCALIT_EINVAL;
    if (njobs == 0) return SCALIT_OK;
    int nmax = 0;
    for (int j = 0; j < njobs; j++) {
        if (jobs[j].n < 1 || !jobs[j].d || !jobs[j].e || !jobs[j].w) return SCALIT_EINVAL;
        nmax = std::max(nmax, jobs[j].n);
    }
    allow_full_lds((const void *)k_tri_eigvals);
    SC_TRY(djobs.alloc(njobs));
    HIP_TRY(hipMemcpyAsync(djobs.p, jobs, sizeof(TriEigJob) * njobs, hipMemcpyHostToDevice, s));
    const size_t lds = nmax <= TRI_LDS_N ? 2 * (size_t)nmax * sizeof(double) : 0;
    hipLaunchKernelGGL(k_tri_eigvals, dim3(njobs), dim3(EIGV_NT), lds, s, djobs.p);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

}  // namespace scalit
