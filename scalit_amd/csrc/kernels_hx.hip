// H*X on the GPU: the MYHX operator of src/iterate/module/osb.hx.h:71-110,
//   Y_b = (RES - s_b) o X_b + sum_l (I x H0_l [x DEP_l] x I) X_b,
// with the per-level mode products of src/libutil/serial/base/matrix.f90:91-155
// (H1X / H2X) and the block loops of src/iterate/libiterate/hpx/hpx.f90:5-35.
//
// Two paths:
//  * generic: any number of levels and DEP pattern; one accumulate-kernel per
//    level (thread per output element).
//  * fused 3-level path for sF=3, sDEP=(F,F,T) -- r, R, theta with
//    DEP(r,R) x L^2_theta -- on FP64 / FP32 MFMA:
//      pass A, one workgroup per (theta point, vector) slab S(n1 x n2) staged
//      in LDS: Y = T1 S + S T2^T + (V - s) o S   (modes 1 and 2 + diagonal)
//      pass B, persistent waves over tiles of 16 grid points p=(r,R) across
//      all theta: Y(p,:) += DEP(p) * L2 X(p,:)    (mode 3)
//    Tiles are computed transposed (D = O^T) so that the 16 lanes of a
//    fragment column run along the contiguous r index: every global load and
//    store of X, Y and V moves 128-byte runs.  Each wave accumulates a block
//    of RB x CB output tiles at once (independent MFMA chains, fragment
//    reuse); the block shape and wave count of each pass are chosen per
//    problem and batch size by timing the candidates on the device (autotune).
#include <cstdlib>
#include <algorithm>
#include <cstring>

#include "device_common.hpp"
#include "outer_mode.hpp"

namespace scalit {

int g_force_generic = 0;
std::atomic<int> g_num_cu{256};   // compute units of the device, set when a context is created


// ---------------------------------------------------------------------------
// generic path
// ---------------------------------------------------------------------------

// Y_b = (RES - s_b) o X_b   (osb.hx.h:82-89)
template <typename T, typename S>
__global__ void k_diag_init(long n, const T *__restrict__ res, const T *__restrict__ shifts, T shift0,
                            const S *__restrict__ x, long ldx, S *__restrict__ y, long ldy, const int *active)
{
    int v = blockIdx.y;
    if (!vec_active(active, v)) return;
    T s = shifts ? shifts[v] : shift0;
    const S *xv = x + (long)v * ldx;
    S *yv = y + (long)v * ldy;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        st_cvt<T, S>(yv, i, (res[i] - s) * ld_cvt<T, S>(xv, i));
}

// Y(k, i, b) += sum_j H(i,j) [DEP(k)] X(k, j, b)     (H1X_XYZ / H1X_DEP)
// transpose: H(j,i) instead (the adjoint).
template <typename T, typename S>
__global__ void k_mode_generic(int nin, int nout, int blk, const T *__restrict__ h, const T *__restrict__ dep,
                               bool transpose, const S *__restrict__ x, long ldx, S *__restrict__ y, long ldy,
                               const int *active)
{
    int v = blockIdx.y;
    if (!vec_active(active, v)) return;
    long n = (long)nin * nout * blk;
    const S *xv = x + (long)v * ldx;
    S *yv = y + (long)v * ldy;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
        int k = (int)(idx % nin);
        long t = idx / nin;
        int i = (int)(t % nout);
        long b = t / nout;
        const S *xb = xv + b * (long)nin * nout + k;
        T acc = 0;
        if (transpose) {
            for (int j = 0; j < nout; j++) acc += h[j + (long)nout * i] * ld_cvt<T, S>(xb, (long)nin * j);
        } else {
            for (int j = 0; j < nout; j++) acc += h[i + (long)nout * j] * ld_cvt<T, S>(xb, (long)nin * j);
        }
        if (dep) acc *= dep[k];
        st_cvt<T, S>(yv, idx, ld_cvt<T, S>(yv, idx) + acc);
    }
}

static int grid_for(long n, int threads)
{
    long g = (n + threads - 1) / threads;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int)g;
}

// sNDVR = T outermost level (H3X, src/libutil/serial/base/matrix.f90:188-203):
//   Y(p, i) += sum_j A(p, i, j) X(p, j)    (trans: A(p, j, i)),
// A = OUTH with the point index p fastest, a dense n x n block per inner point.
// A workgroup takes 16 consecutive points (128-byte runs of A, X and Y) and VB
// vectors; X of its points for every j sits in LDS; thread (pp, ig) owns
// point p0 + pp and rows i = ig + 16 r.  A is streamed once per VB vectors.
template <typename T, typename S, int VB>
__global__ void __launch_bounds__(256) k_h3x(int nin, int n, int nvec, const T *__restrict__ A, int trans,
                                             const S *__restrict__ x, long ldx, S *__restrict__ y, long ldy,
                                             const int *active)
{
    extern __shared__ char smem_raw[];
    T *sX = (T *)smem_raw;   // X(p0 + q, j, v0 + vv) at (j VB + vv) 16 + q
    const int tid = threadIdx.x, pp = tid & 15, ig = tid >> 4;
    const long p0 = (long)blockIdx.x * 16;
    const int v0 = blockIdx.y * VB;
    const long p = p0 + pp;
    const bool pin = p < nin;
    for (int e = tid; e < 16 * n * VB; e += 256) {
        const int q = e & 15, rest = e >> 4, j = rest % n, vv = rest / n, v = v0 + vv;
        const long pq = p0 + q;
        sX[(j * VB + vv) * 16 + q] =
            (pq < nin && v < nvec && vec_active(active, v)) ? ld_cvt<T, S>(x + (long)v * ldx, pq + (long)nin * j) : (T)0;
    }
    __syncthreads();
    constexpr int RI = 8;   // n <= 128
    const int nri = (n + 15) / 16;
    const long si = trans ? (long)nin * n : (long)nin;   // A stride along i
    const long sj = trans ? (long)nin : (long)nin * n;   // A stride along j
    T acc[RI][VB];
#pragma unroll
    for (int r = 0; r < RI; r++)
#pragma unroll
        for (int vv = 0; vv < VB; vv++) acc[r][vv] = (T)0;
    const T *Ap = A + p;
#pragma unroll 2
    for (int j = 0; j < n; j++) {
        T xv[VB];
#pragma unroll
        for (int vv = 0; vv < VB; vv++) xv[vv] = sX[(j * VB + vv) * 16 + pp];
#pragma unroll
        for (int r = 0; r < RI; r++) {
            const int i = ig + 16 * r;
            if (r < nri && i < n && pin) {
                const T a = Ap[si * i + sj * j];
#pragma unroll
                for (int vv = 0; vv < VB; vv++) acc[r][vv] += a * xv[vv];
            }
        }
    }
    if (!pin) return;
#pragma unroll
    for (int vv = 0; vv < VB; vv++) {
        const int v = v0 + vv;
        if (v >= nvec || !vec_active(active, v)) continue;
        S *yv = y + (long)v * ldy + p;
#pragma unroll
        for (int r = 0; r < RI; r++) {
            const int i = ig + 16 * r;
            if (r < nri && i < n) st_cvt<T, S>(yv, (long)nin * i, ld_cvt<T, S>(yv, (long)nin * i) + acc[r][vv]);
        }
    }
}

template <typename T, typename S>
static void launch_h3x(int nin, int n, int nvec, const T *A, bool trans, const S *x, long ldx, S *y, long ldy,
                       const int *active, hipStream_t s)
{
    constexpr int VB = 4;
    const size_t lds = (size_t)16 * n * VB * sizeof(T);
    dim3 g((unsigned)((nin + 15) / 16), (unsigned)((nvec + VB - 1) / VB));
    hipLaunchKernelGGL((k_h3x<T, S, VB>), g, dim3(256), lds, s, nin, n, nvec, A, trans ? 1 : 0, x, ldx, y, ldy,
                       active);
}

// the OUTH layer in the compute type
template <typename T>
static const T *outh_of(const scalit_ctx *c)
{
    if constexpr (sizeof(T) == 8) return (const T *)c->outh.p;
    else return (const T *)c->outh_f.p;
}

template <typename T, typename S>
static int hx_generic(scalit_ctx *c, const T *res, const T *shifts, T shift0, int nvec, const S *x, long ldx, S *y,
                      long ldy, bool transpose, const int *active, hipStream_t s)
{
    const int th = 256;
    dim3 g(grid_for(c->N, th), nvec);
    hipLaunchKernelGGL((k_diag_init<T, S>), g, dim3(th), 0, s, c->N, res, shifts, shift0, x, ldx, y, ldy, active);
    for (int l = 0; l < c->sF; l++) {
        const Level &L = c->lv[l];
        if (c->ndvr && l == c->sF - 1) {
            launch_h3x<T, S>(L.dim, L.n, nvec, outh_of<T>(c), transpose, x, ldx, y, ldy, active, s);
            continue;
        }
        const T *h, *dep = nullptr;
        if constexpr (sizeof(T) == 8) {
            h = (const T *)(transpose ? L.h0t.p : L.h0.p);
            if (L.dep) dep = (const T *)L.depv.p;
        } else {
            h = (const T *)L.h0_f.p;   // the fp32 copies serve the forward product only
            if (L.dep) dep = (const T *)L.depv_f.p;
        }
        hipLaunchKernelGGL((k_mode_generic<T, S>), g, dim3(th), 0, s, L.dim, L.n, L.blk, h, dep, transpose, x, ldx,
                           y, ldy, active);
    }
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// ---------------------------------------------------------------------------
// fused 3-level path
// ---------------------------------------------------------------------------

// Slab s of the persistent pass-A schedule processed by workgroup g of G:
// workgroups are spread round-robin over the 8 XCDs, so XCD x gets the
// contiguous slab range [x S/8, (x+1) S/8) and the vectors sharing a theta
// point (and its slice of V) stay in one XCD's L2.
__device__ __forceinline__ int slab_of(int g, int k, int G, int S)
{
    if ((G & 7) == 0) {
        const int x = g & 7, loc = g >> 3, nloc = G >> 3;
        const int lo = (int)(((long)S * x) / 8), hi = (int)(((long)S * (x + 1)) / 8);
        const int s = lo + loc + k * nloc;
        return s < hi ? s : -1;
    }
    const int s = g + k * G;
    return s < S ? s : -1;
}

// Pass A.  Persistent grid of G workgroups of 64*WAVES threads; workgroup g
// walks its slabs (theta point i3, vector v), slab s = i3 * nvec + v.  LDS:
// S (n1p+1) x n2p, plus T1 n1p x n1p and T2 n2p x n2p when T_LDS, staged once
// per workgroup.  Each wave owns blocks of RB (i2) x CB (i1) output tiles.
// It writes Y = T1 S + S T2^T + (V - s) o S; the mode-3 pass then adds to Y.
// While one slab is computed, the next slab of X and the V values of each
// wave's first output block are loaded into registers, so neither the slab
// start nor the epilogue waits on HBM.
template <typename T, typename S, bool T_LDS, int WAVES, int RB, int CB>
__global__ void __launch_bounds__(64 * WAVES) k_hx_modes12(int n1, int n2, int n3, int nvec, const T *__restrict__ t1,
                                                          const T *__restrict__ t2, const T *__restrict__ res,
                                                          const T *__restrict__ shifts, T shift0,
                                                          const S *__restrict__ x, long ldx, S *__restrict__ y,
                                                          long ldy, const int *__restrict__ active)
{
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr int NT = 64 * WAVES;
    // slab values per thread prefetched in registers (the rest load at slab
    // start): 16 x 512 or 8 x 768 threads cover a 64 x 96 slab
    constexpr int PFMAX = WAVES >= 12 ? 8 : 16;
    extern __shared__ char smem_raw[];
    T *sm = (T *)smem_raw;
    const int n1p = rup16(n1), n2p = rup16(n2);
    const int ldS = n1p + 1;
    T *sS = sm;
    T *sT1 = sS + ldS * n2p;
    T *sT2 = sT1 + n1p * n1p;
    const long P = (long)n1 * n2;
    const int tid = threadIdx.x;
    const int nslab = n3 * nvec;
    const int G = gridDim.x, g = blockIdx.x;
    const int npf = (n1p * n2p + NT - 1) / NT;
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int nrt = n2p / 16, nct = n1p / 16;
    const int nrg = (nrt + RB - 1) / RB, ncg = (nct + CB - 1) / CB;

    if constexpr (T_LDS) {
        if (n1p != n1 || n2p != n2) {
            for (int e = tid; e < n1p * n1p; e += NT)
                if (e % n1p >= n1 || e / n1p >= n1) sT1[e] = (T)0;
            for (int e = tid; e < n2p * n2p; e += NT)
                if (e % n2p >= n2 || e / n2p >= n2) sT2[e] = (T)0;
        }
        const float r1 = 1.0f / (float)n1, r2 = 1.0f / (float)n2;
        stage_lds<NT, 16>(t1, n1 * n1, tid, [&](int e) -> T & {
            const int q = fdiv_small(e, r1);
            return sT1[(e - q * n1) + n1p * q];
        });
        stage_lds<NT, 16>(t2, n2 * n2, tid, [&](int e) -> T & {
            const int q = fdiv_small(e, r2);
            return sT2[(e - q * n2) + n2p * q];
        });
    }
    // first slab of this workgroup that is active
    int k = 0, s = slab_of(g, 0, G, nslab);
    while (s >= 0 && !vec_active(active, s % nvec)) s = slab_of(g, ++k, G, nslab);
    // The prefetches load unconditionally (indices clamped into the slab; the
    // values outside it are replaced by zeros when staged, or never stored), so
    // that the waits for them count exactly and never serialize the loads.
    // Slot j of a thread is element e = tid + j NT of the padded slab
    // (i1 = e % n1p, i2 = e / n1p); its index in the X slab (clamped into the
    // grid, low 16 bits of slot[j]), its LDS index (high 16 bits: the slab
    // holds at most 20,480 values) and two flags (bit j of `inside`: the value
    // lies in the grid, else a zero is staged; bit j of `owned`: the slot lies
    // in the padded slab, else nothing is stored) are the same for every slab,
    // so with 8 slots they are worked out once here.  With 16 (fewer waves)
    // a table would cost 18 registers the 8-wave shapes do not have: there
    // each slot is recomputed per slab, from an opaque copy of the thread
    // index so that the compiler cannot hoist the table back into registers.
    const float rp = 1.0f / (float)n1p;
    auto slot_calc = [&](int t0, int j, unsigned &pk, bool &in, bool &own) {
        const int e = t0 + j * NT;
        const int i2 = fdiv_small(e, rp), i1 = e - i2 * n1p;
        const unsigned g = (unsigned)(min(i1, n1 - 1) + n1 * min(i2, n2 - 1));
        pk = (g & 0xffffu) | ((unsigned)(i1 + ldS * i2) << 16);
        in = i1 < n1 && i2 < n2;
        own = j < npf && e < n1p * n2p;
    };
    constexpr bool TABLE = PFMAX <= 8;
    unsigned slot[TABLE ? PFMAX : 1];
    unsigned inside = 0, owned = 0;
    if constexpr (TABLE) {
#pragma unroll
        for (int j = 0; j < PFMAX; j++) {
            bool in, own;
            slot_calc(tid, j, slot[j], in, own);
            inside |= (unsigned)in << j;
            owned |= (unsigned)own << j;
        }
    }
    T pf[PFMAX];
    // (sl < 0, no next slab: the slots of slab 0 again)
    auto prefetch = [&](int sl) {
        const S *xs = sl >= 0 ? x + (long)(sl % nvec) * ldx + P * (sl / nvec) : x;
        const auto rx = buf_rsrc(xs, P * (long)sizeof(S));
        int t0 = tid;
        if constexpr (!TABLE) asm volatile("" : "+v"(t0));
#pragma unroll
        for (int j = 0; j < PFMAX; j++) {
            unsigned pk;
            if constexpr (TABLE) {
                pk = slot[j];
            } else {
                bool in, own;
                slot_calc(t0, j, pk, in, own);
            }
            pf[j] = buf_ld<T, S>(rx, (int)((pk & 0xffffu) * sizeof(S)), 0);
        }
    };
    // V of the wave's first output block (block index = wave)
    T pv[RB][CB][4];
    auto prefetch_v = [&](int sl) {
        const bool live = sl >= 0;
        const T *rsn = res + (live ? P * (sl / nvec) : 0);
        const int ob = min(wave, nrg * ncg - 1);
        const int rg = ob / ncg, cg = ob % ncg;
#pragma unroll
        for (int r = 0; r < RB; r++)
#pragma unroll
            for (int c = 0; c < CB; c++) {
                const int i1 = live ? min((cg * CB + c) * 16 + lr, n1 - 1) : 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i2 = live ? min((rg * RB + r) * 16 + M::row(lk, q), n2 - 1) : 0;
                    pv[r][c][q] = rsn[i1 + (long)n1 * i2];
                }
            }
    };
    prefetch(s);
    prefetch_v(s);

    while (s >= 0) {
        // the prefetched slab into LDS (padding zero)
        {
            int t0 = tid;
            if constexpr (!TABLE) asm volatile("" : "+v"(t0));
#pragma unroll
            for (int j = 0; j < PFMAX; j++) {
                unsigned pk;
                bool in, own;
                if constexpr (TABLE) {
                    pk = slot[j];
                    in = (inside >> j) & 1u;
                    own = (owned >> j) & 1u;
                } else {
                    slot_calc(t0, j, pk, in, own);
                }
                if (own) sS[pk >> 16] = in ? pf[j] : (T)0;
            }
        }
        if (npf > PFMAX) {
            const S *xs = x + (long)(s % nvec) * ldx + P * (s / nvec);
            for (int e = tid + PFMAX * NT; e < n1p * n2p; e += NT) {
                const int i1 = e % n1p, i2 = e / n1p;
                sS[i1 + ldS * i2] = (i1 < n1 && i2 < n2) ? ld_cvt<T, S>(xs, i1 + (long)n1 * i2) : (T)0;
            }
        }
        __syncthreads();
        const int v = s % nvec, i3 = s / nvec;
        int sn = slab_of(g, ++k, G, nslab);
        while (sn >= 0 && !vec_active(active, sn % nvec)) sn = slab_of(g, ++k, G, nslab);
        prefetch(sn);
        S *ys = y + (long)v * ldy + P * i3;
        const auto ry = buf_rsrc(ys, P * (long)sizeof(S));
        const T *rs = res + P * i3;
        const T sh = shifts ? shifts[v] : shift0;
        // block ob of RB x CB output tiles, V already in pv
        auto block = [&](int ob, bool own) {
            const int rg = ob / ncg, cg = ob % ncg;
            acc_t acc[RB][CB];
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int c = 0; c < CB; c++) acc[r][c] = acc_t{0, 0, 0, 0};
            // Tiles past the edge of an RB x CB block read a clamped (valid) tile;
            // their results are never stored.  No predicate guards a load, and the
            // k loops go in groups of 4 steps (n1p, n2p are multiples of 16) so the
            // fragment reads of a group are all in flight before its MFMAs.
            int rti[RB], cti[CB];
#pragma unroll
            for (int r = 0; r < RB; r++) rti[r] = min(rg * RB + r, nrt - 1) * 16 + lr;
#pragma unroll
            for (int c = 0; c < CB; c++) cti[c] = min(cg * CB + c, nct - 1) * 16 + lr;
            // mode 1: D(i2,i1) += sum_k S(k,i2) T1(i1,k)
            for (int kb = 0; kb < n1p; kb += 16) {
                T a[4][RB], b[4][CB];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int kk = kb + 4 * u + lk;
#pragma unroll
                    for (int r = 0; r < RB; r++) a[u][r] = sS[kk + ldS * rti[r]];
#pragma unroll
                    for (int c = 0; c < CB; c++) {
                        if constexpr (T_LDS) b[u][c] = sT1[cti[c] + n1p * kk];
                        else b[u][c] = (cti[c] < n1 && kk < n1) ? t1[cti[c] + (long)n1 * kk] : (T)0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < RB; r++)
#pragma unroll
                        for (int c = 0; c < CB; c++) acc[r][c] = M::run(a[u][r], b[u][c], acc[r][c]);
            }
            // mode 2: D(i2,i1) += sum_k T2(i2,k) S(i1,k)
            for (int kb = 0; kb < n2p; kb += 16) {
                T a[4][RB], b[4][CB];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int kk = kb + 4 * u + lk;
#pragma unroll
                    for (int r = 0; r < RB; r++) {
                        if constexpr (T_LDS) a[u][r] = sT2[rti[r] + n2p * kk];
                        else a[u][r] = (rti[r] < n2 && kk < n2) ? t2[rti[r] + (long)n2 * kk] : (T)0;
                    }
#pragma unroll
                    for (int c = 0; c < CB; c++) b[u][c] = sS[cti[c] + ldS * kk];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < RB; r++)
#pragma unroll
                        for (int c = 0; c < CB; c++) acc[r][c] = M::run(a[u][r], b[u][c], acc[r][c]);
            }
            // epilogue: + (V - s) S, store (a lane outside the slab, or a wave
            // past the last block, stores to an offset past the resource,
            // which the hardware drops: the store count is the same on every path)
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int c = 0; c < CB; c++) {
                    const int i1 = (cg * CB + c) * 16 + lr;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int i2 = (rg * RB + r) * 16 + M::row(lk, q);
                        const bool in = own && i1 < n1 && i2 < n2;
                        const long o = i1 + (long)n1 * i2;
                        const T sv = sS[min(i1, n1p - 1) + ldS * min(i2, n2p - 1)];
                        buf_st<T, S>(ry, in ? (int)(o * (long)sizeof(S)) : (int)0x80000000, 0,
                                     acc[r][c][q] + (pv[r][c][q] - sh) * sv);
                    }
                }
        };
        // the first block of every wave with the prefetched V (a wave past the
        // last block computes the last one and stores nothing), then the rest,
        // each loading its V first
        block(min(wave, nrg * ncg - 1), wave < nrg * ncg);
        for (int ob = wave + WAVES; ob < nrg * ncg; ob += WAVES) {
            const int rg = ob / ncg, cg = ob % ncg;
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int c = 0; c < CB; c++) {
                    const int i1 = min((cg * CB + c) * 16 + lr, n1 - 1);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int i2 = min((rg * RB + r) * 16 + M::row(lk, q), n2 - 1);
                        pv[r][c][q] = rs[i1 + (long)n1 * i2];
                    }
                }
            block(ob, true);
        }
        prefetch_v(sn);
        __syncthreads();   // every wave is done with sS before the next slab lands
        s = sn;
    }
}

// Pass B (mode 3 x DEP, added to Y) is k_outer_mode of outer_mode.hpp.

static constexpr size_t LDS_MAX = 160 * 1024;

template <typename T>
static size_t modes12_lds(int n1, int n2, bool t_lds)
{
    int n1p = rup16(n1), n2p = rup16(n2);
    size_t s = (size_t)(n1p + 1) * n2p;
    if (t_lds) s += (size_t)n1p * n1p + (size_t)n2p * n2p;
    return s * sizeof(T);
}

template <typename T>
bool fast3_supported(const scalit_ctx *c)
{
    if (c->sF != 3 || c->lv[0].dep || c->lv[1].dep || !(c->lv[2].dep || c->ndvr)) return false;
    if (c->lv[2].n > 128) return false;   // pass B keeps a theta column in registers, H3X 8 row tiles
    if (c->ndvr && (size_t)16 * c->lv[2].n * 4 * sizeof(T) > LDS_MAX) return false;
    return modes12_lds<T>(c->lv[0].n, c->lv[1].n, false) <= LDS_MAX;
}

// ---------------------------------------------------------------- variants
struct VarA {
    int waves, rb, cb;
};
static const VarA kVarA[] = {{4, 1, 1}, {4, 3, 2}, {8, 3, 1}, {8, 2, 2}, {8, 1, 2}, {4, 2, 2}, {12, 2, 1}, {12, 1, 2}};
constexpr int NVA = sizeof(kVarA) / sizeof(kVarA[0]);
constexpr int NVB = 3;   // pass B: 4, 8 or 12 waves per workgroup

template <typename T, typename S, bool TL, int W, int RB, int CB>
static void launch_a1(dim3 g, size_t lds, hipStream_t s, int n1, int n2, int n3, int nvec, const T *t1, const T *t2,
                      const T *res, const T *sh, T sh0, const S *x, long ldx, S *y, long ldy, const int *act)
{
    allow_full_lds((const void *)k_hx_modes12<T, S, TL, W, RB, CB>);
    // persistent grid: a multiple of the 8 XCDs, at most what fits on the CUs
    const int per_cu = std::max(1, std::min((int)(LDS_MAX / std::max<size_t>(lds, 1)), 32 / W));
    int G = std::min((int)g.x, g_num_cu * per_cu);
    if (G >= 8) G &= ~7;
    hipLaunchKernelGGL((k_hx_modes12<T, S, TL, W, RB, CB>), dim3(G), dim3(64 * W), lds, s, n1, n2, n3, nvec, t1,
                       t2, res, sh, sh0, x, ldx, y, ldy, act);
}

template <typename T, typename S, bool TL>
static void launch_a(int va, dim3 g, size_t lds, hipStream_t s, int n1, int n2, int n3, int nvec, const T *t1,
                     const T *t2, const T *res, const T *sh, T sh0, const S *x, long ldx, S *y, long ldy,
                     const int *act)
{
#define LA(W, RB, CB) \
    launch_a1<T, S, TL, W, RB, CB>(g, lds, s, n1, n2, n3, nvec, t1, t2, res, sh, sh0, x, ldx, y, ldy, act)
    switch (va) {
    case 0: LA(4, 1, 1); break;
    case 1: LA(4, 3, 2); break;
    case 2: LA(8, 3, 1); break;
    case 3: LA(8, 2, 2); break;
    case 4: LA(8, 1, 2); break;
    case 5: LA(4, 2, 2); break;
    case 6: LA(12, 2, 1); break;
    default: LA(12, 1, 2); break;
    }
#undef LA
}

template <typename T, typename S>
static void launch_b(int vb, int n1, int n2, int n3, int nvec, const T *l2, const T *dep, const S *x, long ldx, S *y,
                     long ldy, const int *act, hipStream_t s)
{
    const long P = (long)n1 * n2;
    if (vb == 0) launch_outer_mode<T, S, S, 4, true>(P, n3, nvec, l2, dep, x, ldx, y, ldy, act, s);
    else if (vb == 1) launch_outer_mode<T, S, S, 8, true>(P, n3, nvec, l2, dep, x, ldx, y, ldy, act, s);
    else launch_outer_mode<T, S, S, 12, true>(P, n3, nvec, l2, dep, x, ldx, y, ldy, act, s);
}

template <typename T, typename S>
static void pass_a(scalit_ctx *c, int va, const T *t1, const T *t2, const T *res, const T *sh, T sh0, int nvec,
                   const S *x, long ldx, S *y, long ldy, const int *act, hipStream_t s)
{
    const int n1 = c->lv[0].n, n2 = c->lv[1].n, n3 = c->lv[2].n;
    const bool tl = modes12_lds<T>(n1, n2, true) <= LDS_MAX;
    const size_t lds = modes12_lds<T>(n1, n2, tl);
    dim3 ga(n3 * nvec);
    if (tl) launch_a<T, S, true>(va, ga, lds, s, n1, n2, n3, nvec, t1, t2, res, sh, sh0, x, ldx, y, ldy, act);
    else launch_a<T, S, false>(va, ga, lds, s, n1, n2, n3, nvec, t1, t2, res, sh, sh0, x, ldx, y, ldy, act);
}

// ---------------------------------------------------------------- autotune
// Per (precision, batch size) the fastest pass-A and pass-B variants, found by
// timing every candidate on the caller's buffers (outputs overwritten, the
// real call follows).  Skipped while a stream is being captured into a graph.
static int env_variant(const char *name)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : -1;
}

template <typename T, typename S>
static TuneEntry *tune(scalit_ctx *c, int prec, const T *t1, const T *t2, const T *l2, const T *dep, const T *res,
                       const T *sh, T sh0, int nvec, const S *x, long ldx, S *y, long ldy, hipStream_t s)
{
    for (auto &e : c->hx_tune)
        if (e.prec == prec && e.nvec == nvec) return &e;
    int fa = env_variant("SCALIT_HX_VARIANT_A"), fb = env_variant("SCALIT_HX_VARIANT_B");
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(s, &st);
    const int n1 = c->lv[0].n, n2 = c->lv[1].n, n3 = c->lv[2].n;
    TuneEntry e{prec, nvec, fa >= 0 && fa < NVA ? fa : 2, fb >= 0 && fb < NVB ? fb : 0};
    if (st != hipStreamCaptureStatusNone) return nullptr;   // use the defaults, decide later
    if (fa < 0 || fb < 0) {
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return nullptr;
        float best = 1e30f;
        for (int va = 0; va < NVA && fa < 0; va++) {
            for (int w = 0; w < 2; w++) pass_a<T, S>(c, va, t1, t2, res, sh, sh0, nvec, x, ldx, y, ldy, nullptr, s);
            (void)hipEventRecord(e0, s);
            for (int w = 0; w < 3; w++) pass_a<T, S>(c, va, t1, t2, res, sh, sh0, nvec, x, ldx, y, ldy, nullptr, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (hipGetLastError() == hipSuccess && ms < best) {
                best = ms;
                e.va = va;
            }
        }
        best = 1e30f;
        for (int vb = 0; vb < NVB && fb < 0 && l2; vb++) {
            for (int w = 0; w < 2; w++) launch_b<T, S>(vb, n1, n2, n3, nvec, l2, dep, x, ldx, y, ldy, nullptr, s);
            (void)hipEventRecord(e0, s);
            for (int w = 0; w < 3; w++) launch_b<T, S>(vb, n1, n2, n3, nvec, l2, dep, x, ldx, y, ldy, nullptr, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (hipGetLastError() == hipSuccess && ms < best) {
                best = ms;
                e.vb = vb;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    c->hx_tune.push_back(e);
    return &c->hx_tune.back();
}

// outh != nullptr: level 3 is the sNDVR = T layer (H3X, transposed for the
// adjoint when trans_outh) instead of DEP x L2.
template <typename T, typename S>
static int hx_fast3(scalit_ctx *c, int prec, const T *t1, const T *t2, const T *l2, const T *dep, const T *outh,
                    bool trans_outh, const T *res, const T *shifts, T shift0, int nvec, const S *x, long ldx, S *y,
                    long ldy, const int *active, hipStream_t s)
{
    const int n1 = c->lv[0].n, n2 = c->lv[1].n, n3 = c->lv[2].n;
    TuneEntry *te = tune<T, S>(c, prec, t1, t2, l2, dep, res, shifts, shift0, nvec, x, ldx, y, ldy, s);
    const int va = te ? te->va : 2, vb = te ? te->vb : 0;
    // modes 1+2 and the diagonal write Y, then mode 3 adds to it
    pass_a<T, S>(c, va, t1, t2, res, shifts, shift0, nvec, x, ldx, y, ldy, active, s);
    if (outh) launch_h3x<T, S>(n1 * n2, n3, nvec, outh, trans_outh, x, ldx, y, ldy, active, s);
    else launch_b<T, S>(vb, n1, n2, n3, nvec, l2, dep, x, ldx, y, ldy, active, s);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

int hx_apply_f64(scalit_ctx *c, const double *shifts_dev, double shift0, int nvec, const double *x, long ldx,
                 double *y, long ldy, bool transpose, const int *active, hipStream_t s)
{
    if (!c->fast3 || g_force_generic || !fast3_supported<double>(c))
        return hx_generic<double, double>(c, c->res.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, transpose, active,
                                          s);
    const Level *L = c->lv;
    const double *t1 = transpose ? L[0].h0t.p : L[0].h0.p;
    const double *t2 = transpose ? L[1].h0t.p : L[1].h0.p;
    const double *l2 = c->ndvr ? nullptr : (transpose ? L[2].h0t.p : L[2].h0.p);
    return hx_fast3<double, double>(c, SCALIT_FP64, t1, t2, l2, c->ndvr ? nullptr : L[2].depv.p,
                                    c->ndvr ? c->outh.p : nullptr, transpose, c->res.p, shifts_dev, shift0, nvec, x,
                                    ldx, y, ldy, active, s);
}

int hx_apply_f32(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx, float *y,
                 long ldy, const int *active, hipStream_t s)
{
    if (!c->fast3 || g_force_generic || !fast3_supported<float>(c))
        return hx_generic<float, float>(c, c->res_f.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, false, active, s);
    const Level *L = c->lv;
    return hx_fast3<float, float>(c, SCALIT_FP32, L[0].h0_f.p, L[1].h0_f.p, c->ndvr ? nullptr : L[2].h0_f.p,
                                  c->ndvr ? nullptr : L[2].depv_f.p, c->ndvr ? c->outh_f.p : nullptr, false,
                                  c->res_f.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, active, s);
}

int hx_apply_bf16(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const bf16_t *x, long ldx,
                  bf16_t *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->fast3 || g_force_generic || !fast3_supported<float>(c))
        return hx_generic<float, bf16_t>(c, c->res_f.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, false, active, s);
    const Level *L = c->lv;
    return hx_fast3<float, bf16_t>(c, SCALIT_BF16, L[0].h0_f.p, L[1].h0_f.p, c->ndvr ? nullptr : L[2].h0_f.p,
                                   c->ndvr ? nullptr : L[2].depv_f.p, c->ndvr ? c->outh_f.p : nullptr, false,
                                   c->res_f.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, active, s);
}

bool fast3_ok(const scalit_ctx *c) { return fast3_supported<double>(c); }

// Time the pass variants once for this batch size in the given precision
// (scratch vectors allocated here), so that graph recording finds a decision.
int hx_prepare(scalit_ctx *c, int prec, int nvec, hipStream_t s)
{
    if (!c->fast3 || g_force_generic) return SCALIT_OK;
    for (auto &e : c->hx_tune)
        if (e.prec == prec && e.nvec == nvec) return SCALIT_OK;
    const size_t es = prec == SCALIT_FP64 ? 8 : (prec == SCALIT_FP32 ? 4 : 2);
    void *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, es * nvec * c->N));
    if (hipMalloc(&b, es * nvec * c->N) != hipSuccess) {
        (void)hipFree(a);
        return SCALIT_ENOMEM;
    }
    (void)hipMemsetAsync(a, 0, es * nvec * c->N, s);
    int rc;
    if (prec == SCALIT_FP64) rc = hx_apply_f64(c, nullptr, 0.0, nvec, (double *)a, c->N, (double *)b, c->N, false,
                                               nullptr, s);
    else if (prec == SCALIT_FP32) rc = hx_apply_f32(c, nullptr, 0.0f, nvec, (float *)a, c->N, (float *)b, c->N,
                                                    nullptr, s);
    else rc = hx_apply_bf16(c, nullptr, 0.0f, nvec, (bf16_t *)a, c->N, (bf16_t *)b, c->N, nullptr, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int hx_variants(const scalit_ctx *c, int prec, int nvec, int *va, int *vb)
{
    for (auto &e : c->hx_tune)
        if (e.prec == prec && e.nvec == nvec) {
            *va = e.va;
            *vb = e.vb;
            return SCALIT_OK;
        }
    return SCALIT_ESTATE;
}

// Lane-layout probe of the 16x16x4 MFMA builtins: one wave computes
// D = A B with A(i,0) = 100 (i+1), A(i,1) = 1, B(0,j) = 1, B(1,j) = j+1 (other
// k zero) from the fragment mapping the kernels use, so D(i,j) = 100(i+1) +
// (j+1) names its own row and column; out[4 lane + q] = accumulator q.
template <typename T>
__global__ void __launch_bounds__(64) k_mfma_probe(double *out)
{
    using M = Mfma<T>;
    const int lane = threadIdx.x, r = lane & 15, k = lane >> 4;
    T a = (k == 0) ? (T)(100.0 * (r + 1)) : (k == 1 ? (T)1 : (T)0);
    T b = (k == 0) ? (T)1 : (k == 1 ? (T)(r + 1) : (T)0);
    typename M::acc_t acc = {0, 0, 0, 0};
    acc = M::run(a, b, acc);
    for (int q = 0; q < 4; q++) out[4 * lane + q] = (double)acc[q];
}

// MFMA throughput probe: every wave runs `iters` rounds of 8 independent
// 16x16x4 accumulation chains from registers (no memory traffic); the host
// divides 8 * 2048 * iters * waves FLOP by the kernel time.
template <typename T>
__global__ void __launch_bounds__(256) k_mfma_peak(int iters, double *sink)
{
    using M = Mfma<T>;
    typename M::acc_t acc[8];
    T a = (T)(1.0 + 1e-3 * threadIdx.x), b = (T)(1.0 - 1e-3 * threadIdx.x);
#pragma unroll
    for (int i = 0; i < 8; i++) acc[i] = typename M::acc_t{0, 0, 0, 0};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] = M::run(a, b, acc[i]);
    }
    double s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s += (double)(acc[i][0] + acc[i][3]);
    if (s == 12345.678) sink[0] = s;   // keeps the chains live
}

int mfma_peak(int prec, int iters, int blocks, float *ms_out)
{
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double)));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int rep = 0; rep < 2; rep++) {
        (void)hipEventRecord(e0, 0);
        if (prec == SCALIT_FP32) hipLaunchKernelGGL(k_mfma_peak<float>, dim3(blocks), dim3(256), 0, 0, iters, d);
        else hipLaunchKernelGGL(k_mfma_peak<double>, dim3(blocks), dim3(256), 0, 0, iters, d);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
    }
    (void)hipEventElapsedTime(ms_out, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

int mfma_probe(int prec, double *out_host)
{
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, 256 * sizeof(double)));
    if (prec == SCALIT_FP32) hipLaunchKernelGGL(k_mfma_probe<float>, dim3(1), dim3(64), 0, 0, d);
    else hipLaunchKernelGGL(k_mfma_probe<double>, dim3(1), dim3(64), 0, 0, d);
    hipError_t e = hipMemcpy(out_host, d, 256 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIP_TRY(e);
    return SCALIT_OK;
}

// One pass of the fused FP64 path alone with the tuned variant (pass 1: modes
// 1+2+diagonal, pass 2: mode 3), for per-kernel timing.
int hx_pass_f64(scalit_ctx *c, int pass, int nvec, const double *x, long ldx, double *y, long ldy, hipStream_t s)
{
    if (!c->fast3 || c->ndvr || !fast3_supported<double>(c)) return SCALIT_EINVAL;
    const Level *L = c->lv;
    const int n1 = L[0].n, n2 = L[1].n, n3 = L[2].n;
    TuneEntry *te = tune<double, double>(c, SCALIT_FP64, L[0].h0.p, L[1].h0.p, L[2].h0.p, L[2].depv.p, c->res.p,
                                         nullptr, 0.0, nvec, x, ldx, y, ldy, s);
    const int va = te ? te->va : 2, vb = te ? te->vb : 0;
    if (pass == 1)
        pass_a<double, double>(c, va, L[0].h0.p, L[1].h0.p, c->res.p, nullptr, 0.0, nvec, x, ldx, y, ldy, nullptr, s);
    else if (pass == 2)
        launch_b<double, double>(vb, n1, n2, n3, nvec, L[2].h0.p, L[2].depv.p, x, ldx, y, ldy, nullptr, s);
    else
        return SCALIT_EINVAL;
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

}  // namespace scalit
