// P*X on the GPU: the OSB preconditioner of src/iterate/module/osb.px.h:125-219,
//   Y = V (EIG0 - E0)^-1 V^T X,
// V^T applied for levels sF..1 and V for levels 1..sF with the per-block
// rotations of src/iterate/libiterate/hpx/hpx.f90:40-73 (VX / VTX).
//
// Level l >= 2 holds one n_l x n_l rotation per block of the outer indices;
// applying it is a mode product Y(k,i,b) = sum_j A_b(i,j) X(k,j,b) over the
// contiguous inner index k (NIN = n_1...n_{l-1} values).  k_outer_mode
// computes it on MFMA with one grid row per block: the block's rotation is
// staged in LDS, each wave takes 16-point tiles of k (coalesced runs along k)
// with all of j in registers, as the mode-3 pass of H*X does.
//
// Level 1 (the innermost, NIN = 1) holds one n1 x n1 rotation per block of
// all outer indices: VOSB level 1 is N*n1 values, the largest stream of the
// whole iteration.  V1^T, the divide and V1 are adjacent in the sequence, so
// one kernel applies all three per block and VOSB level 1 is read once per
// P*X for all vectors of a batch: on MFMA with the vectors as the 16 columns
// of a tile (batches of 8 and more), or one wave per vector on the VALU.
#include <algorithm>
#include <cstdlib>

#include "device_common.hpp"
#include "outer_mode.hpp"

namespace scalit {

// PX_KERNEL (osb.px.h:156-211) for one value
template <typename T>
__device__ __forceinline__ T px_kernel(int nosb, T y, T eig, T e0, T de, T beta)
{
    T ade = eig - e0;
    if (nosb == SCALIT_TOSBD1) {
        if (fabs(ade) <= de) return ade < (T)0 ? -y * ((T)1 / de) : y * ((T)1 / de);
        return y / ade;
    }
    if (nosb == SCALIT_TOSBD2) {
        T b1 = (T)1 + beta, al = -beta * de, t;
        if (fabs(ade) <= de) t = ade < (T)0 ? (-al + b1 * ade) : (al + b1 * ade);
        else t = ade;
        return y / t;
    }
    return y / ade;
}

// Y(k,i,b) = sum_j V_b(j,i) X(k,j,b)  (transpose, VTX)  or  V_b(i,j) (VX);
// one thread per output value, any n.
template <typename T, typename S, typename SO>
__global__ void k_rotate_generic(int nin, int nout, int blk, const T *__restrict__ vosb, bool transpose,
                                 const S *__restrict__ x, long ldx, SO *__restrict__ y, long ldy, const int *active)
{
    int v = blockIdx.y;
    if (!vec_active(active, v)) return;
    long n = (long)nin * nout * blk;
    const S *xv = x + (long)v * ldx;
    SO *yv = y + (long)v * ldy;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
        int k = (int)(idx % nin);
        long t = idx / nin;
        int i = (int)(t % nout);
        long b = t / nout;
        const T *vb = vosb + b * (long)nout * nout;
        const S *xb = xv + b * (long)nin * nout + k;
        T acc = 0;
        if (transpose)
            for (int j = 0; j < nout; j++) acc += vb[j + (long)nout * i] * ld_cvt<T, S>(xb, (long)nin * j);
        else
            for (int j = 0; j < nout; j++) acc += vb[i + (long)nout * j] * ld_cvt<T, S>(xb, (long)nin * j);
        st_cvt<T, SO>(yv, idx, acc);
    }
}

// Level 1 fused on the VALU, persistent: workgroup g of G (4 waves) takes
// blocks g, g+G, ...; for every vector of a block
//   z = V_b^T y_b ;  z /= kernel(EIG0 - E0) ;  y_b = V_b z.
// While one block is computed, the next block's rotation (up to 16 values per
// thread) and its first vectors (up to 512 values) are loaded into registers,
// so the VOSB stream never waits behind the arithmetic.
template <typename T, typename S, typename SO>
__global__ void __launch_bounds__(256) k_px_level1(int n1, int nblk, int nvec, const T *__restrict__ vosb,
                                                   const T *__restrict__ eig0, const T *__restrict__ shifts, T shift0,
                                                   int nosb, T de, T beta, const S *__restrict__ x, long ldx,
                                                   SO *__restrict__ y, long ldy, const int *active)
{
    extern __shared__ char smem_raw[];
    T *sV = (T *)smem_raw;              // n1 x (n1+1), V(i,j) at i + (n1+1) j
    const int ld = n1 + 1;
    constexpr int VCH = 8;              // vectors per pass
    constexpr int PV = 16;              // V values per thread carried to the next block
    constexpr int PY = 2;               // first-chunk values per thread carried to the next block
    T *sY = sV + (long)n1 * ld;         // n1 x VCH
    T *sZ = sY + n1 * VCH;
    const int tid = threadIdx.x;
    const int G = gridDim.x;
    const int nvv = n1 * n1;
    const int ny0 = n1 * min(VCH, nvec);   // values of the first vector chunk
    T pv[PV], py[PY];
    auto y_of = [&](int bb, int e) -> T {   // value e = i + n1 vv of the first chunk of block bb
        const int i = e % n1, vv = e / n1;
        return vec_active(active, vv) ? ld_cvt<T, S>(x + (long)vv * ldx + (long)bb * n1, i) : (T)0;
    };
    auto load_next = [&](int bb) {
        const T *vb = vosb + (long)bb * nvv;
#pragma unroll
        for (int u = 0; u < PV; u++) {
            const int e = tid + 256 * u;
            pv[u] = e < nvv ? vb[e] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < PY; u++) {
            const int e = tid + 256 * u;
            py[u] = e < ny0 ? y_of(bb, e) : (T)0;
        }
    };
    int b = blockIdx.x;
    if (b < nblk) load_next(b);
    for (; b < nblk; b += G) {
        __syncthreads();   // the previous block is done with sV, sY and sZ
#pragma unroll
        for (int u = 0; u < PV; u++) {
            const int e = tid + 256 * u;
            if (e < nvv) sV[(e % n1) + ld * (e / n1)] = pv[u];
        }
        for (int e = tid + 256 * PV; e < nvv; e += 256) sV[(e % n1) + ld * (e / n1)] = vosb[(long)b * nvv + e];
#pragma unroll
        for (int u = 0; u < PY; u++) {
            const int e = tid + 256 * u;
            if (e < ny0) sY[e] = py[u];
        }
        for (int e = tid + 256 * PY; e < ny0; e += 256) sY[e] = y_of(b, e);
        __syncthreads();
        if (b + G < nblk) load_next(b + G);   // in flight while this block is computed
        const T *eb = eig0 + (long)b * n1;
        for (int v0 = 0; v0 < nvec; v0 += VCH) {
            const int nv = min(VCH, nvec - v0);
            if (v0 > 0) {
                __syncthreads();
                for (int e = tid; e < n1 * nv; e += 256) {
                    int i = e % n1, vv = e / n1;
                    int v = v0 + vv;
                    sY[i + n1 * vv] = vec_active(active, v) ? ld_cvt<T, S>(x + (long)v * ldx + (long)b * n1, i) : (T)0;
                }
                __syncthreads();
            }
            for (int e = tid; e < n1 * nv; e += 256) {
                int i = e % n1, vv = e / n1;
                T acc = 0;
                for (int j = 0; j < n1; j++) acc += sV[j + ld * i] * sY[j + n1 * vv];
                int v = v0 + vv;
                T e0 = shifts ? shifts[v] : shift0;
                sZ[i + n1 * vv] = px_kernel<T>(nosb, acc, eb[i], e0, de, beta);
            }
            __syncthreads();
            for (int e = tid; e < n1 * nv; e += 256) {
                int i = e % n1, vv = e / n1;
                int v = v0 + vv;
                if (!vec_active(active, v)) continue;
                T acc = 0;
                for (int j = 0; j < n1; j++) acc += sV[i + ld * j] * sZ[j + n1 * vv];
                st_cvt<T, SO>(y + (long)v * ldy + (long)b * n1, i, acc);
            }
        }
    }
}

// Level 1 fused on MFMA, n1 <= 112: per block b (one workgroup of 4 waves),
// 16 vectors at a time as the columns of the tiles,
//   Z = V_b^T W (wave r: rows 16r..16r+15), Z /= kernel, Y = V_b Z.
// LDS: V_b (n1p x (n1p+1)), W / Y and Z ((n1p+1) x 16 each).
template <typename T, typename S, typename SO>
__global__ void __launch_bounds__(256) k_px_l1_mfma(int n1, int nvec, const T *__restrict__ vosb,
                                                    const T *__restrict__ eig0, const T *__restrict__ shifts,
                                                    T shift0, int nosb, T de, T beta, const S *__restrict__ x,
                                                    long ldx, SO *__restrict__ y, long ldy, const int *active)
{
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    extern __shared__ char smem_raw[];
    const int n1p = rup16(n1);
    const int ldv = n1p + 1, ldw = n1p + 1;
    T *sV = (T *)smem_raw;          // V(i,j) at i + ldv j
    T *sW = sV + (long)n1p * ldv;   // W(j,c) at j + ldw c; later Y(i,c)
    T *sZ = sW + 16 * ldw;          // Z(j,c)
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int nr = n1p / 16, nk = n1p / 4;
    const T *vb = vosb + (long)b * n1 * n1;
    const T *eb = eig0 + (long)b * n1;
    if (n1p != n1)
        for (int e = tid; e < n1p * n1p; e += 256)
            if (e % n1p >= n1 || e / n1p >= n1) sV[(e % n1p) + ldv * (e / n1p)] = (T)0;
    stage_lds<256, 16>(vb, n1 * n1, tid, [&](int e) -> T & { return sV[(e % n1) + ldv * (e / n1)]; });
    for (int v0 = 0; v0 < nvec; v0 += 16) {
        __syncthreads();   // the previous chunk's Y has been stored
        {
            // 16 vector columns, n1p / 16 <= 7 values per thread, all loads in flight
            constexpr int UW = 7;
            T r[UW];
#pragma unroll
            for (int u = 0; u < UW; u++) {
                const int e = tid + 256 * u;
                const int j = e % n1p, c = e / n1p, v = v0 + c;
                r[u] = (e < n1p * 16 && j < n1 && v < nvec && vec_active(active, v))
                           ? ld_cvt<T, S>(x + (long)v * ldx + (long)b * n1, j)
                           : (T)0;
            }
#pragma unroll
            for (int u = 0; u < UW; u++) {
                const int e = tid + 256 * u;
                if (e < n1p * 16) sW[(e % n1p) + ldw * (e / n1p)] = r[u];
            }
        }
        __syncthreads();
        // Z = V^T W, divided by the kernel; zero outside the block and the batch
        const int vc = v0 + lr;
        const bool vin = vc < nvec && vec_active(active, vc);
        const T e0 = (shifts && vc < nvec) ? shifts[vc] : shift0;
        for (int r = wave; r < nr; r += 4) {
            acc_t acc = acc_t{0, 0, 0, 0};
            for (int s = 0; s < nk; s += 4) {
                T a[4], bb[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = 4 * (s + u) + lk;
                    a[u] = sV[k + ldv * (16 * r + lr)];   // A(i,k) = V(k,i)
                    bb[u] = sW[k + ldw * lr];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) acc = M::run(a[u], bb[u], acc);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = 16 * r + M::row(lk, q);
                sZ[i + ldw * lr] = (vin && i < n1) ? px_kernel<T>(nosb, acc[q], eb[i], e0, de, beta) : (T)0;
            }
        }
        __syncthreads();
        // Y = V Z into sW
        for (int r = wave; r < nr; r += 4) {
            acc_t acc = acc_t{0, 0, 0, 0};
            for (int s = 0; s < nk; s += 4) {
                T a[4], bb[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = 4 * (s + u) + lk;
                    a[u] = sV[(16 * r + lr) + ldv * k];   // A(i,k) = V(i,k)
                    bb[u] = sZ[k + ldw * lr];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) acc = M::run(a[u], bb[u], acc);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) sW[16 * r + M::row(lk, q) + ldw * lr] = acc[q];
        }
        __syncthreads();
        for (int e = tid; e < n1 * 16; e += 256) {
            const int i = e % n1, c = e / n1, v = v0 + c;
            if (v < nvec && vec_active(active, v)) st_cvt<T, SO>(y + (long)v * ldy + (long)b * n1, i, sW[i + ldw * c]);
        }
    }
}

static int pgrid(long n)
{
    long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    return (int)(g < 1 ? 1 : g);
}

static constexpr size_t PX_LDS_MAX = 160 * 1024;

template <typename T>
static size_t l1_mfma_lds(int n1)
{
    const size_t n1p = rup16(n1);
    return (n1p * (n1p + 1) + 2 * 16 * (n1p + 1)) * sizeof(T);
}
template <typename T>
static size_t l1_valu_lds(int n1)
{
    return ((size_t)n1 * (n1 + 1) + 2 * (size_t)n1 * 8) * sizeof(T);
}

// One rotation level, VTX (trans) or VX: MFMA when the block fits (n <= 128),
// else one thread per value.  vosbt holds the per-block transposes.
template <typename T, typename SI, typename SO>
static void rotate(const Level &L, const T *vosb, const T *vosbt, bool trans, int nvec, const SI *x, long ldx, SO *y,
                   long ldy, const int *active, hipStream_t s)
{
    const int nin = L.dim, n = L.n, blk = L.blk;
    if (n > 128 || blk > 65535 || (size_t)rup16(n) * rup16(n) * sizeof(T) > PX_LDS_MAX) {
        dim3 g(pgrid((long)nin * n * blk), nvec);
        hipLaunchKernelGGL((k_rotate_generic<T, SI, SO>), g, dim3(256), 0, s, nin, n, blk, vosb, trans, x, ldx, y, ldy,
                           active);
        return;
    }
    // the persistent outer-mode kernel, one grid row per block: 8 waves per
    // workgroup for batches, 4 for a single vector (fastest measured on the
    // 64x96x80 workload)
    const T *mats = trans ? vosbt : vosb;
    if (nvec >= 4)
        launch_outer_mode<T, SI, SO, 8, false>(nin, n, nvec, mats, nullptr, x, ldx, y, ldy, active, s, blk);
    else
        launch_outer_mode<T, SI, SO, 4, false>(nin, n, nvec, mats, nullptr, x, ldx, y, ldy, active, s, blk);
}

// Level 1: MFMA for batches of 8 and more, VALU otherwise.
template <typename T, typename SI, typename SO>
static void level1(const Level &L, const T *vosb, const T *eig0, const T *shifts, T shift0, int nosb, T de, T beta,
                   int nvec, const SI *x, long ldx, SO *y, long ldy, const int *active, hipStream_t s)
{
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)k_px_level1<T, SI, SO>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)PX_LDS_MAX);
        (void)hipFuncSetAttribute((const void *)k_px_l1_mfma<T, SI, SO>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)PX_LDS_MAX);
        attr = true;
    }
    const int n1 = L.n, nb = L.blk;
    if (nvec >= 8 && n1 <= 112 && l1_mfma_lds<T>(n1) <= PX_LDS_MAX)
        hipLaunchKernelGGL((k_px_l1_mfma<T, SI, SO>), dim3(nb), dim3(256), l1_mfma_lds<T>(n1), s, n1, nvec, vosb, eig0,
                           shifts, shift0, nosb, de, beta, x, ldx, y, ldy, active);
    else {
        // persistent: as many workgroups as are resident at once
        static int per_cu = 0;
        if (per_cu == 0) {
            int nbk = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, k_px_level1<T, SI, SO>, 256, l1_valu_lds<T>(n1)) !=
                    hipSuccess ||
                nbk < 1)
                nbk = 1;
            per_cu = nbk;
        }
        const int G = std::min(nb, g_num_cu * per_cu);
        hipLaunchKernelGGL((k_px_level1<T, SI, SO>), dim3(G), dim3(256), l1_valu_lds<T>(n1), s, n1, nb, nvec, vosb,
                           eig0, shifts, shift0, nosb, de, beta, x, ldx, y, ldy, active);
    }
}

template <typename T, typename S>
static int px_impl(scalit_ctx *c, int nosb, const T *const *vosb, const T *const *vosbt, const T *eig0, T *work,
                   const T *shifts, T shift0, T de, T beta, int nvec, const S *x, long ldx, S *y, long ldy,
                   const int *active, hipStream_t s)
{
    const long N = c->N;
    const int sF = c->sF;
    if (l1_valu_lds<T>(c->lv[0].n) > PX_LDS_MAX) return SCALIT_EINVAL;   // level-1 block too large for LDS
    if (sF == 1) {
        // a single level: input and output are the caller's vectors
        level1<T, S, S>(c->lv[0], vosb[0], eig0, shifts, shift0, nosb, de, beta, nvec, x, ldx, y, ldy, active, s);
        HIP_TRY(hipGetLastError());
        return SCALIT_OK;
    }
    // work holds two ping-pong buffers of nvec * N in the compute type
    T *w0 = work, *w1 = work + (long)nvec * N;
    // V^T for levels sF..2 (1-based); the first reads the input vectors
    rotate<T, S, T>(c->lv[sF - 1], vosb[sF - 1], vosbt[sF - 1], true, nvec, x, ldx, w0, N, active, s);
    T *cur = w0, *dst = w1;
    for (int l = sF - 2; l >= 1; l--) {
        rotate<T, T, T>(c->lv[l], vosb[l], vosbt[l], true, nvec, cur, N, dst, N, active, s);
        std::swap(cur, dst);
    }
    // level 1: V1^T, the kernel, V1
    level1<T, T, T>(c->lv[0], vosb[0], eig0, shifts, shift0, nosb, de, beta, nvec, cur, N, dst, N, active, s);
    std::swap(cur, dst);
    // V for levels 2..sF; the last writes the output vectors
    for (int l = 1; l < sF - 1; l++) {
        rotate<T, T, T>(c->lv[l], vosb[l], vosbt[l], false, nvec, cur, N, dst, N, active, s);
        std::swap(cur, dst);
    }
    rotate<T, T, S>(c->lv[sF - 1], vosb[sF - 1], vosbt[sF - 1], false, nvec, cur, N, y, ldy, active, s);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// dst_b = src_b^T for blk blocks of n x n
template <typename T>
__global__ void k_transpose_blocks(int n, long total, const T *__restrict__ src, T *__restrict__ dst)
{
    const long nn = (long)n * n;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long b = e / nn, r = e % nn;
        const int i = (int)(r % n), j = (int)(r / n);
        dst[b * nn + j + (long)n * i] = src[e];
    }
}

// The V^T rotations read per-block transposes of VOSB levels >= 2, built
// here on first use after the preconditioner changed (kernel buffers of a
// new generation: recorded graphs are re-recorded).
template <typename T>
static int ensure_transposes(scalit_ctx *c, DBuf<T> Level::*vb, DBuf<T> Level::*vt, hipStream_t s)
{
    for (int l = 1; l < c->sF; l++) {
        Level &L = c->lv[l];
        if ((L.*vt).p) continue;
        const long total = (long)L.blk * L.n * L.n;
        SC_TRY((L.*vt).alloc(total));
        hipLaunchKernelGGL(k_transpose_blocks<T>, dim3(pgrid(total)), dim3(256), 0, s, L.n, total, (L.*vb).p,
                           (L.*vt).p);
        HIP_TRY(hipGetLastError());
        c->gen++;
    }
    return SCALIT_OK;
}

int px_apply_f64(scalit_ctx *c, int nosb, const double *shifts_dev, double shift0, int nvec, const double *x,
                 long ldx, double *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    {
        double *old = c->work.p;
        SC_TRY(c->work.alloc((size_t)2 * nvec * c->N));
        if (c->work.p != old) c->gen++;
    }
    SC_TRY(ensure_transposes<double>(c, &Level::vosb, &Level::vosbt, s));
    const double *vosb[SC_FMAX], *vosbt[SC_FMAX];
    for (int l = 0; l < c->sF; l++) {
        vosb[l] = c->lv[l].vosb.p;
        vosbt[l] = c->lv[l].vosbt.p;
    }
    return px_impl<double, double>(c, nosb, vosb, vosbt, c->eig0.p, c->work.p, shifts_dev, shift0, c->de, c->beta,
                                   nvec, x, ldx, y, ldy, active, s);
}

int px_apply_f32(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx, float *y,
                 long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    SC_TRY(precond_to_f32(c));
    {
        float *old = c->work_f.p;
        SC_TRY(c->work_f.alloc((size_t)2 * nvec * c->N));
        if (c->work_f.p != old) c->gen++;
    }
    SC_TRY(ensure_transposes<float>(c, &Level::vosb_f, &Level::vosbt_f, s));
    const float *vosb[SC_FMAX], *vosbt[SC_FMAX];
    for (int l = 0; l < c->sF; l++) {
        vosb[l] = c->lv[l].vosb_f.p;
        vosbt[l] = c->lv[l].vosbt_f.p;
    }
    return px_impl<float, float>(c, SCALIT_TOSB, vosb, vosbt, c->eig0_f.p, c->work_f.p, shifts_dev, shift0, (float)c->de,
                                 (float)c->beta, nvec, x, ldx, y, ldy, active, s);
}

int px_apply_bf16(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const bf16_t *x, long ldx,
                  bf16_t *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    SC_TRY(precond_to_f32(c));
    {
        float *old = c->work_f.p;
        SC_TRY(c->work_f.alloc((size_t)2 * nvec * c->N));
        if (c->work_f.p != old) c->gen++;
    }
    SC_TRY(ensure_transposes<float>(c, &Level::vosb_f, &Level::vosbt_f, s));
    const float *vosb[SC_FMAX], *vosbt[SC_FMAX];
    for (int l = 0; l < c->sF; l++) {
        vosb[l] = c->lv[l].vosb_f.p;
        vosbt[l] = c->lv[l].vosbt_f.p;
    }
    return px_impl<float, bf16_t>(c, SCALIT_TOSB, vosb, vosbt, c->eig0_f.p, c->work_f.p, shifts_dev, shift0, (float)c->de,
                                  (float)c->beta, nvec, x, ldx, y, ldy, active, s);
}

}  // namespace scalit
