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
#include <type_traits>

#include "device_common.hpp"
#include "outer_mode.hpp"

namespace scalit {

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

// Level 1 fused, persistent: workgroup g of G (4 waves) takes blocks g, g+G,
// ...; for every vector of a block
//   z = V_b^T x_b ;  z /= kernel(EIG0 - E0) ;  y_b = V_b z,
// on MFMA with 16 vectors as the columns of the tiles (MF: wave r computes
// rows 16r..16r+15) or on the VALU.
//
// The VOSB stream is read at the rate of a plain copy only if nothing on the
// critical path of a block waits for memory, so: the next block's rotation,
// EIG0 values and first vectors are loaded into registers right after this
// block is staged, with unconditional loads (indices clamped into the block,
// the extra values dropped when staged), and the computation reads LDS only
// (shifts and activity of the vectors are staged once per chunk).  A global
// load in the computation would wait for the whole prefetch (the memory
// counter drains in order).  On the VALU, small batches split every dot
// product over up to 8 threads (partial sums through LDS), so that a block
// costs a few short passes rather than two n1-long chains on one wave.
//
// LDS: V_b (ns x (ns+1), V(i,j) at i + (ns+1) j, zero padded for MF with
// ns = n1 rounded up to 16, else ns = n1), W and Z ((ns+1) x WC each; W holds
// Y after the MF product), partial sums (256), EIG0 of the block (ns), and
// the shifts and activity of the first chunk of vectors and of the current
// later one (2 WC each).  BIG (n1 > 64): the part of a block beyond the
// registers is loaded while it is staged.
template <typename T, typename S, typename SO, bool MF, bool BIG>
__global__ void __launch_bounds__(256) k_px_l1(int n1, int nblk, int nvec, const T *__restrict__ vosb,
                                               const T *__restrict__ eig0, const T *__restrict__ shifts, T shift0,
                                               int nosb, T de, T beta, const S *__restrict__ x, long ldx,
                                               SO *__restrict__ y, long ldy, const int *active)
{
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr int WC = MF ? 16 : 8;   // vectors per chunk
    constexpr int PV = 16;            // rotation values per thread
    constexpr int PY = MF ? 4 : 2;    // first-chunk values per thread
    extern __shared__ char smem_raw[];
    const int ns = MF ? rup16(n1) : n1;
    const int ldv = ns + 1;
    T *sV = (T *)smem_raw;             // V(i,j) at i + ldv j
    T *sW = sV + (long)ns * ldv;       // W(j,c) at j + ldv c
    T *sZ = sW + (long)WC * ldv;       // Z(j,c) at j + ldv c
    T *sP = sZ + (long)WC * ldv;       // partial sums
    T *sE = sP + 256;                  // EIG0(i) of the block
    T *sSh = sE + ns;                  // shifts: first chunk, current later chunk
    int *sAct = (int *)(sSh + 2 * WC);   // activity (0 / 1): first chunk, current later chunk
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int G = gridDim.x;
    const int nvv = n1 * n1;
    const int nw = ns * WC;   // values of one chunk of vectors
    auto vec_info = [&](int v0, int slot) {
        if (tid < WC) {
            const int v = v0 + tid;
            sSh[slot * WC + tid] = (shifts && v < nvec) ? shifts[v] : shift0;
            sAct[slot * WC + tid] = (v < nvec && vec_active(active, v)) ? 1 : 0;
        }
    };
    vec_info(0, 0);
    if (MF && ns != n1)   // the padding is never overwritten
        for (int e = tid; e < ns * ns; e += 256)
            if (e % ns >= n1 || e / ns >= n1) sV[(e % ns) + ldv * (e / ns)] = (T)0;
    // register u holds element tid + 256 u of the block's rotation.  A block
    // past the last loads one value of the last block on every lane, so the
    // loads are the same on every path.
    T pv[PV], py[PY], pe;
    auto load = [&](int bb) {
        const bool live = bb < nblk;
        const long b0 = min(bb, nblk - 1);
        const T *vb = vosb + b0 * nvv;
#pragma unroll
        for (int u = 0; u < PV; u++) pv[u] = vb[live ? min(tid + 256 * u, nvv - 1) : 0];
#pragma unroll
        for (int u = 0; u < PY; u++) {
            const int e = tid + 256 * u;
            const int j = live ? min(e % ns, n1 - 1) : 0, v = live ? min(e / ns, nvec - 1) : 0;
            py[u] = ld_cvt<T, S>(x + (long)v * ldx + b0 * n1, j);
        }
        pe = eig0[b0 * n1 + (live ? min(tid, n1 - 1) : 0)];
    };
    auto stage = [&](int bb) {
#pragma unroll
        for (int u = 0; u < PV; u++) {
            const int e = tid + 256 * u;
            if (e < nvv) sV[(e % n1) + ldv * (e / n1)] = pv[u];
        }
        if constexpr (BIG)
            for (int e = tid + 256 * PV; e < nvv; e += 256) sV[(e % n1) + ldv * (e / n1)] = vosb[(long)bb * nvv + e];
#pragma unroll
        for (int u = 0; u < PY; u++) {
            const int e = tid + 256 * u;
            if (e < nw) {
                const int j = e % ns, c = e / ns;
                sW[j + ldv * c] = (j < n1 && sAct[c]) ? py[u] : (T)0;
            }
        }
        if constexpr (BIG)
            for (int e = tid + 256 * PY; e < nw; e += 256) {
                const int j = e % ns, c = e / ns;
                sW[j + ldv * c] = (j < n1 && sAct[c]) ? ld_cvt<T, S>(x + (long)c * ldx + (long)bb * n1, j) : (T)0;
            }
        if (tid < ns) sE[tid] = tid < n1 ? pe : (T)0;
    };
    // all chunks of vectors of block bb; V, EIG0 and the first chunk are staged
    auto compute = [&](int bb) {
        for (int v0 = 0; v0 < nvec; v0 += WC) {
            const int slot = v0 == 0 ? 0 : 1;
            if (v0 > 0) {
                __syncthreads();   // the previous chunk is done with sW, sZ and slot 1
                vec_info(v0, 1);
                __syncthreads();
                for (int e = tid; e < nw; e += 256) {
                    const int j = e % ns, c = e / ns;
                    sW[j + ldv * c] = (j < n1 && sAct[WC + c])
                                          ? ld_cvt<T, S>(x + (long)(v0 + c) * ldx + (long)bb * n1, j)
                                          : (T)0;
                }
                __syncthreads();
            }
            const T *sh = sSh + slot * WC;
            const int *act = sAct + slot * WC;
            if constexpr (MF) {
                const int nr = ns / 16, nk = ns / 4;
                const bool vin = act[lr] != 0;
                const T e0 = sh[lr];
                // Z = V^T W, divided by the kernel; zero outside the block and the batch
                for (int r = wave; r < nr; r += 4) {
                    acc_t acc = acc_t{0, 0, 0, 0};
                    for (int s = 0; s < nk; s += 4) {
                        T a[4], bv[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int k = 4 * (s + u) + lk;
                            a[u] = sV[k + ldv * (16 * r + lr)];   // A(i,k) = V(k,i)
                            bv[u] = sW[k + ldv * lr];
                        }
#pragma unroll
                        for (int u = 0; u < 4; u++) acc = M::run(a[u], bv[u], acc);
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int i = 16 * r + M::row(lk, q);
                        sZ[i + ldv * lr] = (vin && i < n1) ? px_kernel<T>(nosb, acc[q], sE[i], e0, de, beta) : (T)0;
                    }
                }
                __syncthreads();
                // Y = V Z into sW
                for (int r = wave; r < nr; r += 4) {
                    acc_t acc = acc_t{0, 0, 0, 0};
                    for (int s = 0; s < nk; s += 4) {
                        T a[4], bv[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int k = 4 * (s + u) + lk;
                            a[u] = sV[(16 * r + lr) + ldv * k];   // A(i,k) = V(i,k)
                            bv[u] = sZ[k + ldv * lr];
                        }
#pragma unroll
                        for (int u = 0; u < 4; u++) acc = M::run(a[u], bv[u], acc);
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) sW[16 * r + M::row(lk, q) + ldv * lr] = acc[q];
                }
                __syncthreads();
                for (int e = tid; e < n1 * WC; e += 256) {
                    const int i = e % n1, c = e / n1;
                    if (act[c]) st_cvt<T, SO>(y + (long)(v0 + c) * ldy + (long)bb * n1, i, sW[i + ldv * c]);
                }
            } else {
                const int nv = min(WC, nvec - v0);
                const int no = n1 * nv;   // outputs of a pass
                // np threads per output (a power of two <= 8 with no * np <= 256), each
                // summing a contiguous slice of j
                int np = 1;
                while (np < 8 && 2 * np * no <= 256) np *= 2;
                const int sl = (n1 + np - 1) / np;
                if (np == 1) {
                    for (int e = tid; e < no; e += 256) {
                        const int i = e % n1, c = e / n1;
                        T acc = 0;
#pragma unroll 8
                        for (int j = 0; j < n1; j++) acc += sV[j + ldv * i] * sW[j + ldv * c];
                        sZ[i + ldv * c] = px_kernel<T>(nosb, acc, sE[i], sh[c], de, beta);
                    }
                    __syncthreads();
                    for (int e = tid; e < no; e += 256) {
                        const int i = e % n1, c = e / n1;
                        if (!act[c]) continue;
                        T acc = 0;
#pragma unroll 8
                        for (int j = 0; j < n1; j++) acc += sV[i + ldv * j] * sZ[j + ldv * c];
                        st_cvt<T, SO>(y + (long)(v0 + c) * ldy + (long)bb * n1, i, acc);
                    }
                } else {
                    const int o = tid % no, q = tid / no;   // output, slice
                    const int i = o % n1, c = o / n1;
                    const int j0 = q * sl, j1 = min(n1, j0 + sl);
                    if (q < np) {
                        T acc = 0;
#pragma unroll 8
                        for (int j = j0; j < j1; j++) acc += sV[j + ldv * i] * sW[j + ldv * c];
                        sP[o + no * q] = acc;
                    }
                    __syncthreads();
                    if (tid < no) {
                        T z = 0;
                        for (int k = 0; k < np; k++) z += sP[tid + no * k];
                        sZ[i + ldv * c] = px_kernel<T>(nosb, z, sE[i], sh[c], de, beta);
                    }
                    __syncthreads();
                    if (q < np) {
                        T acc = 0;
#pragma unroll 8
                        for (int j = j0; j < j1; j++) acc += sV[i + ldv * j] * sZ[j + ldv * c];
                        sP[o + no * q] = acc;
                    }
                    __syncthreads();
                    if (tid < no && act[c]) {
                        T yv = 0;
                        for (int k = 0; k < np; k++) yv += sP[tid + no * k];
                        st_cvt<T, SO>(y + (long)(v0 + c) * ldy + (long)bb * n1, i, yv);
                    }
                }
            }
        }
    };
    int b = blockIdx.x;
    load(b);
    for (; b < nblk; b += G) {
        __syncthreads();   // the previous block is done with the LDS (and vec_info is visible)
        stage(b);
        __syncthreads();
        load(b + G);       // in flight while this block is computed
        compute(b);
    }
}

static int pgrid(long n)
{
    long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    return (int)(g < 1 ? 1 : g);
}

static constexpr size_t PX_LDS_MAX = 160 * 1024;

// LDS of k_px_l1 for level-1 blocks of n1 (MF: on MFMA)
template <typename T>
static size_t l1_lds(int n1, bool mf)
{
    const size_t ns = mf ? rup16(n1) : n1, wc = mf ? 16 : 8;
    return (ns * (ns + 1) + 2 * wc * (ns + 1) + 256 + ns + 2 * wc) * sizeof(T) + 2 * wc * sizeof(int);
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

template <typename T, typename SI, typename SO, bool MF, bool BIG>
static void launch_l1(int n1, int nb, const T *vosb, const T *eig0, const T *shifts, T shift0, int nosb, T de, T beta,
                      int nvec, const SI *x, long ldx, SO *y, long ldy, const int *active, hipStream_t s)
{
    // persistent: as many workgroups as are resident at once (per LDS size)
    const size_t lds = l1_lds<T>(n1, MF);
    const int per_cu = resident_per_cu((const void *)k_px_l1<T, SI, SO, MF, BIG>, 256, lds);
    const int G = std::min(nb, g_num_cu * per_cu);
    hipLaunchKernelGGL((k_px_l1<T, SI, SO, MF, BIG>), dim3(G), dim3(256), lds, s, n1, nb, nvec, vosb, eig0, shifts, shift0,
                       nosb, de, beta, x, ldx, y, ldy, active);
}

// Level 1: MFMA for batches of 8 and more, VALU otherwise.
template <typename T, typename SI, typename SO>
static void level1(const Level &L, const T *vosb, const T *eig0, const T *shifts, T shift0, int nosb, T de, T beta,
                   int nvec, const SI *x, long ldx, SO *y, long ldy, const int *active, hipStream_t s)
{
    const int n1 = L.n, nb = L.blk;
    const bool mf = nvec >= 8 && n1 <= 112 && l1_lds<T>(n1, true) <= PX_LDS_MAX;
    if (mf && n1 <= 64)
        launch_l1<T, SI, SO, true, false>(n1, nb, vosb, eig0, shifts, shift0, nosb, de, beta, nvec, x, ldx, y, ldy,
                                          active, s);
    else if (mf)
        launch_l1<T, SI, SO, true, true>(n1, nb, vosb, eig0, shifts, shift0, nosb, de, beta, nvec, x, ldx, y, ldy,
                                         active, s);
    else if (n1 <= 64)
        launch_l1<T, SI, SO, false, false>(n1, nb, vosb, eig0, shifts, shift0, nosb, de, beta, nvec, x, ldx, y, ldy,
                                           active, s);
    else
        launch_l1<T, SI, SO, false, true>(n1, nb, vosb, eig0, shifts, shift0, nosb, de, beta, nvec, x, ldx, y, ldy,
                                          active, s);
}

// Level-1 rotations too large for the fused kernel's LDS (n1 > 256: the
// combined (r, R) layer of an energy-cut basis holds hundreds of points):
//   out_b(i, v) = sum_j A_b(i, j) in_b(j, v)
// for every block b of n x n column-major A (V for the V rotation, the
// block's transpose for V^T), optionally followed by the PX_KERNEL divide
// (DIV).  One thread per output row i, so every load of A is a coalesced run
// down a column; in_b of up to BR_VB vectors is staged in LDS and read by
// broadcast.  Each A value is read once per BR_VB vectors.
constexpr int BR_VB = 16;
template <typename T, typename SI, typename SO, bool DIV>
__global__ void __launch_bounds__(256) k_blockrot(int n, int nblk, int nvec, const T *__restrict__ A,
                                                  const T *__restrict__ eig0, const T *__restrict__ shifts, T shift0,
                                                  int nosb, T de, T beta, const SI *__restrict__ x, long ldx,
                                                  SO *__restrict__ y, long ldy, const int *__restrict__ active)
{
    extern __shared__ char smem_raw[];
    T *sx = (T *)smem_raw;   // in_b(j, v0 + vv) at j + n vv
    const int b = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
    const long base = (long)b * n;
    const T *Ab = A + (long)b * n * n;
    const int ic = min(i, n - 1);   // rows past n compute a valid row and store nothing
    for (int v0 = 0; v0 < nvec; v0 += BR_VB) {
        const int nv = min(BR_VB, nvec - v0);
        __syncthreads();   // the previous pass is done with sx
        for (int e = threadIdx.x; e < n * nv; e += 256) {
            const int j = e % n, vv = e / n, v = v0 + vv;
            sx[e] = vec_active(active, v) ? ld_cvt<T, SI>(x + (long)v * ldx, base + j) : (T)0;
        }
        __syncthreads();
        T acc[BR_VB];
#pragma unroll
        for (int vv = 0; vv < BR_VB; vv++) acc[vv] = (T)0;
        int j = 0;
        for (; j + 8 <= n; j += 8) {
            T a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = Ab[ic + (long)n * (j + u)];
#pragma unroll
            for (int u = 0; u < 8; u++)
#pragma unroll
                for (int vv = 0; vv < BR_VB; vv++)
                    if (vv < nv) acc[vv] += a[u] * sx[(j + u) + n * vv];
        }
        for (; j < n; j++) {
            const T a = Ab[ic + (long)n * j];
#pragma unroll
            for (int vv = 0; vv < BR_VB; vv++)
                if (vv < nv) acc[vv] += a * sx[j + n * vv];
        }
        if (i < n) {
#pragma unroll
            for (int vv = 0; vv < BR_VB; vv++) {
                const int v = v0 + vv;
                if (vv < nv && vec_active(active, v)) {
                    T val = acc[vv];
                    if constexpr (DIV) val = px_kernel<T>(nosb, val, eig0[base + i], shifts ? shifts[v] : shift0, de, beta);
                    st_cvt<T, SO>(y + (long)v * ldy, base + i, val);
                }
            }
        }
    }
}

// Level 1 in two passes through ztmp: Z = V1^T X divided by the kernel, then
// Y = V1 Z (vosbt: the per-block transposes of VOSB level 1).
template <typename T, typename SI, typename SO>
static void level1_big(const Level &L, const T *vosb, const T *vosbt, const T *eig0, const T *shifts, T shift0,
                       int nosb, T de, T beta, int nvec, const SI *x, long ldx, SO *y, long ldy, T *ztmp, long N,
                       const int *active, hipStream_t s)
{
    const int n = L.n, nb = L.blk;
    const dim3 g((n + 255) / 256, nb);
    const size_t lds = (size_t)n * BR_VB * sizeof(T);
    allow_full_lds((const void *)k_blockrot<T, SI, T, true>);
    allow_full_lds((const void *)k_blockrot<T, T, SO, false>);
    hipLaunchKernelGGL((k_blockrot<T, SI, T, true>), g, dim3(256), lds, s, n, nb, nvec, vosbt, eig0, shifts, shift0,
                       nosb, de, beta, x, ldx, ztmp, N, active);
    hipLaunchKernelGGL((k_blockrot<T, T, SO, false>), g, dim3(256), lds, s, n, nb, nvec, vosb, eig0, shifts, shift0,
                       nosb, de, beta, ztmp, N, y, ldy, active);
}

// the fused level-1 kernel holds a block's rotation in LDS; larger ones take level1_big
template <typename T>
static bool l1_big(int n1)
{
    return n1 > 256 || l1_lds<T>(n1, false) > PX_LDS_MAX;
}

template <typename T, typename S>
static int px_impl(scalit_ctx *c, int nosb, const T *const *vosb, const T *const *vosbt, const T *eig0, T *work,
                   const T *shifts, T shift0, T de, T beta, int nvec, const S *x, long ldx, S *y, long ldy,
                   const int *active, hipStream_t s)
{
    const long N = c->N;
    const int sF = c->sF;
    const bool big = l1_big<T>(c->lv[0].n);
    if (big && (long)BR_VB * c->lv[0].n * sizeof(T) > PX_LDS_MAX) return SCALIT_EINVAL;
    const bool win = nosb == SCALIT_TOSBW;
    if (win && (!std::is_same<T, double>::value || !std::is_same<S, double>::value || c->osbw_n <= 0 || c->ndvr))
        return c->osbw_n <= 0 ? SCALIT_ESTATE : SCALIT_EINVAL;
    T *ztmp = work + 2L * nvec * N;   // level1_big's intermediate (work holds 3 nvec N then)
    // TOSBW: the TOSB divide by EIG0 with +inf at the window, the window's
    // components solved beside it
    const T *eig1 = eig0;
    if constexpr (std::is_same<T, double>::value)
        if (win) eig1 = c->osbw_eig0.p;
    const int kern1 = win ? SCALIT_TOSB : nosb;
    auto lev1 = [&](const auto *xin, long ldi, auto *yout, long ldo) {
        using SI = std::remove_const_t<std::remove_pointer_t<decltype(xin)>>;
        using SO = std::remove_pointer_t<decltype(yout)>;
        if constexpr (std::is_same<SI, double>::value && std::is_same<SO, double>::value)
            if (win) osbw_gather(c, nvec, xin, ldi, active, s);
        if (big)
            level1_big<T, SI, SO>(c->lv[0], vosb[0], vosbt[0], eig1, shifts, shift0, kern1, de, beta, nvec, xin, ldi,
                                  yout, ldo, ztmp, N, active, s);
        else
            level1<T, SI, SO>(c->lv[0], vosb[0], eig1, shifts, shift0, kern1, de, beta, nvec, xin, ldi, yout, ldo,
                              active, s);
        if constexpr (std::is_same<SI, double>::value && std::is_same<SO, double>::value)
            if (win) osbw_scatter(c, nvec, yout, ldo, active, s);
    };
    if (sF == 1) {
        // a single level: input and output are the caller's vectors
        lev1(x, ldx, y, ldy);
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
    lev1((const T *)cur, N, dst, N);
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

// The V^T rotations read per-block transposes of VOSB levels >= 2 (and of
// level 1 when it takes level1_big), built
// here on first use after the preconditioner changed (kernel buffers of a
// new generation: recorded graphs are re-recorded).
template <typename T>
static int ensure_transposes(scalit_ctx *c, DBuf<T> Level::*vb, DBuf<T> Level::*vt, hipStream_t s,
                             bool with_l1 = false)
{
    for (int l = (with_l1 || l1_big<T>(c->lv[0].n)) ? 0 : 1; l < c->sF; l++) {
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
        SC_TRY(c->work.alloc((size_t)(l1_big<double>(c->lv[0].n) ? 3 : 2) * nvec * c->N));
        if (c->work.p != old) c->gen++;
    }
    if (nosb == SCALIT_TOSBW && c->osbw_n > 0) {
        double *old = c->osbw_d.p;
        SC_TRY(c->osbw_d.alloc((size_t)nvec * c->osbw_n));
        if (c->osbw_d.p != old) c->gen++;
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

int px_apply_f32(scalit_ctx *c, int nosb, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx,
                 float *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    SC_TRY(precond_to_f32(c));
    {
        float *old = c->work_f.p;
        SC_TRY(c->work_f.alloc((size_t)(l1_big<float>(c->lv[0].n) ? 3 : 2) * nvec * c->N));
        if (c->work_f.p != old) c->gen++;
    }
    SC_TRY(ensure_transposes<float>(c, &Level::vosb_f, &Level::vosbt_f, s));
    const float *vosb[SC_FMAX], *vosbt[SC_FMAX];
    for (int l = 0; l < c->sF; l++) {
        vosb[l] = c->lv[l].vosb_f.p;
        vosbt[l] = c->lv[l].vosbt_f.p;
    }
    return px_impl<float, float>(c, nosb, vosb, vosbt, c->eig0_f.p, c->work_f.p, shifts_dev, shift0, (float)c->de,
                                 (float)c->beta, nvec, x, ldx, y, ldy, active, s);
}

int px_apply_bf16(scalit_ctx *c, int nosb, const float *shifts_dev, float shift0, int nvec, const bf16_t *x,
                  long ldx, bf16_t *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    SC_TRY(precond_to_f32(c));
    {
        float *old = c->work_f.p;
        SC_TRY(c->work_f.alloc((size_t)(l1_big<float>(c->lv[0].n) ? 3 : 2) * nvec * c->N));
        if (c->work_f.p != old) c->gen++;
    }
    SC_TRY(ensure_transposes<float>(c, &Level::vosb_f, &Level::vosbt_f, s));
    const float *vosb[SC_FMAX], *vosbt[SC_FMAX];
    for (int l = 0; l < c->sF; l++) {
        vosb[l] = c->lv[l].vosb_f.p;
        vosbt[l] = c->lv[l].vosbt_f.p;
    }
    return px_impl<float, bf16_t>(c, nosb, vosb, vosbt, c->eig0_f.p, c->work_f.p, shifts_dev, shift0, (float)c->de,
                                  (float)c->beta, nvec, x, ldx, y, ldy, active, s);
}

// The complex kernel of MYPX_DX (src/iterate/module/osb.mypxdx.h): z / d with
// d = EIG0 - E0 (TOSBD1: clamped to +-DE inside the window; TOSBD2:
// alpha + beta (EIG0 - E0) there, alpha = -mBeta DE, beta = mBeta + 1) plus
// i sign A, on split complex vectors in place (ld N per plane).
__global__ void k_cx_div(long N, int nvec, const double *__restrict__ eig0, double e0, int nosb, double de,
                         double mbeta, const double *__restrict__ a, double sign, double *__restrict__ z)
{
    const int v = blockIdx.y;
    double *zr = z + 2L * v * N, *zi = zr + N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (long)gridDim.x * blockDim.x) {
        double ade = eig0[e] - e0;
        if (nosb == SCALIT_TOSBD1) {
            if (fabs(ade) <= de) ade = ade < 0.0 ? -de : de;
        } else if (nosb == SCALIT_TOSBD2) {
            if (fabs(ade) <= de) ade = -mbeta * de + (mbeta + 1.0) * ade;
        }
        const double di = a ? sign * a[e] : 0.0;
        const double xr = zr[e], xi = zi[e];
        if (di == 0.0) {
            zr[e] = xr / ade;
            zi[e] = xi / ade;
        } else {
            // (xr + i xi) / (ade + i di), scaled against overflow as Smith's division
            if (fabs(ade) >= fabs(di)) {
                const double r = di / ade, den = ade + di * r;
                zr[e] = (xr + xi * r) / den;
                zi[e] = (xi - xr * r) / den;
            } else {
                const double r = ade / di, den = di + ade * r;
                zr[e] = (xr * r + xi) / den;
                zi[e] = (xi * r - xr) / den;
            }
        }
    }
}

// TOSBW of a complex P*X: the window's coefficients of both planes before
// the divide (g, nvec x n complex) ...
__global__ void k_cxw_gather(long N, int n, const int *__restrict__ ind, const double *__restrict__ z,
                             double *__restrict__ g)
{
    const int v = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *zv = z + 2L * v * N;
    g[2 * ((long)v * n + k)] = zv[ind[k]];
    g[2 * ((long)v * n + k) + 1] = zv[N + ind[k]];
}

// ... and after it, M g in their place (M = HWR^-1, n x n complex)
__global__ void k_cxw_apply(long N, int n, const int *__restrict__ ind, const dcx *__restrict__ m,
                            const double *__restrict__ g, double *__restrict__ z)
{
    const int v = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *gv = g + 2L * v * n;
    double r = 0.0, i = 0.0;
    for (int j = 0; j < n; j++) {
        const dcx a = m[k + (long)n * j];
        const double br = gv[2 * j], bi = gv[2 * j + 1];
        r += a.re * br - a.im * bi;
        i += a.re * bi + a.im * br;
    }
    double *zv = z + 2L * v * N;
    zv[ind[k]] = r;
    zv[N + ind[k]] = i;
}

int px_apply_cx(scalit_ctx *c, int nosb, int npc, int nvec, const double *x, double *y, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    const bool win = nosb == SCALIT_TOSBW;
    if (nosb != SCALIT_TOSB && nosb != SCALIT_TOSBD1 && nosb != SCALIT_TOSBD2 && !win) return SCALIT_EINVAL;
    // TOSBW divides by EIG0 - E0 even for TAAP0 / TMAP0 while its window
    // matrix takes no E0 there (osb.mypxdx.h:153-160, osb.pij.h:95-105):
    // those two are refused
    if (win && (npc == SCALIT_TAAP0 || npc == SCALIT_TMAP0)) return SCALIT_EINVAL;
    if (win && c->osbw_n <= 0) return SCALIT_ESTATE;
    int rc = SCALIT_OK;
    double sign = 0.0;
    const double *a = cx_imag_diag(c, npc, &sign, &rc);
    SC_TRY(rc);
    const long N = c->N;
    const int sF = c->sF, nr = 2 * nvec;
    const Level &L1 = c->lv[0];
    if ((long)BR_VB * L1.n * sizeof(double) > PX_LDS_MAX) return SCALIT_EINVAL;
    {
        double *old = c->work.p;
        SC_TRY(c->work.alloc((size_t)3 * nr * N));
        if (c->work.p != old) c->gen++;
    }
    SC_TRY(ensure_transposes<double>(c, &Level::vosb, &Level::vosbt, s, true));
    const double e0 = (npc == SCALIT_TA0 || npc == SCALIT_TAAP0 || npc == SCALIT_TMAP0) ? 0.0 : c->e0;
    const dcx *mw = nullptr;
    if (win) {
        SC_TRY(osbw_cx_matrix(c, npc, e0, a, sign, &mw, s));
        double *old = c->osbw_g.p;
        SC_TRY(c->osbw_g.alloc(2 * (size_t)c->osbw_n * nvec));
        if (c->osbw_g.p != old) c->gen++;
    }
    double *w0 = c->work.p, *w1 = w0 + (long)nr * N, *zt = w1 + (long)nr * N;
    // V^T for levels sF..2 (1-based) on the real and imaginary planes
    const double *cur = x;
    double *dst = w0;
    for (int l = sF - 1; l >= 1; l--) {
        rotate<double, double, double>(c->lv[l], c->lv[l].vosb.p, c->lv[l].vosbt.p, true, nr, cur, N, dst, N,
                                       nullptr, s);
        cur = dst;
        dst = (dst == w0) ? w1 : w0;
    }
    // level 1: V1^T, the complex divide, V1
    const int n = L1.n, nb = L1.blk;
    const dim3 g((n + 255) / 256, nb);
    const size_t lds = (size_t)n * BR_VB * sizeof(double);
    allow_full_lds((const void *)k_blockrot<double, double, double, false>);
    hipLaunchKernelGGL((k_blockrot<double, double, double, false>), g, dim3(256), lds, s, n, nb, nr, L1.vosbt.p,
                       c->eig0.p, (const double *)nullptr, 0.0, SCALIT_TOSB, 0.0, 0.0, cur, N, zt, N,
                       (const int *)nullptr);
    const dim3 gw((c->osbw_n + 255) / 256, nvec);
    if (win)
        hipLaunchKernelGGL(k_cxw_gather, gw, dim3(256), 0, s, N, c->osbw_n, c->osbw_ind.p, (const double *)zt,
                           c->osbw_g.p);
    hipLaunchKernelGGL(k_cx_div, dim3((unsigned)std::min<long>((N + 255) / 256, 2048), nvec), dim3(256), 0, s, N,
                       nvec, c->eig0.p, e0, win ? SCALIT_TOSB : nosb, std::fabs(c->de), c->beta, a, sign, zt);
    if (win)
        hipLaunchKernelGGL(k_cxw_apply, gw, dim3(256), 0, s, N, c->osbw_n, c->osbw_ind.p, mw,
                           (const double *)c->osbw_g.p, zt);
    double *out1 = sF == 1 ? y : dst;
    hipLaunchKernelGGL((k_blockrot<double, double, double, false>), g, dim3(256), lds, s, n, nb, nr, L1.vosb.p,
                       c->eig0.p, (const double *)nullptr, 0.0, SCALIT_TOSB, 0.0, 0.0, (const double *)zt, N, out1,
                       N, (const int *)nullptr);
    cur = out1;
    dst = (out1 == w0) ? w1 : w0;
    // V for levels 2..sF; the last writes y
    for (int l = 1; l < sF; l++) {
        double *o = (l == sF - 1) ? y : dst;
        rotate<double, double, double>(c->lv[l], c->lv[l].vosb.p, c->lv[l].vosbt.p, false, nr, cur, N, o, N, nullptr,
                                       s);
        cur = o;
        dst = (o == w0) ? w1 : w0;
    }
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

}  // namespace scalit
