// Product along the outer index of a P x n array, on MFMA:
//   Y(p, i, b) [+]= d(p) * sum_j A_b(i, j) X(p, j, b),   p < P (contiguous), i, j < n,
// for every block b of nblk (A_b at amat + b n^2) and every vector of a batch.
// The mode-3 pass of H*X (one block, A = L^2_theta, d = DEP(r, R), added to Y)
// and every rotation level of P*X (A_b = V_b or V_b^T, d = 1, written) are
// this shape.
//
// Persistent grid (x: workgroups per block, y: block): each workgroup stages
// its block's A in LDS once; each wave then walks its own tiles of 16
// consecutive p of one vector (no workgroup barrier after the staging) and loads the next tile's X and d into registers while the
// current tile's MFMAs run (Y, when adding, at the top of its own tile).  Lane (lr, lk) of a tile holds X(p0 + lr, 4 s + lk)
// for all s (the B fragments); the A fragments come from LDS.  NR = tiles of
// 16 along n (n <= 16 NR, A zero padded), a compile-time count so the register
// arrays hold exactly one column.
#pragma once

#include <algorithm>

#include "device_common.hpp"

namespace scalit {

template <typename T, typename S, typename SO, int WAVES, int NR, bool ACC, bool EX>
__global__ void __launch_bounds__(64 * WAVES) k_outer_mode(long P, int n, int nvec, const T *__restrict__ amat,
                                                          const T *__restrict__ dvec, const S *__restrict__ x,
                                                          long ldx, SO *__restrict__ y, long ldy, const int *active)
{
    // block b of the outer index: its matrix and its P x n slice of every vector
    const int b = blockIdx.y;
    amat += (long)b * n * n;
    x += (long)b * P * n;
    y += (long)b * P * n;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr int np = 16 * NR;
    constexpr int KS = 4 * NR;   // k steps of 4
    extern __shared__ char smem_raw[];
    T *sA = (T *)smem_raw;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: tile bases stay scalar
    const int lr = lane & 15, lk = lane >> 4;
    const long tpv = (P + 15) / 16;
    const long ntile = tpv * nvec;
    const long wstride = (long)gridDim.x * WAVES;
    long t = (long)blockIdx.x * WAVES + wave;

    // X, d and the vector's activity of the current and (PF) the next tile;
    // Y (when adding) of the current tile, loaded at the top of its iteration
    // and used after its MFMAs.  Without PF (12-wave workgroups, whose
    // registers do not hold two X columns) X also loads at the top of its own
    // tile, and the other waves of the SIMD cover the wait.
    //
    // EX (n == 16 NR): raw buffer loads, one resource per tile, one lane
    // offset for all of a column (lanes past the last point read point P-1),
    // and every load and store is issued on every path: the prefetch of a
    // tile past the last reads the last tile again (never used), d and the
    // activity are buffer loads from the caller's arrays or from a constant
    // (no flat access, which would also count against the LDS counter), and
    // a lane that must not store gets an offset past the resource, which the
    // hardware drops.  The wait for one tile's X then counts exactly the
    // memory operations issued after it and never drains the prefetch.
    // Otherwise predicated accesses.
    constexpr bool PF = WAVES < 12;
    T bc[KS], dc, bn[PF ? KS : 1], dn;
    int ac = 0, an = 0;
    T yc[ACC ? NR : 1][4];
    const int ylane = M::row(lk, 0);
    const auto rd = dvec ? buf_rsrc(dvec, P * (long)sizeof(T)) : buf_rsrc(unit_ptr<T>(), sizeof(T));
    const auto ra = active ? buf_rsrc(active, (long)nvec * sizeof(int)) : buf_rsrc(&k_all_on, sizeof(int));
    auto load_x = [&](long tt, T(&b)[KS], T &d, int &a) {
        tt = tt < ntile ? tt : ntile - 1;
        const int v = (int)(tt / tpv);
        const long p0 = (tt % tpv) * 16;
        const long pc = p0 + lr < P ? p0 + lr : P - 1;
        a = buf_ld_i32(ra, 0, active ? v * (int)sizeof(int) : 0);
        d = buf_ld<T, T>(rd, dvec ? (int)(pc * (long)sizeof(T)) : 0, 0);
        const S *xs = x + (long)v * ldx + p0;
        if constexpr (EX) {
            const auto rs = buf_rsrc(xs, (P * n - p0) * (long)sizeof(S));
            const int voff = (int)((pc - p0 + P * lk) * (long)sizeof(S));
#pragma unroll
            for (int s = 0; s < KS; s++) b[s] = buf_ld<T, S>(rs, voff, (int)(P * 4 * s * (long)sizeof(S)));
        } else {
            // predicated on the point only: the activity is applied at the store,
            // so the X loads do not wait for the flag
            const bool pin = p0 + lr < P;
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const int k = 4 * s + lk;
                b[s] = (pin && k < n) ? ld_cvt<T, S>(xs, lr + P * k) : (T)0;
            }
        }
    };
    if (PF) load_x(t, bc, dc, ac);

    if (np != n)
        for (int e = tid; e < np * np; e += 64 * WAVES)
            if (e % np >= n || e / np >= n) sA[e] = (T)0;
    stage_lds<64 * WAVES, 16>(amat, n * n, tid, [&](int e) -> T & { return sA[(e % n) + np * (e / n)]; });
    __syncthreads();

    while (t < ntile) {
        const long tn = t + wstride;
        const int v = (int)(t / tpv);
        const long p0 = (t % tpv) * 16;
        SO *ys = y + (long)v * ldy + p0;
        const auto ry = buf_rsrc(ys, (P * n - p0) * (long)sizeof(SO));
        const long plane = p0 + lr < P ? lr : P - 1 - p0;
        const int yvoff = (int)((plane + P * ylane) * (long)sizeof(SO));
        if constexpr (ACC) {
#pragma unroll
            for (int r = 0; r < NR; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = 16 * r + M::row(lk, q);
                    if constexpr (EX)
                        yc[r][q] = buf_ld<T, SO>(ry, yvoff,
                                                 (int)(P * (16 * r + M::row(0, q)) * (long)sizeof(SO)));
                    else
                        yc[r][q] = (p0 + lr < P && i < n) ? ld_cvt<T, SO>(ys, lr + P * i) : (T)0;
                }
        }
        if constexpr (PF) load_x(tn, bn, dn, an);
        else load_x(t, bc, dc, ac);
        acc_t acc[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) acc[r] = acc_t{0, 0, 0, 0};
        // The A fragments are the same for every tile; an opaque base keeps the
        // compiler from hoisting all KS x NR of them into registers.
        int abase = lr + np * lk;
        asm volatile("" : "+v"(abase));
        // groups of 4 k steps; the scheduling barrier bounds the fragments in flight
#pragma unroll
        for (int s0 = 0; s0 < KS; s0 += 4) {
#pragma unroll
            for (int s = s0; s < s0 + 4; s++)
#pragma unroll
                for (int r = 0; r < NR; r++) acc[r] = M::run(sA[abase + 16 * r + 4 * np * s], bc[s], acc[r]);
            __builtin_amdgcn_sched_barrier(0);
        }
        const bool pin = p0 + lr < P && ac != 0;
        if constexpr (EX) {
            const int so = pin ? yvoff : (int)0x80000000;   // past the resource: dropped
#pragma unroll
            for (int r = 0; r < NR; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    T out = dc * acc[r][q];
                    if constexpr (ACC) out += yc[r][q];
                    buf_st<T, SO>(ry, so, (int)(P * (16 * r + M::row(0, q)) * (long)sizeof(SO)), out);
                }
        } else if (pin) {
#pragma unroll
            for (int r = 0; r < NR; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = 16 * r + M::row(lk, q);
                    T out = dc * acc[r][q];
                    if constexpr (ACC) out += yc[r][q];
                    if (i < n) st_cvt<T, SO>(ys, lr + P * i, out);
                }
        }
        if constexpr (PF) {
#pragma unroll
            for (int s = 0; s < KS; s++) bc[s] = bn[s];
            dc = dn;
            ac = an;
        }
        t = tn;
    }
}

// Launch with WAVES waves per workgroup and as many workgroups as fit on the
// device at once (at most one per 16 tiles' worth of waves).
template <typename T, typename S, typename SO, int WAVES, int NR, bool ACC>
static void launch_outer_mode_nr(long P, int n, int nvec, const T *amat, const T *dvec, const S *x, long ldx, SO *y,
                                 long ldy, const int *active, hipStream_t s, int nblk)
{
    // exact tiles and 32-bit byte offsets within one vector: the buffer path
    const bool ex = n == 16 * NR && P * n * (long)std::max(sizeof(S), sizeof(SO)) < 0x7fffffffL;
    const size_t lds = (size_t)(16 * NR) * (16 * NR) * sizeof(T);
    const void *fn = ex ? (const void *)k_outer_mode<T, S, SO, WAVES, NR, ACC, true>
                        : (const void *)k_outer_mode<T, S, SO, WAVES, NR, ACC, false>;
    const int per_cu = resident_per_cu(fn, 64 * WAVES, lds);
    // workgroups per block: enough for the block's tiles, and all blocks'
    // workgroups together at most what is resident at once
    const long ntile = ((P + 15) / 16) * nvec;
    long G = std::min<long>((ntile + WAVES - 1) / WAVES, std::max<long>(1, (long)g_num_cu * per_cu / nblk));
    if (G < 1) G = 1;
    const dim3 grid((unsigned)G, (unsigned)nblk);
    if (ex)
        hipLaunchKernelGGL((k_outer_mode<T, S, SO, WAVES, NR, ACC, true>), grid, dim3(64 * WAVES), lds, s, P, n, nvec,
                           amat, dvec, x, ldx, y, ldy, active);
    else
        hipLaunchKernelGGL((k_outer_mode<T, S, SO, WAVES, NR, ACC, false>), grid, dim3(64 * WAVES), lds, s, P, n,
                           nvec, amat, dvec, x, ldx, y, ldy, active);
}

// n <= 128, nblk <= 65535 blocks (block b's matrix at amat + b n^2, its data
// at + b P n of each vector): the tile count along n rounded up to a compiled one
template <typename T, typename S, typename SO, int WAVES, bool ACC>
static void launch_outer_mode(long P, int n, int nvec, const T *amat, const T *dvec, const S *x, long ldx, SO *y,
                              long ldy, const int *active, hipStream_t s, int nblk = 1)
{
    const int nr = (n + 15) / 16;
    if (nr <= 2)
        launch_outer_mode_nr<T, S, SO, WAVES, 2, ACC>(P, n, nvec, amat, dvec, x, ldx, y, ldy, active, s, nblk);
    else if (nr <= 4)
        launch_outer_mode_nr<T, S, SO, WAVES, 4, ACC>(P, n, nvec, amat, dvec, x, ldx, y, ldy, active, s, nblk);
    else if (nr == 5)
        launch_outer_mode_nr<T, S, SO, WAVES, 5, ACC>(P, n, nvec, amat, dvec, x, ldx, y, ldy, active, s, nblk);
    else if (nr == 6)
        launch_outer_mode_nr<T, S, SO, WAVES, 6, ACC>(P, n, nvec, amat, dvec, x, ldx, y, ldy, active, s, nblk);
    else
        launch_outer_mode_nr<T, S, SO, WAVES, 8, ACC>(P, n, nvec, amat, dvec, x, ldx, y, ldy, active, s, nblk);
}

}  // namespace scalit
