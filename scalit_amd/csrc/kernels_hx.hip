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
//      pass B, one wave per 16 grid points p=(r,R) across all theta:
//      Y(p,:) += DEP(p) * L2 X(p,:)              (mode 3)
//    Tiles are computed transposed (D = O^T) so that the 16 lanes of a
//    fragment column run along the contiguous r index: every global load and
//    store of X, Y and V moves 128-byte runs.
#include "scalit_impl.hpp"

namespace scalit {

int g_force_generic = 0;

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: A(i,k) comes from lane i + 16 k, B(k,j) from lane j + 16 k;
// accumulator q of lane l holds D(row(l/16, q), l % 16), with the rows
// interleaved for FP64 and blocked for FP32 (checked on the device by
// tests/test_gpu_mfma_layout.py).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    using acc_t = d4_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return lk + 4 * q; }
};
template <>
struct Mfma<float> {
    using acc_t = f4_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};

// storage <-> compute conversions
__device__ __forceinline__ float bf16_to_f32(bf16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ __forceinline__ bf16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);                                   // round to nearest even
    return (bf16_t)(u >> 16);
}
template <typename T, typename S>
__device__ __forceinline__ T ld_cvt(const S *p, long i) { return (T)p[i]; }
template <>
__device__ __forceinline__ float ld_cvt<float, bf16_t>(const bf16_t *p, long i) { return bf16_to_f32(p[i]); }
template <typename T, typename S>
__device__ __forceinline__ void st_cvt(S *p, long i, T v) { p[i] = (S)v; }
template <>
__device__ __forceinline__ void st_cvt<float, bf16_t>(bf16_t *p, long i, float v) { p[i] = f32_to_bf16(v); }

__device__ __forceinline__ bool vec_active(const int *active, int v) { return active == nullptr || active[v] != 0; }

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

template <typename T, typename S>
static int hx_generic(scalit_ctx *c, const T *res, const T *shifts, T shift0, int nvec, const S *x, long ldx, S *y,
                      long ldy, bool transpose, const int *active, hipStream_t s)
{
    const int th = 256;
    dim3 g(grid_for(c->N, th), nvec);
    hipLaunchKernelGGL((k_diag_init<T, S>), g, dim3(th), 0, s, c->N, res, shifts, shift0, x, ldx, y, ldy, active);
    for (int l = 0; l < c->sF; l++) {
        const Level &L = c->lv[l];
        const T *h, *dep = nullptr;
        if constexpr (sizeof(T) == 8) {
            h = (const T *)(transpose ? L.h0t.p : L.h0.p);
            if (L.dep) dep = (const T *)L.depv.p;
        } else {
            h = (const T *)L.h0_f.p;   // fp32 copies are symmetric-agnostic: only the forward product
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
__host__ __device__ static inline int rup16(int n) { return (n + 15) & ~15; }

// XCD-aware remap: consecutive workgroup ids land on different XCDs (8 on
// MI355X); give each XCD a contiguous range of logical slabs so that the
// vectors sharing one theta point (and its slice of V) share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk)
{
    if ((nblk & 7) != 0) return bid;
    return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

// Pass A.  Grid: n3 * nvec workgroups of 256 threads (4 waves).
// LDS: S (n1p+1) x n2p, and T1 n1p x n1p, T2 n2p x n2p when T_LDS.
template <typename T, typename S, bool T_LDS>
__global__ void __launch_bounds__(256) k_hx_modes12(int n1, int n2, int n3, int nvec, const T *__restrict__ t1,
                                                   const T *__restrict__ t2, const T *__restrict__ res,
                                                   const T *__restrict__ shifts, T shift0, const S *__restrict__ x,
                                                   long ldx, S *__restrict__ y, long ldy, const int *active)
{
    using M = Mfma<T>;
    extern __shared__ char smem_raw[];
    T *sm = (T *)smem_raw;
    const int n1p = rup16(n1), n2p = rup16(n2);
    const int ldS = n1p + 1;
    T *sS = sm;
    T *sT1 = sS + ldS * n2p;
    T *sT2 = sT1 + n1p * n1p;

    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int v = logical % nvec;
    const int i3 = logical / nvec;
    if (!vec_active(active, v)) return;
    const long P = (long)n1 * n2;
    const S *xs = x + (long)v * ldx + P * i3;
    S *ys = y + (long)v * ldy + P * i3;
    const T *rs = res + P * i3;
    const T sh = shifts ? shifts[v] : shift0;
    const int tid = threadIdx.x;

    // stage the slab (coalesced along r), zero the padding
    for (int e = tid; e < n1p * n2p; e += 256) {
        int i1 = e % n1p, i2 = e / n1p;
        T val = 0;
        if (i1 < n1 && i2 < n2) val = ld_cvt<T, S>(xs, i1 + (long)n1 * i2);
        sS[i1 + ldS * i2] = val;
    }
    if constexpr (T_LDS) {
        for (int e = tid; e < n1p * n1p; e += 256) {
            int a = e % n1p, b = e / n1p;
            sT1[e] = (a < n1 && b < n1) ? t1[a + (long)n1 * b] : (T)0;
        }
        for (int e = tid; e < n2p * n2p; e += 256) {
            int a = e % n2p, b = e / n2p;
            sT2[e] = (a < n2 && b < n2) ? t2[a + (long)n2 * b] : (T)0;
        }
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int nrt = n2p / 16, nct = n1p / 16;
    for (int t = wave; t < nrt * nct; t += 4) {
        const int r0 = (t / nct) * 16;   // i2 rows of the transposed tile
        const int c0 = (t % nct) * 16;   // i1 columns
        typename M::acc_t acc = {0, 0, 0, 0};
        // mode 1: D(i2,i1) += sum_k S(k,i2) T1(i1,k)
        for (int k0 = 0; k0 < n1p; k0 += 4) {
            T a = sS[(k0 + lk) + ldS * (r0 + lr)];
            T b;
            if constexpr (T_LDS) b = sT1[(c0 + lr) + n1p * (k0 + lk)];
            else b = (c0 + lr < n1 && k0 + lk < n1) ? t1[(c0 + lr) + (long)n1 * (k0 + lk)] : (T)0;
            acc = M::run(a, b, acc);
        }
        // mode 2: D(i2,i1) += sum_k T2(i2,k) S(i1,k)
        for (int k0 = 0; k0 < n2p; k0 += 4) {
            T a;
            if constexpr (T_LDS) a = sT2[(r0 + lr) + n2p * (k0 + lk)];
            else a = (r0 + lr < n2 && k0 + lk < n2) ? t2[(r0 + lr) + (long)n2 * (k0 + lk)] : (T)0;
            T b = sS[(c0 + lr) + ldS * (k0 + lk)];
            acc = M::run(a, b, acc);
        }
        // epilogue: + (V - s) S, store
        const int i1 = c0 + lr;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i2 = r0 + M::row(lk, q);
            if (i1 < n1 && i2 < n2) {
                long o = i1 + (long)n1 * i2;
                T sv = sS[i1 + ldS * i2];
                st_cvt<T, S>(ys, o, acc[q] + (rs[o] - sh) * sv);
            }
        }
    }
}

// Pass B.  Grid: ceil(P / (16*WAVES)) x nvec workgroups of 64*WAVES threads.
// LDS: L2 n3p x n3p.
template <typename T, typename S, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_hx_mode3(int n1, int n2, int n3, const T *__restrict__ l2,
                                                        const T *__restrict__ dep, const S *__restrict__ x, long ldx,
                                                        S *__restrict__ y, long ldy, const int *active)
{
    using M = Mfma<T>;
    extern __shared__ char smem_raw[];
    T *sL = (T *)smem_raw;
    const int n3p = rup16(n3);
    const int v = blockIdx.y;
    if (!vec_active(active, v)) return;
    const long P = (long)n1 * n2;
    const S *xv = x + (long)v * ldx;
    S *yv = y + (long)v * ldy;
    const int tid = threadIdx.x;
    for (int e = tid; e < n3p * n3p; e += 64 * WAVES) {
        int a = e % n3p, b = e / n3p;
        sL[e] = (a < n3 && b < n3) ? l2[a + (long)n3 * b] : (T)0;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const long p0 = ((long)blockIdx.x * WAVES + wave) * 16;
    if (p0 >= P) return;
    const long p = p0 + lr;
    const bool pin = p < P;
    // B fragments X^T(k, p) for all k: lane holds k = k0 + lk
    constexpr int KMAX = 128 / 4;
    T bfr[KMAX];
    const int nk = n3p / 4;
#pragma unroll
    for (int s = 0; s < KMAX; s++) {
        if (s < nk) {
            int k = 4 * s + lk;
            bfr[s] = (pin && k < n3) ? ld_cvt<T, S>(xv, p + P * k) : (T)0;
        }
    }
    const T dp = pin ? dep[p] : (T)0;
    for (int r0 = 0; r0 < n3p; r0 += 16) {
        typename M::acc_t acc = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KMAX; s++) {
            if (s < nk) {
                T a = sL[(r0 + lr) + n3p * (4 * s + lk)];
                acc = M::run(a, bfr[s], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i3 = r0 + M::row(lk, q);
            if (pin && i3 < n3) {
                long o = p + P * i3;
                st_cvt<T, S>(yv, o, ld_cvt<T, S>(yv, o) + dp * acc[q]);
            }
        }
    }
}

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
    if (c->sF != 3 || c->lv[0].dep || c->lv[1].dep || !c->lv[2].dep) return false;
    if (c->lv[2].n > 128) return false;   // pass B keeps all of theta in registers
    return modes12_lds<T>(c->lv[0].n, c->lv[1].n, false) <= LDS_MAX;
}

template <typename T, typename S>
static int hx_fast3(scalit_ctx *c, const T *t1, const T *t2, const T *l2, const T *dep, const T *res,
                    const T *shifts, T shift0, int nvec, const S *x, long ldx, S *y, long ldy, const int *active,
                    hipStream_t s)
{
    const int n1 = c->lv[0].n, n2 = c->lv[1].n, n3 = c->lv[2].n;
    const bool tl = modes12_lds<T>(n1, n2, true) <= LDS_MAX;
    size_t lds = modes12_lds<T>(n1, n2, tl);
    dim3 ga(n3 * nvec);
    if (tl)
        hipLaunchKernelGGL((k_hx_modes12<T, S, true>), ga, dim3(256), lds, s, n1, n2, n3, nvec, t1, t2, res, shifts,
                           shift0, x, ldx, y, ldy, active);
    else
        hipLaunchKernelGGL((k_hx_modes12<T, S, false>), ga, dim3(256), lds, s, n1, n2, n3, nvec, t1, t2, res, shifts,
                           shift0, x, ldx, y, ldy, active);
    constexpr int WB = 4;
    const long P = (long)n1 * n2;
    dim3 gb((unsigned)((P + 16 * WB - 1) / (16 * WB)), nvec);
    size_t ldsb = (size_t)rup16(n3) * rup16(n3) * sizeof(T);
    hipLaunchKernelGGL((k_hx_mode3<T, S, WB>), gb, dim3(64 * WB), ldsb, s, n1, n2, n3, l2, dep, x, ldx, y, ldy,
                       active);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

static bool lds_attr_done[4] = {false, false, false, false};
template <typename T, typename S>
static void ensure_lds_attr(int slot)
{
    if (lds_attr_done[slot]) return;
    (void)hipFuncSetAttribute((const void *)k_hx_modes12<T, S, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)LDS_MAX);
    (void)hipFuncSetAttribute((const void *)k_hx_modes12<T, S, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)LDS_MAX);
    (void)hipFuncSetAttribute((const void *)k_hx_mode3<T, S, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)LDS_MAX);
    lds_attr_done[slot] = true;
}

int hx_apply_f64(scalit_ctx *c, const double *shifts_dev, double shift0, int nvec, const double *x, long ldx,
                 double *y, long ldy, bool transpose, const int *active, hipStream_t s)
{
    if (!c->fast3 || g_force_generic || !fast3_supported<double>(c))
        return hx_generic<double, double>(c, c->res.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, transpose, active,
                                          s);
    ensure_lds_attr<double, double>(0);
    const Level *L = c->lv;
    const double *t1 = transpose ? L[0].h0t.p : L[0].h0.p;
    const double *t2 = transpose ? L[1].h0t.p : L[1].h0.p;
    const double *l2 = transpose ? L[2].h0t.p : L[2].h0.p;
    return hx_fast3<double, double>(c, t1, t2, l2, L[2].depv.p, c->res.p, shifts_dev, shift0, nvec, x, ldx, y, ldy,
                                    active, s);
}

int hx_apply_f32(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx, float *y,
                 long ldy, const int *active, hipStream_t s)
{
    if (!c->fast3 || g_force_generic || !fast3_supported<float>(c))
        return hx_generic<float, float>(c, c->res_f.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, false, active, s);
    ensure_lds_attr<float, float>(1);
    const Level *L = c->lv;
    return hx_fast3<float, float>(c, L[0].h0_f.p, L[1].h0_f.p, L[2].h0_f.p, L[2].depv_f.p, c->res_f.p, shifts_dev,
                                  shift0, nvec, x, ldx, y, ldy, active, s);
}

int hx_apply_bf16(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const bf16_t *x, long ldx,
                  bf16_t *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->fast3 || g_force_generic || !fast3_supported<float>(c))
        return hx_generic<float, bf16_t>(c, c->res_f.p, shifts_dev, shift0, nvec, x, ldx, y, ldy, false, active, s);
    ensure_lds_attr<float, bf16_t>(2);
    const Level *L = c->lv;
    return hx_fast3<float, bf16_t>(c, L[0].h0_f.p, L[1].h0_f.p, L[2].h0_f.p, L[2].depv_f.p, c->res_f.p, shifts_dev,
                                   shift0, nvec, x, ldx, y, ldy, active, s);
}

bool fast3_ok(const scalit_ctx *c) { return fast3_supported<double>(c); }

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

// One pass of the fused FP64 path alone (pass 1: modes 1+2+diagonal, pass 2:
// mode 3), for per-kernel timing.
int hx_pass_f64(scalit_ctx *c, int pass, int nvec, const double *x, long ldx, double *y, long ldy, hipStream_t s)
{
    if (!c->fast3 || !fast3_supported<double>(c)) return SCALIT_EINVAL;
    ensure_lds_attr<double, double>(0);
    const Level *L = c->lv;
    const int n1 = L[0].n, n2 = L[1].n, n3 = L[2].n;
    if (pass == 1) {
        const bool tl = modes12_lds<double>(n1, n2, true) <= LDS_MAX;
        size_t lds = modes12_lds<double>(n1, n2, tl);
        if (tl)
            hipLaunchKernelGGL((k_hx_modes12<double, double, true>), dim3(n3 * nvec), dim3(256), lds, s, n1, n2, n3,
                               nvec, L[0].h0.p, L[1].h0.p, c->res.p, (const double *)nullptr, 0.0, x, ldx, y, ldy,
                               (const int *)nullptr);
        else
            hipLaunchKernelGGL((k_hx_modes12<double, double, false>), dim3(n3 * nvec), dim3(256), lds, s, n1, n2, n3,
                               nvec, L[0].h0.p, L[1].h0.p, c->res.p, (const double *)nullptr, 0.0, x, ldx, y, ldy,
                               (const int *)nullptr);
    } else if (pass == 2) {
        constexpr int WB = 4;
        const long P = (long)n1 * n2;
        dim3 gb((unsigned)((P + 16 * WB - 1) / (16 * WB)), nvec);
        size_t ldsb = (size_t)rup16(n3) * rup16(n3) * sizeof(double);
        hipLaunchKernelGGL((k_hx_mode3<double, double, WB>), gb, dim3(64 * WB), ldsb, s, n1, n2, n3, L[2].h0.p,
                           L[2].depv.p, x, ldx, y, ldy, (const int *)nullptr);
    } else {
        return SCALIT_EINVAL;
    }
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

}  // namespace scalit
