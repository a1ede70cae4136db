// P*X on the GPU: the OSB preconditioner of src/iterate/module/osb.px.h:125-219,
//   Y = V (EIG0 - E0)^-1 V^T X,
// V^T applied for levels sF..1 and V for levels 1..sF with the per-block
// rotations of src/iterate/libiterate/hpx/hpx.f90:40-73 (VX / VTX).
//
// Level 1 (the innermost, NIN = 1) holds one n1 x n1 rotation per (R, theta)
// block: VOSB level 1 is N*n1 values, the largest stream of the whole
// iteration.  V1^T, the divide and V1 are adjacent in the sequence, so one
// kernel applies all three per block: VOSB level 1 is read once per P*X
// instead of twice, and once for all vectors of a batch.
#include "scalit_impl.hpp"

namespace scalit {

__device__ __forceinline__ float bf16_to_f32_px(bf16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ __forceinline__ bf16_t f32_to_bf16_px(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
template <typename T, typename S>
__device__ __forceinline__ T pld(const S *p, long i) { return (T)p[i]; }
template <>
__device__ __forceinline__ float pld<float, bf16_t>(const bf16_t *p, long i) { return bf16_to_f32_px(p[i]); }
template <typename T, typename S>
__device__ __forceinline__ void pst(S *p, long i, T v) { p[i] = (S)v; }
template <>
__device__ __forceinline__ void pst<float, bf16_t>(bf16_t *p, long i, float v) { p[i] = f32_to_bf16_px(v); }

__device__ __forceinline__ bool px_active(const int *active, int v) { return active == nullptr || active[v] != 0; }

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

// Y(k,i,b) = sum_j V_b(j,i) X(k,j,b)  (transpose, VTX)  or  V_b(i,j) (VX)
template <typename T, typename S, typename SO>
__global__ void k_rotate_generic(int nin, int nout, int blk, const T *__restrict__ vosb, bool transpose,
                                 const S *__restrict__ x, long ldx, SO *__restrict__ y, long ldy, const int *active)
{
    int v = blockIdx.y;
    if (!px_active(active, v)) return;
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
            for (int j = 0; j < nout; j++) acc += vb[j + (long)nout * i] * pld<T, S>(xb, (long)nin * j);
        else
            for (int j = 0; j < nout; j++) acc += vb[i + (long)nout * j] * pld<T, S>(xb, (long)nin * j);
        pst<T, SO>(yv, idx, acc);
    }
}

// Level 1 fused: per block b (one workgroup), for every vector:
//   z = V_b^T y_b ;  z /= kernel(EIG0 - E0) ;  y_b = V_b z
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
    T *sY = sV + (long)n1 * ld;         // n1 x VCH
    T *sZ = sY + n1 * VCH;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const T *vb = vosb + (long)b * n1 * n1;
    for (int e = tid; e < n1 * n1; e += 256) {
        int i = e % n1, j = e / n1;
        sV[i + ld * j] = vb[e];
    }
    const T *eb = eig0 + (long)b * n1;
    for (int v0 = 0; v0 < nvec; v0 += VCH) {
        const int nv = min(VCH, nvec - v0);
        __syncthreads();
        for (int e = tid; e < n1 * nv; e += 256) {
            int i = e % n1, vv = e / n1;
            int v = v0 + vv;
            sY[i + n1 * vv] = px_active(active, v) ? pld<T, S>(x + (long)v * ldx + (long)b * n1, i) : (T)0;
        }
        __syncthreads();
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
            if (!px_active(active, v)) continue;
            T acc = 0;
            for (int j = 0; j < n1; j++) acc += sV[i + ld * j] * sZ[j + n1 * vv];
            pst<T, SO>(y + (long)v * ldy + (long)b * n1, i, acc);
        }
    }
}

static int pgrid(long n)
{
    long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    return (int)(g < 1 ? 1 : g);
}

template <typename T, typename S>
static int px_impl(scalit_ctx *c, int nosb, const T *const *vosb, const T *eig0, T *work, const T *shifts, T shift0,
                   T de, T beta, int nvec, const S *x, long ldx, S *y, long ldy, const int *active, hipStream_t s)
{
    const long N = c->N;
    const int sF = c->sF;
    // work holds two ping-pong buffers of nvec * N in the compute type
    T *w0 = work, *w1 = work + (long)nvec * N;
    dim3 g(pgrid(N), nvec);
    const T *cur = nullptr;
    long ldc = N;
    // V^T for levels sF..2 (1-based): first reads the input vectors
    bool first = true;
    T *dst = w0;
    for (int l = sF - 1; l >= 1; l--) {
        const Level &L = c->lv[l];
        if (first)
            hipLaunchKernelGGL((k_rotate_generic<T, S, T>), g, dim3(256), 0, s, L.dim, L.n, L.blk, vosb[l], true, x,
                               ldx, dst, N, active);
        else
            hipLaunchKernelGGL((k_rotate_generic<T, T, T>), g, dim3(256), 0, s, L.dim, L.n, L.blk, vosb[l], true, cur,
                               ldc, dst, N, active);
        first = false;
        cur = dst;
        ldc = N;
        dst = (dst == w0) ? w1 : w0;
    }
    // level 1 fused V1^T, kernel, V1
    const int n1 = c->lv[0].n, nb = c->lv[0].blk;
    size_t lds = ((size_t)n1 * (n1 + 1) + 2 * (size_t)n1 * 8) * sizeof(T);
    if (sF == 1) {
        // a single level: input and output are the caller's vectors
        hipLaunchKernelGGL((k_px_level1<T, S, S>), dim3(nb), dim3(256), lds, s, n1, nb, nvec, vosb[0], eig0, shifts,
                           shift0, nosb, de, beta, x, ldx, y, ldy, active);
        HIP_TRY(hipGetLastError());
        return SCALIT_OK;
    }
    hipLaunchKernelGGL((k_px_level1<T, T, T>), dim3(nb), dim3(256), lds, s, n1, nb, nvec, vosb[0], eig0, shifts,
                       shift0, nosb, de, beta, cur, ldc, dst, N, active);
    cur = dst;
    dst = (dst == w0) ? w1 : w0;
    // V for levels 2..sF; the last writes the output vectors
    for (int l = 1; l < sF; l++) {
        const Level &L = c->lv[l];
        if (l == sF - 1)
            hipLaunchKernelGGL((k_rotate_generic<T, T, S>), g, dim3(256), 0, s, L.dim, L.n, L.blk, vosb[l], false, cur,
                               N, y, ldy, active);
        else {
            hipLaunchKernelGGL((k_rotate_generic<T, T, T>), g, dim3(256), 0, s, L.dim, L.n, L.blk, vosb[l], false, cur,
                               N, dst, N, active);
            cur = dst;
            dst = (dst == w0) ? w1 : w0;
        }
    }
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

static bool px_attr_done = false;
static void px_attrs()
{
    if (px_attr_done) return;
    const int m = 160 * 1024;
    (void)hipFuncSetAttribute((const void *)k_px_level1<double, double, double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, m);
    (void)hipFuncSetAttribute((const void *)k_px_level1<float, float, float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, m);
    (void)hipFuncSetAttribute((const void *)k_px_level1<float, bf16_t, bf16_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, m);
    px_attr_done = true;
}

int px_apply_f64(scalit_ctx *c, int nosb, const double *shifts_dev, double shift0, int nvec, const double *x,
                 long ldx, double *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    px_attrs();
    {
        double *old = c->work.p;
        SC_TRY(c->work.alloc((size_t)2 * nvec * c->N));
        if (c->work.p != old) c->gen++;
    }
    const double *vosb[SC_FMAX];
    for (int l = 0; l < c->sF; l++) vosb[l] = c->lv[l].vosb.p;
    return px_impl<double, double>(c, nosb, vosb, c->eig0.p, c->work.p, shifts_dev, shift0, c->de, c->beta, nvec, x,
                                   ldx, y, ldy, active, s);
}

int px_apply_f32(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx, float *y,
                 long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    px_attrs();
    SC_TRY(precond_to_f32(c));
    {
        float *old = c->work_f.p;
        SC_TRY(c->work_f.alloc((size_t)2 * nvec * c->N));
        if (c->work_f.p != old) c->gen++;
    }
    const float *vosb[SC_FMAX];
    for (int l = 0; l < c->sF; l++) vosb[l] = c->lv[l].vosb_f.p;
    return px_impl<float, float>(c, SCALIT_TOSB, vosb, c->eig0_f.p, c->work_f.p, shifts_dev, shift0, (float)c->de,
                                 (float)c->beta, nvec, x, ldx, y, ldy, active, s);
}

int px_apply_bf16(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const bf16_t *x, long ldx,
                  bf16_t *y, long ldy, const int *active, hipStream_t s)
{
    if (!c->has_precond) return SCALIT_ESTATE;
    px_attrs();
    SC_TRY(precond_to_f32(c));
    {
        float *old = c->work_f.p;
        SC_TRY(c->work_f.alloc((size_t)2 * nvec * c->N));
        if (c->work_f.p != old) c->gen++;
    }
    const float *vosb[SC_FMAX];
    for (int l = 0; l < c->sF; l++) vosb[l] = c->lv[l].vosb_f.p;
    return px_impl<float, bf16_t>(c, SCALIT_TOSB, vosb, c->eig0_f.p, c->work_f.p, shifts_dev, shift0, (float)c->de,
                                  (float)c->beta, nvec, x, ldx, y, ldy, active, s);
}

}  // namespace scalit
