// Vector kernels of the Krylov solvers.
//
// Reductions are deterministic: a kernel writes NPART per-block partial sums
// per vector (fixed grid, fixed tree inside a block) and every consumer adds
// the partials in index order.  The QMR recurrence of
// src/iterate/libiterate/qmr/qmr.f90:5-175 is split into fused
// reduce-then-update kernels whose scalars live in device memory (struct
// QmrScal), so that an iteration needs no host round trip and can be recorded
// once in a HIP graph.  Every kernel of an iteration skips vectors whose
// status is no longer "running", which reproduces the reference's `exit`
// semantics when more iterations are replayed than a vector needs.
#include "scalit_impl.hpp"
#include "solver_state.hpp"

namespace scalit {

__device__ __forceinline__ float v_bf16_to_f32(bf16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ __forceinline__ bf16_t v_f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
template <typename S>
__device__ __forceinline__ double lv(const S *p, long i) { return (double)p[i]; }
template <>
__device__ __forceinline__ double lv<bf16_t>(const bf16_t *p, long i) { return (double)v_bf16_to_f32(p[i]); }
template <typename S>
__device__ __forceinline__ void sv(S *p, long i, double v) { p[i] = (S)v; }
template <>
__device__ __forceinline__ void sv<bf16_t>(bf16_t *p, long i, double v) { p[i] = v_f32_to_bf16((float)v); }

// deterministic block reduction over 256 threads (warp shuffles + LDS)
__device__ __forceinline__ double bsum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) r = (red[0] + red[1]) + (red[2] + red[3]);
    return r;   // valid on thread 0
}

// partials per vector written by a grid sized for vectors of length n (the
// host's nparts); the slots past them hold zeros
__device__ __forceinline__ int nparts_dev(long n)
{
    const long g = (n + 1023) / 1024;
    return (int)(g < 1 ? 1 : (g > NPART ? NPART : g));
}

// The sum of one vector's partials, by a whole wave (all 64 lanes call it
// together): lane l adds partials l, l+64, ... in order, then an xor butterfly
// combines the lanes.  IEEE addition is commutative, so every lane -- and
// every wave and workgroup that sums the same partials -- gets the identical
// value; summing the zero tail as well changes nothing.
__device__ __forceinline__ double wave_sum_parts(const double *p, int g)
{
    const int lane = threadIdx.x & 63;
    double s = 0;
    for (int i = lane; i < g; i += 64) s += p[i];
    for (int o = 1; o < 64; o <<= 1) s += __shfl_xor(s, o);
    return s;
}
__device__ __forceinline__ double sum_parts_n(const double *p, long n) { return wave_sum_parts(p, nparts_dev(n)); }
__device__ __forceinline__ double sum_parts(const double *p) { return wave_sum_parts(p, NPART); }

#define GSTRIDE(i, n) for (long i = blockIdx.x * 256L + threadIdx.x; i < (n); i += (long)gridDim.x * 256L)

// The QMR kernels walk the same elements in the same order as GSTRIDE, in
// chunks of QU per thread: every input of a chunk is loaded (from an index
// clamped into the vector) before any output of it is stored, so a thread has
// QU loads of each vector in flight.  Written element by element, each load
// would wait for the store before it (the compiler cannot tell that the
// vectors do not overlap), which serialises a thread's memory round trips.
constexpr int QU = 4;
#define QCHUNK(base, n) for (long base = blockIdx.x * 256L + threadIdx.x; base < (n); base += (long)gridDim.x * 256L * QU)
#define QELEM(t, tc, base, e, n)                                                                                  \
    const long t = (base) + (long)(e) * gridDim.x * 256L;                                                         \
    const long tc = t < (n) ? t : (n) - 1
// the value a store to S reads back as
template <typename S>
__device__ __forceinline__ double rt(double v)
{
    S tmp;
    sv<S>(&tmp, 0, v);
    return lv<S>(&tmp, 0);
}

// A grid of nparts(n) <= NPART workgroups writes partials 0..gridDim.x-1;
// workgroup 0 zeroes the rest, so every consumer still adds all NPART slots in
// index order (deterministic) while small vectors launch few workgroups.
__device__ __forceinline__ void zero_tail(double *pv)
{
    if (blockIdx.x == 0)
        for (int i = gridDim.x + threadIdx.x; i < NPART; i += blockDim.x) pv[i] = 0.0;
}
static unsigned nparts(long n)
{
    const long g = (n + 1023) / 1024;
    return (unsigned)(g < 1 ? 1 : (g > NPART ? NPART : g));
}

// ---------------------------------------------------------------- generic
__global__ void __launch_bounds__(256) k_dot_part(long n, const double *x, long ldx, const double *y, long ldy,
                                                  double *part)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    const double *xv = x + (long)v * ldx, *yv = y + (long)v * ldy;
    double s = 0;
    GSTRIDE(i, n) s += xv[i] * yv[i];
    s = bsum(s, red);
    if (threadIdx.x == 0) part[(long)v * NPART + blockIdx.x] = s;
    zero_tail(part + (long)v * NPART);
}

// one wave per vector
__global__ void k_reduce_parts(int nvec, const double *part, double *out)
{
    const int v = blockIdx.x;
    const double r = sum_parts(part + (long)v * NPART);
    if (threadIdx.x == 0) out[v] = r;
}

void launch_dot_partials(int nvec, long n, const double *x, long ldx, const double *y, long ldy, double *part,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_part, dim3(nparts(n), nvec), dim3(256), 0, s, n, x, ldx, y, ldy, part);
}

void launch_reduce_partials(int nvec, const double *part, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_parts, dim3(nvec), dim3(64), 0, s, nvec, part, out);
}

int dot_host(int nvec, long n, const double *x, long ldx, const double *y, long ldy, double *out_host,
             double *part_dev, double *red_dev, hipStream_t s)
{
    launch_dot_partials(nvec, n, x, ldx, y, ldy, part_dev, s);
    launch_reduce_partials(nvec, part_dev, red_dev, s);
    HIP_TRY(hipMemcpyAsync(out_host, red_dev, sizeof(double) * nvec, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

// y = a x + b y  (per-vector scalars on the host side: a, b broadcast)
__global__ void k_axpby(long n, double a, const double *x, long ldx, double b, double *y, long ldy)
{
    const int v = blockIdx.y;
    GSTRIDE(i, n) y[(long)v * ldy + i] = a * x[(long)v * ldx + i] + b * y[(long)v * ldy + i];
}

// casts between storage types
template <typename A, typename B>
__global__ void k_cast(long n, const A *x, long ldx, B *y, long ldy)
{
    const int v = blockIdx.y;
    GSTRIDE(i, n) sv<B>(y + (long)v * ldy, i, lv<A>(x + (long)v * ldx, i));
}

void launch_cast_d2f(int nvec, long n, const double *x, long ldx, float *y, long ldy, hipStream_t s)
{
    hipLaunchKernelGGL((k_cast<double, float>), dim3(nparts(n), nvec), dim3(256), 0, s, n, x, ldx, y, ldy);
}
void launch_cast_f2d(int nvec, long n, const float *x, long ldx, double *y, long ldy, hipStream_t s)
{
    hipLaunchKernelGGL((k_cast<float, double>), dim3(nparts(n), nvec), dim3(256), 0, s, n, x, ldx, y, ldy);
}
void launch_cast_d2b(int nvec, long n, const double *x, long ldx, bf16_t *y, long ldy, hipStream_t s)
{
    hipLaunchKernelGGL((k_cast<double, bf16_t>), dim3(nparts(n), nvec), dim3(256), 0, s, n, x, ldx, y, ldy);
}
void launch_cast_b2d(int nvec, long n, const bf16_t *x, long ldx, double *y, long ldy, hipStream_t s)
{
    hipLaunchKernelGGL((k_cast<bf16_t, double>), dim3(nparts(n), nvec), dim3(256), 0, s, n, x, ldx, y, ldy);
}
void launch_axpby(int nvec, long n, double a, const double *x, long ldx, double b, double *y, long ldy,
                  hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, dim3(nparts(n), nvec), dim3(256), 0, s, n, a, x, ldx, b, y, ldy);
}

// ---------------------------------------------------------------- QMR
// Vectors of one QMR: nvec columns each of R, UX, UR, Q, S, V, Z, HX0 with
// leading dimension ld (storage type S); B and X are the caller's (leading
// dimension ldb / ldx, storage type S).
//
// V and Z are stored as the operators leave them: the reference divides V by
// VNRM and Z by ZVRAT and then by ZNRM in place (qmr.f90: the VNRM / ZNRM
// scaling at the top of the loop and Z = Z / ZVRAT after the preconditioner);
// here the kernels that read them apply those divisions to the loaded values
// instead, which saves the two passes that only rescaled them.  V.Z is summed
// while Z is read for Z.Z, as (V/VNRMPO).(Z/ZVRAT), and divided by ZNRM once
// ZNRM is known.  ZVRAT_NEXT holds the ratio of the running iteration (the
// reference's ZVRAT after its update); the end of the iteration rolls it into
// ZVRAT.

__device__ __forceinline__ bool running(const QmrScal *q, int v) { return q[v].status == QMR_RUNNING; }

// init part 1: R = V = B, X = 0 (ntype <= 0); partial B.B
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_init(long n, int zero_x, QmrVecs<S> w, double *part)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    const S *b = w.b + (long)v * w.ldb;
    S *x = w.x + (long)v * w.ldx;
    S *R = w.R + (long)v * w.ld, *V = w.V + (long)v * w.ld;
    double s = 0;
    QCHUNK(base, n)
    {
        S bv[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            bv[e] = b[tc];
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) {
                R[t] = bv[e];
                V[t] = bv[e];
                if (zero_x) x[t] = (S)0;
                const double d = lv<S>(bv, e);
                s += d * d;
            }
        }
    }
    s = bsum(s, red);
    if (threadIdx.x == 0) part[(long)v * NPART + blockIdx.x] = s;
    zero_tail(part + (long)v * NPART);
}

// init part 2 (after Z = P B): partials of Z.Z and (V/VNRM).Z (ZVRAT = 1)
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_init_z(long n, QmrVecs<S> w, const double *part_bb, double *part,
                                                    double *part_vz)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    const S *Z = w.Z + (long)v * w.ld, *V = w.V + (long)v * w.ld;
    const double vnrm = sqrt(sum_parts(part_bb + (long)v * NPART));
    double s = 0, svz = 0;
    GSTRIDE(i, n)
    {
        double d = lv<S>(Z, i);
        s += d * d;
        svz += (lv<S>(V, i) / vnrm) * d;
    }
    s = bsum(s, red);
    if (threadIdx.x == 0) part[(long)v * NPART + blockIdx.x] = s;
    zero_tail(part + (long)v * NPART);
    svz = bsum(svz, red);
    if (threadIdx.x == 0) part_vz[(long)v * NPART + blockIdx.x] = svz;
    zero_tail(part_vz + (long)v * NPART);
}

// init scalars: BNRM, VNRM, ZNRM, GAMAMO=1, ETA=-1, ZVRAT=1 (one wave per vector)
__global__ void k_qmr_init_scal(int nvec, const double *part_bb, const double *part_zz, QmrScal *q, int *active,
                                int qmrmax)
{
    const int v = blockIdx.x;
    const double bb = sum_parts(part_bb + (long)v * NPART), zz = sum_parts(part_zz + (long)v * NPART);
    if (threadIdx.x != 0) return;
    QmrScal &s = q[v];
    s.bnrm = sqrt(bb);
    s.vnrm = s.bnrm;
    s.znrm = sqrt(zz);
    s.gamamo = 1.0;
    s.eta = -1.0;
    s.zvrat = 1.0;
    s.thetamo = 0.0;
    s.qsdot = 0.0;
    s.ret = 0;
    s.iter = 1;
    s.res = 0.0;
    s.status = (qmrmax >= 1) ? QMR_RUNNING : QMR_EXIT;
    active[v] = s.status == QMR_RUNNING;
}

// k2 + k3: exit check on VNRM / ZNRM, ZVRAT_NEXT = ZVRAT ZNRM / VNRM,
// VZDOT = (V/VNRM).(Z/ZVRAT) / ZNRM from the partials of the previous step,
// breakdown check, Q = Z/ZVRAT/ZNRM - (VNRM VZDOT / QSDOT) Q
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_b(long n, QmrVecs<S> w, QmrScal *q, int *active, const double *part_vz)
{
    const int v = blockIdx.y;
    if (!running(q, v)) return;
    const double vnrm = q[v].vnrm, znrm = q[v].znrm, zvrat = q[v].zvrat;
    if (vnrm == 0.0 || znrm == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            q[v].status = QMR_EXIT;
            active[v] = 0;
        }
        return;
    }
    const double vzdot = sum_parts_n(part_vz + (long)v * NPART, n) / znrm;
    if (vzdot == 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            q[v].status = QMR_EXIT;
            active[v] = 0;
        }
        return;
    }
    S *Q = w.Q + (long)v * w.ld;
    const S *Z = w.Z + (long)v * w.ld;
    const bool first = q[v].iter == 1;
    const double c = first ? 0.0 : vnrm * vzdot / q[v].qsdot;
    QCHUNK(base, n)
    {
        double z[QU], qo[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            z[e] = lv<S>(Z, tc);
            qo[e] = first ? 0.0 : lv<S>(Q, tc);
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) sv<S>(Q, t, first ? z[e] / zvrat / znrm : z[e] / zvrat / znrm - c * qo[e]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        q[v].vzdot = vzdot;
        q[v].zvrat_next = zvrat * znrm / vnrm;
    }
}

// k4: S = ZVRAT * S, partial Q.S
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_c(long n, QmrVecs<S> w, const QmrScal *q, double *part)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    if (!running(q, v)) return;
    const double zvrat = q[v].zvrat_next;
    S *Sv = w.Sv + (long)v * w.ld;
    const S *Q = w.Q + (long)v * w.ld;
    double s = 0;
    QCHUNK(base, n)
    {
        double sv0[QU], qv[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            sv0[e] = lv<S>(Sv, tc);
            qv[e] = lv<S>(Q, tc);
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) {
                const double a = zvrat * sv0[e];
                sv<S>(Sv, t, a);
                s += qv[e] * rt<S>(a);
            }
        }
    }
    s = bsum(s, red);
    if (threadIdx.x == 0) part[(long)v * NPART + blockIdx.x] = s;
    zero_tail(part + (long)v * NPART);
}

// k5: QSDOT, BETA, breakdown, V = S - BETA V/VNRM, partial V.V
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_d(long n, QmrVecs<S> w, QmrScal *q, int *active, const double *part_in,
                                               double *part)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    if (!running(q, v)) return;
    const double qsdot = sum_parts_n(part_in + (long)v * NPART, n);
    const double beta = qsdot / q[v].vzdot;
    if (qsdot == 0.0 || beta == 0.0) {
        __syncthreads();
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            q[v].status = QMR_EXIT;
            active[v] = 0;
        }
        return;
    }
    const double vnrm = q[v].vnrm;
    S *V = w.V + (long)v * w.ld;
    const S *Sv = w.Sv + (long)v * w.ld;
    double s = 0;
    QCHUNK(base, n)
    {
        double sv0[QU], vv[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            sv0[e] = lv<S>(Sv, tc);
            vv[e] = lv<S>(V, tc);
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) {
                const double a = sv0[e] - beta * (vv[e] / vnrm);
                sv<S>(V, t, a);
                const double ar = rt<S>(a);
                s += ar * ar;
            }
        }
    }
    s = bsum(s, red);
    zero_tail(part + (long)v * NPART);
    if (threadIdx.x == 0) {
        part[(long)v * NPART + blockIdx.x] = s;
        if (blockIdx.x == 0) {
            q[v].qsdot = qsdot;
            q[v].beta = beta;
        }
    }
}

// k6 (after Z = P V): VNRMPO; partials of (Z/ZVRAT).(Z/ZVRAT) and
// (V/VNRMPO).(Z/ZVRAT), V and Z left as they are
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_e(long n, QmrVecs<S> w, QmrScal *q, const double *part_vv, double *part,
                                               double *part_vz)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    if (!running(q, v)) return;
    const double zvrat = q[v].zvrat_next;
    const double vnrmpo = sqrt(sum_parts_n(part_vv + (long)v * NPART, n));
    const S *Z = w.Z + (long)v * w.ld, *V = w.V + (long)v * w.ld;
    double s = 0, svz = 0;
    QCHUNK(base, n)
    {
        double zv[QU], vv[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            zv[e] = lv<S>(Z, tc);
            vv[e] = lv<S>(V, tc);
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) {
                const double z = zv[e] / zvrat;
                s += z * z;
                svz += (vv[e] / vnrmpo) * z;
            }
        }
    }
    s = bsum(s, red);
    zero_tail(part + (long)v * NPART);
    if (threadIdx.x == 0) part[(long)v * NPART + blockIdx.x] = s;
    svz = bsum(svz, red);
    zero_tail(part_vz + (long)v * NPART);
    if (threadIdx.x == 0) {
        part_vz[(long)v * NPART + blockIdx.x] = svz;
        if (blockIdx.x == 0) q[v].vnrmpo = vnrmpo;
    }
}

// The end of an iteration (residual, convergence, roll of the two-level
// variables), one wave per vector.  It stays a kernel of its own: ending the
// iteration in the last workgroup of the kernel before it would need a
// device-scope release and acquire per workgroup, which on this GPU write back
// and invalidate the XCD's L2 each time (measured: +0.5 ms per iteration at
// 16 vectors of 491,520 points).
struct QmrEnd {
    int conv, qmrmax;
    double tol;
    const double *hh, *rr, *uxux, *xx, *urur;
};

// one wave; gama, theta, eta, znrm: this iteration's values (k_qmr_f)
__device__ void qmr_end_vector(long n, const QmrEnd &e, QmrScal *q, int *active, int v, double gama, double theta,
                               double eta, double znrm)
{
    QmrScal &s = q[v];
    double res;
    if (e.conv == 1) {
        double xnrm = sqrt(sum_parts_n(e.hh + (long)v * NPART, n));
        res = sqrt(sum_parts_n(e.rr + (long)v * NPART, n)) / xnrm;
    } else if (e.conv == 2) {
        double rr = sum_parts_n(e.rr + (long)v * NPART, n);
        double r1 = sqrt(rr) / s.bnrm;
        double r2 = sqrt(sum_parts_n(e.uxux + (long)v * NPART, n) / sum_parts_n(e.xx + (long)v * NPART, n));
        double r3 = sqrt(sum_parts_n(e.urur + (long)v * NPART, n) / rr);
        res = r1 > r2 ? r1 : r2;
        res = res > r3 ? res : r3;
    } else {
        res = sqrt(sum_parts_n(e.rr + (long)v * NPART, n)) / s.bnrm;
    }
    if (threadIdx.x != 0) return;
    s.res = res;
    if (fabs(res) < e.tol) {
        s.ret = s.iter;
        s.status = QMR_CONVERGED;
        active[v] = 0;
        return;
    }
    s.ret = -s.iter;
    s.zvrat = s.zvrat_next;
    s.vnrm = s.vnrmpo;
    s.gamamo = gama;
    s.thetamo = theta;
    s.eta = eta;
    s.znrm = znrm;
    s.iter += 1;
    if (s.iter > e.qmrmax) {
        s.status = QMR_EXIT;
        active[v] = 0;
    }
}

// k7: ZNRM, THETA, GAMA, ETA, breakdown on GAMA; UX, UR, X, R updates and the
// partial sums the convergence test of type |nType| = 0 / 2 needs.
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_f(long n, int conv, QmrVecs<S> w, QmrScal *q, int *active,
                                               const double *part_zz, double *part_rr, double *part_uxux,
                                               double *part_xx, double *part_urur)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    if (!running(q, v)) return;
    const QmrScal &o = q[v];
    const double znrm = sqrt(sum_parts_n(part_zz + (long)v * NPART, n));
    const double theta = o.vnrmpo / (o.gamamo * fabs(o.beta));
    const double gama = 1.0 / sqrt(1.0 + theta * theta);
    const double eta = -o.eta * o.vnrm * gama * gama / (o.beta * o.gamamo * o.gamamo);
    if (gama == 0.0) {
        __syncthreads();
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            q[v].status = QMR_EXIT;
            active[v] = 0;
        }
        return;
    }
    const bool first = o.iter == 1;
    const double a = eta * o.zvrat_next;
    const double tg = (o.thetamo * gama) * (o.thetamo * gama);
    S *UX = w.UX + (long)v * w.ld, *UR = w.UR + (long)v * w.ld, *R = w.R + (long)v * w.ld;
    const S *Q = w.Q + (long)v * w.ld, *Sv = w.Sv + (long)v * w.ld;
    S *X = w.x + (long)v * w.ldx;
    double srr = 0, sux = 0, sxx = 0, sur = 0;
    QCHUNK(base, n)
    {
        double qv[QU], sv0[QU], uxo[QU], uro[QU], xo[QU], ro[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            qv[e] = lv<S>(Q, tc);
            sv0[e] = lv<S>(Sv, tc);
            uxo[e] = first ? 0.0 : lv<S>(UX, tc);
            uro[e] = first ? 0.0 : lv<S>(UR, tc);
            xo[e] = lv<S>(X, tc);
            ro[e] = lv<S>(R, tc);
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) {
                const double ux = first ? a * qv[e] : a * qv[e] + tg * uxo[e];
                const double ur = first ? eta * sv0[e] : eta * sv0[e] + tg * uro[e];
                sv<S>(UX, t, ux);
                sv<S>(UR, t, ur);
                const double uxr = rt<S>(ux), urr = rt<S>(ur);
                const double xn = xo[e] + uxr, rn = ro[e] - urr;
                sv<S>(X, t, xn);
                sv<S>(R, t, rn);
                if (conv != 1) {
                    const double r = rt<S>(rn);
                    srr += r * r;
                    if (conv == 2) {
                        const double xx = rt<S>(xn);
                        sux += uxr * uxr;
                        sxx += xx * xx;
                        sur += urr * urr;
                    }
                }
            }
        }
    }
    if (conv != 1) {
        srr = bsum(srr, red);
        if (threadIdx.x == 0) part_rr[(long)v * NPART + blockIdx.x] = srr;
        zero_tail(part_rr + (long)v * NPART);
        if (conv == 2) {
            sux = bsum(sux, red);
            if (threadIdx.x == 0) part_uxux[(long)v * NPART + blockIdx.x] = sux;
            zero_tail(part_uxux + (long)v * NPART);
            sxx = bsum(sxx, red);
            if (threadIdx.x == 0) part_xx[(long)v * NPART + blockIdx.x] = sxx;
            zero_tail(part_xx + (long)v * NPART);
            sur = bsum(sur, red);
            if (threadIdx.x == 0) part_urur[(long)v * NPART + blockIdx.x] = sur;
            zero_tail(part_urur + (long)v * NPART);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        q[v].znrm_next = znrm;
        q[v].theta = theta;
        q[v].gama = gama;
        q[v].eta_next = eta;
    }
}

// k8 (conv type 1, after HX0 = H X): partial HX0.HX0 and (HX0-B).(HX0-B)
template <typename S>
__global__ void __launch_bounds__(256) k_qmr_g(long n, QmrVecs<S> w, const QmrScal *q, double *part_hh,
                                               double *part_rr)
{
    __shared__ double red[4];
    const int v = blockIdx.y;
    if (!running(q, v)) return;
    const S *H = w.HX0 + (long)v * w.ld;
    const S *b = w.b + (long)v * w.ldb;
    double s1 = 0, s2 = 0;
    QCHUNK(base, n)
    {
        double hv[QU], bv[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            hv[e] = lv<S>(H, tc);
            bv[e] = lv<S>(b, tc);
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) {
                const double r = hv[e] - bv[e];
                s1 += hv[e] * hv[e];
                s2 += r * r;
            }
        }
    }
    s1 = bsum(s1, red);
    if (threadIdx.x == 0) part_hh[(long)v * NPART + blockIdx.x] = s1;
    zero_tail(part_hh + (long)v * NPART);
    s2 = bsum(s2, red);
    if (threadIdx.x == 0) part_rr[(long)v * NPART + blockIdx.x] = s2;
    zero_tail(part_rr + (long)v * NPART);
}

// k9: the end of the iteration
__global__ void k_qmr_h(long n, QmrEnd e, QmrScal *q, int *active)
{
    const int v = blockIdx.x;
    if (q[v].status != QMR_RUNNING) return;
    qmr_end_vector(n, e, q, active, v, q[v].gama, q[v].theta, q[v].eta_next, q[v].znrm_next);
}

template <typename S>
void launch_qmr_init(int nvec, long n, int zero_x, const QmrVecs<S> &w, double *part, hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_init<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, zero_x, w, part);
}
template <typename S>
void launch_qmr_init_z(int nvec, long n, const QmrVecs<S> &w, const double *pbb, double *part, double *pvz,
                       hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_init_z<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, w, pbb, part, pvz);
}
void launch_qmr_init_scal(int nvec, const double *pbb, const double *pzz, QmrScal *q, int *active, int qmrmax,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_qmr_init_scal, dim3(nvec), dim3(64), 0, s, nvec, pbb, pzz, q, active, qmrmax);
}
template <typename S>
void launch_qmr_b(int nvec, long n, const QmrVecs<S> &w, QmrScal *q, int *active, const double *part, hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_b<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, w, q, active, part);
}
template <typename S>
void launch_qmr_c(int nvec, long n, const QmrVecs<S> &w, const QmrScal *q, double *part, hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_c<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, w, q, part);
}
template <typename S>
void launch_qmr_d(int nvec, long n, const QmrVecs<S> &w, QmrScal *q, int *active, const double *pin, double *part,
                  hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_d<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, w, q, active, pin, part);
}
template <typename S>
void launch_qmr_e(int nvec, long n, const QmrVecs<S> &w, QmrScal *q, const double *pvv, double *part, double *pvz,
                  hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_e<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, w, q, pvv, part, pvz);
}
template <typename S>
void launch_qmr_f(int nvec, long n, int conv, const QmrVecs<S> &w, QmrScal *q, int *active, const double *pzz,
                  double *prr, double *pux, double *pxx, double *pur, hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_f<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, conv, w, q, active, pzz, prr, pux, pxx,
                       pur);
}
template <typename S>
void launch_qmr_g(int nvec, long n, const QmrVecs<S> &w, const QmrScal *q, double *phh, double *prr, hipStream_t s)
{
    hipLaunchKernelGGL((k_qmr_g<S>), dim3(nparts(n), nvec), dim3(256), 0, s, n, w, q, phh, prr);
}
void launch_qmr_h(long n, int nvec, int conv, int qmrmax, double tol, QmrScal *q, int *active, const double *phh,
                  const double *prr, const double *pux, const double *pxx, const double *pur, hipStream_t s)
{
    hipLaunchKernelGGL(k_qmr_h, dim3(nvec), dim3(64), 0, s, n, QmrEnd{conv, qmrmax, tol, phh, prr, pux, pxx, pur}, q,
                       active);
}

#define QMR_INST(S)                                                                                               \
    template void launch_qmr_init<S>(int, long, int, const QmrVecs<S> &, double *, hipStream_t);                  \
    template void launch_qmr_init_z<S>(int, long, const QmrVecs<S> &, const double *, double *, double *,          \
                                       hipStream_t);                                                             \
    template void launch_qmr_b<S>(int, long, const QmrVecs<S> &, QmrScal *, int *, const double *, hipStream_t);  \
    template void launch_qmr_c<S>(int, long, const QmrVecs<S> &, const QmrScal *, double *, hipStream_t);         \
    template void launch_qmr_d<S>(int, long, const QmrVecs<S> &, QmrScal *, int *, const double *, double *,      \
                                  hipStream_t);                                                                  \
    template void launch_qmr_e<S>(int, long, const QmrVecs<S> &, QmrScal *, const double *, double *, double *,  \
                                  hipStream_t);                                                                  \
    template void launch_qmr_f<S>(int, long, int, const QmrVecs<S> &, QmrScal *, int *, const double *, double *, \
                                  double *, double *, double *, hipStream_t);                                    \
    template void launch_qmr_g<S>(int, long, const QmrVecs<S> &, const QmrScal *, double *, double *, hipStream_t);
QMR_INST(double)
QMR_INST(float)
QMR_INST(bf16_t)

// ---------------------------------------------------------------- MGS / PIST
// One modified Gram-Schmidt step (src/iterate/libiterate/orth/gsorth.f90:60-63)
// for nw independent windows: w -= d_i v_i with d_i from the partials of the
// previous launch, then the partials of w . v_{i+1}.  Window k's basis is
// vj + k * ldw (columns of length n, stride ldv); w_k at w + k * ldwk.
__global__ void __launch_bounds__(256) k_mgs_step(long n, int i, int inext, const double *vj, long ldv, long ldw,
                                                  double *w, long ldwk, const double *part_in, double *part_out,
                                                  const int *active)
{
    __shared__ double red[4];
    const int k = blockIdx.y;
    if (active && !active[k]) return;
    const double *V = vj + (long)k * ldw;
    double *W = w + (long)k * ldwk;
    double s = 0;
    if (i >= 0) {
        const double d = sum_parts_n(part_in + (long)k * NPART, n);
        const double *vi = V + (long)i * ldv;
        // with no next column the partial is of the updated W with itself
        const double *vn = inext >= 0 ? V + (long)inext * ldv : nullptr;
        QCHUNK(base, n)
        {
            double wv[QU], iv[QU], nv[QU];
#pragma unroll
            for (int e = 0; e < QU; e++) {
                QELEM(t, tc, base, e, n);
                wv[e] = W[tc];
                iv[e] = vi[tc];
                nv[e] = vn ? vn[tc] : 0.0;
            }
#pragma unroll
            for (int e = 0; e < QU; e++) {
                QELEM(t, tc, base, e, n);
                (void)tc;
                if (t < n) {
                    const double a = wv[e] - d * iv[e];
                    W[t] = a;
                    s += a * (vn ? nv[e] : a);
                }
            }
        }
    } else {
        const double *vn = V + (long)inext * ldv;
        GSTRIDE(t, n) s += W[t] * vn[t];
    }
    s = bsum(s, red);
    if (threadIdx.x == 0) part_out[(long)k * NPART + blockIdx.x] = s;
    zero_tail(part_out + (long)k * NPART);
}

// normalise: dst = w / ||w|| (norm from partials); status[k] = 0 when the norm is 0
__global__ void __launch_bounds__(256) k_normalize(long n, const double *w, long ldwk, double *dst, long ldd,
                                                   const double *part, int *ok, const int *active)
{
    const int k = blockIdx.y;
    if (active && !active[k]) return;
    const double nrm = sqrt(sum_parts_n(part + (long)k * NPART, n));
    if (blockIdx.x == 0 && threadIdx.x == 0 && ok) ok[k] = nrm != 0.0;
    if (nrm == 0.0) return;
    const double *W = w + (long)k * ldwk;
    double *D = dst + (long)k * ldd;
    QCHUNK(base, n)
    {
        double wv[QU];
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            wv[e] = W[tc];
        }
#pragma unroll
        for (int e = 0; e < QU; e++) {
            QELEM(t, tc, base, e, n);
            (void)tc;
            if (t < n) D[t] = wv[e] / nrm;
        }
    }
}

void launch_mgs_step(int nw, long n, int i, int inext, const double *vj, long ldv, long ldw, double *w, long ldwk,
                     const double *pin, double *pout, const int *active, hipStream_t s)
{
    hipLaunchKernelGGL(k_mgs_step, dim3(nparts(n), nw), dim3(256), 0, s, n, i, inext, vj, ldv, ldw, w, ldwk, pin, pout,
                       active);
}

void launch_normalize(int nw, long n, const double *w, long ldwk, double *dst, long ldd, const double *part, int *ok,
                      const int *active, hipStream_t s)
{
    hipLaunchKernelGGL(k_normalize, dim3(nparts(n), nw), dim3(256), 0, s, n, w, ldwk, dst, ldd, part, ok, active);
}

// MGS_ORTH (gsorth.f90:42-80) of a short vector in one workgroup per window:
// W is copied to C, its norm checked, the projections on V(:,0..jnew-1) taken
// one after the other (each dot product reduced in the workgroup, then the
// update), and C normalised; ok = 0 on a zero norm before or after.  The
// vector lives in registers (PER values per thread), so a projection is one
// read of V(:,i) and one workgroup reduction instead of a launch of its own.
// W may be C itself: all of W is read before C is written.
constexpr int MGS_NT = 1024;
template <int PER>
__global__ void __launch_bounds__(MGS_NT) k_mgs_one(long n, int jnew, const double *__restrict__ vj, long ldv,
                                                    long ldw, const double *w, long ldwk, double *col,
                                                    const int *__restrict__ active, int *__restrict__ ok,
                                                    double *__restrict__ nrm_out)
{
    __shared__ double red[2][MGS_NT / 64];
    const int k = blockIdx.x;
    if (active && !active[k]) return;
    const double *V = vj + (long)k * ldw;
    const double *W = w + (long)k * ldwk;
    double *C = col + (long)k * ldw;
    const int tid = threadIdx.x;
    int par = 0;
    // the sum over the workgroup, the same on every thread; two buffers, so
    // one barrier per reduction suffices
    auto wg_sum = [&](double v) -> double {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((tid & 63) == 0) red[par][tid >> 6] = v;
        __syncthreads();
        double r = 0;
#pragma unroll
        for (int q = 0; q < MGS_NT / 64; q++) r += red[par][q];
        par ^= 1;
        return r;
    };
    double a[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const long t = tid + (long)e * MGS_NT;
        a[e] = t < n ? W[t] : 0.0;
    }
    double s = 0;
#pragma unroll
    for (int e = 0; e < PER; e++) s += a[e] * a[e];
    const double normw = sqrt(wg_sum(s));
    if (normw != 0.0) {
        for (int i = 0; i < jnew; i++) {
            const double *vi = V + (long)i * ldv;
            double b[PER];
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const long t = tid + (long)e * MGS_NT;
                b[e] = t < n ? vi[t] : 0.0;
            }
            double d = 0;
#pragma unroll
            for (int e = 0; e < PER; e++) d += a[e] * b[e];
            d = wg_sum(d);
#pragma unroll
            for (int e = 0; e < PER; e++) a[e] -= d * b[e];
        }
        s = 0;
#pragma unroll
        for (int e = 0; e < PER; e++) s += a[e] * a[e];
    }
    const double nrm = normw != 0.0 ? sqrt(wg_sum(s)) : 0.0;
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const long t = tid + (long)e * MGS_NT;
        if (t < n) C[t] = nrm != 0.0 ? a[e] / nrm : a[e];
    }
    if (tid == 0) {
        ok[k] = nrm != 0.0;
        if (nrm_out) nrm_out[k] = nrm;
    }
}

bool launch_mgs_one(int nw, long n, int jnew, const double *vj, long ldv, long ldw, const double *w, long ldwk,
                    double *col, const int *active, int *ok, double *nrm_out, hipStream_t s)
{
    if (n <= 8L * MGS_NT)
        hipLaunchKernelGGL(k_mgs_one<8>, dim3(nw), dim3(MGS_NT), 0, s, n, jnew, vj, ldv, ldw, w, ldwk, col, active, ok,
                           nrm_out);
    else if (n <= 16L * MGS_NT)
        hipLaunchKernelGGL(k_mgs_one<16>, dim3(nw), dim3(MGS_NT), 0, s, n, jnew, vj, ldv, ldw, w, ldwk, col, active,
                           ok, nrm_out);
    else if (n <= 20L * MGS_NT)
        hipLaunchKernelGGL(k_mgs_one<20>, dim3(nw), dim3(MGS_NT), 0, s, n, jnew, vj, ldv, ldw, w, ldwk, col, active,
                           ok, nrm_out);
    else
        return false;
    return true;
}

// Multi-dot: part[k][i][blk] = partial of <V_k(:, i0+i), w_k> for i < m
// (the Rayleigh-Ritz column of pistH0, pistHij.f90:375-381).  blockIdx.z picks
// a group of MD_CG columns: a thread keeps one running sum per column of its
// group while it walks its points once, and the group's sums share one
// workgroup reduction.
constexpr int MD_CG = 8;
__global__ void __launch_bounds__(256) k_multidot(long n, int i0, int m, const double *vj, long ldv, long ldw,
                                                  const double *w, long ldwk, double *part, const int *active)
{
    __shared__ double red[4][MD_CG];
    const int k = blockIdx.y;
    if (active && !active[k]) return;
    const int c0 = blockIdx.z * MD_CG;
    const int nc = m - c0 < MD_CG ? m - c0 : MD_CG;
    const double *V = vj + (long)k * ldw + (long)(i0 + c0) * ldv;
    const double *W = w + (long)k * ldwk;
    double s[MD_CG];
#pragma unroll
    for (int c = 0; c < MD_CG; c++) s[c] = 0.0;
    GSTRIDE(t, n)
    {
        const double x = W[t];
#pragma unroll
        for (int c = 0; c < MD_CG; c++)
            if (c < nc) s[c] += V[(long)c * ldv + t] * x;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < MD_CG; c++) {
        for (int o = 32; o > 0; o >>= 1) s[c] += __shfl_down(s[c], o, 64);
        if (lane == 0) red[wv][c] = s[c];
    }
    __syncthreads();
    if (threadIdx.x < nc) {
        const int c = threadIdx.x;
        part[((long)k * m + c0 + c) * NPART + blockIdx.x] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
    }
    for (int c = 0; c < nc; c++) zero_tail(part + ((long)k * m + c0 + c) * NPART);
}

// hmat_k(i0+i, j) = hmat_k(j, i0+i) = sum of partials   (upper and lower)
__global__ void k_multidot_fin(int nw, int m, int i0, int j, const double *part, double *hmat, int ldh, long ldhk,
                               const int *active)
{
    const int t = blockIdx.x;   // one wave per entry
    const int k = t / m, i = t % m;
    if (active && !active[k]) return;
    const double s = sum_parts(part + ((long)k * m + i) * NPART);
    if (threadIdx.x != 0) return;
    double *H = hmat + (long)k * ldhk;
    H[(i0 + i) + (long)ldh * j] = s;
    H[j + (long)ldh * (i0 + i)] = s;
}

void launch_multidot(int nw, long n, int i0, int m, const double *vj, long ldv, long ldw, const double *w, long ldwk,
                     double *part, const int *active, hipStream_t s)
{
    if (m < 1) return;
    hipLaunchKernelGGL(k_multidot, dim3(nparts(n), nw, (m + MD_CG - 1) / MD_CG), dim3(256), 0, s, n, i0, m, vj, ldv, ldw,
                       w, ldwk, part, active);
}

void launch_multidot_fin(int nw, int m, int i0, int j, const double *part, double *hmat, int ldh, long ldhk,
                         const int *active, hipStream_t s)
{
    hipLaunchKernelGGL(k_multidot_fin, dim3(nw * m), dim3(64), 0, s, nw, m, i0, j, part, hmat, ldh, ldhk, active);
}

// Ritz vectors: psi(:, c) = sum_i V(:, i) Z(i, col[c])   (one window)
__global__ void k_ritz(long n, int m, int nc, const double *vj, long ldv, const double *z, int ldz, const int *cols,
                       double *psi, long ldp)
{
    const int c = blockIdx.y;
    const int col = cols[c];
    GSTRIDE(t, n)
    {
        double s = 0;
        for (int i = 0; i < m; i++) s += vj[t + (long)ldv * i] * z[i + (long)ldz * col];
        psi[t + (long)ldp * c] = s;
    }
}

void launch_ritz(long n, int m, int nc, const double *vj, long ldv, const double *z, int ldz, const int *cols,
                 double *psi, long ldp, hipStream_t s)
{
    hipLaunchKernelGGL(k_ritz, dim3(nparts(n), nc), dim3(256), 0, s, n, m, nc, vj, ldv, z, ldz, cols, psi, ldp);
}

// Ritz vectors of several windows in one launch: blockIdx.y picks the job
// psi = sum_i vj(:, i) z(i, col) for i < m
__global__ void k_ritz_batch(long n, long ldv, const RitzJob *jobs)
{
    const RitzJob jb = jobs[blockIdx.y];
    GSTRIDE(t, n)
    {
        double s = 0;
        for (int i = 0; i < jb.m; i++) s += jb.vj[t + ldv * i] * jb.z[i + (long)jb.ldz * jb.col];
        jb.psi[t] = s;
    }
}

void launch_ritz_batch(long n, long ldv, int njobs, const RitzJob *jobs_dev, hipStream_t s)
{
    if (njobs > 0) hipLaunchKernelGGL(k_ritz_batch, dim3(nparts(n), njobs), dim3(256), 0, s, n, ldv, jobs_dev);
}

// Hellmann-Feynman: part[(k*np + j)][blk] = partial of sum psi_k^2 dvdp_j
__global__ void __launch_bounds__(256) k_hf_part(long n, int np, const double *psi, long ld, const double *dvdp,
                                                 long lddv, double *part)
{
    __shared__ double red[4];
    const int k = blockIdx.y;
    const double *p = psi + (long)k * ld;
    for (int j = 0; j < np; j++) {
        const double *g = dvdp + (long)j * lddv;
        double s = 0;
        GSTRIDE(t, n) s += p[t] * p[t] * g[t];
        s = bsum(s, red);
        if (threadIdx.x == 0) part[((long)k * np + j) * NPART + blockIdx.x] = s;
        zero_tail(part + ((long)k * np + j) * NPART);
    }
}

void launch_hf_part(int nk, long n, int np, const double *psi, long ld, const double *dvdp, long lddv, double *part,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_hf_part, dim3(nparts(n), nk), dim3(256), 0, s, n, np, psi, ld, dvdp, lddv, part);
}

// dv += sum_b x_b o g_b
__global__ void k_dv_accum(long n, int nvec, const double *x, const double *g, long ld, double *dv)
{
    GSTRIDE(t, n)
    {
        double s = dv[t];
        for (int b = 0; b < nvec; b++) s += x[t + (long)ld * b] * g[t + (long)ld * b];
        dv[t] = s;
    }
}

void launch_dv_accum(long n, int nvec, const double *x, const double *g, long ld, double *dv, hipStream_t s)
{
    hipLaunchKernelGGL(k_dv_accum, dim3(nparts(n)), dim3(256), 0, s, n, nvec, x, g, ld, dv);
}

}  // namespace scalit
