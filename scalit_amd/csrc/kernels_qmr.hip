// Persistent QMR for small operators: the whole iteration loop of
// src/iterate/libiterate/qmr/qmr.f90:56-175 in one launch, for the fused
// 3-level operator (H*X as in osb.hx.h:71-110, the OSB preconditioner P*X as
// in osb.px.h:125-219 with the kernel chosen by sOSB).
//
// On grids of a few ten thousand points the launch-per-step iteration is
// about sixteen kernels of a few microseconds each, none of which fills the
// device.  Here G workgroups per right-hand side stay resident for the whole
// solve and pass the phases of an iteration to each other through a grid
// barrier (an arrival counter per right-hand side, waits bounded by a wall
// clock limit).  Each phase works on one of two partitions of the grid:
//
//   slabs   - the n3 theta planes (n1 x n2 points): modes 1 and 2 of H*X, the
//             diagonal term, and P*X levels 2 and 1 (whose blocks lie inside
//             a plane);
//   pencils - chunks of pc (r, R) points with all their n3 values: mode 3 of
//             H*X (x DEP) and the level-3 rotations of P*X.
//
// One iteration (conv type 1, two H*X and one P*X) is seven phases:
//   1 slab   Q = Z/ZVRAT/ZNRM - c Q;  S = (modes 1, 2, diagonal) Q
//   2 pencil S += DEP o (H3 Q);  S = ZVRAT S;  Q.S
//   3 pencil V = S - BETA V/VNRM;  V.V;  Z = V3^T V
//   4 slab   Z = V2 V1 D^-1 V1^T V2^T Z  (D: the sOSB kernel of EIG0 - E0)
//   5 pencil Z = V3 Z;  Z.Z, V.Z
//   6 slab   UX, UR, X, R;  HX0 = (modes 1, 2, diagonal) X
//   7 pencil HX0 += DEP o (H3 X);  HX0.HX0, (HX0-B).(HX0-B)
// Every workgroup adds the G partial sums of a dot product in the same fixed
// order, so all of them take the same branch of the recurrence; the
// arithmetic of each vector value is the launch-per-step path's (the dot
// products add their terms in another order).
#include <cstdlib>

#include "device_common.hpp"
#include "solver_state.hpp"

namespace scalit {

constexpr int QP_NT = 1024;    // threads per workgroup
constexpr int QP_GMAX = 64;    // workgroups per right-hand side at most (one wave adds the partials)
constexpr int QP_NMAX = 48;    // largest level size
enum { QS_QS, QS_VV, QS_ZZ, QS_VZ, QS_RR, QS_UX, QS_XX, QS_UR, QS_HH, QS_R1, QS_SLOTS };

struct QpArgs {
    int n1, n2, n3, pc, kb, nosb, conv, qmrmax, G;   // kb: level-1 blocks staged in LDS at a time
    long n;
    double de, beta, tol;
    const double *h1, *h2, *h3, *dep, *res, *eig0, *v1, *v2, *v3;
    QmrVecs<double> w;
    QmrScal *q;
    int *active;
    const double *shifts;    // per-vector shift E0
    const double *part_vz;   // NPART partial sums per vector of (V/VNRM).Z from the initialisation
    double *part;            // nvec x QS_SLOTS x G partial sums
    unsigned *bar;           // one arrival counter per vector, zero at launch
    int *err;                // set when a wait gives up
    long wait_ticks;         // wall-clock ticks a barrier waits at most
};

// LDS doubles besides the level-1 stage, and the stage per block (the
// block's columns padded to n1 + 1 values, so that a wave reading a row of
// V1^T spreads over the banks)
static size_t qmr_persist_lds_base(int n1, int n2, int n3, int pc)
{
    const size_t P = (size_t)n1 * n2;
    return (size_t)n1 * n1 + 2 * (size_t)n2 * n2 + 2 * (size_t)n3 * n3 + 2 * P + (size_t)pc * n3 + QP_NT / 64 + 1;
}
static size_t qmr_persist_lds(int n1, int n2, int n3, int pc, int kb)
{
    return sizeof(double) * (qmr_persist_lds_base(n1, n2, n3, pc) + (size_t)kb * n1 * (n1 + 1));
}

struct QpLds {
    double *H1, *H2, *H3, *V3, *blk, *sA, *sB, *pen, *red;   // red: QP_NT / 64 sums, then the barrier flag
    double *v1s;                                               // kb level-1 blocks
};

__device__ __forceinline__ double qp_wave_sum(double v)
{
    // xor butterfly: every lane ends with the same value (a + b == b + a)
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// the workgroup's sum of one value per thread (the same on every thread)
__device__ __forceinline__ double qp_block_sum(double v, double *red)
{
    v = qp_wave_sum(v);
    __syncthreads();   // red is free
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < QP_NT / 64; k++) r += red[k];
    return r;
}

// the sum of the G partials of one slot, the same on every thread of every
// workgroup of the vector
__device__ __forceinline__ double qp_grid_sum(const double *part, int slot, int G)
{
    const int lane = threadIdx.x & 63;
    return qp_wave_sum(lane < G ? part[slot * G + lane] : 0.0);
}

__device__ __forceinline__ void qp_put(double *part, int slot, int G, double v, double *red)
{
    v = qp_block_sum(v, red);
    if (threadIdx.x == 0) part[slot * G + blockIdx.x] = v;
}

// Grid barrier among the G workgroups of one vector.  False when the wait
// gave up (every later barrier of this vector then gives up too, so every
// workgroup leaves the loop).
__device__ bool qp_sync(unsigned *bar, unsigned &target, unsigned G, long wait_ticks, int *err, int *ok_s)
{
    __syncthreads();
    target += G;
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        int ok = 1;
        const long t0 = (long)wall_clock64();
        while ((int)(__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
            if ((long)wall_clock64() - t0 > wait_ticks) {
                ok = 0;
                atomicOr(err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
        *ok_s = ok;
    }
    __syncthreads();
    return *ok_s != 0;
}

// out(slab) = (RES - sh) x + H1 x (mode 1) + H2 x (mode 2), x = the slab in sA
__device__ __forceinline__ void qp_slab_hx12(const QpArgs &a, const QpLds &L, long o, double sh, double *out)
{
    const int n1 = a.n1, n2 = a.n2, P = n1 * n2;
    const float r1 = 1.0f / (float)n1;
    for (int e = threadIdx.x; e < P; e += QP_NT) {
        const int i2 = fdiv_small(e, r1), i1 = e - i2 * n1;
        double acc = (a.res[o + e] - sh) * L.sA[e];
#pragma unroll 4
        for (int j = 0; j < n1; j++) acc += L.H1[i1 + n1 * j] * L.sA[j + n1 * i2];
#pragma unroll 4
        for (int j = 0; j < n2; j++) acc += L.H2[i2 + n2 * j] * L.sA[i1 + n1 * j];
        out[o + e] = acc;
    }
}

// pen(pp, j) = x(p0 + pp, j) for the np points of one pencil chunk
__device__ __forceinline__ void qp_load_pencil(const QpArgs &a, const QpLds &L, const double *x, int p0, int np)
{
    const int P = a.n1 * a.n2, pc = a.pc;
    const float rp = 1.0f / (float)np;
    for (int e = threadIdx.x; e < np * a.n3; e += QP_NT) {
        const int j = fdiv_small(e, rp), pp = e - j * np;
        L.pen[pp + pc * j] = x[p0 + pp + (long)P * j];
    }
}

__global__ void __launch_bounds__(QP_NT) k_qmr_persist(QpArgs a)
{
    extern __shared__ double qp_sm[];
    const int v = blockIdx.y, g = blockIdx.x, G = a.G, tid = threadIdx.x;
    QmrScal *qv = a.q + v;
    QmrScal s = *qv;
    // the same for every workgroup of the vector: nothing writes the state
    // before all of them have passed the first barrier below
    if (s.status != QMR_RUNNING) return;
    const int n1 = a.n1, n2 = a.n2, n3 = a.n3, P = n1 * n2, pc = a.pc;
    const int nch = (P + pc - 1) / pc;
    QpLds L;
    L.H1 = qp_sm;
    L.H2 = L.H1 + n1 * n1;
    L.H3 = L.H2 + n2 * n2;
    L.V3 = L.H3 + n3 * n3;
    L.blk = L.V3 + n3 * n3;
    L.sA = L.blk + n2 * n2;
    L.sB = L.sA + P;
    L.pen = L.sB + P;
    L.red = L.pen + pc * n3;
    int *ok_s = (int *)(L.red + QP_NT / 64);
    L.v1s = L.red + QP_NT / 64 + 1;
    for (int e = tid; e < n1 * n1; e += QP_NT) L.H1[e] = a.h1[e];
    for (int e = tid; e < n2 * n2; e += QP_NT) L.H2[e] = a.h2[e];
    for (int e = tid; e < n3 * n3; e += QP_NT) {
        L.H3[e] = a.h3[e];
        L.V3[e] = a.v3[e];
    }

    const QmrVecs<double> &w = a.w;
    const double *B = w.b + (long)v * w.ldb;
    double *X = w.x + (long)v * w.ldx;
    double *R = w.R + (long)v * w.ld, *UX = w.UX + (long)v * w.ld, *UR = w.UR + (long)v * w.ld;
    double *Q = w.Q + (long)v * w.ld, *Sv = w.Sv + (long)v * w.ld, *V = w.V + (long)v * w.ld;
    double *Z = w.Z + (long)v * w.ld, *HX0 = w.HX0 + (long)v * w.ld;
    double *part = a.part + (long)v * QS_SLOTS * G;
    unsigned *bar = a.bar + v;
    const double sh = a.shifts[v];
    const int conv = a.conv;

    // (V/VNRM).Z of the initialisation from its NPART partials, one wave
    double vz_sum;
    {
        const int lane = tid & 63;
        double t = 0.0;
        for (int k = lane; k < NPART; k += 64) t += a.part_vz[(long)v * NPART + k];
        vz_sum = qp_wave_sum(t);
    }
    unsigned target = 0;
    int status = QMR_RUNNING;
    // every workgroup has read the state (and its matrices are in LDS)
    bool alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s);

    while (alive) {
        const bool first = s.iter == 1;
        // ---- phase 1 (slabs): exit and breakdown checks, Q, modes 1 + 2 of H Q
        if (s.vnrm == 0.0 || s.znrm == 0.0) {
            status = QMR_EXIT;
            break;
        }
        const double vzdot = vz_sum / s.znrm;
        if (vzdot == 0.0) {
            status = QMR_EXIT;
            break;
        }
        const double cq = first ? 0.0 : s.vnrm * vzdot / s.qsdot;
        const double zvrat = s.zvrat, znrm = s.znrm;
        const double zvrat_next = zvrat * znrm / s.vnrm;
        s.vzdot = vzdot;
        s.zvrat_next = zvrat_next;
        for (int sl = g; sl < n3; sl += G) {
            const long o = (long)sl * P;
            for (int e = tid; e < P; e += QP_NT) {
                const double z = Z[o + e];
                const double qn = first ? z / zvrat / znrm : z / zvrat / znrm - cq * Q[o + e];
                Q[o + e] = qn;
                L.sA[e] = qn;
            }
            __syncthreads();
            qp_slab_hx12(a, L, o, sh, Sv);
            __syncthreads();
        }
        if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;

        // ---- phase 2 (pencils): mode 3 of H Q, S = ZVRAT S, Q.S
        {
            double qs = 0.0;
            for (int c = g; c < nch; c += G) {
                const int p0 = c * pc, np = min(pc, P - p0);
                qp_load_pencil(a, L, Q, p0, np);
                __syncthreads();
                const float rp = 1.0f / (float)np;
                for (int e = tid; e < np * n3; e += QP_NT) {
                    const int i = fdiv_small(e, rp), pp = e - i * np;
                    double acc = 0.0;
#pragma unroll 4
                    for (int j = 0; j < n3; j++) acc += L.H3[i + n3 * j] * L.pen[pp + pc * j];
                    const long idx = p0 + pp + (long)P * i;
                    const double sv = zvrat_next * (Sv[idx] + a.dep[p0 + pp] * acc);
                    Sv[idx] = sv;
                    qs += L.pen[pp + pc * i] * sv;
                }
                __syncthreads();
            }
            qp_put(part, QS_QS, G, qs, L.red);
        }
        if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;

        // ---- phase 3 (pencils): BETA, V = S - BETA V/VNRM, V.V, Z = V3^T V
        const double qsdot = qp_grid_sum(part, QS_QS, G);
        const double beta = qsdot / vzdot;
        if (qsdot == 0.0 || beta == 0.0) {
            status = QMR_EXIT;
            break;
        }
        s.qsdot = qsdot;
        s.beta = beta;
        {
            double vv = 0.0;
            for (int c = g; c < nch; c += G) {
                const int p0 = c * pc, np = min(pc, P - p0);
                const float rp = 1.0f / (float)np;
                for (int e = tid; e < np * n3; e += QP_NT) {
                    const int j = fdiv_small(e, rp), pp = e - j * np;
                    const long idx = p0 + pp + (long)P * j;
                    const double vn = Sv[idx] - beta * (V[idx] / s.vnrm);
                    V[idx] = vn;
                    L.pen[pp + pc * j] = vn;
                    vv += vn * vn;
                }
                __syncthreads();
                for (int e = tid; e < np * n3; e += QP_NT) {
                    const int i = fdiv_small(e, rp), pp = e - i * np;
                    double acc = 0.0;
#pragma unroll 4
                    for (int j = 0; j < n3; j++) acc += L.V3[j + n3 * i] * L.pen[pp + pc * j];
                    Z[p0 + pp + (long)P * i] = acc;
                }
                __syncthreads();
            }
            qp_put(part, QS_VV, G, vv, L.red);
        }
        if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;

        // ---- phase 4 (slabs): levels 2 and 1 of P*X on the rotated Z
        {
            const float r1 = 1.0f / (float)n1;
            const int ld1 = n1 + 1, bs1 = n1 * ld1, kb = a.kb;
            for (int sl = g; sl < n3; sl += G) {
                const long o = (long)sl * P;
                const double *A2 = a.v2 + (long)sl * n2 * n2;
                for (int e = tid; e < P; e += QP_NT) L.sA[e] = Z[o + e];
                for (int e = tid; e < n2 * n2; e += QP_NT) L.blk[e] = A2[e];
                __syncthreads();
                for (int e = tid; e < P; e += QP_NT) {   // V2^T
                    const int i = fdiv_small(e, r1), i1 = e - i * n1;
                    double acc = 0.0;
#pragma unroll 4
                    for (int j = 0; j < n2; j++) acc += L.blk[j + n2 * i] * L.sA[i1 + n1 * j];
                    L.sB[e] = acc;
                }
                // level 1, kb blocks (consecutive in memory) at a time: sB
                // column i2 -> sA column i2 (V1^T, the kernel) -> sB column i2 (V1)
                for (int b0 = 0; b0 < n2; b0 += kb) {
                    const int nb = min(kb, n2 - b0);
                    const double *A1 = a.v1 + ((long)sl * n2 + b0) * n1 * n1;
                    __syncthreads();   // sB complete; the stage is free
                    for (int e = tid; e < nb * n1 * n1; e += QP_NT) {
                        const int c = fdiv_small(e, r1), r = e - c * n1;   // c = block * n1 + column
                        L.v1s[c * ld1 + r] = A1[e];
                    }
                    __syncthreads();
                    for (int e = tid; e < nb * n1; e += QP_NT) {
                        const int bl = fdiv_small(e, r1), i1 = e - bl * n1, i2 = b0 + bl;
                        const double *A = L.v1s + bl * bs1;
                        double acc = 0.0;
#pragma unroll 4
                        for (int j = 0; j < n1; j++) acc += A[i1 * ld1 + j] * L.sB[j + n1 * i2];
                        L.sA[i1 + n1 * i2] = px_kernel<double>(a.nosb, acc, a.eig0[o + i1 + n1 * i2], sh, a.de, a.beta);
                    }
                    __syncthreads();
                    for (int e = tid; e < nb * n1; e += QP_NT) {
                        const int bl = fdiv_small(e, r1), i1 = e - bl * n1, i2 = b0 + bl;
                        const double *A = L.v1s + bl * bs1;
                        double acc = 0.0;
#pragma unroll 4
                        for (int j = 0; j < n1; j++) acc += A[j * ld1 + i1] * L.sA[j + n1 * i2];
                        L.sB[i1 + n1 * i2] = acc;
                    }
                }
                __syncthreads();
                for (int e = tid; e < P; e += QP_NT) {   // V2
                    const int i = fdiv_small(e, r1), i1 = e - i * n1;
                    double acc = 0.0;
#pragma unroll 4
                    for (int j = 0; j < n2; j++) acc += L.blk[i + n2 * j] * L.sB[i1 + n1 * j];
                    Z[o + e] = acc;
                }
                __syncthreads();
            }
        }
        if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;

        // ---- phase 5 (pencils): Z = V3 Z, Z.Z and V.Z over ZVRAT
        const double vnrmpo = sqrt(qp_grid_sum(part, QS_VV, G));
        s.vnrmpo = vnrmpo;
        {
            double zz = 0.0, vz = 0.0;
            for (int c = g; c < nch; c += G) {
                const int p0 = c * pc, np = min(pc, P - p0);
                qp_load_pencil(a, L, Z, p0, np);
                __syncthreads();
                const float rp = 1.0f / (float)np;
                for (int e = tid; e < np * n3; e += QP_NT) {
                    const int i = fdiv_small(e, rp), pp = e - i * np;
                    double acc = 0.0;
#pragma unroll 4
                    for (int j = 0; j < n3; j++) acc += L.V3[i + n3 * j] * L.pen[pp + pc * j];
                    const long idx = p0 + pp + (long)P * i;
                    Z[idx] = acc;
                    const double z = acc / zvrat_next;
                    zz += z * z;
                    vz += (V[idx] / vnrmpo) * z;
                }
                __syncthreads();
            }
            qp_put(part, QS_ZZ, G, zz, L.red);
            qp_put(part, QS_VZ, G, vz, L.red);
        }
        if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;

        // ---- phase 6 (slabs): GAMA, ETA, the UX / UR / X / R updates, modes 1 + 2 of H X
        const double znrm_next = sqrt(qp_grid_sum(part, QS_ZZ, G));
        const double vz_next = qp_grid_sum(part, QS_VZ, G);
        const double theta = vnrmpo / (s.gamamo * fabs(beta));
        const double gama = 1.0 / sqrt(1.0 + theta * theta);
        const double eta = -s.eta * s.vnrm * gama * gama / (beta * s.gamamo * s.gamamo);
        if (gama == 0.0) {
            status = QMR_EXIT;
            break;
        }
        {
            const double ae = eta * zvrat_next;
            const double tg = (s.thetamo * gama) * (s.thetamo * gama);
            double rr = 0.0, uxux = 0.0, xx = 0.0, urur = 0.0;
            for (int sl = g; sl < n3; sl += G) {
                const long o = (long)sl * P;
                for (int e = tid; e < P; e += QP_NT) {
                    const long t = o + e;
                    const double ux = first ? ae * Q[t] : ae * Q[t] + tg * UX[t];
                    const double ur = first ? eta * Sv[t] : eta * Sv[t] + tg * UR[t];
                    UX[t] = ux;
                    UR[t] = ur;
                    const double xn = X[t] + ux, rn = R[t] - ur;
                    X[t] = xn;
                    R[t] = rn;
                    rr += rn * rn;
                    uxux += ux * ux;
                    xx += xn * xn;
                    urur += ur * ur;
                    L.sA[e] = xn;
                }
                __syncthreads();
                if (conv == 1) qp_slab_hx12(a, L, o, sh, HX0);
                __syncthreads();
            }
            if (conv != 1) qp_put(part, QS_RR, G, rr, L.red);
            if (conv == 2) {
                qp_put(part, QS_UX, G, uxux, L.red);
                qp_put(part, QS_XX, G, xx, L.red);
                qp_put(part, QS_UR, G, urur, L.red);
            }
        }
        if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;

        // ---- phase 7 (pencils, conv type 1): mode 3 of H X, HX0.HX0, (HX0-B).(HX0-B)
        if (conv == 1) {
            double hh = 0.0, r1 = 0.0;
            for (int c = g; c < nch; c += G) {
                const int p0 = c * pc, np = min(pc, P - p0);
                qp_load_pencil(a, L, X, p0, np);
                __syncthreads();
                const float rp = 1.0f / (float)np;
                for (int e = tid; e < np * n3; e += QP_NT) {
                    const int i = fdiv_small(e, rp), pp = e - i * np;
                    double acc = 0.0;
#pragma unroll 4
                    for (int j = 0; j < n3; j++) acc += L.H3[i + n3 * j] * L.pen[pp + pc * j];
                    const long idx = p0 + pp + (long)P * i;
                    const double hv = HX0[idx] + a.dep[p0 + pp] * acc;
                    HX0[idx] = hv;
                    const double r = hv - B[idx];
                    hh += hv * hv;
                    r1 += r * r;
                }
                __syncthreads();
            }
            qp_put(part, QS_HH, G, hh, L.red);
            qp_put(part, QS_R1, G, r1, L.red);
            if (!(alive = qp_sync(bar, target, G, a.wait_ticks, a.err, ok_s))) break;
        }

        // ---- the end of the iteration (every workgroup alike)
        double res;
        if (conv == 1) {
            const double xnrm = sqrt(qp_grid_sum(part, QS_HH, G));
            res = sqrt(qp_grid_sum(part, QS_R1, G)) / xnrm;
        } else if (conv == 2) {
            const double rr = qp_grid_sum(part, QS_RR, G);
            const double q1 = sqrt(rr) / s.bnrm;
            const double q2 = sqrt(qp_grid_sum(part, QS_UX, G) / qp_grid_sum(part, QS_XX, G));
            const double q3 = sqrt(qp_grid_sum(part, QS_UR, G) / rr);
            res = q1 > q2 ? q1 : q2;
            res = res > q3 ? res : q3;
        } else {
            res = sqrt(qp_grid_sum(part, QS_RR, G)) / s.bnrm;
        }
        s.res = res;
        s.theta = theta;
        s.gama = gama;
        s.znrm_next = znrm_next;
        s.eta_next = eta;
        if (fabs(res) < a.tol) {
            s.ret = s.iter;
            status = QMR_CONVERGED;
            break;
        }
        s.ret = -s.iter;
        s.zvrat = zvrat_next;
        s.vnrm = vnrmpo;
        s.gamamo = gama;
        s.thetamo = theta;
        s.eta = eta;
        s.znrm = znrm_next;
        vz_sum = vz_next;
        s.iter += 1;
        if (s.iter > a.qmrmax) {
            status = QMR_EXIT;
            break;
        }
    }
    if (g == 0 && tid == 0) {
        s.status = alive ? status : QMR_EXIT;   // a wait that gave up ends the vector
        *qv = s;
        a.active[v] = 0;
    }
}

// Whether the persistent solver can run this operator: the fused 3-level
// layout (sF = 3, DEP on level 3 only, no sNDVR layer) with its
// preconditioner, levels of at most QP_NMAX points, at most 65,536 points.
bool qmr_persist_ok(const scalit_ctx *c)
{
    if (!c->fast3 || c->ndvr || c->sF != 3 || !c->has_precond || !c->has_res) return false;
    for (int l = 0; l < 3; l++)
        if (c->lv[l].n > QP_NMAX || c->lv[l].dep != (l == 2) || !c->lv[l].vosb.p) return false;
    return c->N <= 65536;
}

static int env_or(const char *name, int def)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : def;
}

// Run the iterations for nvec right-hand sides after the initialisation
// kernels (scalars in q, (V/VNRM).Z partials in part_vz).  Returns
// QP_NOT_LAUNCHED (nothing launched) when the G workgroups of every vector
// cannot all be resident at once, SCALIT_EHIP when a barrier wait gave up.
int qmr_persist_solve(scalit_ctx *c, int nvec, const QmrVecs<double> &w, QmrScal *q, int *active,
                      const double *shifts, const double *part_vz, int conv, int qmrmax, double tol, QmrPersistBufs &pb,
                      hipStream_t s)
{
    QpArgs a;
    a.n1 = c->lv[0].n;
    a.n2 = c->lv[1].n;
    a.n3 = c->lv[2].n;
    const int P = a.n1 * a.n2;
    int G = env_or("SCALIT_QMR_PERSIST_G", std::max(16, a.n3));
    G = std::max(1, std::min(G, QP_GMAX));
    a.pc = (P + G - 1) / G;
    // as many level-1 blocks per stage as fit in 150 KB of LDS
    const size_t base = qmr_persist_lds_base(a.n1, a.n2, a.n3, a.pc);
    const size_t per_blk = (size_t)a.n1 * (a.n1 + 1);
    if ((base + per_blk) * sizeof(double) > 150 * 1024) return QP_NOT_LAUNCHED;
    a.kb = (int)std::min<size_t>((size_t)a.n2, (150 * 1024 / sizeof(double) - base) / per_blk);
    const size_t lds = qmr_persist_lds(a.n1, a.n2, a.n3, a.pc, a.kb);
    allow_full_lds((const void *)k_qmr_persist);
    const int per_cu = resident_per_cu((const void *)k_qmr_persist, QP_NT, lds);
    const int ncu = g_num_cu.load();
    if (ncu <= 0 || (long)G * nvec > (long)per_cu * ncu) return QP_NOT_LAUNCHED;
    a.G = G;
    a.n = c->N;
    a.nosb = c->osb;
    a.de = c->de;
    a.beta = c->beta;
    a.conv = conv;
    a.qmrmax = qmrmax;
    a.tol = tol;
    a.h1 = c->lv[0].h0.p;
    a.h2 = c->lv[1].h0.p;
    a.h3 = c->lv[2].h0.p;
    a.dep = c->lv[2].depv.p;
    a.res = c->res.p;
    a.eig0 = c->eig0.p;
    a.v1 = c->lv[0].vosb.p;
    a.v2 = c->lv[1].vosb.p;
    a.v3 = c->lv[2].vosb.p;
    a.w = w;
    a.q = q;
    a.active = active;
    a.shifts = shifts;
    a.part_vz = part_vz;
    SC_TRY(pb.part.alloc((size_t)nvec * QS_SLOTS * G));
    SC_TRY(pb.bar.alloc(nvec));
    SC_TRY(pb.err.alloc(1));
    a.part = pb.part.p;
    a.bar = pb.bar.p;
    a.err = pb.err.p;
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    a.wait_ticks = (long)std::max(khz, 1000) * 1000L * 2;   // two seconds
    HIP_TRY(hipMemsetAsync(pb.bar.p, 0, sizeof(unsigned) * nvec, s));
    HIP_TRY(hipMemsetAsync(pb.err.p, 0, sizeof(int), s));
    void *args[] = {&a};
    HIP_TRY(hipLaunchCooperativeKernel((const void *)k_qmr_persist, dim3(G, nvec), dim3(QP_NT), args, (unsigned)lds,
                                       s));
    HIP_TRY(hipMemcpyAsync(&pb.herr, pb.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return pb.herr ? SCALIT_EHIP : SCALIT_OK;
}

}  // namespace scalit
