// Persistent solvers for small operators: the whole iteration loop of QMR
// (src/iterate/libiterate/qmr/qmr.f90:56-175) in one launch, and the Krylov
// steps of pistH0 (src/iterate/libiterate/pist/pistHij.f90:70-113: a QMR
// solve per new Krylov vector, then MGS_ORTH, gsorth.f90:42-72) in one launch
// per block of steps.  Both run on the fused 3-level operator (H*X as in
// osb.hx.h:71-110, the OSB preconditioner P*X as in osb.px.h:125-219 with the
// kernel chosen by sOSB) in FP64.
//
// On grids of a few ten thousand points a launch-per-step QMR iteration is
// about sixteen kernels of a few microseconds each, none of which fills the
// device, and every Krylov step adds host round trips.  Here G workgroups per
// right-hand side (per PIST window) stay resident for the whole call and pass
// the phases to each other through a grid barrier (an arrival counter per
// right-hand side, waits bounded by a wall-clock limit).  Each phase works on
// one of two partitions of the grid:
//
//   slabs   - the n3 theta planes (n1 x n2 points): modes 1 and 2 of H*X, the
//             diagonal term, and P*X levels 2 and 1 (whose blocks lie inside
//             a plane);
//   pencils - chunks of pc (r, R) points with all their n3 values: mode 3 of
//             H*X (x DEP), the level-3 rotations of P*X, and the Gram-Schmidt
//             vector work (each workgroup owns at most one chunk).
//
// One QMR iteration (conv type 1, two H*X and one P*X) is seven phases:
//   1 slab   Q = Z/ZVRAT/ZNRM - c Q;  S = (modes 1, 2, diagonal) Q
//   2 pencil S += DEP o (H3 Q);  S = ZVRAT S;  Q.S
//   3 pencil V = S - BETA V/VNRM;  V.V;  Z = V3^T V
//   4 slab   Z = V2 V1 D^-1 V1^T V2^T Z  (D: the sOSB kernel of EIG0 - E0)
//   5 pencil Z = V3 Z;  Z.Z, V.Z
//   6 slab   UX, UR, X, R;  HX0 = (modes 1, 2, diagonal) X
//   7 pencil HX0 += DEP o (H3 X);  HX0.HX0, (HX0-B).(HX0-B)
// and phases 1 and 2 of the next iteration share the barrier intervals of 6
// and 7, so that an iteration costs five grid barriers.
// Every workgroup adds the G partial sums of a dot product in the same fixed
// order, so all of them take the same branch of the recurrence; the
// arithmetic of each vector value is the launch-per-step path's (the dot
// products add their terms in another order).
//
// The Gram-Schmidt step of the PIST kernel is MGS_ORTH in its
// low-synchronisation form (the inverse compact WY form of modified
// Gram-Schmidt: Swirydowicz, Langou, Ananthan, Yang and Thomas, "Low
// synchronization Gram-Schmidt and generalized minimal residual algorithms",
// Numer. Linear Algebra Appl. 28 (2021) e2343).  MGS projects w on v_1..v_m in
// turn, h_i = v_i . (w - sum_{k<i} h_k v_k) = c_i - sum_{k<i} (v_i . v_k) h_k
// with c = V^T w, so h solves (I + L) h = c, L the strictly lower part of
// V^T V; the result is w - V h.  One pass computes c and the newest row of L
// (the Gram matrix is kept from step to step), every workgroup solves the
// small triangular system, and a second pass forms w - V h: two grid
// barriers instead of one reduction per basis vector, and the same
// projections as MGS in exact arithmetic.
#include <cstdlib>

#include "device_common.hpp"
#include "solver_state.hpp"

namespace scalit {

constexpr int QP_NT = 1024;    // threads per workgroup
constexpr int QP_NW = QP_NT / 64;
constexpr int QP_GMAX = 64;    // workgroups per right-hand side at most (one wave adds the partials)
constexpr int QP_NMAX = 48;    // largest level size
enum { QS_QS, QS_VV, QS_ZZ, QS_VZ, QS_RR, QS_UX, QS_XX, QS_UR, QS_HH, QS_R1, QS_BB, QS_SLOTS };

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

// the PIST steps: columns j = j_from..j_to of each window's basis, column j
// from the QMR solve with column j - 1 as the right-hand side
struct PpArgs {
    QpArgs q;
    double *vj;              // window v's basis at vj + v * ldw, column j at + j * ldv
    long ldv, ldw;
    double *gram;            // window v's Gram matrix (ldg x ldg, v_i . v_k at i + ldg k for k < i)
    int ldg;
    int j_from, j_to, orth, zero_x;
    int *qret;               // QMR return of window v's solve for column j at v * ldg + j
    int *wret;               // per window: 1 when every step ran, else the pistH0 value that ended it (<= 0)
    double *opart;           // per window (2 ldg + 2) x G partial sums of the Gram-Schmidt step
};

// LDS doubles besides the level-1 stage, and the stage per block (the
// block's columns padded to n1 + 1 values, so that a wave reading a row of
// V1^T spreads over the banks)
static size_t qmr_persist_lds_base(int n1, int n2, int n3, int pc)
{
    const size_t P = (size_t)n1 * n2;
    return (size_t)n1 * n1 + 2 * (size_t)n2 * n2 + 2 * (size_t)n3 * n3 + 2 * P + (size_t)pc * n3 + QP_NW + 1;
}
static size_t qmr_persist_lds(int n1, int n2, int n3, int pc, int kb)
{
    return sizeof(double) * (qmr_persist_lds_base(n1, n2, n3, pc) + (size_t)kb * n1 * (n1 + 1));
}

struct QpLds {
    double *H1, *H2, *H3, *V3, *blk, *sA, *sB, *pen, *red;   // red: QP_NW sums, then the barrier flag
    double *v1s;                                               // kb level-1 blocks (the Gram-Schmidt scratch)
};

// one workgroup's view of its right-hand side
struct QpW {
    const QpArgs *a;
    QpLds L;
    int v, g, G, n1, n2, n3, P, pc, nch;
    double sh;
    const double *B;
    double *X, *R, *UX, *UR, *Q, *Sv, *V, *Z, *HX0, *part;
    unsigned *bar;
    unsigned target;
    int *ok_s;
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
    for (int k = 0; k < QP_NW; k++) r += red[k];
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
// workgroup leaves its loops).
__device__ bool qp_sync(QpW &W)
{
    __syncthreads();
    W.target += W.G;
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(W.bar, 1u);
        int ok = 1;
        const long t0 = (long)wall_clock64();
        while ((int)(__hip_atomic_load(W.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - W.target) < 0) {
            if ((long)wall_clock64() - t0 > W.a->wait_ticks) {
                ok = 0;
                atomicOr(W.a->err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
        *W.ok_s = ok;
    }
    __syncthreads();
    return *W.ok_s != 0;
}

// the LDS layout, the operator's small matrices loaded
__device__ void qp_setup(QpW &W, const QpArgs &a, double *sm)
{
    W.a = &a;
    W.v = blockIdx.y;
    W.g = blockIdx.x;
    W.G = a.G;
    W.n1 = a.n1;
    W.n2 = a.n2;
    W.n3 = a.n3;
    W.P = a.n1 * a.n2;
    W.pc = a.pc;
    W.nch = (W.P + W.pc - 1) / W.pc;
    W.target = 0;
    QpLds &L = W.L;
    L.H1 = sm;
    L.H2 = L.H1 + W.n1 * W.n1;
    L.H3 = L.H2 + W.n2 * W.n2;
    L.V3 = L.H3 + W.n3 * W.n3;
    L.blk = L.V3 + W.n3 * W.n3;
    L.sA = L.blk + W.n2 * W.n2;
    L.sB = L.sA + W.P;
    L.pen = L.sB + W.P;
    L.red = L.pen + W.pc * W.n3;
    W.ok_s = (int *)(L.red + QP_NW);
    L.v1s = L.red + QP_NW + 1;
    const int tid = threadIdx.x;
    for (int e = tid; e < W.n1 * W.n1; e += QP_NT) L.H1[e] = a.h1[e];
    for (int e = tid; e < W.n2 * W.n2; e += QP_NT) L.H2[e] = a.h2[e];
    for (int e = tid; e < W.n3 * W.n3; e += QP_NT) {
        L.H3[e] = a.h3[e];
        L.V3[e] = a.v3[e];
    }
    const QmrVecs<double> &w = a.w;
    const int v = W.v;
    W.B = w.b + (long)v * w.ldb;
    W.X = w.x + (long)v * w.ldx;
    W.R = w.R + (long)v * w.ld;
    W.UX = w.UX + (long)v * w.ld;
    W.UR = w.UR + (long)v * w.ld;
    W.Q = w.Q + (long)v * w.ld;
    W.Sv = w.Sv + (long)v * w.ld;
    W.V = w.V + (long)v * w.ld;
    W.Z = w.Z + (long)v * w.ld;
    W.HX0 = w.HX0 + (long)v * w.ld;
    W.part = a.part + (long)v * QS_SLOTS * W.G;
    W.bar = a.bar + v;
    W.sh = a.shifts[v];
}

// out(slab) = (RES - sh) x + H1 x (mode 1) + H2 x (mode 2), x = the slab in sA
__device__ __forceinline__ void qp_slab_hx12(const QpW &W, long o, double *out)
{
    const int n1 = W.n1, n2 = W.n2, P = W.P;
    const double *res = W.a->res;
    const float r1 = 1.0f / (float)n1;
    for (int e = threadIdx.x; e < P; e += QP_NT) {
        const int i2 = fdiv_small(e, r1), i1 = e - i2 * n1;
        double acc = (res[o + e] - W.sh) * W.L.sA[e];
#pragma unroll 4
        for (int j = 0; j < n1; j++) acc += W.L.H1[i1 + n1 * j] * W.L.sA[j + n1 * i2];
#pragma unroll 4
        for (int j = 0; j < n2; j++) acc += W.L.H2[i2 + n2 * j] * W.L.sA[i1 + n1 * j];
        out[o + e] = acc;
    }
}

// pen(pp, j) = x(p0 + pp, j) for the np points of one pencil chunk
__device__ __forceinline__ void qp_load_pencil(const QpW &W, const double *x, int p0, int np)
{
    const float rp = 1.0f / (float)np;
    for (int e = threadIdx.x; e < np * W.n3; e += QP_NT) {
        const int j = fdiv_small(e, rp), pp = e - j * np;
        W.L.pen[pp + W.pc * j] = x[p0 + pp + (long)W.P * j];
    }
}

// y(pencil point pp, i) = sum_j A(i, j) pen(pp, j) (A = V3 or H3 when !trans,
// V3^T when trans), for the element e of the chunk
__device__ __forceinline__ double qp_pencil_dot(const QpW &W, const double *A, bool trans, int pp, int i)
{
    const int n3 = W.n3, pc = W.pc;
    double acc = 0.0;
    if (trans) {
#pragma unroll 4
        for (int j = 0; j < n3; j++) acc += A[j + n3 * i] * W.L.pen[pp + pc * j];
    } else {
#pragma unroll 4
        for (int j = 0; j < n3; j++) acc += A[i + n3 * j] * W.L.pen[pp + pc * j];
    }
    return acc;
}

// P*X levels 2 and 1 on the own slabs of Z, in place
__device__ void qp_px_slabs(QpW &W)
{
    const QpArgs &a = *W.a;
    QpLds &L = W.L;
    const int n1 = W.n1, n2 = W.n2, n3 = W.n3, P = W.P, tid = threadIdx.x;
    const float r1 = 1.0f / (float)n1;
    const int ld1 = n1 + 1, bs1 = n1 * ld1, kb = a.kb;
    for (int sl = W.g; sl < n3; sl += W.G) {
        const long o = (long)sl * P;
        const double *A2 = a.v2 + (long)sl * n2 * n2;
        for (int e = tid; e < P; e += QP_NT) L.sA[e] = W.Z[o + e];
        for (int e = tid; e < n2 * n2; e += QP_NT) L.blk[e] = A2[e];
        __syncthreads();
        for (int e = tid; e < P; e += QP_NT) {   // V2^T
            const int i = fdiv_small(e, r1), i1 = e - i * n1;
            double acc = 0.0;
#pragma unroll 4
            for (int j = 0; j < n2; j++) acc += L.blk[j + n2 * i] * L.sA[i1 + n1 * j];
            L.sB[e] = acc;
        }
        // level 1, kb blocks (consecutive in memory) at a time: sB column
        // i2 -> sA column i2 (V1^T, the kernel) -> sB column i2 (V1)
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
                L.sA[i1 + n1 * i2] = px_kernel<double>(a.nosb, acc, a.eig0[o + i1 + n1 * i2], W.sh, a.de, a.beta);
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
            W.Z[o + e] = acc;
        }
        __syncthreads();
    }
}

// Q = Z/ZVRAT/ZNRM - c Q on the own slabs, then S = (modes 1, 2, diagonal) Q
// (phase 1 of an iteration)
__device__ void qp_phase_q(QpW &W, bool first, double zvrat, double znrm, double cq)
{
    const int g = W.g, G = W.G, n3 = W.n3, P = W.P, tid = threadIdx.x;
    for (int sl = g; sl < n3; sl += G) {
        const long o = (long)sl * P;
        for (int e = tid; e < P; e += QP_NT) {
            const double z = W.Z[o + e];
            const double qn = first ? z / zvrat / znrm : z / zvrat / znrm - cq * W.Q[o + e];
            W.Q[o + e] = qn;
            W.L.sA[e] = qn;
        }
        __syncthreads();
        qp_slab_hx12(W, o, W.Sv);
        __syncthreads();
    }
}

// S += DEP o (H3 Q), S = ZVRAT_NEXT S on the own pencils; returns the partial
// Q.S (phase 2 of an iteration)
__device__ double qp_phase_s(QpW &W, double zvrat_next)
{
    const int g = W.g, G = W.G, n3 = W.n3, P = W.P, pc = W.pc, nch = W.nch, tid = threadIdx.x;
    double qs = 0.0;
    for (int c = g; c < nch; c += G) {
        const int p0 = c * pc, np = min(pc, P - p0);
        qp_load_pencil(W, W.Q, p0, np);
        __syncthreads();
        const float rp = 1.0f / (float)np;
        for (int e = tid; e < np * n3; e += QP_NT) {
            const int i = fdiv_small(e, rp), pp = e - i * np;
            const double acc = qp_pencil_dot(W, W.L.H3, false, pp, i);
            const long idx = p0 + pp + (long)P * i;
            const double sv = zvrat_next * (W.Sv[idx] + W.a->dep[p0 + pp] * acc);
            W.Sv[idx] = sv;
            qs += W.L.pen[pp + pc * i] * sv;
        }
        __syncthreads();
    }
    return qs;
}

// The QMR iterations from the state s (qmr.f90:56-175), vz_sum = the
// (V/VNRM).Z of the step before.  status: QMR_CONVERGED or QMR_EXIT at the
// end.  False when a barrier wait gave up.
//
// Phases 1 and 2 of iteration i+1 need nothing that phases 6 and 7 of
// iteration i produce (only the convergence test does), so they run in the
// same barrier intervals as 6 (slabs) and 7 (pencils): five barriers per
// iteration.  When iteration i ends the solve, what they wrote (Q, S and a
// partial sum) is scratch that nothing reads.
__device__ bool qp_qmr_loop(QpW &W, QmrScal &s, double vz_sum, int &status)
{
    const QpArgs &a = *W.a;
    QpLds &L = W.L;
    const int g = W.g, G = W.G, n3 = W.n3, P = W.P, pc = W.pc, nch = W.nch, tid = threadIdx.x;
    const int conv = a.conv;
    double *X = W.X, *R = W.R, *UX = W.UX, *UR = W.UR, *Q = W.Q, *Sv = W.Sv, *V = W.V, *Z = W.Z, *HX0 = W.HX0;
    double *part = W.part;
    status = QMR_RUNNING;
    // ---- iteration 1, phases 1 and 2: the exit and breakdown checks, Q, H Q
    if (s.vnrm == 0.0 || s.znrm == 0.0) {
        status = QMR_EXIT;
        return true;
    }
    double vzdot = vz_sum / s.znrm;
    if (vzdot == 0.0) {
        status = QMR_EXIT;
        return true;
    }
    double zvrat_next = s.zvrat * s.znrm / s.vnrm;
    s.vzdot = vzdot;
    s.zvrat_next = zvrat_next;
    qp_phase_q(W, s.iter == 1, s.zvrat, s.znrm, s.iter == 1 ? 0.0 : s.vnrm * vzdot / s.qsdot);
    if (!qp_sync(W)) return false;
    qp_put(part, QS_QS, G, qp_phase_s(W, zvrat_next), L.red);
    if (!qp_sync(W)) return false;

    for (;;) {
        const bool first = s.iter == 1;
        // ---- phase 3 (pencils): BETA, V = S - BETA V/VNRM, V.V, Z = V3^T V
        const double qsdot = qp_grid_sum(part, QS_QS, G);
        const double beta = qsdot / vzdot;
        if (qsdot == 0.0 || beta == 0.0) {
            status = QMR_EXIT;
            return true;
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
                    Z[p0 + pp + (long)P * i] = qp_pencil_dot(W, L.V3, true, pp, i);
                }
                __syncthreads();
            }
            qp_put(part, QS_VV, G, vv, L.red);
        }
        if (!qp_sync(W)) return false;

        // ---- phase 4 (slabs): levels 2 and 1 of P*X on the rotated Z
        qp_px_slabs(W);
        if (!qp_sync(W)) return false;

        // ---- phase 5 (pencils): Z = V3 Z, Z.Z and V.Z over ZVRAT
        const double vnrmpo = sqrt(qp_grid_sum(part, QS_VV, G));
        s.vnrmpo = vnrmpo;
        {
            double zz = 0.0, vz = 0.0;
            for (int c = g; c < nch; c += G) {
                const int p0 = c * pc, np = min(pc, P - p0);
                qp_load_pencil(W, Z, p0, np);
                __syncthreads();
                const float rp = 1.0f / (float)np;
                for (int e = tid; e < np * n3; e += QP_NT) {
                    const int i = fdiv_small(e, rp), pp = e - i * np;
                    const double acc = qp_pencil_dot(W, L.V3, false, pp, i);
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
        if (!qp_sync(W)) return false;

        // ---- GAMA, ETA; the next iteration's checks and scalars
        const double znrm_next = sqrt(qp_grid_sum(part, QS_ZZ, G));
        const double vz_next = qp_grid_sum(part, QS_VZ, G);
        const double theta = vnrmpo / (s.gamamo * fabs(beta));
        const double gama = 1.0 / sqrt(1.0 + theta * theta);
        const double eta = -s.eta * s.vnrm * gama * gama / (beta * s.gamamo * s.gamamo);
        if (gama == 0.0) {
            status = QMR_EXIT;
            return true;
        }
        const double vzdot_n = (vnrmpo == 0.0 || znrm_next == 0.0) ? 0.0 : vz_next / znrm_next;
        const bool next_ok = vzdot_n != 0.0;   // VNRM, ZNRM and VZDOT of iteration i+1 nonzero
        const double zvrat_next_n = next_ok ? zvrat_next * znrm_next / vnrmpo : 0.0;
        const double cq_n = next_ok ? vnrmpo * vzdot_n / qsdot : 0.0;

        // ---- phase 6 (slabs): the UX / UR / X / R updates, modes 1 + 2 of H X;
        // then phase 1 of iteration i+1
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
                if (conv == 1) qp_slab_hx12(W, o, HX0);
                __syncthreads();
            }
            if (conv != 1) qp_put(part, QS_RR, G, rr, L.red);
            if (conv == 2) {
                qp_put(part, QS_UX, G, uxux, L.red);
                qp_put(part, QS_XX, G, xx, L.red);
                qp_put(part, QS_UR, G, urur, L.red);
            }
            // (Q and S of iteration i are read above, before this overwrites them)
            if (next_ok) qp_phase_q(W, false, zvrat_next, znrm_next, cq_n);
        }
        if (!qp_sync(W)) return false;

        // ---- phase 7 (pencils, conv type 1): mode 3 of H X, HX0.HX0,
        // (HX0-B).(HX0-B); then phase 2 of iteration i+1
        {
            if (conv == 1) {
                double hh = 0.0, r1 = 0.0;
                for (int c = g; c < nch; c += G) {
                    const int p0 = c * pc, np = min(pc, P - p0);
                    qp_load_pencil(W, X, p0, np);
                    __syncthreads();
                    const float rp = 1.0f / (float)np;
                    for (int e = tid; e < np * n3; e += QP_NT) {
                        const int i = fdiv_small(e, rp), pp = e - i * np;
                        const double acc = qp_pencil_dot(W, L.H3, false, pp, i);
                        const long idx = p0 + pp + (long)P * i;
                        const double hv = HX0[idx] + a.dep[p0 + pp] * acc;
                        HX0[idx] = hv;
                        const double r = hv - W.B[idx];
                        hh += hv * hv;
                        r1 += r * r;
                    }
                    __syncthreads();
                }
                qp_put(part, QS_HH, G, hh, L.red);
                qp_put(part, QS_R1, G, r1, L.red);
            }
            if (next_ok) qp_put(part, QS_QS, G, qp_phase_s(W, zvrat_next_n), L.red);
        }
        if (!qp_sync(W)) return false;

        // ---- the end of iteration i (every workgroup alike)
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
            return true;
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
            return true;
        }
        // the checks at the top of iteration i+1
        if (!next_ok) {
            status = QMR_EXIT;
            return true;
        }
        vzdot = vzdot_n;
        zvrat_next = zvrat_next_n;
        s.vzdot = vzdot;
        s.zvrat_next = zvrat_next;
    }
}

__global__ void __launch_bounds__(QP_NT) k_qmr_persist(QpArgs a)
{
    extern __shared__ double qp_sm[];
    QmrScal *qv = a.q + blockIdx.y;
    QmrScal s = *qv;
    // the same for every workgroup of the vector: nothing writes the state
    // before all of them have passed the first barrier below
    if (s.status != QMR_RUNNING) return;
    QpW W;
    qp_setup(W, a, qp_sm);
    // (V/VNRM).Z of the initialisation from its NPART partials, one wave
    double vz_sum;
    {
        double t = 0.0;
        for (int k = threadIdx.x & 63; k < NPART; k += 64) t += a.part_vz[(long)W.v * NPART + k];
        vz_sum = qp_wave_sum(t);
    }
    int status = QMR_EXIT;
    // every workgroup has read the state (and its matrices are in LDS)
    bool alive = qp_sync(W);
    if (alive) alive = qp_qmr_loop(W, s, vz_sum, status);
    if (W.g == 0 && threadIdx.x == 0) {
        s.status = alive ? status : QMR_EXIT;   // a wait that gave up ends the vector
        *qv = s;
        a.active[W.v] = 0;
    }
}

// ---------------------------------------------------------------- the PIST steps

// the initialisation of a QMR solve of (H - E0) X = B (qmr.f90:36-50): R = V
// = B, X = 0 (zero_x), Z = P B, and the scalars.  Three phases.
__device__ bool qp_qmr_init(QpW &W, bool zero_x, int qmrmax, QmrScal &s, double &vz_sum)
{
    QpLds &L = W.L;
    const int g = W.g, G = W.G, n3 = W.n3, P = W.P, pc = W.pc, nch = W.nch, tid = threadIdx.x;
    // I1 (pencils): R, V, X; B.B; Z = V3^T B
    {
        double bb = 0.0;
        for (int c = g; c < nch; c += G) {
            const int p0 = c * pc, np = min(pc, P - p0);
            const float rp = 1.0f / (float)np;
            for (int e = tid; e < np * n3; e += QP_NT) {
                const int j = fdiv_small(e, rp), pp = e - j * np;
                const long idx = p0 + pp + (long)P * j;
                const double b = W.B[idx];
                W.R[idx] = b;
                W.V[idx] = b;
                if (zero_x) W.X[idx] = 0.0;
                bb += b * b;
                L.pen[pp + pc * j] = b;
            }
            __syncthreads();
            for (int e = tid; e < np * n3; e += QP_NT) {
                const int i = fdiv_small(e, rp), pp = e - i * np;
                W.Z[p0 + pp + (long)P * i] = qp_pencil_dot(W, L.V3, true, pp, i);
            }
            __syncthreads();
        }
        qp_put(W.part, QS_BB, G, bb, L.red);
    }
    if (!qp_sync(W)) return false;
    // I2 (slabs): levels 2 and 1
    qp_px_slabs(W);
    if (!qp_sync(W)) return false;
    // I3 (pencils): Z = V3 Z; Z.Z, (V/VNRM).Z
    const double bnrm = sqrt(qp_grid_sum(W.part, QS_BB, G));
    {
        double zz = 0.0, vz = 0.0;
        for (int c = g; c < nch; c += G) {
            const int p0 = c * pc, np = min(pc, P - p0);
            qp_load_pencil(W, W.Z, p0, np);
            __syncthreads();
            const float rp = 1.0f / (float)np;
            for (int e = tid; e < np * n3; e += QP_NT) {
                const int i = fdiv_small(e, rp), pp = e - i * np;
                const double d = qp_pencil_dot(W, L.V3, false, pp, i);
                const long idx = p0 + pp + (long)P * i;
                W.Z[idx] = d;
                zz += d * d;
                vz += (W.V[idx] / bnrm) * d;
            }
            __syncthreads();
        }
        qp_put(W.part, QS_ZZ, G, zz, L.red);
        qp_put(W.part, QS_VZ, G, vz, L.red);
    }
    if (!qp_sync(W)) return false;
    s.bnrm = bnrm;
    s.vnrm = bnrm;
    s.znrm = sqrt(qp_grid_sum(W.part, QS_ZZ, G));
    vz_sum = qp_grid_sum(W.part, QS_VZ, G);
    s.gamamo = 1.0;
    s.eta = -1.0;
    s.zvrat = 1.0;
    s.thetamo = 0.0;
    s.qsdot = 0.0;
    s.ret = 0;
    s.iter = 1;
    s.res = 0.0;
    s.status = qmrmax >= 1 ? QMR_RUNNING : QMR_EXIT;
    return true;
}

// element t of the workgroup's pencil chunk (p0, np) as a grid index
__device__ __forceinline__ long qp_elem(const QpW &W, int p0, float rp, int np, int t)
{
    const int j = fdiv_small(t, rp), pp = t - j * np;
    return p0 + pp + (long)W.P * j;
}

// c_i = v_i . w (i < j) into slot i and, with gj, v_{j-1} . v_i (i < j - 1)
// into slot ldg + i: one wave per basis vector, partial sums over the
// workgroup's chunk (w and v_{j-1} staged in wl / vl)
__device__ void qp_gs_dots(const QpW &W, const double *vjv, long ldv, int j, bool gj, const double *wl,
                           const double *vl, int p0, int np, double *op, int ldg)
{
    constexpr int K = 16;   // loads in flight per lane
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, ne = np * W.n3;
    const float rp = 1.0f / (float)(np > 0 ? np : 1);
    for (int i = wave; i < j; i += QP_NW) {
        const double *vi = vjv + (long)i * ldv;
        const bool wantg = gj && i < j - 1;
        double ac = 0.0, ag = 0.0;
        for (int t0 = 0; t0 < ne; t0 += 64 * K) {
            double x[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int t = t0 + lane + 64 * k;
                x[k] = t < ne ? vi[qp_elem(W, p0, rp, np, t)] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int t = t0 + lane + 64 * k;
                if (t < ne) {
                    ac += x[k] * wl[t];
                    if (wantg) ag += x[k] * vl[t];
                }
            }
        }
        ac = qp_wave_sum(ac);
        ag = qp_wave_sum(ag);
        if (lane == 0) {
            op[i * W.G + W.g] = ac;
            if (wantg) op[(ldg + i) * W.G + W.g] = ag;
        }
    }
}

// the sums of slots [s0, s0 + cnt) over the G workgroups into dst (LDS), one
// wave per slot
__device__ void qp_gs_reduce(const QpW &W, const double *op, int s0, int cnt, double *dst)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = wave; k < cnt; k += QP_NW) {
        const double r = qp_wave_sum(lane < W.G ? op[(s0 + k) * W.G + lane] : 0.0);
        if (lane == 0) dst[k] = r;
    }
}

// h = (I + L)^-1 c by forward substitution in wave 0 (L column-packed in Lp:
// column k holds rows k+1..j-1 from offset k (j-1) - k (k-1) / 2); c in cs,
// h into hs.  j <= 192.
__device__ void qp_gs_solve(int j, const double *Lp, const double *cs, double *hs)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    double c0 = lane < j ? cs[lane] : 0.0;
    double c1 = lane + 64 < j ? cs[lane + 64] : 0.0;
    double c2 = lane + 128 < j ? cs[lane + 128] : 0.0;
    for (int k = 0; k < j; k++) {
        const int src = k & 63, tk = k >> 6;
        const double hk = __shfl(tk == 0 ? c0 : (tk == 1 ? c1 : c2), src);
        const double *col = Lp + (long)k * (j - 1) - (long)k * (k - 1) / 2 - (k + 1);   // col[i] = L(i, k)
        if (lane > k && lane < j) c0 -= col[lane] * hk;
        if (lane + 64 > k && lane + 64 < j) c1 -= col[lane + 64] * hk;
        if (lane + 128 > k && lane + 128 < j) c2 -= col[lane + 128] * hk;
    }
    if (lane < j) hs[lane] = c0;
    if (lane + 64 < j) hs[lane + 64] = c1;
    if (lane + 128 < j) hs[lane + 128] = c2;
}

// MGS_ORTH (gsorth.f90:42-72) of w = X against columns 0..j-1 into column j
// of window v's basis, in the low-synchronisation form; a second pass on the
// normalised result when orth = 2.  ok = 0 on a zero norm (before or after
// the projections).  The own pencil chunk's elements only; the scratch of
// LDS: wl = pen, vl = sA/sB, the stage = c, g, h and the packed L.
__device__ bool qp_gs(QpW &W, const PpArgs &pa, int j, int &ok)
{
    QpLds &L = W.L;
    const int g = W.g, G = W.G, tid = threadIdx.x, ldg = pa.ldg;
    const bool own = g < W.nch;
    const int p0 = own ? g * W.pc : 0, np = own ? min(W.pc, W.P - p0) : 0, ne = np * W.n3;
    const float rp = 1.0f / (float)(np > 0 ? np : 1);
    double *vjv = pa.vj + (long)W.v * pa.ldw;
    double *gram = pa.gram + (long)W.v * ldg * ldg;
    double *op = pa.opart + (long)W.v * (2 * ldg + 2) * G;
    double *wl = L.pen, *vl = L.sA;
    double *cs = L.v1s, *gs = cs + ldg, *hs = gs + ldg, *Lp = hs + ldg;
    ok = 1;
    // O1: stage w and v_{j-1}; c = V^T w, the Gram row of v_{j-1}, w.w
    double ww = 0.0;
    for (int t = tid; t < ne; t += QP_NT) {
        const long e = qp_elem(W, p0, rp, np, t);
        const double x = W.X[e];
        wl[t] = x;
        ww += x * x;
        if (j >= 2) vl[t] = vjv[(long)(j - 1) * pa.ldv + e];
    }
    __syncthreads();
    qp_gs_dots(W, vjv, pa.ldv, j, j >= 2, wl, vl, p0, np, op, ldg);
    qp_put(op, 2 * ldg, G, ww, L.red);
    if (!qp_sync(W)) return false;
    // O2: the sums; L; h; w1 = w - V h; w1.w1 (and V^T w1 for a second pass)
    qp_gs_reduce(W, op, 0, j, cs);
    if (j >= 2) qp_gs_reduce(W, op, ldg, j - 1, gs);
    const double wsum = qp_grid_sum(op, 2 * ldg, G);
    __syncthreads();
    if (wsum == 0.0) {   // normw == 0
        ok = 0;
        return true;
    }
    if (g == 0)
        for (int k = tid; k < j - 1; k += QP_NT) gram[(j - 1) + (long)ldg * k] = gs[k];
    // the packed L: column k, rows k+1..j-1 (row j-1 from gs)
    for (int k = tid >> 6; k < j - 1; k += QP_NW) {
        double *col = Lp + (long)k * (j - 1) - (long)k * (k - 1) / 2 - (k + 1);
        for (int i = k + 1 + (tid & 63); i < j; i += 64) col[i] = (i == j - 1) ? gs[k] : gram[i + (long)ldg * k];
    }
    __syncthreads();
    qp_gs_solve(j, Lp, cs, hs);
    __syncthreads();
    double nn = 0.0;
    for (int t = tid; t < ne; t += QP_NT) {
        const long e = qp_elem(W, p0, rp, np, t);
        double acc = wl[t];
#pragma unroll 16
        for (int i = 0; i < j; i++) acc -= hs[i] * vjv[(long)i * pa.ldv + e];
        wl[t] = acc;
        nn += acc * acc;
    }
    __syncthreads();
    const bool twice = pa.orth >= 2;
    if (twice) qp_gs_dots(W, vjv, pa.ldv, j, false, wl, vl, p0, np, op, ldg);
    qp_put(op, 2 * ldg + 1, G, nn, L.red);
    if (!qp_sync(W)) return false;
    // O3: the norm; one pass: the new column
    const double nrm1 = sqrt(qp_grid_sum(op, 2 * ldg + 1, G));
    if (nrm1 == 0.0) {
        ok = 0;
        return true;
    }
    if (!twice) {
        for (int t = tid; t < ne; t += QP_NT) vjv[(long)j * pa.ldv + qp_elem(W, p0, rp, np, t)] = wl[t] / nrm1;
        __syncthreads();
        return true;
    }
    // second pass on w1 / |w1|: its projections are V^T w1 / |w1|
    qp_gs_reduce(W, op, 0, j, cs);
    __syncthreads();
    for (int k = tid; k < j; k += QP_NT) cs[k] /= nrm1;
    __syncthreads();
    qp_gs_solve(j, Lp, cs, hs);
    __syncthreads();
    nn = 0.0;
    for (int t = tid; t < ne; t += QP_NT) {
        const long e = qp_elem(W, p0, rp, np, t);
        double acc = wl[t] / nrm1;
#pragma unroll 16
        for (int i = 0; i < j; i++) acc -= hs[i] * vjv[(long)i * pa.ldv + e];
        wl[t] = acc;
        nn += acc * acc;
    }
    qp_put(op, 2 * ldg + 1, G, nn, L.red);
    if (!qp_sync(W)) return false;
    const double nrm2 = sqrt(qp_grid_sum(op, 2 * ldg + 1, G));
    if (nrm2 == 0.0) {
        ok = 0;
        return true;
    }
    for (int t = tid; t < ne; t += QP_NT) vjv[(long)j * pa.ldv + qp_elem(W, p0, rp, np, t)] = wl[t] / nrm2;
    __syncthreads();
    return true;
}

__global__ void __launch_bounds__(QP_NT) k_pist_persist(PpArgs pa)
{
    extern __shared__ double qp_sm[];
    const QpArgs &a = pa.q;
    if (!a.active[blockIdx.y]) return;   // written by the host before the launch
    QpW W;
    qp_setup(W, a, qp_sm);
    const int v = W.v;
    double *vjv = pa.vj + (long)v * pa.ldw;
    int wret = 1;
    bool alive = qp_sync(W);
    for (int j = pa.j_from; alive && j <= pa.j_to; j++) {
        W.B = vjv + (long)(j - 1) * pa.ldv;   // LinSolv(N, VJ(:, J), WJ, RES)
        QmrScal s;
        double vz_sum = 0.0;
        int status = QMR_EXIT;
        alive = qp_qmr_init(W, pa.zero_x != 0, a.qmrmax, s, vz_sum);
        if (alive && s.status == QMR_RUNNING) alive = qp_qmr_loop(W, s, vz_sum, status);
        if (!alive) break;
        if (W.g == 0 && threadIdx.x == 0) {
            pa.qret[(long)v * pa.ldg + j] = s.ret;
            if (j == pa.j_to || s.ret <= 0) a.q[v] = s;
        }
        if (s.ret <= 0) {   // pistHij.f90:105-108
            wret = s.ret;
            break;
        }
        int ok = 1;
        alive = qp_gs(W, pa, j, ok);
        if (alive && !ok) {   // pistHij.f90:110-112
            wret = -j;
            break;
        }
    }
    if (W.g == 0 && threadIdx.x == 0) pa.wret[v] = wret;   // a wait that gave up is reported through err
}

// ---------------------------------------------------------------- host side

// Whether the persistent solvers take this operator: the fused 3-level
// layout (sF = 3, DEP on level 3 only, no sNDVR layer) with its
// preconditioner, levels of at most QP_NMAX points, at most 65,536 points and
// theta planes of at most 1,536 points.  Past that a plane's phase is longer than
// the launches it saves (measured: 81 against 113 us per QMR iteration at
// 32x40x24, 151 against 121 at 40x48x32).
bool qmr_persist_ok(const scalit_ctx *c)
{
    if (!c->fast3 || c->ndvr || c->sF != 3 || !c->has_precond || !c->has_res) return false;
    if (c->osb == SCALIT_TOSBW) return false;   // the window solve couples level-1 blocks
    for (int l = 0; l < 3; l++)
        if (c->lv[l].n > QP_NMAX || c->lv[l].dep != (l == 2) || !c->lv[l].vosb.p) return false;
    return c->N <= 65536 && (long)c->lv[0].n * c->lv[1].n <= 1536;
}

static int env_or(const char *name, int def)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : def;
}

// the launch shape and the operator part of the arguments; false when the
// G workgroups of nvec vectors cannot all be resident at once
static bool qp_shape(scalit_ctx *c, int nvec, QpArgs &a, size_t &lds)
{
    a.n1 = c->lv[0].n;
    a.n2 = c->lv[1].n;
    a.n3 = c->lv[2].n;
    const int P = a.n1 * a.n2;
    int G = env_or("SCALIT_QMR_PERSIST_G", std::max(16, a.n3));
    G = std::max(1, std::min(G, QP_GMAX));
    const int ncu = g_num_cu.load();
    if (ncu <= 0) return false;
    // with 150 KB of LDS a workgroup has a CU to itself
    G = std::min(G, ncu / std::max(nvec, 1));
    if (G < 1) return false;
    a.pc = (P + G - 1) / G;
    // as many level-1 blocks per stage as fit in 150 KB of LDS
    const size_t base = qmr_persist_lds_base(a.n1, a.n2, a.n3, a.pc);
    const size_t per_blk = (size_t)a.n1 * (a.n1 + 1);
    if ((base + per_blk) * sizeof(double) > 150 * 1024) return false;
    a.kb = (int)std::min<size_t>((size_t)a.n2, (150 * 1024 / sizeof(double) - base) / per_blk);
    lds = qmr_persist_lds(a.n1, a.n2, a.n3, a.pc, a.kb);
    allow_full_lds((const void *)k_qmr_persist);
    allow_full_lds((const void *)k_pist_persist);
    const int per_cu = std::min(resident_per_cu((const void *)k_qmr_persist, QP_NT, lds),
                                resident_per_cu((const void *)k_pist_persist, QP_NT, lds));
    if ((long)G * nvec > (long)per_cu * ncu) return false;
    a.G = G;
    a.n = c->N;
    a.nosb = c->osb;
    a.de = c->de;
    a.beta = c->beta;
    a.h1 = c->lv[0].h0.p;
    a.h2 = c->lv[1].h0.p;
    a.h3 = c->lv[2].h0.p;
    a.dep = c->lv[2].depv.p;
    a.res = c->res.p;
    a.eig0 = c->eig0.p;
    a.v1 = c->lv[0].vosb.p;
    a.v2 = c->lv[1].vosb.p;
    a.v3 = c->lv[2].vosb.p;
    return true;
}

static int qp_buffers(scalit_ctx *c, int nvec, QpArgs &a, QmrPersistBufs &pb, hipStream_t s)
{
    SC_TRY(pb.part.alloc((size_t)nvec * QS_SLOTS * a.G));
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
    return SCALIT_OK;
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
    size_t lds = 0;
    if (!qp_shape(c, nvec, a, lds)) return QP_NOT_LAUNCHED;
    a.conv = conv;
    a.qmrmax = qmrmax;
    a.tol = tol;
    a.w = w;
    a.q = q;
    a.active = active;
    a.shifts = shifts;
    a.part_vz = part_vz;
    SC_TRY(qp_buffers(c, nvec, a, pb, s));
    void *args[] = {&a};
    HIP_TRY(hipLaunchCooperativeKernel((const void *)k_qmr_persist, dim3(a.G, nvec), dim3(QP_NT), args, (unsigned)lds,
                                       s));
    HIP_TRY(hipMemcpyAsync(&pb.herr, pb.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return pb.herr ? SCALIT_EHIP : SCALIT_OK;
}

// The largest basis the PIST kernel orthogonalises against: the Gram-Schmidt
// scratch (3 ldg values and the packed lower triangle) lives in the level-1
// stage, and the triangular solve holds up to 192 values per wave.
int pist_persist_max_cols(scalit_ctx *c, int nw, int ldg)
{
    QpArgs a;
    size_t lds = 0;
    if (!qmr_persist_ok(c) || !qp_shape(c, nw, a, lds)) return 0;
    // the chunk's elements must fit the scratch (pen, sA + sB)
    if ((long)a.pc * a.n3 > 2L * a.n1 * a.n2) return 0;
    const long stage = (long)a.kb * a.n1 * (a.n1 + 1) - 3L * ldg;
    int j = 0;
    while (j < 192 && (long)(j + 1) * j / 2 <= stage) j++;
    return j;
}

int pist_persist_steps(scalit_ctx *c, int nw, const PistPersistCall &pc, QmrVecs<double> w, QmrScal *q,
                       int *active, const double *shifts, QmrPersistBufs &pb, hipStream_t s)
{
    PpArgs pa;
    QpArgs &a = pa.q;
    size_t lds = 0;
    if (!qp_shape(c, nw, a, lds)) return QP_NOT_LAUNCHED;
    a.conv = pc.conv;
    a.qmrmax = pc.qmrmax;
    a.tol = pc.tol;
    w.b = nullptr;   // each step's right-hand side is a basis column
    w.ldb = 0;
    a.w = w;
    a.q = q;
    a.active = active;
    a.shifts = shifts;
    a.part_vz = nullptr;
    SC_TRY(qp_buffers(c, nw, a, pb, s));
    SC_TRY(pb.opart.alloc((size_t)nw * (2 * pc.ldg + 2) * a.G));
    pa.vj = pc.vj;
    pa.ldv = pc.ldv;
    pa.ldw = pc.ldw;
    pa.gram = pc.gram;
    pa.ldg = pc.ldg;
    pa.j_from = pc.j_from;
    pa.j_to = pc.j_to;
    pa.orth = pc.orth;
    pa.zero_x = pc.zero_x;
    pa.qret = pc.qret;
    pa.wret = pc.wret;
    pa.opart = pb.opart.p;
    void *args[] = {&pa};
    HIP_TRY(hipLaunchCooperativeKernel((const void *)k_pist_persist, dim3(a.G, nw), dim3(QP_NT), args, (unsigned)lds,
                                       s));
    HIP_TRY(hipMemcpyAsync(&pb.herr, pb.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return pb.herr ? SCALIT_EHIP : SCALIT_OK;
}

}  // namespace scalit
