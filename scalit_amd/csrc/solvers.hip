// Krylov solvers on the context's operator, everything resident on the GPU.
//
//  * QMR (src/iterate/libiterate/qmr/qmr.f90:5-175) for nvec right-hand sides
//    at once, each with its own shift; one iteration (2 H*X, 1 P*X, the fused
//    vector/reduction kernels) is recorded once in a HIP graph and replayed,
//    with a host status check every few replays.
//  * mixed precision: the same QMR on fp32 or bf16-stored vectors with fp32
//    MFMA operators, wrapped in fp64 iterative refinement.
//  * PIST_CONV / pistH0 (src/iterate/libiterate/pist/pist.f90:7-98,
//    pistHij.f90:328-383) for nwin shift windows in lock-step; MGS_ORTH
//    (src/iterate/libiterate/orth/gsorth.f90:42-72) and the Rayleigh-Ritz
//    columns on the GPU in fp64; the small eigensolves (the reference's
//    DSYEV('N')) on the GPU; window selection over the <= M_MAX Ritz values
//    on the host, as the reference's root rank does.
//  * Lan_Conv / lanH0 (src/iterate/libiterate/lanczos/lan.f90:7-105,
//    lanHij.f90:6-45): fp64 Lanczos with full re-orthogonalisation.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "scalit_impl.hpp"
#include "solver_state.hpp"

namespace scalit {

int ensure_f32(scalit_ctx *c);

// ---------------------------------------------------------------- helpers
// src/libutil/serial/base/util.f90:669-717 (first index wins ties)
static void get_window0(double e0, int n, const double *vec, int m, int *ind, double *out)
{
    if (n < m) {
        for (int i = 0; i < n; i++) {
            out[i] = vec[i];
            ind[i] = i;
        }
        for (int i = std::max(n - 1, 0); i < m; i++) {
            out[i] = vec[n - 1];
            ind[i] = n - 1;
        }
        return;
    }
    std::vector<char> label(n, 1);
    for (int i = 0; i < m; i++) {
        int index = -1;
        double val0 = 0;
        for (int j = 0; j < n; j++) {
            if (!label[j]) continue;
            double a = std::fabs(vec[j] - e0);
            if (index < 0 || val0 > a) {
                index = j;
                val0 = a;
            }
        }
        ind[i] = index;
        out[i] = vec[index];
        label[index] = 0;
    }
}
// util.f90:84-100
static double get_diff_max(int n, const double *a, const double *b)
{
    double m = std::fabs(a[0] - b[0]);
    for (int i = 1; i < n; i++) m = std::max(m, std::fabs(a[i] - b[i]));
    return m;
}
// util.f90:124-149
static double get_sep_max(int n1, const double *v1, int n2, const double *v2)
{
    double g = 0;
    for (int i = 0; i < n1; i++) {
        double mm = std::fabs(v1[i] - v2[0]);
        for (int j = 1; j < n2; j++) mm = std::min(mm, std::fabs(v1[i] - v2[j]));
        if (i == 0 || g < mm) g = mm;
    }
    return g;
}

// ---------------------------------------------------------------- operators per storage type
template <typename S>
struct Ops;
template <>
struct Ops<double> {
    static int hx(scalit_ctx *c, const double *sh, int nv, const double *x, long ldx, double *y, long ldy,
                  const int *act, hipStream_t s)
    {
        return hx_apply_f64(c, sh, c->e0, nv, x, ldx, y, ldy, false, act, s);
    }
    static int px(scalit_ctx *c, const double *sh, int nv, const double *x, long ldx, double *y, long ldy,
                  const int *act, hipStream_t s)
    {
        return px_apply_f64(c, c->osb, sh, c->e0, nv, x, ldx, y, ldy, act, s);
    }
};
template <>
struct Ops<float> {
    static int hx(scalit_ctx *c, const float *sh, int nv, const float *x, long ldx, float *y, long ldy,
                  const int *act, hipStream_t s)
    {
        return hx_apply_f32(c, sh, (float)c->e0, nv, x, ldx, y, ldy, act, s);
    }
    static int px(scalit_ctx *c, const float *sh, int nv, const float *x, long ldx, float *y, long ldy,
                  const int *act, hipStream_t s)
    {
        return px_apply_f32(c, c->osb, sh, (float)c->e0, nv, x, ldx, y, ldy, act, s);
    }
};
template <>
struct Ops<bf16_t> {
    static int hx(scalit_ctx *c, const float *sh, int nv, const bf16_t *x, long ldx, bf16_t *y, long ldy,
                  const int *act, hipStream_t s)
    {
        return hx_apply_bf16(c, sh, (float)c->e0, nv, x, ldx, y, ldy, act, s);
    }
    static int px(scalit_ctx *c, const float *sh, int nv, const bf16_t *x, long ldx, bf16_t *y, long ldy,
                  const int *act, hipStream_t s)
    {
        return px_apply_bf16(c, c->osb, sh, (float)c->e0, nv, x, ldx, y, ldy, act, s);
    }
};
template <typename S>
struct ShiftT {
    using type = float;
};
template <>
struct ShiftT<double> {
    using type = double;
};

// ---------------------------------------------------------------- QMR workspace
template <typename S>
struct QmrWS {
    int nvec = 0;
    long n = 0, ld = 0;
    DBuf<S> vecs;        // B, X, R, UX, UR, Q, S, V, Z, HX0
    DBuf<double> parts;  // 10 partial arrays
    DBuf<QmrScal> scal;
    DBuf<int> active;
    DBuf<typename ShiftT<S>::type> shifts;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_conv = -99, g_qmrmax = -1;
    double g_tol = -1;
    long g_gen = -1;   // operator buffer generation the graph was recorded with
    QmrPersistBufs persist;
    std::vector<QmrScal> hscal;
    std::vector<int> hact, hones;
    std::vector<typename ShiftT<S>::type> hshift;

    S *col(int which) { return vecs.p + (long)which * nvec * ld; }
    S *B() { return col(0); }
    S *X() { return col(1); }
    double *part(int k) { return parts.p + (long)k * nvec * NPART; }
    QmrVecs<S> view()
    {
        QmrVecs<S> w;
        w.b = B();
        w.x = X();
        w.ldb = ld;
        w.ldx = ld;
        w.R = col(2);
        w.UX = col(3);
        w.UR = col(4);
        w.Q = col(5);
        w.Sv = col(6);
        w.V = col(7);
        w.Z = col(8);
        w.HX0 = col(9);
        w.ld = ld;
        return w;
    }
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        g_conv = -99;
    }
    int ensure(int nv, long nn)
    {
        if (nv == nvec && nn == n && vecs.p) return SCALIT_OK;
        drop_graph();
        nvec = nv;
        n = nn;
        ld = (nn + 63) & ~63L;
        SC_TRY(vecs.alloc((size_t)10 * nv * ld));
        SC_TRY(parts.alloc((size_t)10 * nv * NPART));
        SC_TRY(scal.alloc(nv));
        SC_TRY(active.alloc(nv));
        SC_TRY(shifts.alloc(nv));
        hscal.resize(nv);
        hact.resize(nv);
        return SCALIT_OK;
    }
    void release()
    {
        drop_graph();
        vecs.release();
        parts.release();
        scal.release();
        active.release();
        shifts.release();
        persist.part.release();
        persist.bar.release();
        persist.err.release();
        nvec = 0;
    }
};

// Operator callbacks supplied by a caller of the Fortran layer: the vectors
// stay on the GPU and only the operator applications round-trip to the host.
struct HostCb {
    scalit_fop hx = nullptr, px = nullptr;
    std::vector<double> hin, hout;
};

template <typename S>
static int apply_host(HostCb *cb, scalit_fop f, QmrWS<S> &w, const S *x, S *y, hipStream_t s)
{
    if constexpr (!std::is_same<S, double>::value) {
        return SCALIT_EINVAL;
    } else {
        const int n = (int)w.n;
        cb->hin.resize(n);
        cb->hout.resize(n);
        for (int v = 0; v < w.nvec; v++) {
            HIP_TRY(hipMemcpyAsync(cb->hin.data(), x + (long)v * w.ld, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            f(&n, cb->hin.data(), cb->hout.data());
            HIP_TRY(hipMemcpyAsync(y + (long)v * w.ld, cb->hout.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        return SCALIT_OK;
    }
}

template <typename S>
static int op_hx(scalit_ctx *c, QmrWS<S> &w, HostCb *cb, const S *x, S *y, const int *act, hipStream_t s)
{
    if (cb) return apply_host<S>(cb, cb->hx, w, x, y, s);
    return Ops<S>::hx(c, w.shifts.p, w.nvec, x, w.ld, y, w.ld, act, s);
}
template <typename S>
static int op_px(scalit_ctx *c, QmrWS<S> &w, HostCb *cb, const S *x, S *y, const int *act, hipStream_t s)
{
    if (cb) return apply_host<S>(cb, cb->px, w, x, y, s);
    return Ops<S>::px(c, w.shifts.p, w.nvec, x, w.ld, y, w.ld, act, s);
}

// one QMR iteration, launched (or recorded) on stream s
template <typename S>
static int qmr_iteration(scalit_ctx *c, QmrWS<S> &w, int conv, int qmrmax, double tol, hipStream_t s,
                         HostCb *cb = nullptr)
{
    const int nv = w.nvec;
    const long n = w.n;
    QmrVecs<S> v = w.view();
    QmrScal *q = w.scal.p;
    int *act = w.active.p;
    launch_qmr_b<S>(nv, n, v, q, act, w.part(9), s);
    SC_TRY(op_hx<S>(c, w, cb, v.Q, v.Sv, act, s));
    launch_qmr_c<S>(nv, n, v, q, w.part(1), s);
    launch_qmr_d<S>(nv, n, v, q, act, w.part(1), w.part(2), s);
    SC_TRY(op_px<S>(c, w, cb, v.V, v.Z, act, s));
    launch_qmr_e<S>(nv, n, v, q, w.part(2), w.part(3), w.part(9), s);
    launch_qmr_f<S>(nv, n, conv, v, q, act, w.part(3), w.part(4), w.part(5), w.part(6), w.part(7), s);
    if (conv == 1) {
        SC_TRY(op_hx<S>(c, w, cb, v.x, v.HX0, act, s));
        launch_qmr_g<S>(nv, n, v, q, w.part(8), w.part(4), s);
    }
    launch_qmr_h(w.n, nv, conv, qmrmax, tol, q, act, w.part(8), w.part(4), w.part(5), w.part(6), w.part(7), s);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// QMR iterations are replayed from a recorded HIP graph unless
// SCALIT_QMR_GRAPH=0; the host reads the convergence flags every
// SCALIT_QMR_CHECK (default 4) iterations.
static int env_int(const char *name, int def)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : def;
}
int g_qmr_use_graph = env_int("SCALIT_QMR_GRAPH", 1);
int g_qmr_check_every = std::max(1, env_int("SCALIT_QMR_CHECK", 4));
// FP64 solves on operators the persistent solver takes run it unless
// SCALIT_QMR_PERSIST=0
int g_qmr_persist = env_int("SCALIT_QMR_PERSIST", 1);
static bool qmr_persist_wanted(const scalit_ctx *c)
{
    return (c->qmr_persist < 0 ? g_qmr_persist : c->qmr_persist) != 0;
}

// Solve (H - shift_v) x_v = b_v for the nvec columns already in w.B(); the
// solution lands in w.X().  shifts: host, nvec values (the QMR shift E0 per
// vector).  ret / res: host, per vector.
// The shifts and activity flags of a solve, and its initialisation
// (qmr.f90:36-50), on stream s.
template <typename S>
static int qmr_prelude(scalit_ctx *c, QmrWS<S> &w, int ntype, int qmrmax, const double *shifts, hipStream_t s,
                       HostCb *cb = nullptr)
{
    const int nv = w.nvec;
    const long n = w.n;
    // the host sources live in the workspace, so the copies need no wait here:
    // they are rewritten only by the next solve, after this one's last wait
    w.hshift.resize(nv);
    for (int i = 0; i < nv; i++) w.hshift[i] = (typename ShiftT<S>::type)shifts[i];
    HIP_TRY(hipMemcpyAsync(w.shifts.p, w.hshift.data(), sizeof(w.hshift[0]) * nv, hipMemcpyHostToDevice, s));
    w.hones.assign(nv, 1);
    HIP_TRY(hipMemcpyAsync(w.active.p, w.hones.data(), sizeof(int) * nv, hipMemcpyHostToDevice, s));
    QmrVecs<S> v = w.view();
    launch_qmr_init<S>(nv, n, ntype <= 0 ? 1 : 0, v, w.part(0), s);
    SC_TRY(op_px<S>(c, w, cb, v.b, v.Z, w.active.p, s));
    launch_qmr_init_z<S>(nv, n, v, w.part(0), w.part(1), w.part(9), s);
    launch_qmr_init_scal(nv, w.part(0), w.part(1), w.scal.p, w.active.p, qmrmax, s);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

template <typename S>
static int qmr_solve(scalit_ctx *c, QmrWS<S> &w, int ntype, int qmrmax, double tol, const double *shifts, int *ret,
                     double *res, hipStream_t s, HostCb *cb = nullptr)
{
    const int nv = w.nvec;
    const int conv = std::abs(ntype) == 1 ? 1 : (std::abs(ntype) == 2 ? 2 : 0);
    SC_TRY(qmr_prelude<S>(c, w, ntype, qmrmax, shifts, s, cb));
    QmrVecs<S> v = w.view();
    if constexpr (std::is_same<S, double>::value) {
        if (cb == nullptr && qmr_persist_wanted(c) && qmrmax >= 1 && qmr_persist_ok(c)) {
            const int rc = qmr_persist_solve(c, nv, v, w.scal.p, w.active.p, w.shifts.p, w.part(9), conv, qmrmax, tol,
                                             w.persist, s);
            if (rc != QP_NOT_LAUNCHED) {
                SC_TRY(rc);
                HIP_TRY(hipMemcpyAsync(w.hscal.data(), w.scal.p, sizeof(QmrScal) * nv, hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                for (int i = 0; i < nv; i++) {
                    if (ret) ret[i] = w.hscal[i].ret;
                    if (res) res[i] = w.hscal[i].res;
                }
                return SCALIT_OK;
            }
        }
    }
    const bool use_graph = g_qmr_use_graph != 0 && cb == nullptr;
    if (cb == nullptr) {
        const int prec = std::is_same<S, double>::value ? SCALIT_FP64
                                                        : (std::is_same<S, float>::value ? SCALIT_FP32 : SCALIT_BF16);
        SC_TRY(hx_prepare(c, prec, nv, s));
    }
    if (use_graph && (w.g_conv != conv || w.g_qmrmax != qmrmax || w.g_tol != tol || w.g_gen != c->gen || !w.exec)) {
        w.drop_graph();
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = qmr_iteration<S>(c, w, conv, qmrmax, tol, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != SCALIT_OK) return rc;
        HIP_TRY(e);
        w.graph = g;
        HIP_TRY(hipGraphInstantiate(&w.exec, g, nullptr, nullptr, 0));
        w.g_conv = conv;
        w.g_qmrmax = qmrmax;
        w.g_tol = tol;
        w.g_gen = c->gen;
    }
    int done_iters = 0;
    while (done_iters < qmrmax) {
        int k = std::min(g_qmr_check_every, qmrmax - done_iters);
        for (int i = 0; i < k; i++) {
            if (use_graph) HIP_TRY(hipGraphLaunch(w.exec, s));
            else SC_TRY(qmr_iteration<S>(c, w, conv, qmrmax, tol, s, cb));
        }
        done_iters += k;
        // the flags and the scalars together: after the last check the
        // results are already on the host
        HIP_TRY(hipMemcpyAsync(w.hact.data(), w.active.p, sizeof(int) * nv, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(w.hscal.data(), w.scal.p, sizeof(QmrScal) * nv, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        bool any = false;
        for (int i = 0; i < nv; i++) any = any || w.hact[i];
        if (!any) break;
    }
    if (done_iters == 0) {
        HIP_TRY(hipMemcpyAsync(w.hscal.data(), w.scal.p, sizeof(QmrScal) * nv, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    for (int i = 0; i < nv; i++) {
        if (ret) ret[i] = w.hscal[i].ret;
        if (res) res[i] = w.hscal[i].res;
    }
    return SCALIT_OK;
}

// ---------------------------------------------------------------- per-context solver state
// A recorded multi-problem QMR iteration (scalit_qmr_multi), kept by the
// first job's context for the same jobs, workspaces and options.
struct MultiGraph {
    std::vector<long> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
};

struct PistState {
    MultiGraph mg;
    QmrWS<double> q64;
    QmrWS<float> q32;
    QmrWS<bf16_t> q16;
    DBuf<double> vj, hmat, w, eigw, zw, part, part2, tmp, r64, x64, psi_z, zsh, bb, shd;
    DBuf<int> ok, active, cols;
    DBuf<EigvalJob> ejobs;
    DBuf<EigvecJob> vjobs;
    DBuf<RitzJob> rjobs;
    DBuf<double> gpart, gred;   // Hellmann-Feynman partial sums and results
    // Lanczos (kept between calls: a device free would synchronise the whole device)
    DBuf<double> lvj, lalpha, lbeta, lw, lpart, lpart2, lnrm, lev, lzero;
    DBuf<TriEigJob> lejobs;
    DBuf<int> lokv;
    int nwin = 0, m_max = 0, m_last = 0;
    std::vector<int> last_m;
    std::vector<std::vector<int>> last_ind;   // window index of each returned eigenvalue
    std::vector<int> last_cnt;                // Krylov size of each window's last eigensolve
    std::vector<int> last_ret;                // each window's PIST_CONV return code
    long ldv = 0;
    // the persistent Krylov steps: the basis columns' Gram matrix per window,
    // valid while every step of the run went through them
    DBuf<double> gram;
    DBuf<int> qret, wret;
    bool gram_valid = false;
    std::vector<int> hqret, hwret;
    // QMR parameters of the OSB_QMR callback
    int qmr_type = -1, qmr_max = 1000;
    double qmr_tol = 1e-3;
    void release_multi() { mg.drop(); }
    void release()
    {
        q64.release();
        q32.release();
        q16.release();
        vj.release();
        hmat.release();
        w.release();
        eigw.release();
        zw.release();
        part.release();
        part2.release();
        tmp.release();
        r64.release();
        x64.release();
        psi_z.release();
        zsh.release();
        bb.release();
        shd.release();
        ok.release();
        active.release();
        cols.release();
        ejobs.release();
        vjobs.release();
        rjobs.release();
        gpart.release();
        gred.release();
        for (DBuf<double> *b : {&lvj, &lalpha, &lbeta, &lw, &lpart, &lpart2, &lnrm, &lev, &lzero})
            b->release();
        lejobs.release();
        lokv.release();
        gram.release();
        qret.release();
        wret.release();
    }
};

static PistState &pstate(scalit_ctx *c)
{
    if (!c->pist) c->pist = new PistState();
    return *c->pist;
}

void release_solvers(scalit_ctx *c)
{
    if (c->pist) {
        c->pist->release_multi();
        c->pist->release();
        delete c->pist;
        c->pist = nullptr;
    }
}

// ---------------------------------------------------------------- public QMR
int qmr_f64(scalit_ctx *c, int ntype, int qmrmax, double tol, int nvec, const double *b, double *x, long ld,
            const double *shifts, int *ret, double *res, hipStream_t s)
{
    PistState &P = pstate(c);
    SC_TRY(P.q64.ensure(nvec, c->N));
    auto &w = P.q64;
    HIP_TRY(hipMemcpy2DAsync(w.B(), w.ld * 8, b, ld * 8, c->N * 8, nvec, hipMemcpyDeviceToDevice, s));
    if (ntype > 0) HIP_TRY(hipMemcpy2DAsync(w.X(), w.ld * 8, x, ld * 8, c->N * 8, nvec, hipMemcpyDeviceToDevice, s));
    std::vector<double> sh(nvec, c->e0);
    if (shifts)
        for (int i = 0; i < nvec; i++) sh[i] = shifts[i];
    SC_TRY(qmr_solve<double>(c, w, ntype, qmrmax, tol, sh.data(), ret, res, s));
    HIP_TRY(hipMemcpy2DAsync(x, ld * 8, w.X(), w.ld * 8, c->N * 8, nvec, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

// Several operators' FP64 QMR solves (one per context, all on one device)
// in lock step: one iteration of every solve, each on its context's stream
// (a branch per problem, forked from and joined to the first context's
// stream), is recorded as one HIP graph, and each graph launch advances all
// of them.  A converged or broken-down right-hand side stops updating (its
// activity flag), so problems that finish early cost empty launches only
// until the last one converges.
int qmr_multi_f64(int njob, scalit_qmr_job *jobs, int ntype, int qmrmax, double tol)
{
    const int conv = std::abs(ntype) == 1 ? 1 : (std::abs(ntype) == 2 ? 2 : 0);
    std::vector<QmrWS<double> *> ws(njob);
    for (int j = 0; j < njob; j++) {
        scalit_ctx *c = jobs[j].ctx;
        hipStream_t s = c->stream;
        PistState &P = pstate(c);
        SC_TRY(P.q64.ensure(jobs[j].nvec, c->N));
        auto &w = P.q64;
        ws[j] = &w;
        HIP_TRY(hipMemcpy2DAsync(w.B(), w.ld * 8, jobs[j].b, jobs[j].ld * 8, c->N * 8, jobs[j].nvec,
                                 hipMemcpyDeviceToDevice, s));
        if (ntype > 0)
            HIP_TRY(hipMemcpy2DAsync(w.X(), w.ld * 8, jobs[j].x, jobs[j].ld * 8, c->N * 8, jobs[j].nvec,
                                     hipMemcpyDeviceToDevice, s));
        std::vector<double> sh(jobs[j].nvec, c->e0);
        if (jobs[j].shifts)
            for (int i = 0; i < jobs[j].nvec; i++) sh[i] = jobs[j].shifts[i];
        SC_TRY(qmr_prelude<double>(c, w, ntype, qmrmax, sh.data(), s));
        SC_TRY(hx_prepare(c, SCALIT_FP64, jobs[j].nvec, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    hipStream_t s0 = jobs[0].ctx->stream;
    // the recorded iteration is reused while the jobs, their workspaces, the
    // operators' buffers and the options are those it was recorded with
    MultiGraph &mg = pstate(jobs[0].ctx).mg;
    std::vector<long> key{(long)conv, (long)qmrmax, 0L};
    std::memcpy(&key[2], &tol, sizeof(double));
    for (int j = 0; j < njob; j++) {
        const auto &w = *ws[j];
        key.insert(key.end(), {jobs[j].ctx->uid, jobs[j].ctx->gen, (long)w.nvec, (long)(intptr_t)w.vecs.p,
                               (long)(intptr_t)w.parts.p, (long)(intptr_t)w.scal.p, (long)(intptr_t)w.active.p,
                               (long)(intptr_t)w.shifts.p});
    }
    int rc = SCALIT_OK;
    if (!mg.exec || mg.key != key) {
        mg.drop();
        std::vector<hipEvent_t> ev(njob, nullptr);
        for (auto &e : ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                e = nullptr;
                rc = SCALIT_EHIP;
            }
        if (rc == SCALIT_OK && hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess)
            rc = SCALIT_EHIP;
        if (rc == SCALIT_OK) {
            (void)hipEventRecord(ev[0], s0);
            for (int j = 0; j < njob && rc == SCALIT_OK; j++) {
                hipStream_t s = jobs[j].ctx->stream;
                if (j > 0) (void)hipStreamWaitEvent(s, ev[0], 0);
                rc = qmr_iteration<double>(jobs[j].ctx, *ws[j], conv, qmrmax, tol, s);
                if (j > 0) {
                    (void)hipEventRecord(ev[j], s);
                    (void)hipStreamWaitEvent(s0, ev[j], 0);
                }
            }
            const hipError_t e = hipStreamEndCapture(s0, &mg.graph);
            if (rc == SCALIT_OK && e != hipSuccess) rc = SCALIT_EHIP;
            if (rc == SCALIT_OK && hipGraphInstantiate(&mg.exec, mg.graph, nullptr, nullptr, 0) != hipSuccess)
                rc = SCALIT_EHIP;
        }
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (rc == SCALIT_OK) mg.key = key;
        else mg.drop();
    }
    hipGraphExec_t exec = mg.exec;
    for (int done = 0; rc == SCALIT_OK && done < qmrmax;) {
        const int k = std::min(g_qmr_check_every, qmrmax - done);
        for (int i = 0; i < k && rc == SCALIT_OK; i++)
            if (hipGraphLaunch(exec, s0) != hipSuccess) rc = SCALIT_EHIP;
        done += k;
        bool any = false;
        for (int j = 0; j < njob && rc == SCALIT_OK; j++) {
            auto &w = *ws[j];
            if (hipMemcpyAsync(w.hact.data(), w.active.p, sizeof(int) * w.nvec, hipMemcpyDeviceToHost, s0) !=
                    hipSuccess ||
                hipMemcpyAsync(w.hscal.data(), w.scal.p, sizeof(QmrScal) * w.nvec, hipMemcpyDeviceToHost, s0) !=
                    hipSuccess)
                rc = SCALIT_EHIP;
        }
        if (rc == SCALIT_OK && hipStreamSynchronize(s0) != hipSuccess) rc = SCALIT_EHIP;
        for (int j = 0; j < njob; j++)
            for (int i = 0; i < ws[j]->nvec; i++) any = any || ws[j]->hact[i];
        if (!any) break;
    }
    if (rc == SCALIT_OK && qmrmax < 1)
        for (int j = 0; j < njob && rc == SCALIT_OK; j++) {
            auto &w = *ws[j];
            if (hipMemcpyAsync(w.hscal.data(), w.scal.p, sizeof(QmrScal) * w.nvec, hipMemcpyDeviceToHost, s0) !=
                hipSuccess)
                rc = SCALIT_EHIP;
        }
    for (int j = 0; j < njob && rc == SCALIT_OK; j++) {
        auto &w = *ws[j];
        if (hipMemcpy2DAsync(jobs[j].x, jobs[j].ld * 8, w.X(), w.ld * 8, jobs[j].ctx->N * 8, w.nvec,
                             hipMemcpyDeviceToDevice, s0) != hipSuccess)
            rc = SCALIT_EHIP;
    }
    if (rc == SCALIT_OK && hipStreamSynchronize(s0) != hipSuccess) rc = SCALIT_EHIP;
    SC_TRY(rc);
    for (int j = 0; j < njob; j++) {
        auto &w = *ws[j];
        for (int i = 0; i < w.nvec; i++) {
            jobs[j].ret[i] = w.hscal[i].ret;
            if (jobs[j].res) jobs[j].res[i] = w.hscal[i].res;
        }
        jobs[j].status = SCALIT_OK;
    }
    return SCALIT_OK;
}

// QMR with caller-supplied host callbacks (Fortran layer): b, x host arrays
int qmr_host_cb(scalit_ctx *c, int ntype, int qmrmax, double tol, long n, const double *b, double *x,
                scalit_fop hx, scalit_fop px, int *ret, double *res)
{
    PistState &P = pstate(c);
    hipStream_t s = c->stream;
    SC_TRY(P.q64.ensure(1, n));
    auto &w = P.q64;
    HIP_TRY(hipMemcpyAsync(w.B(), b, 8 * n, hipMemcpyHostToDevice, s));
    if (ntype > 0) HIP_TRY(hipMemcpyAsync(w.X(), x, 8 * n, hipMemcpyHostToDevice, s));
    HostCb cb;
    cb.hx = hx;
    cb.px = px;
    double sh = 0.0;
    SC_TRY(qmr_solve<double>(c, w, ntype, qmrmax, tol, &sh, ret, res, s, &cb));
    HIP_TRY(hipMemcpyAsync(x, w.X(), 8 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

// MGS_ORTH on host arrays: VJ (n x (m+1)) staged to the GPU
int mgs_host(scalit_ctx *c, long n, const double *wj, int m, double *vj, int *ok)
{
    hipStream_t s = c->stream;
    const long ld = n;
    DBuf<double> v, pa, pb, nr;
    SC_TRY(v.alloc((size_t)(m + 1) * ld));
    SC_TRY(pa.alloc(NPART));
    SC_TRY(pb.alloc(NPART));
    SC_TRY(nr.alloc(1));
    DBuf<int> okd;
    SC_TRY(okd.alloc(1));
    HIP_TRY(hipMemcpyAsync(v.p, vj, 8 * (size_t)m * n, hipMemcpyHostToDevice, s));
    double *col = v.p + (long)m * ld;
    HIP_TRY(hipMemcpyAsync(col, wj, 8 * n, hipMemcpyHostToDevice, s));
    double nw = 0;
    int okh = 0;
    if (launch_mgs_one(1, n, m, v.p, ld, 0, col, 0, col, nullptr, okd.p, nullptr, s)) {
        nw = 1.0;   // a zero norm of WJ shows in ok
    } else {
        launch_dot_partials(1, n, col, ld, col, ld, pa.p, s);
        launch_reduce_partials(1, pa.p, nr.p, s);
        HIP_TRY(hipMemcpyAsync(&nw, nr.p, 8, hipMemcpyDeviceToHost, s));
        if (m > 0) launch_mgs_step(1, n, -1, 0, v.p, ld, 0, col, 0, nullptr, pa.p, nullptr, s);
        double *p0 = pa.p, *p1 = pb.p;
        for (int i = 0; i < m; i++) {
            int inext = (i + 1 < m) ? i + 1 : -1;
            launch_mgs_step(1, n, i, inext, v.p, ld, 0, col, 0, p0, p1, nullptr, s);
            std::swap(p0, p1);
        }
        if (m == 0) launch_dot_partials(1, n, col, ld, col, ld, p0, s);
        launch_normalize(1, n, col, ld, col, ld, p0, okd.p, nullptr, s);
    }
    HIP_TRY(hipMemcpyAsync(&okh, okd.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(vj + (long)m * n, col, 8 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *ok = (nw != 0.0) && okh;
    v.release();
    pa.release();
    pb.release();
    nr.release();
    okd.release();
    return SCALIT_OK;
}

// Copy na columns of n doubles, column i from src + sidx[i] sld to
// dst + didx[i] dld (a null index list means i); runs of consecutive columns
// go as one 2-D copy (one launch) while both pitches stay below 2 GiB.
static int copy_cols(long n, int na, double *dst, long dld, const int *didx, const double *src, long sld,
                     const int *sidx, hipStream_t s)
{
    auto si = [&](int k) { return sidx ? sidx[k] : k; };
    auto di = [&](int k) { return didx ? didx[k] : k; };
    const bool rect = 8 * sld < 0x7fffffffL && 8 * dld < 0x7fffffffL;
    for (int i = 0; i < na;) {
        int j = i + 1;
        while (rect && j < na && si(j) == si(j - 1) + 1 && di(j) == di(j - 1) + 1) j++;
        if (j - i == 1)
            HIP_TRY(hipMemcpyAsync(dst + (long)di(i) * dld, src + (long)si(i) * sld, 8 * n, hipMemcpyDeviceToDevice,
                                   s));
        else
            HIP_TRY(hipMemcpy2DAsync(dst + (long)di(i) * dld, 8 * dld, src + (long)si(i) * sld, 8 * sld, 8 * n, j - i,
                                     hipMemcpyDeviceToDevice, s));
        i = j;
    }
    return SCALIT_OK;
}

// Mixed precision: inner QMR in fp32 / bf16 on the residual, outer fp64
// refinement x += d, r = b - (H - E0) x, until the reference's type-1
// criterion ||(H-E0)x - b|| / ||(H-E0)x|| < tol (or ||r||/||b|| for other
// types) holds or outer_max refinements are spent.  ret = total inner
// iterations (negated when not converged).  The inner solves stop on the
// recurrence residual (type 0): a true residual formed in fp32 stalls near
// eps32 * cond, far above the inner tolerance for the near-singular shifts
// PIST uses, while the outer fp64 residual decides convergence anyway.
int qmr_mixed(scalit_ctx *c, int prec, int ntype, int qmrmax, double tol, int inner_max, double inner_tol, int nvec,
              const double *b, long ldb, double *x, long ldx, const double *shifts, int *ret, double *res,
              hipStream_t s)
{
    PistState &P = pstate(c);
    SC_TRY(ensure_f32(c));
    const long N = c->N;
    SC_TRY(P.r64.alloc((size_t)nvec * N));
    SC_TRY(P.tmp.alloc((size_t)nvec * N));
    SC_TRY(P.part.alloc((size_t)2 * nvec * NPART));
    SC_TRY(P.part2.alloc((size_t)2 * nvec));
    std::vector<double> sh(nvec, c->e0);
    if (shifts)
        for (int i = 0; i < nvec; i++) sh[i] = shifts[i];
    DBuf<double> &shd = P.shd;
    SC_TRY(shd.alloc(nvec));
    HIP_TRY(hipMemcpyAsync(shd.p, sh.data(), 8 * nvec, hipMemcpyHostToDevice, s));
    std::vector<int> total(nvec, 0), conv(nvec, 0), rin(nvec);
    std::vector<double> bn(nvec), rr(nvec), hh(nvec), rsn(nvec);
    // ||b||
    SC_TRY(dot_host(nvec, N, b, ldb, b, ldb, bn.data(), P.part.p, P.part2.p, s));
    // x = 0, r = b
    HIP_TRY(hipMemset2DAsync(x, 8 * ldx, 0, 8 * N, nvec, s));
    SC_TRY(copy_cols(N, nvec, P.r64.p, N, nullptr, b, ldb, nullptr, s));
    const int outer_max = std::max(1, qmrmax);
    int outer = 0;
    for (; outer < outer_max; outer++) {
        // inner solve on the residual
        if (prec == SCALIT_BF16) {
            SC_TRY(P.q16.ensure(nvec, N));
            launch_cast_d2b(nvec, N, P.r64.p, N, P.q16.B(), P.q16.ld, s);
            SC_TRY(qmr_solve<bf16_t>(c, P.q16, 0, inner_max, inner_tol, sh.data(), rin.data(), nullptr, s));
            launch_cast_b2d(nvec, N, P.q16.X(), P.q16.ld, P.tmp.p, N, s);
        } else {
            SC_TRY(P.q32.ensure(nvec, N));
            launch_cast_d2f(nvec, N, P.r64.p, N, P.q32.B(), P.q32.ld, s);
            SC_TRY(qmr_solve<float>(c, P.q32, 0, inner_max, inner_tol, sh.data(), rin.data(), nullptr, s));
            launch_cast_f2d(nvec, N, P.q32.X(), P.q32.ld, P.tmp.p, N, s);
        }
        for (int v = 0; v < nvec; v++) {
            total[v] += std::abs(rin[v]);
            if (!conv[v]) launch_axpby(1, N, 1.0, P.tmp.p + (long)v * N, N, 1.0, x + (long)v * ldx, ldx, s);
        }
        // fp64 residual: hx = (H - E0) x ;  r = b - hx
        SC_TRY(hx_apply_f64(c, shd.p, c->e0, nvec, x, ldx, P.tmp.p, N, false, nullptr, s));
        SC_TRY(dot_host(nvec, N, P.tmp.p, N, P.tmp.p, N, hh.data(), P.part.p, P.part2.p, s));
        SC_TRY(copy_cols(N, nvec, P.r64.p, N, nullptr, b, ldb, nullptr, s));
        launch_axpby(nvec, N, -1.0, P.tmp.p, N, 1.0, P.r64.p, N, s);
        SC_TRY(dot_host(nvec, N, P.r64.p, N, P.r64.p, N, rr.data(), P.part.p, P.part2.p, s));
        bool all = true;
        for (int v = 0; v < nvec; v++) {
            double r = std::abs(ntype) == 1 ? std::sqrt(rr[v]) / std::sqrt(hh[v]) : std::sqrt(rr[v] / bn[v]);
            rsn[v] = r;
            if (std::fabs(r) < tol) conv[v] = 1;
            all = all && conv[v];
        }
        if (all) break;
    }
    for (int v = 0; v < nvec; v++) {
        ret[v] = conv[v] ? std::max(total[v], 1) : -std::max(total[v], 1);
        if (res) res[v] = rsn[v];
    }
    return SCALIT_OK;
}

// ---------------------------------------------------------------- PIST
struct PistParams {
    double etol;
    int ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, prec;
    double qmrtol;
};

// LinSolv for the active windows: W_k = (H - E0_k)^-1 V_k(:, j) ; returns per-window codes

static int pist_linsolv(scalit_ctx *c, PistState &P, const PistParams &pp, int nw, const double *e0, int jcol,
                        const std::vector<int> &act, std::vector<int> &num, std::vector<int> &stats,
                        hipStream_t s)
{
    const long N = c->N;
    std::vector<int> idx;
    for (int k = 0; k < nw; k++)
        if (act[k]) idx.push_back(k);
    const int na = (int)idx.size();
    if (na == 0) return SCALIT_OK;
    std::vector<double> sh(na);
    for (int i = 0; i < na; i++) sh[i] = e0[idx[i]];
    std::vector<int> r(na);
    if (pp.prec == SCALIT_FP64) {
        SC_TRY(P.q64.ensure(na, N));
        SC_TRY(copy_cols(N, na, P.q64.B(), P.q64.ld, nullptr, P.vj.p + (long)jcol * P.ldv, P.ldv * pp.m_max,
                         idx.data(), s));
        SC_TRY(qmr_solve<double>(c, P.q64, pp.qmr_type, pp.qmrmax, pp.qmrtol, sh.data(), r.data(), nullptr, s));
        SC_TRY(copy_cols(N, na, P.w.p, P.ldv, idx.data(), P.q64.X(), P.q64.ld, nullptr, s));
    } else {
        SC_TRY(P.x64.alloc((size_t)na * N));
        SC_TRY(P.tmp.alloc((size_t)na * N));
        // gather the right-hand sides
        DBuf<double> &bb = P.bb;
        SC_TRY(bb.alloc((size_t)na * N));
        SC_TRY(copy_cols(N, na, bb.p, N, nullptr, P.vj.p + (long)jcol * P.ldv, P.ldv * pp.m_max, idx.data(), s));
        std::vector<double> rs(na);
        // inner tolerance: the requested one, floored at 1e-3 (2e-3 for BF16
        // storage).  Each FP64 refinement step then gains about three digits,
        // and a shift near an eigenvalue (an ill-conditioned H - E0, as PIST
        // windows place them) cannot stall the FP32 recurrence above a floor
        // that its rounding never reaches.
        double itol = std::max(pp.qmrtol, pp.prec == SCALIT_BF16 ? 2e-3 : 1e-3);
        SC_TRY(qmr_mixed(c, pp.prec, pp.qmr_type, 8, pp.qmrtol, pp.qmrmax, itol, na, bb.p, N, P.x64.p, N, sh.data(),
                         r.data(), rs.data(), s));
        SC_TRY(copy_cols(N, na, P.w.p, P.ldv, idx.data(), P.x64.p, N, nullptr, s));
    }
    for (int i = 0; i < na; i++) {
        num[idx[i]] = r[i];
        if (r[i] > 0) stats[0] += r[i];
        stats[1] += 1;
    }
    return SCALIT_OK;
}

// MGS_ORTH for the active windows: VJ_k(:, jnew) from W_k against columns 0..jnew-1.
// ok[k] = 0 on a zero norm (before or after the projections).  With
// c->pist_orth == 2 the new column is orthogonalised against the previous ones
// a second time and normalised again: one MGS pass loses orthogonality once
// the Krylov vectors of a long solve become nearly dependent (the spectral
// transform has converged towards the lowest levels), and Rayleigh-Ritz on a
// basis that is no longer orthonormal gives spurious values; "twice is
// enough" (Parlett) restores it.  One pass is the reference's MGS_ORTH.
static int pist_mgs(scalit_ctx *c, PistState &P, int nw, int m_max, int jnew, const int *act_dev, std::vector<int> &ok,
                    hipStream_t s)
{
    const long N = c->N;
    const long ldw = P.ldv * m_max;
    double *col = P.vj.p + (long)jnew * P.ldv;   // window 0's column jnew; window k at + k*ldw
    const bool twice = c->pist_orth >= 2;
    if (launch_mgs_one(nw, N, jnew, P.vj.p, P.ldv, ldw, P.w.p, P.ldv, col, act_dev, P.ok.p, nullptr, s)) {
        // (the kernel holds its input in registers before writing: in place is fine)
        if (twice) launch_mgs_one(nw, N, jnew, P.vj.p, P.ldv, ldw, col, ldw, col, act_dev, P.ok.p + nw, nullptr, s);
        std::vector<int> both(2 * (size_t)nw, 1);
        HIP_TRY(hipMemcpyAsync(both.data(), P.ok.p, sizeof(int) * (twice ? 2 : 1) * nw, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int k = 0; k < nw; k++) ok[k] = both[k] && (!twice || both[nw + k]);
        return SCALIT_OK;
    }
    // VJ(:, M0) = WJ ; normw = ||WJ||
    SC_TRY(copy_cols(N, nw, col, ldw, nullptr, P.w.p, P.ldv, nullptr, s));
    launch_dot_partials(nw, N, P.w.p, P.ldv, P.w.p, P.ldv, P.part.p, s);
    launch_reduce_partials(nw, P.part.p, P.tmp.p, s);
    std::vector<double> nrm(nw);
    HIP_TRY(hipMemcpyAsync(nrm.data(), P.tmp.p, 8 * nw, hipMemcpyDeviceToHost, s));
    // projections: launch i computes d_i from the previous partials and the next partial
    launch_mgs_step(nw, N, -1, 0, P.vj.p, P.ldv, ldw, col, ldw, nullptr, P.part.p, act_dev, s);
    double *pa = P.part.p, *pb = P.part2.p;
    for (int i = 0; i < jnew; i++) {
        int inext = (i + 1 < jnew) ? i + 1 : -1;
        launch_mgs_step(nw, N, i, inext, P.vj.p, P.ldv, ldw, col, ldw, pa, pb, act_dev, s);
        std::swap(pa, pb);
    }
    launch_normalize(nw, N, col, ldw, col, ldw, pa, P.ok.p, act_dev, s);
    if (twice) {
        // the second pass on the normalised column, in place
        launch_mgs_step(nw, N, -1, 0, P.vj.p, P.ldv, ldw, col, ldw, nullptr, P.part.p, act_dev, s);
        pa = P.part.p;
        pb = P.part2.p;
        for (int i = 0; i < jnew; i++) {
            int inext = (i + 1 < jnew) ? i + 1 : -1;
            launch_mgs_step(nw, N, i, inext, P.vj.p, P.ldv, ldw, col, ldw, pa, pb, act_dev, s);
            std::swap(pa, pb);
        }
        launch_normalize(nw, N, col, ldw, col, ldw, pa, P.ok.p + nw, act_dev, s);
    }
    std::vector<int> both(2 * (size_t)nw, 1);
    HIP_TRY(hipMemcpyAsync(both.data(), P.ok.p, sizeof(int) * (twice ? 2 : 1) * nw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < nw; k++) ok[k] = both[k] && (!twice || both[nw + k]) && nrm[k] != 0.0;
    return SCALIT_OK;
}

// The Krylov steps j = j_from..j_to of pist_h0 in one persistent launch
// (kernels_qmr.hip), with the same per-window outcome and counts as the loop
// of pist_linsolv / pist_mgs calls.  QP_NOT_LAUNCHED when the windows do not
// fit the device at once.
static int pist_steps_persistent(scalit_ctx *c, PistState &P, const PistParams &pp, int nw, const double *e0,
                                 int j_from, int j_to, std::vector<int> &act, std::vector<int> &ret,
                                 std::vector<int> &stats, hipStream_t s)
{
    const long N = c->N;
    const int ldg = pp.m_max;
    QmrWS<double> &q = P.q64;
    SC_TRY(q.ensure(nw, N));
    q.hshift.assign(e0, e0 + nw);
    HIP_TRY(hipMemcpyAsync(q.shifts.p, q.hshift.data(), 8 * (size_t)nw, hipMemcpyHostToDevice, s));
    SC_TRY(P.gram.alloc((size_t)nw * ldg * ldg));
    SC_TRY(P.qret.alloc((size_t)nw * ldg));
    SC_TRY(P.wret.alloc(nw));
    HIP_TRY(hipMemsetAsync(P.qret.p, 0x80, sizeof(int) * (size_t)nw * ldg, s));   // "not run"
    PistPersistCall pc;
    pc.vj = P.vj.p;
    pc.ldv = P.ldv;
    pc.ldw = P.ldv * pp.m_max;
    pc.gram = P.gram.p;
    pc.ldg = ldg;
    pc.j_from = j_from;
    pc.j_to = j_to;
    pc.orth = c->pist_orth >= 2 ? 2 : 1;
    pc.zero_x = pp.qmr_type <= 0 ? 1 : 0;
    pc.conv = std::abs(pp.qmr_type) == 1 ? 1 : (std::abs(pp.qmr_type) == 2 ? 2 : 0);
    pc.qmrmax = pp.qmrmax;
    pc.tol = pp.qmrtol;
    pc.qret = P.qret.p;
    pc.wret = P.wret.p;
    const int rc = pist_persist_steps(c, nw, pc, q.view(), q.scal.p, P.active.p, q.shifts.p, q.persist, s);
    if (rc != SCALIT_OK) return rc;
    P.hqret.resize((size_t)nw * ldg);
    P.hwret.resize(nw);
    HIP_TRY(hipMemcpyAsync(P.hqret.data(), P.qret.p, sizeof(int) * (size_t)nw * ldg, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(P.hwret.data(), P.wret.p, sizeof(int) * nw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int not_run = (int)0x80808080u;
    for (int k = 0; k < nw; k++) {
        if (!act[k]) continue;
        for (int j = j_from; j <= j_to; j++) {
            const int r = P.hqret[(size_t)k * ldg + j];
            if (r == not_run) break;
            stats[1] += 1;
            if (r > 0) stats[0] += r;
        }
        if (P.hwret[k] != 1) {
            ret[k] = P.hwret[k];
            act[k] = 0;
        }
    }
    HIP_TRY(hipMemcpyAsync(P.active.p, act.data(), sizeof(int) * nw, hipMemcpyHostToDevice, s));
    return SCALIT_OK;
}

// pistH0 for nw windows; ret[k] gets the reference's return value.
static int pist_h0(scalit_ctx *c, PistState &P, const PistParams &pp, int nw, const double *e0, int m_low, int m_high,
                   std::vector<int> &act, std::vector<int> &ret, std::vector<int> &stats, hipStream_t s)
{
    const long N = c->N;
    const int m_max = pp.m_max;
    const long ldw = P.ldv * m_max;
    for (int k = 0; k < nw; k++)
        if (act[k]) ret[k] = m_high;
    int m_start = (m_low == 1) ? 1 : m_low - 1;
    std::vector<int> num(nw, 0), ok(nw, 1);
    int rc = QP_NOT_LAUNCHED;
    if (m_start <= m_high - 1 && P.gram_valid && pp.prec == SCALIT_FP64 && qmr_persist_wanted(c) &&
        m_high <= pist_persist_max_cols(c, nw, m_max))
        rc = pist_steps_persistent(c, P, pp, nw, e0, m_start, m_high - 1, act, ret, stats, s);
    if (rc != QP_NOT_LAUNCHED) SC_TRY(rc);
    else if (m_start <= m_high - 1) P.gram_valid = false;   // later steps cannot take the persistent path in this run
    for (int j = m_start; rc == QP_NOT_LAUNCHED && j <= m_high - 1; j++) {
        SC_TRY(pist_linsolv(c, P, pp, nw, e0, j - 1, act, num, stats, s));
        for (int k = 0; k < nw; k++)
            if (act[k] && num[k] <= 0) {
                ret[k] = num[k];
                act[k] = 0;
            }
        HIP_TRY(hipMemcpyAsync(P.active.p, act.data(), sizeof(int) * nw, hipMemcpyHostToDevice, s));
        SC_TRY(pist_mgs(c, P, nw, m_max, j, P.active.p, ok, s));
        for (int k = 0; k < nw; k++)
            if (act[k] && !ok[k]) {
                ret[k] = -j;
                act[k] = 0;
            }
        HIP_TRY(hipMemcpyAsync(P.active.p, act.data(), sizeof(int) * nw, hipMemcpyHostToDevice, s));
    }
    // Rayleigh-Ritz columns j = m_low..m_high: H0X = H (no shift)
    SC_TRY(P.part.alloc((size_t)nw * m_high * NPART));
    std::vector<double> zero(nw, 0.0);
    DBuf<double> &z = P.zsh;   // kept: a device free would synchronise the whole device
    SC_TRY(z.alloc(nw));
    HIP_TRY(hipMemcpyAsync(z.p, zero.data(), 8 * nw, hipMemcpyHostToDevice, s));
    for (int j = m_low; j <= m_high; j++) {
        SC_TRY(hx_apply_f64(c, z.p, 0.0, nw, P.vj.p + (long)(j - 1) * P.ldv, ldw, P.w.p, P.ldv, false, P.active.p, s));
        launch_multidot(nw, N, 0, m_high, P.vj.p, P.ldv, ldw, P.w.p, P.ldv, P.part.p, P.active.p, s);
        launch_multidot_fin(nw, m_high, 0, j - 1, P.part.p, P.hmat.p, m_max, (long)m_max * m_max, P.active.p, s);
        stats[3] += 1;
    }
    HIP_TRY(hipStreamSynchronize(s));
    SC_TRY(P.part.alloc((size_t)2 * nw * NPART));
    return SCALIT_OK;
}

int pist_conv(scalit_ctx *c, int nw, const double *e0, const PistParams &pp, const double *x, long ldx, double *eig,
              int *ret_out, double *res_out, int *stats_out, hipStream_t s)
{
    const long N = c->N;
    const int m = pp.m, m_max = pp.m_max;
    for (int k = 0; k < nw; k++) {
        ret_out[k] = 0;
        if (res_out) res_out[k] = 0;
    }
    if (pp.stepeig >= m_max || pp.startcnt > m_max || m > m_max) return SCALIT_OK;   // PIST_CONV = 0
    if (!c->has_precond) return SCALIT_ESTATE;
    PistState &P = pstate(c);
    P.ldv = (N + 63) & ~63L;
    SC_TRY(P.vj.alloc((size_t)nw * m_max * P.ldv));
    SC_TRY(P.hmat.alloc((size_t)nw * m_max * m_max));
    SC_TRY(P.w.alloc((size_t)nw * P.ldv));
    SC_TRY(P.part.alloc((size_t)2 * nw * NPART));
    SC_TRY(P.part2.alloc((size_t)2 * nw * NPART));
    SC_TRY(P.tmp.alloc((size_t)std::max<long>(nw, 64)));
    SC_TRY(P.ok.alloc(2 * (size_t)nw));   // the zero-norm flags of both orthogonalisation passes
    SC_TRY(P.active.alloc(nw));
    SC_TRY(P.eigw.alloc((size_t)2 * nw * m_max));
    SC_TRY(P.zw.alloc(small_sym_eig_work(m_max)));
    if (pp.prec != SCALIT_FP64) SC_TRY(ensure_f32(c));
    HIP_TRY(hipMemsetAsync(P.hmat.p, 0, 8 * (size_t)nw * m_max * m_max, s));
    const long ldw = P.ldv * m_max;
    std::vector<int> act(nw, 1), stats(4, 0), ret(nw, 0), done(nw, 0);
    std::vector<double> tmpeig(m_max), oldeig(m), eigk(m), res(nw, 0);
    std::vector<int> ind(m);
    P.last_ind.assign(nw, std::vector<int>(m, 0));
    P.last_cnt.assign(nw, 0);
    P.last_ret.assign(nw, 0);
    P.nwin = nw;
    P.m_max = m_max;
    P.m_last = m;
    // VJ(:,1) = X / ||X||  (pistHij.f90:345-352)
    {
        // read straight from the caller's start vectors (no scratch copy: a
        // device free would synchronise the whole device)
        launch_dot_partials(nw, N, x, ldx, x, ldx, P.part.p, s);
        launch_normalize(nw, N, x, ldx, P.vj.p, ldw, P.part.p, P.ok.p, nullptr, s);
        std::vector<int> ok(nw);
        HIP_TRY(hipMemcpyAsync(ok.data(), P.ok.p, sizeof(int) * nw, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int k = 0; k < nw; k++)
            if (!ok[k]) {
                act[k] = 0;
                done[k] = 1;
                ret_out[k] = -std::max({pp.startcnt, m, pp.stepeig + 1});
            }
    }
    HIP_TRY(hipMemcpyAsync(P.active.p, act.data(), sizeof(int) * nw, hipMemcpyHostToDevice, s));
    P.gram_valid = true;   // no Gram rows are needed before the first step
    int cnt_low = 1;
    int cnt_high = std::max({pp.startcnt, m, pp.stepeig + 1});
    while (cnt_high <= m_max) {
        std::vector<int> hret(nw, 0);
        SC_TRY(pist_h0(c, P, pp, nw, e0, cnt_low, cnt_high, act, hret, stats, s));
        bool any = false;
        // DSYEV on the leading (cnt_high - stepeig) and cnt_high blocks of
        // every window still running, all in one launch
        std::vector<int> wins;
        std::vector<EigvalJob> jobs;
        for (int k = 0; k < nw; k++) {
            if (done[k]) continue;
            if (hret[k] < 0 || !act[k]) {
                ret_out[k] = -cnt_high;
                done[k] = 1;
                act[k] = 0;
                continue;
            }
            const double *hk = P.hmat.p + (long)k * m_max * m_max;
            double *wk = P.eigw.p + (long)2 * wins.size() * m_max;
            jobs.push_back({hk, wk, cnt_high - pp.stepeig, m_max});
            jobs.push_back({hk, wk + m_max, cnt_high, m_max});
            wins.push_back(k);
        }
        std::vector<double> weig((size_t)2 * wins.size() * m_max);
        if (!wins.empty()) {
            SC_TRY(small_sym_eigvals_batch((int)jobs.size(), jobs.data(), P.ejobs, P.zw, s));
            HIP_TRY(hipMemcpyAsync(weig.data(), P.eigw.p, 8 * weig.size(), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        for (size_t wi = 0; wi < wins.size(); wi++) {
            const int k = wins[wi];
            int num = cnt_high - pp.stepeig;
            get_window0(e0[k], num, weig.data() + 2 * wi * m_max, m, ind.data(), oldeig.data());
            num = cnt_high;
            std::memcpy(tmpeig.data(), weig.data() + (2 * wi + 1) * m_max, 8 * (size_t)num);
            get_window0(e0[k], num, tmpeig.data(), m, ind.data(), eigk.data());
            double r = pp.ntype <= 0 ? get_diff_max(m, eigk.data(), oldeig.data())
                                     : get_sep_max(m, oldeig.data(), cnt_high, tmpeig.data());
            res[k] = r;
            std::memcpy(eig + (long)k * m, eigk.data(), 8 * m);
            P.last_ind[k].assign(ind.begin(), ind.end());
            P.last_cnt[k] = cnt_high;
            if (r <= pp.etol) {
                ret_out[k] = cnt_high;
                done[k] = 1;
                act[k] = 0;
            } else {
                ret_out[k] = -cnt_high;
                any = true;
            }
        }
        HIP_TRY(hipMemcpyAsync(P.active.p, act.data(), sizeof(int) * nw, hipMemcpyHostToDevice, s));
        if (!any) break;
        cnt_low = cnt_high + 1;
        if (cnt_low >= m_max) {
            for (int k = 0; k < nw; k++)
                if (!done[k]) ret_out[k] = -m_max;
            break;
        }
        cnt_high = std::min(cnt_high + pp.stepcnt, m_max);
    }
    for (int k = 0; k < nw; k++) {
        // Reorder('A') (pist.f90:96), keeping the window indices aligned
        std::vector<std::pair<double, int>> pr(m);
        for (int i = 0; i < m; i++) pr[i] = {eig[(long)k * m + i], P.last_ind[k][i]};
        std::stable_sort(pr.begin(), pr.end(),
                         [](const std::pair<double, int> &a, const std::pair<double, int> &b) { return a.first < b.first; });
        for (int i = 0; i < m; i++) {
            eig[(long)k * m + i] = pr[i].first;
            P.last_ind[k][i] = pr[i].second;
        }
        if (res_out) res_out[k] = res[k];
        P.last_ret[k] = ret_out[k];
    }
    if (stats_out) {
        stats_out[0] = stats[0];
        stats_out[1] = stats[1];
        stats_out[2] = cnt_high > m_max ? m_max : cnt_high;
        stats_out[3] = stats[3];
    }
    return SCALIT_OK;
}

// Ritz vectors of window k for its m returned eigenvalues (unit norm)
int pist_vectors(scalit_ctx *c, int k, double *psi, long ldp, hipStream_t s)
{
    if (!c->pist || k < 0 || k >= c->pist->nwin) return SCALIT_EINVAL;
    PistState &P = *c->pist;
    const int cnt = P.last_cnt[k], m = P.m_last, m_max = P.m_max;
    if (cnt <= 0 || P.last_ret[k] <= 0) return SCALIT_ESTATE;   // unconverged: no Ritz vectors
    SC_TRY(P.psi_z.alloc((size_t)cnt * cnt));
    SC_TRY(P.eigw.alloc((size_t)m_max));
    SC_TRY(P.zw.alloc(small_sym_eig_work(m_max)));
    SC_TRY(P.cols.alloc(m));
    const double *hk = P.hmat.p + (long)k * m_max * m_max;
    SC_TRY(small_sym_eig(cnt, hk, m_max, P.eigw.p, P.psi_z.p, cnt, P.zw.p, s));
    HIP_TRY(hipMemcpyAsync(P.cols.p, P.last_ind[k].data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
    launch_ritz(c->N, cnt, m, P.vj.p + (long)k * P.ldv * m_max, P.ldv, P.psi_z.p, cnt, P.cols.p, psi, ldp, s);
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

// Ritz vectors of every window for its m returned eigenvalues, window k's at
// psi + (k m + i) ldp: one eigenvector launch and one Ritz launch for all
int pist_vectors_all(scalit_ctx *c, double *psi, long ldp, int ncols, hipStream_t s)
{
    if (!c->pist || c->pist->nwin < 1) return SCALIT_EINVAL;
    PistState &P = *c->pist;
    const int nw = P.nwin, m = P.m_last, m_max = P.m_max;
    if ((long)nw * m > ncols) return SCALIT_EINVAL;   // psi must hold every window's m vectors
    int cmax = 0;
    for (int k = 0; k < nw; k++) {
        // a window that stopped unconverged has no trustworthy Ritz vectors
        if (P.last_cnt[k] <= 0 || P.last_ret[k] <= 0) return SCALIT_ESTATE;
        cmax = std::max(cmax, P.last_cnt[k]);
    }
    SC_TRY(P.psi_z.alloc((size_t)nw * cmax * cmax));
    SC_TRY(P.eigw.alloc((size_t)nw * cmax));
    std::vector<EigvecJob> ej(nw);
    std::vector<RitzJob> rj((size_t)nw * m);
    for (int k = 0; k < nw; k++) {
        const int cnt = P.last_cnt[k];
        double *z = P.psi_z.p + (long)k * cmax * cmax;
        ej[k] = {P.hmat.p + (long)k * m_max * m_max, P.eigw.p + (long)k * cmax, z, cnt, m_max, cnt};
        for (int i = 0; i < m; i++)
            rj[(size_t)k * m + i] = {P.vj.p + (long)k * P.ldv * m_max, z, psi + ((long)k * m + i) * ldp, cnt, cnt,
                                     P.last_ind[k][i]};
    }
    SC_TRY(small_sym_eig_batch(nw, ej.data(), P.vjobs, P.zw, s));
    SC_TRY(P.rjobs.alloc(rj.size()));
    HIP_TRY(hipMemcpyAsync(P.rjobs.p, rj.data(), sizeof(RitzJob) * rj.size(), hipMemcpyHostToDevice, s));
    launch_ritz_batch(c->N, P.ldv, (int)rj.size(), P.rjobs.p, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));   // the host job arrays die here
    return SCALIT_OK;
}

// ---------------------------------------------------------------- Lanczos
__global__ void k_lan_sub_beta_alpha_part(long n, int j, const double *beta, const double *vprev, const double *vj,
                                          double *w, double *part)
{
    __shared__ double red[4];
    double b = (j > 0) ? beta[j - 1] : 0.0;
    double s = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        double a = w[i];
        if (j > 0) a = a - b * vprev[i];
        w[i] = a;
        s += a * vj[i];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_lan_sub_alpha(long n, int j, const double *part, double *alpha, const double *vj, double *w,
                                double *dst, double *part_out)
{
    __shared__ double red[4];
    double a = 0;
    for (int i = 0; i < NPART; i++) a += part[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) alpha[j] = a;
    double s = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        double v = w[i] - a * vj[i];
        w[i] = v;
        dst[i] = v;
        s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) part_out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_store_norm(const double *part, double *out, int idx)
{
    double a = 0;
    for (int i = 0; i < NPART; i++) a += part[i];
    out[idx] = sqrt(a);
}

int lan_conv(scalit_ctx *c, double e0, double etol, int ntype, const double *x, int startcnt, int stepcnt,
             int stepeig, int m, int m_max, double *eig, double *res, hipStream_t s)
{
    const long N = c->N;
    if (stepeig >= m_max || startcnt > m_max || m > m_max) return 0;
    PistState &P = pstate(c);
    const long ldv = (N + 63) & ~63L;
    DBuf<double> &vj = P.lvj, &alpha = P.lalpha, &beta = P.lbeta, &w = P.lw, &part = P.lpart, &part2 = P.lpart2,
                 &nrm = P.lnrm, &ev = P.lev, &zero = P.lzero;
    SC_TRY(vj.alloc((size_t)(m_max + 1) * ldv));
    SC_TRY(alpha.alloc(m_max));
    SC_TRY(beta.alloc(m_max));
    SC_TRY(w.alloc(ldv));
    SC_TRY(part.alloc(NPART));
    SC_TRY(part2.alloc(NPART));
    SC_TRY(nrm.alloc(4));
    SC_TRY(ev.alloc(2 * (size_t)m_max));
    DBuf<TriEigJob> &ejobs = P.lejobs;
    DBuf<int> &okv = P.lokv;   // per Lanczos step: the re-orthogonalised vector has a nonzero norm
    SC_TRY(okv.alloc((size_t)m_max + 1));
    HIP_TRY(hipMemsetAsync(alpha.p, 0, 8 * m_max, s));
    HIP_TRY(hipMemsetAsync(beta.p, 0, 8 * m_max, s));
    int ret = 0;
    std::vector<double> w_h(2 * (size_t)m_max), oldeig(m), eigk(m);
    std::vector<int> ind(m);
    // VJ(:,1) = X/||X||
    launch_dot_partials(1, N, x, ldv, x, ldv, part.p, s);
    launch_normalize(1, N, x, ldv, vj.p, ldv, part.p, nullptr, nullptr, s);
    double xn = 0;
    {
        launch_reduce_partials(1, part.p, nrm.p, s);
        HIP_TRY(hipMemcpyAsync(&xn, nrm.p, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    int cnt_low = 1, cnt_high = std::max({startcnt, m, stepeig + 1});
    SC_TRY(zero.alloc(1));
    HIP_TRY(hipMemsetAsync(zero.p, 0, 8, s));
    // the buffers stay with the context for the next call
    auto fail = [](int r) { return r; };
    while (cnt_high <= m_max) {
        // lanH0 (lanHij.f90:6-45): no steps unless 1 <= CNT_LOW < CNT_HIGH, and
        // then Lan_Conv returns -(CNT_LOW-1) (lan.f90:47-50); a zero start
        // vector makes the first block fail the same way
        if (cnt_low < 1 || cnt_low >= cnt_high) return fail(-(cnt_low - 1));
        if (cnt_low == 1 && xn == 0.0) return fail(0);
        bool bad = false, deferred = false;
        for (int j = cnt_low; j <= cnt_high && !bad; j++) {
            double *vjj = vj.p + (long)(j - 1) * ldv;
            double *vnext = vj.p + (long)j * ldv;
            SC_TRY(hx_apply_f64(c, zero.p, 0.0, 1, vjj, ldv, w.p, ldv, false, nullptr, s));
            hipLaunchKernelGGL(k_lan_sub_beta_alpha_part, dim3(NPART), dim3(256), 0, s, N, j - 1, beta.p,
                               j > 1 ? vj.p + (long)(j - 2) * ldv : vjj, vjj, w.p, part.p);
            hipLaunchKernelGGL(k_lan_sub_alpha, dim3(NPART), dim3(256), 0, s, N, j - 1, part.p, alpha.p, vjj, w.p, vnext,
                               part2.p);
            // LAN_MGS_ORTH in one workgroup when the vector fits it: BETA(J) and
            // the flag of step j land on the device and are read once after the
            // steps (a failed step makes this block fail with the same code, so
            // the steps after it only compute values that are never used)
            if (launch_mgs_one(1, N, j, vj.p, ldv, 0, vnext, 0, vnext, nullptr, okv.p + j, beta.p + (j - 1), s)) {
                deferred = true;
                continue;
            }
            hipLaunchKernelGGL(k_store_norm, dim3(1), dim3(1), 0, s, part2.p, nrm.p, 0);
            // LAN_MGS_ORTH: against columns 1..j
            launch_mgs_step(1, N, -1, 0, vj.p, ldv, 0, vnext, 0, nullptr, part.p, nullptr, s);
            double *pa = part.p, *pb = part2.p;
            for (int i = 0; i < j; i++) {
                int inext = (i + 1 < j) ? i + 1 : -1;
                launch_mgs_step(1, N, i, inext, vj.p, ldv, 0, vnext, 0, pa, pb, nullptr, s);
                std::swap(pa, pb);
            }
            hipLaunchKernelGGL(k_store_norm, dim3(1), dim3(1), 0, s, pa, beta.p, j - 1);
            launch_normalize(1, N, vnext, 0, vnext, 0, pa, nullptr, nullptr, s);
            double nn[1];
            HIP_TRY(hipMemcpyAsync(nn, nrm.p, 8, hipMemcpyDeviceToHost, s));
            double bj = 0;
            HIP_TRY(hipMemcpyAsync(&bj, beta.p + (j - 1), 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (nn[0] == 0.0 || bj == 0.0) bad = true;
        }
        if (deferred) {
            std::vector<int> okh(cnt_high - cnt_low + 1);
            HIP_TRY(hipMemcpyAsync(okh.data(), okv.p + cnt_low, sizeof(int) * okh.size(), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (int v : okh)
                if (!v) bad = true;
        }
        if (bad) return fail(-(cnt_low - 1));
        // DSTEV of the leading (cnt_high - stepeig) and cnt_high blocks of the
        // tridiagonal (ALPHA, BETA), lan.f90:57-69, in one launch
        const TriEigJob jobs[2] = {{alpha.p, beta.p, ev.p, cnt_high - stepeig},
                                   {alpha.p, beta.p, ev.p + m_max, cnt_high}};
        SC_TRY(tridiag_eigvals_batch(2, jobs, ejobs, s));
        HIP_TRY(hipMemcpyAsync(w_h.data(), ev.p, 8 * 2 * (size_t)m_max, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        get_window0(e0, cnt_high - stepeig, w_h.data(), m, ind.data(), oldeig.data());
        const int num = cnt_high;
        std::memmove(w_h.data(), w_h.data() + m_max, 8 * (size_t)num);
        get_window0(e0, num, w_h.data(), m, ind.data(), eigk.data());
        double r = ntype <= 0 ? get_diff_max(m, eigk.data(), oldeig.data())
                              : get_sep_max(m, oldeig.data(), cnt_high, w_h.data());
        *res = r;
        std::memcpy(eig, eigk.data(), 8 * m);
        if (r < etol) {
            ret = cnt_high;
            break;
        }
        ret = -cnt_high;
        cnt_low = cnt_high + 1;
        if (cnt_low >= m_max) {
            ret = -m_max;
            break;
        }
        cnt_high = std::min(cnt_high + stepcnt, m_max);
    }
    std::sort(eig, eig + m);
    return fail(ret);
}

// ---------------------------------------------------------------- gradients
int eig_gradients(scalit_ctx *c, int nk, const double *psi, long ld, int np, const double *dvdp, long lddv,
                  double *out, hipStream_t s)
{
    // kept in the context: a device free would synchronise the whole device
    PistState &P = pstate(c);
    SC_TRY(P.gpart.alloc((size_t)nk * np * NPART));
    SC_TRY(P.gred.alloc((size_t)nk * np));
    launch_hf_part(nk, c->N, np, psi, ld, dvdp, lddv, P.gpart.p, s);
    launch_reduce_partials(nk * np, P.gpart.p, P.gred.p, s);
    HIP_TRY(hipMemcpyAsync(out, P.gred.p, 8 * (size_t)nk * np, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

}  // namespace scalit
