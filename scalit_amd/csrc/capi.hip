// C ABI of libscalit_hip.so (declarations and reference citations in
// include/scalit.h).  No exception crosses the ABI; every entry point returns
// a status code or the reference function's own return value.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

#include "scalit_impl.hpp"
#include "solver_state.hpp"

namespace scalit {
int qmr_multi_f64(int njob, scalit_qmr_job *jobs, int ntype, int qmrmax, double tol);
int qmr_f64(scalit_ctx *c, int ntype, int qmrmax, double tol, int nvec, const double *b, double *x, long ld,
            const double *shifts, int *ret, double *res, hipStream_t s);
int qmr_mixed(scalit_ctx *c, int prec, int ntype, int qmrmax, double tol, int inner_max, double inner_tol, int nvec,
              const double *b, long ldb, double *x, long ldx, const double *shifts, int *ret, double *res,
              hipStream_t s);
struct PistParams {
    double etol;
    int ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, prec;
    double qmrtol;
};
int pist_conv(scalit_ctx *c, int nw, const double *e0, const PistParams &pp, const double *x, long ldx, double *eig,
              int *ret_out, double *res_out, int *stats_out, hipStream_t s);
int pist_vectors(scalit_ctx *c, int k, double *psi, long ldp, hipStream_t s);
int pist_vectors_all(scalit_ctx *c, double *psi, long ldp, int ncols, hipStream_t s);
int lan_conv(scalit_ctx *c, double e0, double etol, int ntype, const double *x, int startcnt, int stepcnt,
             int stepeig, int m, int m_max, double *eig, double *res, hipStream_t s);
int eig_gradients(scalit_ctx *c, int nk, const double *psi, long ld, int np, const double *dvdp, long lddv,
                  double *out, hipStream_t s);
void release_solvers(scalit_ctx *c);
int qmr_host_cb(scalit_ctx *c, int ntype, int qmrmax, double tol, long n, const double *b, double *x,
                scalit_fop hx, scalit_fop px, int *ret, double *res);
int mgs_host(scalit_ctx *c, long n, const double *wj, int m, double *vj, int *ok);
int hx_pass_f64(scalit_ctx *c, int pass, int nvec, const double *x, long ldx, double *y, long ldy, hipStream_t s);
int mfma_probe(int prec, double *out_host);
int mfma_peak(int prec, int iters, int blocks, float *ms_out);

__global__ void k_d2f(long n, const double *a, float *b)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

static int to_f32(DBuf<double> &src, DBuf<float> &dst, long n, hipStream_t s)
{
    if (!src.p || n == 0) return SCALIT_OK;
    SC_TRY(dst.alloc(n));
    hipLaunchKernelGGL(k_d2f, dim3(1024), dim3(256), 0, s, n, src.p, dst.p);
    HIP_TRY(hipGetLastError());
    return SCALIT_OK;
}

// fp32 copies of the operator (H0, DEP, RES) for the mixed-precision paths
int ensure_f32(scalit_ctx *c)
{
    bool changed = false;
    for (int l = 0; l < c->sF; l++) {
        Level &L = c->lv[l];
        if (!L.h0_f.p && L.has_h0) {
            SC_TRY(to_f32(L.h0, L.h0_f, (long)L.n * L.n, c->stream));
            changed = true;
        }
        if (L.dep && !L.depv_f.p) {
            SC_TRY(to_f32(L.depv, L.depv_f, L.dim, c->stream));
            changed = true;
        }
    }
    if (!c->res_f.p) {
        SC_TRY(to_f32(c->res, c->res_f, c->N, c->stream));
        changed = true;
    }
    if (c->ndvr && !c->outh_f.p) {
        SC_TRY(to_f32(c->outh, c->outh_f, c->N * (long)c->lv[c->sF - 1].n, c->stream));
        changed = true;
    }
    if (c->has_precond) SC_TRY(precond_to_f32(c));
    if (changed) c->gen++;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SCALIT_OK;
}

int precond_to_f32(scalit_ctx *c)
{
    if (c->eig0_f.p) return SCALIT_OK;
    for (int l = 0; l < c->sF; l++)
        SC_TRY(to_f32(c->lv[l].vosb, c->lv[l].vosb_f, (long)c->lv[l].blk * c->lv[l].n * c->lv[l].n, c->stream));
    SC_TRY(to_f32(c->eig0, c->eig0_f, c->N, c->stream));
    c->gen++;
    return SCALIT_OK;
}

// Launch-shape caches shared by every context and host thread (several
// problems may run on their own threads at once): one mutex guards both.
static std::mutex g_launch_mu;
static std::set<std::pair<int, const void *>> g_lds_raised;
static std::map<std::tuple<int, const void *, int, size_t>, int> g_resident;

static void allow_full_lds_locked(int dev, const void *fn)
{
    if (!g_lds_raised.insert({dev, fn}).second) return;
    // the CU's 160 KB less the kernel's static LDS; a refusal leaves the
    // default limit (the launch then reports what does not fit) and is
    // cleared here, so that it is not reported by a later, unrelated call
    hipFuncAttributes fa{};
    size_t stat = 0;
    if (hipFuncGetAttributes(&fa, fn) == hipSuccess) stat = fa.sharedSizeBytes;
    else (void)hipGetLastError();
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(160 * 1024 - stat)) != hipSuccess)
        (void)hipGetLastError();
}

void allow_full_lds(const void *fn)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_launch_mu);
    allow_full_lds_locked(dev, fn);
}

int resident_per_cu(const void *fn, int threads, size_t lds)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_launch_mu);
    allow_full_lds_locked(dev, fn);
    const auto key = std::make_tuple(dev, fn, threads, lds);
    auto it = g_resident.find(key);
    if (it != g_resident.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds) != hipSuccess || nb < 1) nb = 1;
    g_resident.emplace(key, nb);
    return nb;
}

static void drop_f32(scalit_ctx *c)
{
    for (int l = 0; l < c->sF; l++) {
        c->lv[l].h0_f.release();
        c->lv[l].depv_f.release();
        c->lv[l].vosb_f.release();
        c->lv[l].vosbt_f.release();
    }
    c->res_f.release();
    c->eig0_f.release();
    c->outh_f.release();
    c->osbw_n = 0;   // a TOSBW window belongs to the operator it was built on
    c->gen++;
}

}  // namespace scalit

using namespace scalit;

// NULL is the legacy default stream (what torch reports for its default stream)
static hipStream_t pick(scalit_ctx *c, scalit_stream s)
{
    (void)c;
    return (hipStream_t)s;
}

#define CTX_GUARD(c)                                    \
    do {                                                \
        if (!(c)) return SCALIT_EINVAL;                 \
        if (hipSetDevice((c)->device) != hipSuccess) return SCALIT_EHIP; \
    } while (0)

extern "C" {

int scalit_create(scalit_ctx **out, int sF, const int *sN, const int *sDEP, int device)
{
    if (!out || sF < 1 || sF > SC_FMAX || !sN || !sDEP) return SCALIT_EINVAL;
    *out = nullptr;
    for (int l = 0; l < sF; l++)
        if (sN[l] < 1) return SCALIT_EINVAL;
    if (sDEP[0]) return SCALIT_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SCALIT_EHIP;
    if (device < 0 || device >= ndev) return SCALIT_EINVAL;
    HIP_TRY(hipSetDevice(device));
    scalit_ctx *c = new scalit_ctx();
    static std::atomic<long> next_uid{1};
    c->uid = next_uid++;
    c->device = device;
    c->sF = sF;
    long dim = 1;
    for (int l = 0; l < sF; l++) {
        c->lv[l].n = sN[l];
        c->lv[l].dep = sDEP[l] != 0;
        c->lv[l].dim = (int)dim;
        dim *= sN[l];
    }
    long blk = 1;
    for (int l = sF - 1; l >= 0; l--) {
        c->lv[l].blk = (int)blk;
        blk *= sN[l];
    }
    c->N = dim;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SCALIT_EHIP;
    }
    c->fast3 = fast3_ok(c);
    // the fixed scratch of the no-allocation entry point: its address on this device
    double *scratch_dev = nullptr;
    if (int r = noalloc_scratch(device, &scratch_dev); r != SCALIT_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return r;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        g_num_cu = ncu;
    *out = c;
    return SCALIT_OK;
}

int scalit_destroy(scalit_ctx *c)
{
    if (!c) return SCALIT_EINVAL;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    release_solvers(c);
    release_complex(c);
    for (int l = 0; l < c->sF; l++) {
        Level &L = c->lv[l];
        L.h0.release();
        L.h0t.release();
        L.depv.release();
        L.vosb.release();
        L.vosbt.release();
        L.h0_f.release();
        L.depv_f.release();
        L.vosb_f.release();
        L.vosbt_f.release();
    }
    c->res.release();
    c->eig0.release();
    c->res_f.release();
    c->outh.release();
    c->outh_f.release();
    c->eig0_f.release();
    c->work.release();
    c->work_f.release();
    c->shifts_d.release();
    c->shifts_f.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SCALIT_OK;
}

int scalit_get_len(const scalit_ctx *c, long *n)
{
    if (!c || !n) return SCALIT_EINVAL;
    *n = c->N;
    return SCALIT_OK;
}

int scalit_set_h0(scalit_ctx *c, int level, const double *h0)
{
    CTX_GUARD(c);
    if (level < 1 || level > c->sF || !h0) return SCALIT_EINVAL;
    Level &L = c->lv[level - 1];
    const int n = L.n;
    SC_TRY(L.h0.alloc((size_t)n * n));
    SC_TRY(L.h0t.alloc((size_t)n * n));
    std::vector<double> t((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) t[i + (size_t)n * j] = h0[j + (size_t)n * i];
    HIP_TRY(hipMemcpy(L.h0.p, h0, sizeof(double) * n * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(L.h0t.p, t.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
    L.has_h0 = true;
    drop_f32(c);
    return SCALIT_OK;
}

int scalit_set_dep(scalit_ctx *c, int level, const double *dep)
{
    CTX_GUARD(c);
    if (level < 1 || level > c->sF || !dep) return SCALIT_EINVAL;
    Level &L = c->lv[level - 1];
    if (!L.dep) return SCALIT_EINVAL;
    SC_TRY(L.depv.alloc(L.dim));
    HIP_TRY(hipMemcpy(L.depv.p, dep, sizeof(double) * L.dim, hipMemcpyHostToDevice));
    L.has_dep = true;
    drop_f32(c);
    return SCALIT_OK;
}

int scalit_set_outh(scalit_ctx *c, const double *outh)
{
    CTX_GUARD(c);
    if (!outh || c->sF < 2) return SCALIT_EINVAL;
    Level &L = c->lv[c->sF - 1];
    const size_t n = (size_t)c->N * L.n;   // outLen = sN(sF) * myLen (osb.init.h:19)
    SC_TRY(c->outh.alloc(n));
    HIP_TRY(hipMemcpy(c->outh.p, outh, sizeof(double) * n, hipMemcpyHostToDevice));
    c->ndvr = true;
    L.dep = false;   // sNDVR forces sDEP(sF) = F (osb.print.h:204)
    c->fast3 = fast3_ok(c);
    drop_f32(c);
    return SCALIT_OK;
}

int scalit_set_res(scalit_ctx *c, const double *res)
{
    CTX_GUARD(c);
    if (!res) return SCALIT_EINVAL;
    SC_TRY(c->res.alloc(c->N));
    HIP_TRY(hipMemcpy(c->res.p, res, sizeof(double) * c->N, hipMemcpyHostToDevice));
    c->has_res = true;
    drop_f32(c);
    return SCALIT_OK;
}

int scalit_set_shift(scalit_ctx *c, double e0, double de, double beta)
{
    if (!c) return SCALIT_EINVAL;
    c->e0 = e0;
    c->de = std::fabs(de);   // sOSBW%mDE = ABS(sOSBW%mDE) (mosb.print.h:254)
    c->beta = beta;
    return SCALIT_OK;
}

int scalit_set_qmr_persistent(scalit_ctx *c, int on)
{
    if (!c || on < -1 || on > 1) return SCALIT_EINVAL;
    c->qmr_persist = on;
    return SCALIT_OK;
}

int scalit_set_osb(scalit_ctx *c, int nosb)
{
    if (!c || nosb < SCALIT_TOSB || nosb > SCALIT_TOSBW) return SCALIT_EINVAL;
    c->osb = nosb;
    return SCALIT_OK;
}

static int operator_ready(const scalit_ctx *c)
{
    if (!c->has_res) return SCALIT_ESTATE;
    for (int l = 0; l < c->sF; l++) {
        if (c->ndvr && l == c->sF - 1) continue;   // the OUTH layer needs no H0
        if (!c->lv[l].has_h0) return SCALIT_ESTATE;
        if (c->lv[l].dep && !c->lv[l].has_dep) return SCALIT_ESTATE;
    }
    return SCALIT_OK;
}

int scalit_block_diag(scalit_ctx *c, int bjmax, double bjtol)
{
    CTX_GUARD(c);
    if (!c->scx) SC_TRY(operator_ready(c));   // the sCX set-up checks its complex data itself
    if (bjmax < 0) return SCALIT_EINVAL;
    int r = block_diag(c, bjmax, bjtol);
    c->gen++;
    return r;
}

int scalit_set_precond(scalit_ctx *c, const double *const *vosb, const double *eig0)
{
    CTX_GUARD(c);
    if (!vosb || !eig0) return SCALIT_EINVAL;
    for (int l = 0; l < c->sF; l++) {
        Level &L = c->lv[l];
        size_t n = (size_t)L.blk * L.n * L.n;
        if (!vosb[l]) return SCALIT_EINVAL;
        SC_TRY(L.vosb.alloc(n));
        HIP_TRY(hipMemcpy(L.vosb.p, vosb[l], sizeof(double) * n, hipMemcpyHostToDevice));
        L.vosb_f.release();
        L.vosbt.release();
        L.vosbt_f.release();
    }
    SC_TRY(c->eig0.alloc(c->N));
    HIP_TRY(hipMemcpy(c->eig0.p, eig0, sizeof(double) * c->N, hipMemcpyHostToDevice));
    c->eig0_f.release();
    c->has_precond = true;
    c->osbw_n = 0;
    c->gen++;
    return SCALIT_OK;
}

int scalit_get_precond(scalit_ctx *c, double *const *vosb, double *eig0)
{
    CTX_GUARD(c);
    if (!c->has_precond) return SCALIT_ESTATE;
    HIP_TRY(hipDeviceSynchronize());
    for (int l = 0; l < c->sF; l++) {
        Level &L = c->lv[l];
        if (vosb && vosb[l])
            HIP_TRY(hipMemcpy(vosb[l], L.vosb.p, sizeof(double) * (size_t)L.blk * L.n * L.n, hipMemcpyDeviceToHost));
    }
    if (eig0) HIP_TRY(hipMemcpy(eig0, c->eig0.p, sizeof(double) * c->N, hipMemcpyDeviceToHost));
    return SCALIT_OK;
}

int scalit_hx(scalit_ctx *c, int ntype, int nvec, const double *x, long ldx, double *y, long ldy,
              const double *shifts, scalit_stream stream)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (nvec < 1 || !x || !y || ldx < c->N || ldy < c->N || x == y) return SCALIT_EINVAL;
    double s0 = (ntype == SCALIT_TA0) ? 0.0 : c->e0;
    return hx_apply_f64(c, shifts, s0, nvec, x, ldx, y, ldy, false, nullptr, pick(c, stream));
}

int scalit_hx_lp(scalit_ctx *c, int prec, int ntype, int nvec, const void *x, long ldx, void *y, long ldy,
                 const float *shifts, scalit_stream stream)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (nvec < 1 || !x || !y || ldx < c->N || ldy < c->N || x == y) return SCALIT_EINVAL;
    SC_TRY(ensure_f32(c));
    float s0 = (ntype == SCALIT_TA0) ? 0.0f : (float)c->e0;
    if (prec == SCALIT_FP32)
        return hx_apply_f32(c, shifts, s0, nvec, (const float *)x, ldx, (float *)y, ldy, nullptr, pick(c, stream));
    if (prec == SCALIT_BF16)
        return hx_apply_bf16(c, shifts, s0, nvec, (const bf16_t *)x, ldx, (bf16_t *)y, ldy, nullptr, pick(c, stream));
    return SCALIT_EINVAL;
}

int scalit_hx_adjoint(scalit_ctx *c, int ntype, int nvec, const double *x, const double *g, long ld, double *dx,
                      double *dv, scalit_stream stream)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (nvec < 1 || !g || !dx || ld < c->N || g == dx) return SCALIT_EINVAL;
    hipStream_t s = pick(c, stream);
    double s0 = (ntype == SCALIT_TA0) ? 0.0 : c->e0;
    SC_TRY(hx_apply_f64(c, nullptr, s0, nvec, g, ld, dx, ld, true, nullptr, s));
    if (dv) {
        if (!x) return SCALIT_EINVAL;
        launch_dv_accum(c->N, nvec, x, g, ld, dv, s);
        HIP_TRY(hipGetLastError());
    }
    return SCALIT_OK;
}

int scalit_px(scalit_ctx *c, int nosb, int npc, int nvec, const double *x, long ldx, double *y, long ldy,
              const double *shifts, scalit_stream stream)
{
    CTX_GUARD(c);
    if (!c->has_precond) return SCALIT_ESTATE;
    if (nvec < 1 || !x || !y || ldx < c->N || ldy < c->N || x == y) return SCALIT_EINVAL;
    if (nosb < SCALIT_TOSB || nosb > SCALIT_TOSBW) return SCALIT_EINVAL;
    double e0 = (npc == SCALIT_TA0) ? 0.0 : c->e0;
    return px_apply_f64(c, nosb, shifts, e0, nvec, x, ldx, y, ldy, nullptr, pick(c, stream));
}

int scalit_px_lp(scalit_ctx *c, int prec, int npc, int nvec, const void *x, long ldx, void *y, long ldy,
                 const float *shifts, scalit_stream stream)
{
    CTX_GUARD(c);
    if (!c->has_precond) return SCALIT_ESTATE;
    if (nvec < 1 || !x || !y || ldx < c->N || ldy < c->N || x == y) return SCALIT_EINVAL;
    SC_TRY(ensure_f32(c));
    float e0 = (npc == SCALIT_TA0) ? 0.0f : (float)c->e0;
    if (prec == SCALIT_FP32)
        return px_apply_f32(c, SCALIT_TOSB, shifts, e0, nvec, (const float *)x, ldx, (float *)y, ldy, nullptr, pick(c, stream));
    if (prec == SCALIT_BF16)
        return px_apply_bf16(c, SCALIT_TOSB, shifts, e0, nvec, (const bf16_t *)x, ldx, (bf16_t *)y, ldy, nullptr, pick(c, stream));
    return SCALIT_EINVAL;
}

int scalit_qmr(scalit_ctx *c, int ntype, int qmrmax, double qmrtol, int nvec, const double *b, double *x, long ld,
               const double *shifts, int *ret, double *res, scalit_stream stream)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (!c->has_precond) return SCALIT_ESTATE;
    if (nvec < 1 || !b || !x || !ret || ld < c->N || qmrmax < 0) return SCALIT_EINVAL;
    (void)stream;
    HIP_TRY(hipDeviceSynchronize());   // order after whatever the caller enqueued on any stream
    return qmr_f64(c, ntype, qmrmax, qmrtol, nvec, b, x, ld, shifts, ret, res, c->stream);
}

int scalit_qmr_multi(int njob, scalit_qmr_job *jobs, int ntype, int qmrmax, double qmrtol)
{
    if (njob < 1 || !jobs || qmrmax < 0) return SCALIT_EINVAL;
    for (int j = 0; j < njob; j++) {
        scalit_qmr_job &J = jobs[j];
        J.status = SCALIT_EINVAL;
        if (!J.ctx || J.nvec < 1 || !J.b || !J.x || !J.ret || J.ld < J.ctx->N) return SCALIT_EINVAL;
        if (J.ctx->device != jobs[0].ctx->device) return SCALIT_EINVAL;
        for (int i = 0; i < j; i++)
            if (jobs[i].ctx == J.ctx) return SCALIT_EINVAL;   // one job per context: each owns its workspace
        SC_TRY(operator_ready(J.ctx));
        if (!J.ctx->has_precond) return (J.status = SCALIT_ESTATE);
    }
    HIP_TRY(hipSetDevice(jobs[0].ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    return qmr_multi_f64(njob, jobs, ntype, qmrmax, qmrtol);
}

int scalit_qmr_mixed(scalit_ctx *c, int prec, int qmrmax, double qmrtol, int inner_max, double inner_tol, int nvec,
                     const double *b, double *x, long ld, const double *shifts, int *ret, double *res,
                     scalit_stream stream)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (!c->has_precond) return SCALIT_ESTATE;
    if (prec != SCALIT_FP32 && prec != SCALIT_BF16) return SCALIT_EINVAL;
    if (nvec < 1 || !b || !x || !ret || ld < c->N || qmrmax < 1 || inner_max < 1) return SCALIT_EINVAL;
    (void)stream;
    HIP_TRY(hipDeviceSynchronize());   // order after whatever the caller enqueued on any stream
    return qmr_mixed(c, prec, -1, qmrmax, qmrtol, inner_max, inner_tol, nvec, b, ld, x, ld, shifts, ret, res,
                     c->stream);
}

int scalit_pist_conv(scalit_ctx *c, double e0, double etol, int ntype, const double *x, int startcnt, int stepcnt,
                     int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int prec, double *eig,
                     double *res, int *stats)
{
    int ret = 0;
    double r = 0;
    int rc = scalit_pist_conv_batched(c, 1, &e0, etol, ntype, x, c ? c->N : 0, startcnt, stepcnt, stepeig, m, m_max,
                                      qmr_type, qmrmax, qmrtol, prec, eig, &ret, &r, stats);
    if (rc != SCALIT_OK) return rc;   // the reference's parameter errors come back as ret = 0
    if (res) *res = r;
    return ret;
}

int scalit_pist_conv_batched(scalit_ctx *c, int nwin, const double *e0, double etol, int ntype, const double *x,
                             long ld, int startcnt, int stepcnt, int stepeig, int m, int m_max, int qmr_type,
                             int qmrmax, double qmrtol, int prec, double *eig, int *ret, double *res, int *stats)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (nwin < 1 || !e0 || !x || !eig || !ret || m < 1 || m_max < 1 || ld < c->N) return SCALIT_EINVAL;
    if (prec != SCALIT_FP64 && prec != SCALIT_FP32 && prec != SCALIT_BF16) return SCALIT_EINVAL;
    PistParams pp{etol, ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, prec, qmrtol};
    HIP_TRY(hipDeviceSynchronize());
    return pist_conv(c, nwin, e0, pp, x, ld, eig, ret, res, stats, c->stream);
}

int scalit_pist_conv_multi(int njob, scalit_pist_job *jobs, double etol, int ntype, int startcnt, int stepcnt,
                           int stepeig, int m, int m_max, int qmr_type, int qmrmax, double qmrtol, int prec)
{
    if (njob < 1 || !jobs || m < 1 || m_max < 1) return SCALIT_EINVAL;
    if (prec != SCALIT_FP64 && prec != SCALIT_FP32 && prec != SCALIT_BF16) return SCALIT_EINVAL;
    for (int j = 0; j < njob; j++) {
        scalit_pist_job &J = jobs[j];
        J.status = SCALIT_EINVAL;
        if (!J.ctx || J.nwin < 1 || !J.e0 || !J.x || !J.eig || !J.ret || J.ld < J.ctx->N) return SCALIT_EINVAL;
        for (int i = 0; i < j; i++)
            if (jobs[i].ctx == J.ctx) return SCALIT_EINVAL;   // one job per context: each owns its stream and state
        if (hipSetDevice(J.ctx->device) != hipSuccess) return SCALIT_EHIP;
        SC_TRY(operator_ready(J.ctx));
    }
    PistParams pp{etol, ntype, startcnt, stepcnt, stepeig, m, m_max, qmr_type, qmrmax, prec, qmrtol};
    for (int j = 0; j < njob; j++) {
        HIP_TRY(hipSetDevice(jobs[j].ctx->device));
        HIP_TRY(hipDeviceSynchronize());
    }
    // one host thread per problem, each on its context's stream: the problems'
    // kernels share the device while every solve keeps its own control flow
    std::vector<std::thread> th;
    for (int j = 0; j < njob; j++)
        th.emplace_back([&jobs, j, &pp]() {
            scalit_pist_job &J = jobs[j];
            scalit_ctx *c = J.ctx;
            if (hipSetDevice(c->device) != hipSuccess) {
                J.status = SCALIT_EHIP;
                return;
            }
            J.status = pist_conv(c, J.nwin, J.e0, pp, J.x, J.ld, J.eig, J.ret, J.res, J.stats, c->stream);
        });
    for (auto &t : th) t.join();
    for (int j = 0; j < njob; j++)
        if (jobs[j].status != SCALIT_OK) return jobs[j].status;
    return SCALIT_OK;
}

int scalit_pist_vectors(scalit_ctx *c, int win, double *psi, long ld)
{
    CTX_GUARD(c);
    if (!psi || ld < c->N) return SCALIT_EINVAL;
    return pist_vectors(c, win, psi, ld, c->stream);
}

int scalit_pist_vectors_all(scalit_ctx *c, double *psi, long ld, int ncols)
{
    CTX_GUARD(c);
    if (!psi || ld < c->N || ncols < 1) return SCALIT_EINVAL;
    return pist_vectors_all(c, psi, ld, ncols, c->stream);
}

int scalit_lan_conv(scalit_ctx *c, double e0, double etol, int ntype, const double *x, int startcnt, int stepcnt,
                    int stepeig, int m, int m_max, double *eig, double *res)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (!x || !eig || !res || m < 1) return SCALIT_EINVAL;
    HIP_TRY(hipDeviceSynchronize());
    return lan_conv(c, e0, etol, ntype, x, startcnt, stepcnt, stepeig, m, m_max, eig, res, c->stream);
}

int scalit_set_pist_orth(scalit_ctx *c, int passes)
{
    if (!c || passes < 1 || passes > 2) return SCALIT_EINVAL;
    c->pist_orth = passes;
    return SCALIT_OK;
}

int scalit_eig_gradients(scalit_ctx *c, int nk, const double *psi, long ld, int np, const double *dvdp, long lddv,
                         double *out)
{
    CTX_GUARD(c);
    if (nk < 1 || np < 1 || !psi || !dvdp || !out || ld < c->N || lddv < c->N) return SCALIT_EINVAL;
    HIP_TRY(hipDeviceSynchronize());
    return eig_gradients(c, nk, psi, ld, np, dvdp, lddv, out, c->stream);
}

// ---------------------------------------------------------------------------
// Fortran drop-in layer: host arrays, by-reference arguments, bound context.
// ---------------------------------------------------------------------------
static scalit_ctx *g_bound = nullptr;
static int g_qmr_type = -1, g_qmr_max = 1000;
static double g_qmr_tol = 1e-3;
// host staging of the bound context
static double *g_dx = nullptr, *g_dy = nullptr;
static long g_dn = 0;

static int stage(long n)
{
    if (n <= g_dn) return SCALIT_OK;
    if (g_dx) (void)hipFree(g_dx);
    if (g_dy) (void)hipFree(g_dy);
    g_dx = g_dy = nullptr;
    g_dn = 0;
    if (hipMalloc((void **)&g_dx, 8 * n) != hipSuccess) return SCALIT_ENOMEM;
    if (hipMalloc((void **)&g_dy, 8 * n) != hipSuccess) return SCALIT_ENOMEM;
    g_dn = n;
    return SCALIT_OK;
}

int scalit_bind(scalit_ctx *c)
{
    g_bound = c;
    return SCALIT_OK;
}

void scalit_set_qmr_params(int qmr_type, int qmrmax, double qmrtol)
{
    g_qmr_type = qmr_type;
    g_qmr_max = qmrmax;
    g_qmr_tol = qmrtol;
}

// host-array operator application; on failure y is filled with NaN so that a
// caller without a status channel cannot mistake it for a result
static void host_apply(int which, const int *n, const double *x, double *y)
{
    scalit_ctx *c = g_bound;
    const long nn = n ? *n : 0;
    int rc = SCALIT_EINVAL;
    if (c && nn == c->N && hipSetDevice(c->device) == hipSuccess && stage(nn) == SCALIT_OK) {
        rc = SCALIT_OK;
        if (hipMemcpy(g_dx, x, 8 * nn, hipMemcpyHostToDevice) != hipSuccess) rc = SCALIT_EHIP;
        if (rc == SCALIT_OK) {
            switch (which) {
            case 0: rc = scalit_hx(c, SCALIT_TA1, 1, g_dx, nn, g_dy, nn, nullptr, nullptr); break;   // HX (sHC=TA1)
            case 1: rc = scalit_hx(c, SCALIT_TA0, 1, g_dx, nn, g_dy, nn, nullptr, nullptr); break;   // HX0
            case 2: rc = scalit_px(c, SCALIT_TOSB, SCALIT_TA1, 1, g_dx, nn, g_dy, nn, nullptr, nullptr); break;
            }
        }
        if (rc == SCALIT_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = SCALIT_EHIP;
        if (rc == SCALIT_OK && hipMemcpy(y, g_dy, 8 * nn, hipMemcpyDeviceToHost) != hipSuccess) rc = SCALIT_EHIP;
    }
    if (rc != SCALIT_OK) {
        std::fprintf(stderr, "scalit: operator callback failed (%d)\n", rc);
        for (long i = 0; i < nn; i++) y[i] = std::nan("");
    }
}

void hx_(const int *n, const double *x, double *y) { host_apply(0, n, x, y); }
void ehx_(const int *n, const double *x, double *y) { host_apply(0, n, x, y); }
void hx0_(const int *n, const double *x, double *y) { host_apply(1, n, x, y); }
void px_(const int *n, const double *x, double *y) { host_apply(2, n, x, y); }
void epx_(const int *n, const double *x, double *y) { host_apply(2, n, x, y); }

int osb_qmr_(const int *n, const double *b, double *x, double *eres)
{
    int ntype = g_qmr_type, mx = g_qmr_max;
    return qmr_(&ntype, &mx, &g_qmr_tol, n, b, hx_, px_, x, eres);
}

// QMR on host arrays.  With this library's own callbacks the vectors live on
// the GPU for the whole solve; other callbacks run a host restatement of the
// same recurrence around them (they may be anything the caller links).
int qmr_(const int *ntype, const int *qmrmax, const double *qmrtol, const int *n, const double *b, scalit_fop hx,
         scalit_fop px, double *x, double *res)
{
    const long nn = *n;
    const bool own = (hx == hx_ || hx == ehx_) && (px == px_ || px == epx_) && g_bound && g_bound->N == nn;
    if (own) {
        scalit_ctx *c = g_bound;
        if (hipSetDevice(c->device) != hipSuccess || stage(nn) != SCALIT_OK) return -1;
        if (hipMemcpy(g_dx, b, 8 * nn, hipMemcpyHostToDevice) != hipSuccess) return -1;
        if (*ntype > 0 && hipMemcpy(g_dy, x, 8 * nn, hipMemcpyHostToDevice) != hipSuccess) return -1;
        int ret = 0;
        double r = 0;
        if (scalit_qmr(c, *ntype, *qmrmax, *qmrtol, 1, g_dx, g_dy, nn, nullptr, &ret, &r, nullptr) != SCALIT_OK)
            return -1;
        if (hipMemcpy(x, g_dy, 8 * nn, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        if (ret != 0) *res = r;
        return ret;
    }
    // foreign callbacks: the recurrence stays on the GPU, each operator
    // application round-trips its vector to the caller's host callback
    scalit_ctx *c = g_bound;
    if (!c || c->N != nn) {
        std::fprintf(stderr, "scalit: qmr_ needs a bound context of the same length (scalit_bind)\n");
        return -1;
    }
    if (hipSetDevice(c->device) != hipSuccess) return -1;
    int ret = 0;
    double r = 0;
    if (qmr_host_cb(c, *ntype, *qmrmax, *qmrtol, nn, b, x, hx, px, &ret, &r) != SCALIT_OK) return -1;
    if (ret != 0) *res = r;
    return ret;
}

int pist_conv_(const double *e0, const double *etol, const int *ntype, const int *n, const double *x,
               const int *startcnt, const int *stepcnt, const int *stepeig, const int *m, const int *m_max,
               scalit_fop h0x, scalit_flinsolv linsolv, double *eig, double *res)
{
    scalit_ctx *c = g_bound;
    // GPU-resident when the callbacks are this library's operator and OSB QMR
    if (!c || c->N != *n || h0x != hx0_ || linsolv != osb_qmr_) {
        std::fprintf(stderr, "scalit: pist_conv_ needs the bound context's hx0_ and osb_qmr_ callbacks\n");
        return 0;
    }
    if (stage(*n) != SCALIT_OK || hipMemcpy(g_dx, x, 8 * (size_t)*n, hipMemcpyHostToDevice) != hipSuccess) return 0;
    c->e0 = *e0;
    return scalit_pist_conv(c, *e0, *etol, *ntype, g_dx, *startcnt, *stepcnt, *stepeig, *m, *m_max, g_qmr_type,
                            g_qmr_max, g_qmr_tol, SCALIT_FP64, eig, res, nullptr);
}

int lan_conv_(const double *e0, const double *etol, const int *ntype, const int *n, const double *x,
              const int *startcnt, const int *stepcnt, const int *stepeig, const int *m, const int *m_max,
              scalit_fop hx, double *eig, double *res)
{
    scalit_ctx *c = g_bound;
    if (!c || c->N != *n || (hx != hx0_)) {
        std::fprintf(stderr, "scalit: lan_conv_ needs the bound context's hx0_ callback\n");
        return 0;
    }
    if (stage(*n) != SCALIT_OK || hipMemcpy(g_dx, x, 8 * (size_t)*n, hipMemcpyHostToDevice) != hipSuccess) return 0;
    return scalit_lan_conv(c, *e0, *etol, *ntype, g_dx, *startcnt, *stepcnt, *stepeig, *m, *m_max, eig, res);
}

// MGS_ORTH on host arrays (gsorth.f90:42-72); vj is N x (M+1), column-major.
// Runs on the bound context's GPU.
int mgs_orth_(const int *n, const double *wj, const int *m, double *vj)
{
    scalit_ctx *c = g_bound;
    if (!c || hipSetDevice(c->device) != hipSuccess) {
        std::fprintf(stderr, "scalit: mgs_orth_ needs a bound context (scalit_bind)\n");
        return 0;
    }
    int ok = 0;
    if (mgs_host(c, *n, wj, *m, vj, &ok) != SCALIT_OK) return 0;
    return ok;
}

// test hook: force the generic (any-level) H*X path
int scalit_debug_force_generic(int on)
{
    g_force_generic = on;
    return SCALIT_OK;
}

int scalit_debug_hx_variant(scalit_ctx *c, int prec, int nvec, int *va, int *vb)
{
    CTX_GUARD(c);
    if (!va || !vb || nvec < 1) return SCALIT_EINVAL;
    SC_TRY(hx_prepare(c, prec, nvec, nullptr));
    return hx_variants(c, prec, nvec, va, vb);
}

int scalit_debug_mfma_peak(int prec, int iters, int blocks, float *ms)
{
    if (!ms || iters < 1 || blocks < 1) return SCALIT_EINVAL;
    return mfma_peak(prec, iters, blocks, ms);
}

int scalit_debug_sym_eig(int n, const double *a, int lda, double *w, double *z, int ldz)
{
    if (n < 1 || !a || !w || lda < n || (z && ldz < n)) return SCALIT_EINVAL;
    DBuf<double> da, dw, dz, work;
    auto run = [&]() -> int {
        SC_TRY(da.alloc((size_t)n * n));
        SC_TRY(dw.alloc(n));
        SC_TRY(work.alloc(small_sym_eig_work(n)));
        if (z) SC_TRY(dz.alloc((size_t)n * n));
        for (int j = 0; j < n; j++)
            HIP_TRY(hipMemcpy(da.p + (long)j * n, a + (long)j * lda, sizeof(double) * n, hipMemcpyHostToDevice));
        SC_TRY(small_sym_eig(n, da.p, n, dw.p, z ? dz.p : nullptr, n, work.p, nullptr));
        HIP_TRY(hipMemcpy(w, dw.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (z)
            for (int j = 0; j < n; j++)
                HIP_TRY(hipMemcpy(z + (long)j * ldz, dz.p + (long)j * n, sizeof(double) * n, hipMemcpyDeviceToHost));
        return SCALIT_OK;
    };
    const int rc = run();
    da.release();
    dw.release();
    dz.release();
    work.release();
    return rc;
}

int scalit_debug_sym_eigvals_batch(int njobs, const int *ns, const double *a, int lda, double *w)
{
    if (njobs < 1 || !ns || !a || !w || lda < 1) return SCALIT_EINVAL;
    for (int j = 0; j < njobs; j++)
        if (ns[j] < 1 || ns[j] > lda) return SCALIT_EINVAL;
    const size_t per = (size_t)lda * lda;
    DBuf<double> da, dw, work;
    DBuf<EigvalJob> djobs;
    auto run = [&]() -> int {
        SC_TRY(da.alloc(per * njobs));
        SC_TRY(dw.alloc((size_t)lda * njobs));
        HIP_TRY(hipMemcpy(da.p, a, sizeof(double) * per * njobs, hipMemcpyHostToDevice));
        std::vector<EigvalJob> jobs(njobs);
        for (int j = 0; j < njobs; j++) jobs[j] = {da.p + per * j, dw.p + (size_t)lda * j, ns[j], lda};
        SC_TRY(small_sym_eigvals_batch(njobs, jobs.data(), djobs, work, nullptr));
        HIP_TRY(hipMemcpy(w, dw.p, sizeof(double) * lda * njobs, hipMemcpyDeviceToHost));
        return SCALIT_OK;
    };
    const int rc = run();
    da.release();
    dw.release();
    work.release();
    djobs.release();
    return rc;
}

int scalit_debug_mfma_probe(int prec, double *out)
{
    if (!out) return SCALIT_EINVAL;
    return mfma_probe(prec, out);
}

int scalit_profile_hx_pass(scalit_ctx *c, int pass, int nvec, const double *x, long ldx, double *y, long ldy,
                           scalit_stream stream)
{
    CTX_GUARD(c);
    SC_TRY(operator_ready(c));
    if (nvec < 1 || !x || !y || ldx < c->N || ldy < c->N || x == y) return SCALIT_EINVAL;
    return hx_pass_f64(c, pass, nvec, x, ldx, y, ldy, pick(c, stream));
}

}  // extern "C"
