// Internal definitions shared by the HIP sources of libscalit_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/scalit.h"

#define SC_FMAX SCALIT_FMAX

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            std::fprintf(stderr, "scalit: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                             \
            return SCALIT_EHIP;                                                                 \
        }                                                                                       \
    } while (0)

#define SC_TRY(expr)             \
    do {                         \
        int r_ = (expr);         \
        if (r_ != SCALIT_OK) return r_; \
    } while (0)

namespace scalit {

using bf16_t = uint16_t;

// A device buffer owned by the context (or by a scope: it is freed when the
// owner goes, on every return path).  Movable, not copyable.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    int alloc(size_t cnt) {
        if (cnt <= n && p) return SCALIT_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (cnt == 0) return SCALIT_OK;
        if (hipMalloc((void **)&p, cnt * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return SCALIT_ENOMEM;
        }
        n = cnt;
        return SCALIT_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Level {
    int n = 0;          // sN(level)
    bool dep = false;   // sDEP(level)
    int dim = 1;        // myDim(level): product of inner sizes (NIN)
    int blk = 1;        // myBlk(level): product of outer sizes (BLK)
    DBuf<double> h0, h0t, depv, vosb;   // h0t: transposed copy for the adjoint
    DBuf<float> h0_f, depv_f, vosb_f;
    // per-block transposes of vosb for levels >= 2 (the V^T rotations of P*X),
    // built on first use and released whenever vosb changes
    DBuf<double> vosbt;
    DBuf<float> vosbt_f;
    bool has_h0 = false, has_dep = false;
    // complex level matrix and DEP (sCX = T), (re, im) adjacent: n x n
    // column-major and myDim values
    DBuf<double> h0cx, depcx;
    bool has_h0cx = false, has_depcx = false;
};

// a complex value stored as the reference stores double complex
struct dcx {
    double re, im;
};

struct PistState;   // defined in solvers.hip
struct CxState;     // defined in complex.hip

// autotuned H*X pass variants for one (precision, batch size)
struct TuneEntry {
    int prec, nvec, va, vb;
};

}  // namespace scalit

struct scalit_ctx {
    int device = 0;
    int sF = 0;
    long N = 0;
    scalit::Level lv[SC_FMAX];
    scalit::DBuf<double> res, eig0;
    scalit::DBuf<float> res_f, eig0_f;
    bool has_res = false, has_precond = false;
    double e0 = 0.0, de = 1.0, beta = 0.0;
    hipStream_t stream = nullptr;
    // work buffers
    scalit::DBuf<double> work;      // generic-path ping-pong / mode-product temporaries
    scalit::DBuf<float> work_f;
    scalit::DBuf<double> shifts_d;  // per-vector shift staging
    scalit::DBuf<float> shifts_f;
    // sNDVR = T: the outermost level is the dense per-point layer OUTH
    // (myDim(sF) x n x n, point index fastest) instead of H0 [x DEP]
    bool ndvr = false;
    scalit::DBuf<double> outh;
    scalit::DBuf<float> outh_f;
    // sCX = T: the complex level matrices (Level::h0cx / depcx, and OUTHCX,
    // myDim(sF) x n x n complex) replace the real ones in the complex
    // operators and the OSB set-up (scalit_set_scx)
    scalit::DBuf<double> outhcx;
    bool has_outhcx = false, scx = false;
    // fused 3-level fast path is available (sF==3, sDEP=(F,F,T) or (F,F) + OUTH, sizes fit)
    bool fast3 = false;
    scalit::PistState *pist = nullptr;
    // orthogonalisation passes per PIST Krylov vector: 1 = the reference's
    // MGS_ORTH, 2 = once more against the basis (scalit_set_pist_orth)
    int pist_orth = 1;
    // the preconditioner kernel of the solvers' P*X (the reference's sOSB:
    // TOSB, TOSBD1 or TOSBD2, with DE and beta from scalit_set_shift)
    int osb = 0;
    // the TOSBW window (scalit_set_osbw): grid indices of its states
    // (ascending), the starts of its runs of states in one level-1 block
    // (osbw_nruns + 1), the inverse of its shifted Hamiltonian (osbw_n x
    // osbw_n), EIG0 with +inf at the window (the divide then leaves the
    // window's components 0) and the window's part of P*X per vector
    // (osbw_d, nvec x osbw_n); osbw_n = 0 when no window is set up
    scalit::DBuf<int> osbw_ind, osbw_runs;
    scalit::DBuf<double> osbw_m, osbw_eig0, osbw_d;
    // complex jobs with TOSBW (real H0): the window's H_W (EIG0 on the
    // diagonal, unshifted; n x n column-major) and indices on the host, and
    // per preconditioner type the inverse of H_W + diag(-E0 +- i A_W)
    // (n x n complex), rebuilt when E0, the absorbing potentials or the
    // window change; osbw_g holds a complex P*X's gathered window values
    struct OsbwCx {
        int npc = -1;
        double e0 = 0.0;
        long apv = -1, wv = -1;
        scalit::DBuf<double> m;
    };
    std::vector<double> osbw_hw_host, osbw_hwi_host;   // hwi: imaginary part (sCX), else empty
    std::vector<int> osbw_ind_host;
    OsbwCx osbw_cx[4];
    int osbw_cx_next = 0;
    scalit::DBuf<double> osbw_g;
    long ap_ver = 0, osbw_ver = 0;
    int osbw_n = 0, osbw_nruns = 0;
    double osbw_shift = 0.0;
    // the absorbing potentials AP, APP, APR of resonance and CRP jobs (N
    // values each) and the complex solvers' work buffers
    scalit::DBuf<double> ap[3];
    bool has_ap[3] = {false, false, false};
    scalit::CxState *cx = nullptr;
    // FP64 QMR on small fused 3-level operators: 1 = the persistent solver,
    // 0 = the launch-per-step iteration, -1 = as SCALIT_QMR_PERSIST says
    // (default 1)
    int qmr_persist = -1;
    // bumped whenever a device buffer that operator kernels read is
    // (re)allocated: recorded graphs of older generations are re-recorded
    long gen = 0;
    // unique over the process's contexts (a recorded multi-problem graph
    // names its contexts by it)
    long uid = 0;
    std::vector<scalit::TuneEntry> hx_tune;
};

namespace scalit {

// ---------------------------------------------------------------- kernels
// H*X (kernels_hx.hip)
int hx_apply_f64(scalit_ctx *c, const double *shifts_dev, double shift0, int nvec, const double *x, long ldx,
                 double *y, long ldy, bool transpose, const int *active, hipStream_t s);
int hx_apply_f32(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx,
                 float *y, long ldy, const int *active, hipStream_t s);
int hx_apply_bf16(scalit_ctx *c, const float *shifts_dev, float shift0, int nvec, const bf16_t *x, long ldx,
                  bf16_t *y, long ldy, const int *active, hipStream_t s);
// force the generic (any sF) path, for testing
extern int g_force_generic;
extern std::atomic<int> g_num_cu;
// Raise kernel fn's dynamic-LDS limit to the CU's 160 KB on the current device
// (once per kernel and device), and return how many workgroups of `threads`
// threads using `lds` bytes are resident per CU (cached per kernel, device,
// threads and lds; at least 1).  Safe to call from several host threads.
int resident_per_cu(const void *fn, int threads, size_t lds);
void allow_full_lds(const void *fn);
// choose the fused-path kernel variants for (prec, nvec) before graph capture
int hx_prepare(scalit_ctx *c, int prec, int nvec, hipStream_t s);
int hx_variants(const scalit_ctx *c, int prec, int nvec, int *va, int *vb);

// P*X (kernels_px.hip)
int px_apply_f64(scalit_ctx *c, int nosb, const double *shifts_dev, double shift0, int nvec, const double *x,
                 long ldx, double *y, long ldy, const int *active, hipStream_t s);
int px_apply_f32(scalit_ctx *c, int nosb, const float *shifts_dev, float shift0, int nvec, const float *x, long ldx,
                 float *y, long ldy, const int *active, hipStream_t s);
int px_apply_bf16(scalit_ctx *c, int nosb, const float *shifts_dev, float shift0, int nvec, const bf16_t *x,
                  long ldx, bf16_t *y, long ldy, const int *active, hipStream_t s);
int precond_to_f32(scalit_ctx *c);
// TOSBW around the level-1 rotation (osbw.hip): osbw_gather forms
// d = (H_W - E0)^-1 (V1^T x)(window) from level 1's input x; the level-1 pass
// then divides by osbw_eig0 (the window's components become 0) and
// osbw_scatter adds V1(:, window) d to its output
void osbw_gather(scalit_ctx *c, int nvec, const double *x, long ldx, const int *active, hipStream_t s);
// Complex vectors in split layout: vector v's real part at x + 2 v N, its
// imaginary part at x + (2 v + 1) N.  px_apply_cx is MYPX_DX (kernels_px.hip);
// cx_imag_diag returns the +-AP / APP / APR diagonal an nType or nPC adds
// (nullptr when none, SCALIT_ESTATE in *rc when that array is not set).
int px_apply_cx(scalit_ctx *c, int nosb, int npc, int nvec, const double *x, double *y, hipStream_t s);
int hx_apply_cxh(scalit_ctx *c, int ntype, int nvec, const double *x, double *y, hipStream_t s);
int osbw_cx_matrix(scalit_ctx *c, int npc, double e0, const double *a, double sign, const dcx **m, hipStream_t s);
bool real_levels_ready(const scalit_ctx *c);
const double *cx_imag_diag(const scalit_ctx *c, int ntype, double *sign, int *rc);
void release_complex(scalit_ctx *c);
void osbw_scatter(scalit_ctx *c, int nvec, double *y, long ldy, const int *active, hipStream_t s);

// OSB set-up (kernels_diag.hip)
int block_diag(scalit_ctx *c, int bjmax, double bjtol);
// symmetric eigen-solve of a small dense matrix on the GPU (one workgroup).
// a: n x n device, lda; w: n device (ascending); z: eigenvectors (device, ldz) or nullptr
int small_sym_eig(int n, const double *a, int lda, double *w, double *z, int ldz, double *work, hipStream_t s);
size_t small_sym_eig_work(int n);
// eigenvalues only of several small symmetric matrices in one launch (one
// workgroup each): the leading n x n block of a (device, leading dimension
// lda, upper triangle read), eigenvalues ascending into w (device).  The jobs
// are copied to djobs on stream s; the host array must stay valid until the
// stream has passed the launch.  work grows to njobs n_max (n_max + 1) doubles
// when n_max exceeds the in-LDS size.
struct EigvalJob {
    const double *a;
    double *w;
    int n, lda;
};
int small_sym_eigvals_batch(int njobs, const EigvalJob *jobs, DBuf<EigvalJob> &djobs, DBuf<double> &work,
                            hipStream_t s);
// eigenvalues, ascending, of several symmetric tridiagonal matrices in one
// launch (one workgroup each): diagonal d[0..n), off-diagonal e[0..n-1), into
// w (device).  Same lifetime rule for the host jobs.
struct TriEigJob {
    const double *d, *e;
    double *w;
    int n;
};
int tridiag_eigvals_batch(int njobs, const TriEigJob *jobs, DBuf<TriEigJob> &djobs, hipStream_t s);
// eigenvalues and eigenvectors of several small symmetric matrices in one
// launch (cyclic Jacobi, one workgroup each): z gets the eigenvectors (n x n,
// ldz) in the order of w.  Same lifetime rule for the host jobs; work grows to
// njobs small_sym_eig_work(n_max) doubles.
struct EigvecJob {
    const double *a;
    double *w, *z;
    int n, lda, ldz;
};
int small_sym_eig_batch(int njobs, const EigvecJob *jobs, DBuf<EigvecJob> &djobs, DBuf<double> &work,
                        hipStream_t s);

// vector helpers (kernels_vec.hip)
constexpr int NPART = 512;   // partial sums per vector in the deterministic reductions
void launch_dot_partials(int nvec, long n, const double *x, long ldx, const double *y, long ldy, double *part,
                         hipStream_t s);
void launch_reduce_partials(int nvec, const double *part, double *out, hipStream_t s);   // out[v] = sum
// device address of the no-allocation entry point's fixed scratch on `device`
// (current), looked up once per device (noalloc.hip)
int noalloc_scratch(int device, double **ptr);

int dot_host(int nvec, long n, const double *x, long ldx, const double *y, long ldy, double *out_host,
             double *part_dev, double *red_dev, hipStream_t s);

}  // namespace scalit
