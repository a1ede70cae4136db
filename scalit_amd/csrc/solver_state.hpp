// Device-resident scalar state of the QMR recurrence (one per right-hand side)
// and the launchers of the solver kernels (kernels_vec.hip).
#pragma once
#include "scalit_impl.hpp"

namespace scalit {

enum { QMR_RUNNING = 0, QMR_CONVERGED = 1, QMR_EXIT = 2 };

struct QmrScal {
    double zvrat, zvrat_next, vnrm, znrm, znrm_next, vzdot, qsdot, beta, vnrmpo, theta, thetamo, gama, gamamo, eta,
        eta_next, bnrm, res;
    int status, ret, iter, pad;
};

// Vectors of one (batched) QMR: nvec columns each of R, UX, UR, Q, S, V, Z,
// HX0 with leading dimension ld; B and X with their own leading dimensions.
template <typename S>
struct QmrVecs {
    const S *b;
    S *x;
    long ldb, ldx;
    S *R, *UX, *UR, *Q, *Sv, *V, *Z, *HX0;
    long ld;
};

template <typename S>
void launch_qmr_init(int nvec, long n, int zero_x, const QmrVecs<S> &w, double *part, hipStream_t s);
template <typename S>
void launch_qmr_init_z(int nvec, long n, const QmrVecs<S> &w, const double *pbb, double *part, double *pvz,
                       hipStream_t s);
void launch_qmr_init_scal(int nvec, const double *pbb, const double *pzz, QmrScal *q, int *active, int qmrmax,
                          hipStream_t s);
template <typename S>
void launch_qmr_b(int nvec, long n, const QmrVecs<S> &w, QmrScal *q, int *active, const double *part, hipStream_t s);
template <typename S>
void launch_qmr_c(int nvec, long n, const QmrVecs<S> &w, const QmrScal *q, double *part, hipStream_t s);
template <typename S>
void launch_qmr_d(int nvec, long n, const QmrVecs<S> &w, QmrScal *q, int *active, const double *pin, double *part,
                  hipStream_t s);
template <typename S>
void launch_qmr_e(int nvec, long n, const QmrVecs<S> &w, QmrScal *q, const double *pvv, double *part, double *pvz,
                  hipStream_t s);
template <typename S>
void launch_qmr_f(int nvec, long n, int conv, const QmrVecs<S> &w, QmrScal *q, int *active, const double *pzz,
                  double *prr, double *pux, double *pxx, double *pur, hipStream_t s);
template <typename S>
void launch_qmr_g(int nvec, long n, const QmrVecs<S> &w, const QmrScal *q, double *phh, double *prr, hipStream_t s);
void launch_qmr_h(long n, int nvec, int conv, int qmrmax, double tol, QmrScal *q, int *active, const double *phh,
                  const double *prr, const double *pux, const double *pxx, const double *pur, hipStream_t s);

void launch_cast_d2f(int nvec, long n, const double *x, long ldx, float *y, long ldy, hipStream_t s);
void launch_cast_f2d(int nvec, long n, const float *x, long ldx, double *y, long ldy, hipStream_t s);
void launch_cast_d2b(int nvec, long n, const double *x, long ldx, bf16_t *y, long ldy, hipStream_t s);
void launch_cast_b2d(int nvec, long n, const bf16_t *x, long ldx, double *y, long ldy, hipStream_t s);
void launch_axpby(int nvec, long n, double a, const double *x, long ldx, double b, double *y, long ldy,
                  hipStream_t s);

// MGS_ORTH of W (ldwk apart per window) into column col (ldw apart) in one
// launch, one workgroup per active window; ok[k] = 0 on a zero norm before or
// after the projections, nrm_out[k] (when not null) = the norm after them;
// false (nothing launched) when n is too long for a workgroup's registers
bool launch_mgs_one(int nw, long n, int jnew, const double *vj, long ldv, long ldw, const double *w, long ldwk,
                    double *col, const int *active, int *ok, double *nrm_out, hipStream_t s);
void launch_mgs_step(int nw, long n, int i, int inext, const double *vj, long ldv, long ldw, double *w, long ldwk,
                     const double *pin, double *pout, const int *active, hipStream_t s);
void launch_normalize(int nw, long n, const double *w, long ldwk, double *dst, long ldd, const double *part, int *ok,
                      const int *active, hipStream_t s);
void launch_multidot(int nw, long n, int i0, int m, const double *vj, long ldv, long ldw, const double *w, long ldwk,
                     double *part, const int *active, hipStream_t s);
void launch_multidot_fin(int nw, int m, int i0, int j, const double *part, double *hmat, int ldh, long ldhk,
                         const int *active, hipStream_t s);
void launch_ritz(long n, int m, int nc, const double *vj, long ldv, const double *z, int ldz, const int *cols,
                 double *psi, long ldp, hipStream_t s);
// one Ritz vector per job (window's Krylov basis vj, eigenvector column col
// of its z), all in one launch; jobs_dev: device array of njobs
struct RitzJob {
    const double *vj;
    const double *z;
    double *psi;
    int m, ldz, col;
};
void launch_ritz_batch(long n, long ldv, int njobs, const RitzJob *jobs_dev, hipStream_t s);
void launch_hf_part(int nk, long n, int np, const double *psi, long ld, const double *dvdp, long lddv, double *part,
                    hipStream_t s);
void launch_dv_accum(long n, int nvec, const double *x, const double *g, long ld, double *dv, hipStream_t s);

bool fast3_ok(const scalit_ctx *c);

// The persistent QMR solver (kernels_qmr.hip): every iteration of the solve
// in one launch for the fused 3-level operator on small grids (FP64).
struct QmrPersistBufs {
    DBuf<double> part;   // partial sums of the dot products
    DBuf<double> opart;  // partial sums of the Gram-Schmidt step
    DBuf<unsigned> bar;  // barrier arrival counters
    DBuf<int> err;       // a barrier wait gave up
    int herr = 0;
};
constexpr int QP_NOT_LAUNCHED = 1;
bool qmr_persist_ok(const scalit_ctx *c);
int qmr_persist_solve(scalit_ctx *c, int nvec, const QmrVecs<double> &w, QmrScal *q, int *active,
                      const double *shifts, const double *part_vz, int conv, int qmrmax, double tol, QmrPersistBufs &pb,
                      hipStream_t s);
// The Krylov steps of pistH0 for nw windows in one persistent launch: for
// j = j_from..j_to, the QMR solve with basis column j - 1 as its right-hand
// side, then MGS_ORTH of the solution into column j (orth passes).  Per
// window: qret[v * ldg + j] = the QMR return of step j (steps not run keep
// what the caller put there), wret[v] = 1 when every step ran, else the
// pistH0 value that ended the window.  gram: per window ldg x ldg, the basis
// columns' dot products, extended by one row per step.
struct PistPersistCall {
    double *vj;
    long ldv, ldw;
    double *gram;
    int ldg, j_from, j_to, orth, zero_x, conv, qmrmax;
    double tol;
    int *qret, *wret;
};
int pist_persist_max_cols(scalit_ctx *c, int nw, int ldg);
int pist_persist_steps(scalit_ctx *c, int nw, const PistPersistCall &pc, QmrVecs<double> w, QmrScal *q,
                       int *active, const double *shifts, QmrPersistBufs &pb, hipStream_t s);

}  // namespace scalit
