// Sequential truncation of a direct-product DVR basis, the contraction step
// of ScalIT's Hamiltonian generators, with its dense eigensolves on the GPU:
//
//  * scalit_podvr: the potential-optimised DVR of one coordinate,
//    DVR_Step2 + DVR_Step3 (src/ham/libham/dvr/dvr.f90:84-160): diagonalise
//    the 1-D Hamiltonian h = T + diag(v) on the primitive grid x, keep its m
//    lowest eigenvectors U, diagonalise the position operator in that basis,
//    X = U^T diag(x) U = W diag(x0) W^T, and return the new grid x0 with the
//    1-D Hamiltonian H0 = W^T diag(E) W on it.  DVR_Step3 leaves the signs of
//    W to DSYEV; a sign flip of column a flips row and column a of H0, an
//    equivalent operator (D H0 D with D = diag(+-1), the same spectrum with
//    every diagonal potential).  Here each column of U and of W has its
//    largest-magnitude entry positive (U's signs pass into W's rows and so
//    into which entry of a column of W is normalised), a choice that, unlike DIAG's
//    first-full-row rule (src/libutil/serial/base/mm.f90:66-84), does not
//    hang on entries at rounding level.
//  * scalit_combine_2d: the energy-cut combined layer of two coordinates,
//    saveH0CBALL (src/ham/ja3/module/ja3.h0:154-199): of the n1 x n2 product
//    points keep those with the lowest v1(i) + v2(j) and build the layer's
//    Hamiltonian H0(a,b) = [j_a == j_b] H1(i_a,i_b) + [i_a == i_b] H2(j_a,j_b).
//    The reference keeps every point with v1 + v2 <= Ecutoff; here the cut is
//    placed so that exactly `nkeep` points stay (ties broken in the
//    reference's loop order), and the kept points are listed in that order
//    (j outer, i inner).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "scalit_impl.hpp"

namespace scalit {

// eigenvalues (ascending) and eigenvectors of a symmetric n x n host matrix
// on the current device
static int dense_eig(int n, const std::vector<double> &a, std::vector<double> &w, std::vector<double> &z,
                     hipStream_t s)
{
    DBuf<double> da, dw, dz, work;
    SC_TRY(da.alloc((size_t)n * n));
    SC_TRY(dw.alloc(n));
    SC_TRY(dz.alloc((size_t)n * n));
    SC_TRY(work.alloc(small_sym_eig_work(n)));
    HIP_TRY(hipMemcpyAsync(da.p, a.data(), 8 * (size_t)n * n, hipMemcpyHostToDevice, s));
    SC_TRY(small_sym_eig(n, da.p, n, dw.p, dz.p, n, work.p, s));
    w.resize(n);
    z.resize((size_t)n * n);
    HIP_TRY(hipMemcpyAsync(w.data(), dw.p, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(z.data(), dz.p, 8 * (size_t)n * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}

// every column of the n x n column-major z with its largest-magnitude entry
// positive (the first such entry on a tie)
static void column_signs(int n, std::vector<double> &z)
{
    for (int j = 0; j < n; j++) {
        double *c = z.data() + (long)n * j;
        int im = 0;
        for (int i = 1; i < n; i++)
            if (std::fabs(c[i]) > std::fabs(c[im])) im = i;
        if (c[im] < 0.0)
            for (int i = 0; i < n; i++) c[i] = -c[i];
    }
}

}  // namespace scalit

using namespace scalit;

extern "C" int scalit_podvr(int n, const double *x, const double *t, const double *v, int m, double *x0, double *h0,
                            double *e)
{
    if (n < 1 || n > 1024 || m < 1 || m > n || !x || !t || !v || !x0 || !h0) return SCALIT_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SCALIT_EHIP;
    // DVR_Step2: h = T + diag(V(x)), all eigenpairs
    std::vector<double> h((size_t)n * n), eh, u;
    for (long k = 0; k < (long)n * n; k++) h[k] = t[k];
    for (int i = 0; i < n; i++) h[i + (long)n * i] += v[i];
    SC_TRY(dense_eig(n, h, eh, u, nullptr));
    // the signs of U's columns carry into the sign choice of W's below
    column_signs(n, u);
    // DVR_Step3: X(I,J) = sum_k U(k,I) x(k) U(k,J) over the m lowest
    std::vector<double> xm((size_t)m * m), w, z;
    for (int i = 0; i < m; i++)
        for (int j = i; j < m; j++) {
            double acc = 0;
            for (int k = 0; k < n; k++) acc += u[k + (long)n * i] * x[k] * u[k + (long)n * j];
            xm[i + (long)m * j] = acc;
            xm[j + (long)m * i] = acc;
        }
    SC_TRY(dense_eig(m, xm, w, z, nullptr));
    column_signs(m, z);
    // H0(I,J) = sum_k W(k,I) E(k) W(k,J)
    for (int i = 0; i < m; i++) {
        x0[i] = w[i];
        for (int j = i; j < m; j++) {
            double acc = 0;
            for (int k = 0; k < m; k++) acc += z[k + (long)m * i] * eh[k] * z[k + (long)m * j];
            h0[i + (long)m * j] = acc;
            h0[j + (long)m * i] = acc;
        }
    }
    if (e)
        for (int k = 0; k < m; k++) e[k] = eh[k];
    return SCALIT_OK;
}

extern "C" int scalit_combine_2d(int n1, const double *h1, const double *v1, int n2, const double *h2,
                                 const double *v2, int nkeep, int *idx1, int *idx2, double *h0c)
{
    if (n1 < 1 || n2 < 1 || nkeep < 1 || (long)nkeep > (long)n1 * n2 || !h1 || !v1 || !h2 || !v2 || !idx1 || !idx2 ||
        !h0c)
        return SCALIT_EINVAL;
    // the product points in the reference's loop order: j (level 2) outer, i inner
    const long np = (long)n1 * n2;
    std::vector<long> ord(np);
    std::iota(ord.begin(), ord.end(), 0L);
    std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) {
        return v1[a % n1] + v2[a / n1] < v1[b % n1] + v2[b / n1];
    });
    std::vector<long> kept(ord.begin(), ord.begin() + nkeep);
    std::sort(kept.begin(), kept.end());   // back to loop order
    for (int a = 0; a < nkeep; a++) {
        idx1[a] = (int)(kept[a] % n1);
        idx2[a] = (int)(kept[a] / n1);
    }
    for (int a = 0; a < nkeep; a++) {
        for (int b = a + 1; b < nkeep; b++) {
            double hv = 0.0;
            if (idx1[a] == idx1[b]) hv += h2[idx2[a] + (long)n2 * idx2[b]];
            if (idx2[a] == idx2[b]) hv += h1[idx1[a] + (long)n1 * idx1[b]];
            h0c[a + (long)nkeep * b] = hv;
            h0c[b + (long)nkeep * a] = hv;
        }
        h0c[a + (long)nkeep * a] = h1[idx1[a] + (long)n1 * idx1[a]] + h2[idx2[a] + (long)n2 * idx2[a]];
    }
    return SCALIT_OK;
}
