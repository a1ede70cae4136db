// Hellmann-Feynman gradients without any device allocation, for callers that
// cannot call hipMalloc (code under HIP graph capture, serving loops,
// harnesses that forbid allocation).  No counterpart in the reference.
//
// The partial sums and the reduced results of scalit_eig_gradients live in a
// file-scope device array, `scratch`, one copy per GPU for the life of the
// process.  Host code never passes the array's name to a kernel (in host code
// it is only a stand-in symbol): its device address is looked up once per
// device with hipGetSymbolAddress, with that device current, and that address
// is what the kernels receive.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two host threads, or two streams in flight) race on it and must be
// serialised by the caller.  The array keeps the last call's contents; the
// partial-sum kernel writes every slot the reduction reads, so nothing needs
// clearing between calls.
#include <mutex>

#include "scalit_impl.hpp"
#include "solver_state.hpp"

namespace scalit {

// (psi_k, dV/dp_j) pairs one call may ask for: nk * np <= NOALLOC_MAX_PAIRS.
// Scratch: NPART partials plus one result per pair, 8.4 MB per GPU.
constexpr long NOALLOC_MAX_PAIRS = 2048;
constexpr long NOALLOC_SCRATCH = NOALLOC_MAX_PAIRS * (NPART + 1);

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH];

namespace {
constexpr int kMaxDev = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDev] = {};
}  // namespace

// Device address of `scratch` on `device` (made current by the caller),
// looked up on the first call for that device.
int noalloc_scratch(int device, double **ptr)
{
    if (device < 0 || device >= kMaxDev) return SCALIT_EINVAL;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "scalit: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
                         hipGetErrorString(e));
            return SCALIT_EHIP;
        }
        g_scratch_ptr[device] = (double *)p;
    }
    *ptr = g_scratch_ptr[device];
    return SCALIT_OK;
}

}  // namespace scalit

extern "C" int scalit_eig_gradients_noalloc(scalit_ctx *c, int nk, const double *psi, long ld, int np,
                                            const double *dvdp, long lddv, double *out)
{
    using namespace scalit;
    if (!c) return SCALIT_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return SCALIT_EHIP;
    if (nk < 1 || np < 1 || !psi || !dvdp || !out || ld < c->N || lddv < c->N) return SCALIT_EINVAL;
    if ((long)nk * np > NOALLOC_MAX_PAIRS) {
        std::fprintf(stderr,
                     "scalit_eig_gradients_noalloc: %d levels x %d parameters exceeds the %ld pairs its "
                     "fixed scratch holds\n",
                     nk, np, NOALLOC_MAX_PAIRS);
        return SCALIT_EINVAL;
    }
    double *part = nullptr;
    SC_TRY(noalloc_scratch(c->device, &part));
    double *red = part + NOALLOC_MAX_PAIRS * NPART;
    HIP_TRY(hipDeviceSynchronize());
    hipStream_t s = c->stream;
    launch_hf_part(nk, c->N, np, psi, ld, dvdp, lddv, part, s);
    launch_reduce_partials(nk * np, part, red, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, red, 8 * (size_t)nk * np, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SCALIT_OK;
}
