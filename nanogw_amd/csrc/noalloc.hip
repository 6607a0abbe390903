// The entry point that allocates nothing: the QP Newton solve of
// ngw_qp_newton (SIGMA/printsigma.F90:97-128) on device-resident inputs and
// outputs, with the kernel's spline work arrays in a file-scope device array
// instead of a context buffer.  For callers that cannot allocate device
// memory.  Every GPU the library runs on holds its own copy of the array
// until the process ends; all calls on one GPU share it.

#include "ngw_kernels.h"

#include <mutex>

namespace ngw {

namespace {

// qp_newton_kernel's work arrays: 5 nen doubles per QP state (x, f, Im f and
// the two spline vectors).  Every entry a call reads is written earlier in
// the same call, so nothing has to be cleared between calls.
constexpr long QP_SCRATCH_DOUBLES = 5L * NGW_QP_NOALLOC_MAX_POINTS;
static __device__ __align__(256) double scratch[QP_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};  // device address of each GPU's copy

}  // namespace

// Looks up this GPU's address of `scratch` once per device; `device` must be
// current.  The array's host-side name is only a stand-in: kernels get the
// address found here.
int noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return fail(NGW_ERR_ARG, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return NGW_OK;
    int cur = -1;
    NGW_HIP(hipGetDevice(&cur));
    if (cur != device) return fail(NGW_ERR_STATE, "noalloc_setup: the device is not current");
    void* p = nullptr;
    NGW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<double*>(p);
    return NGW_OK;
}

}  // namespace ngw

using namespace ngw;

extern "C" int ngw_qp_newton_dev_noalloc(ngw_context* ctx, int nqp, int nen, double deltae_ev, const double* d_in,
                                         double* d_out)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (nqp <= 0 || nen < 4) return fail(NGW_ERR_ARG, "need nqp > 0 and nen >= 4");
    if (!d_in || !d_out) return fail(NGW_ERR_ARG, "null device array");
    if ((long)nqp * nen > NGW_QP_NOALLOC_MAX_POINTS)
        return fail(NGW_ERR_ARG, "nqp * nen = " + std::to_string((long)nqp * nen) +
                                     " exceeds NGW_QP_NOALLOC_MAX_POINTS = " +
                                     std::to_string((long)NGW_QP_NOALLOC_MAX_POINTS));
    double* scr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < MAX_DEVICES) scr = g_scratch[ctx->device];
    }
    if (!scr) return fail(NGW_ERR_STATE, "no scratch address for the context's device");
    NGW_HIP(hipSetDevice(ctx->device));
    return qp_newton(ctx->stream, nqp, nen, deltae_ev, d_in, d_out, scr);
}
