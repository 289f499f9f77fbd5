// alt3_cg without device allocations: cg_quadratic (THERMACOND/cg.f90:19-69) through the same
// kernels as alt3_cg, with A, b, x, the CG vectors, the dot-product partials and the state
// flags staged in a file-scope device array instead of buffers allocated per call.
//
// Layout of the array (doubles): work [4n] | red [2*256 + 16] | state [1, two ints] | x [n] |
// b [n] | A [n*n] (b and A only when the caller passes them; otherwise the context's
// collision matrix and right-hand side are used in place).
#include <mutex>

#include "alt3_impl.h"

namespace {

constexpr size_t kScratchDoubles = (size_t)6 << 20;   // 48 MiB on every device the library runs on
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices];   // device address of scratch, per device (hipGetSymbolAddress)
constexpr size_t kRed = 2 * 256 + 16;

}  // namespace

namespace alt3 {

// once per device, with that device current: the device's own copy of scratch (in host code the
// array's name is only a stand-in, so every use goes through this address)
int scratch_setup()
{
    int dev = 0;
    ALT3_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ALT3_E_ARG, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return ALT3_OK;
    void *p = nullptr;
    ALT3_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<double *>(p);
    return ALT3_OK;
}

}  // namespace alt3

using namespace alt3;

extern "C" int alt3_cg_static(alt3_ctx *ctx, int n, const double *A, const double *b, double *x, int max_iter,
                              int ndig, double *funcmin, int *iters)
{
    if (!ctx || !x || n <= 0) return fail(ALT3_E_ARG, "bad argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    if ((A == nullptr) != (b == nullptr)) return fail(ALT3_E_ARG, "A and b must both be given or both be NULL");
    const size_t nn = (size_t)n;
    const size_t need = 4 * nn + kRed + 1 + nn + (A ? nn + nn * nn : 0);
    if (need > kScratchDoubles)
        return fail(ALT3_E_ARG, "alt3_cg_static: n = " + std::to_string(n) + " needs " + std::to_string(need) +
                                    " doubles of static scratch, more than its " + std::to_string(kScratchDoubles) +
                                    " (n <= 2505 with A given)");
    double *ws = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) ws = g_scratch[ctx->device];
    }
    if (!ws) return fail(ALT3_E_STATE, "static scratch not set up on this device");
    double *work = ws, *red = work + 4 * nn;
    int *state = reinterpret_cast<int *>(red + kRed);
    double *dx = red + kRed + 1, *db = dx + nn, *dA = db + nn;
    const double *pA, *pb;
    if (!A) {
        if (!ctx->have_cm) return fail(ALT3_E_STATE, "alt3_cg_static(A=NULL) needs alt3_collision first");
        if (ctx->ibz_lo != 0 || ctx->ibz_hi != ctx->nibz - 1) return fail(ALT3_E_STATE, "collision matrix does not cover the IBZ");
        if (n != ctx->nibz * ctx->nb * 3) return fail(ALT3_E_ARG, "n must be nibz*nb*3");
        pA = ctx->d_cm.p;
        pb = ctx->d_rhs.p;
    } else {
        ALT3_HIP(hipMemcpyAsync(dA, A, sizeof(double) * nn * nn, hipMemcpyHostToDevice, ctx->stream));
        ALT3_HIP(hipMemcpyAsync(db, b, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
        pA = dA;
        pb = db;
    }
    ALT3_HIP(hipMemcpyAsync(dx, x, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    ALT3_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    ALT3_TRY(cg_device_ws(ctx, n, pA, pb, dx, max_iter, ndig, funcmin, iters, A == nullptr,
                          A == nullptr && ctx->cm_fp32 ? ctx->d_cmf.p : nullptr, work, red, state));
    ALT3_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ALT3_HIP(hipMemcpyAsync(x, dx, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    ALT3_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    return ALT3_OK;
}
