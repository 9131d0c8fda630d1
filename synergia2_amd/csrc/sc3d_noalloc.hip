// sc3d_noalloc.hip -- sc3d_moments_noalloc: the bunch moments of sc3d_moments
// for callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc).  It takes no
// context (creating one allocates) and keeps its chunk partials in a
// file-scope device array instead of a context buffer.
//
// The array lives once per GPU for the life of the process.  Its device
// address is looked up with hipGetSymbolAddress on the first call on each
// device (with that device current) and kept per device; host code passes
// that address to the kernels, never the array's name.  Every call on a GPU
// shares the array: the partial-sum kernels write every chunk slot they read
// before reading it, so nothing needs clearing between calls, and two calls
// that overlap on one GPU (different streams, or threads) race on it.
#include <mutex>

#include "sc3d_ctx.h"

using namespace sc3d;

namespace {

// Moments of at most 2^26 particles: chunks of 4096, 21 partial sums each,
// then the 6 + 21 results -- 2.6 MiB per GPU.
constexpr int64_t kNoAllocMaxParticles = int64_t(1) << 26;
constexpr int64_t kNoAllocChunks = kNoAllocMaxParticles / 4096;
constexpr int64_t kNoAllocScratch = kNoAllocChunks * 21 + 27;

static __device__ __align__(256) double scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once).
int scratch_ptr(double** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(SC3D_ERR_HIP, "sc3d_moments_noalloc: no usable HIP device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SC3D_ERR_HIP, std::string("sc3d_moments_noalloc: hipGetSymbolAddress "
                                                  "failed: ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return SC3D_OK;
}

}  // namespace

extern "C" int sc3d_moments_noalloc(const sc3d_bunch* b, double mean[6], double mom2[36],
                                    void* stream)
{
    if (!b || !mean || !mom2) return fail(SC3D_ERR_ARG, "sc3d_moments_noalloc: null argument");
    const int64_t n = b->n_active;
    if (n < 0 || b->stride < n || (n > 0 && (!b->parts || !b->masks)))
        return fail(SC3D_ERR_ARG, "sc3d_moments_noalloc: bad bunch");
    if (n > kNoAllocMaxParticles)
        return fail(SC3D_ERR_SIZE, "sc3d_moments_noalloc: more than 2^26 active particles (" +
                                       std::to_string(n) + "); use sc3d_moments");
    double* P = nullptr;
    if (int rc = scratch_ptr(&P)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* out = P + kNoAllocChunks * 21;  // 27 results after the partials
    launch_moments(b->parts, b->stride, b->masks, n, b->total_num, P, out, st);
    if (hipError_t e = hipGetLastError(); e != hipSuccess)
        return fail(SC3D_ERR_HIP, std::string("sc3d_moments_noalloc: launch failed: ") +
                                      hipGetErrorString(e));
    double h[27];
    if (hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(SC3D_ERR_HIP, "sc3d_moments_noalloc: copy of the results failed");
    for (int j = 0; j < 6; ++j) mean[j] = h[j];
    int q = 6;
    for (int j = 0; j < 6; ++j)
        for (int k = 0; k <= j; ++k) {
            mom2[j * 6 + k] = h[q];
            mom2[k * 6 + j] = h[q];
            ++q;
        }
    return SC3D_OK;
}
