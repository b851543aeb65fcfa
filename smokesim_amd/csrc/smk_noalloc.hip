// smk_noalloc.hip -- the fixed staging array of smk_download_field_noalloc.
//
// smk_download_field converts a field to host values in a device buffer it
// allocates for every call.  The no-allocation variant stages through this
// file-scope array instead: each GPU holds its own copy (48 MiB) from the
// library's first use of it on that device until the process ends.  Host
// code never passes the array's name to a kernel or a copy (on the host it
// is only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device and kept in a table.
#include <hip/hip_runtime.h>

#include <mutex>

#include "smk_kernels.h"

namespace smk {

// [0, kNoAllocMaxValues) doubles: the field as doubles (the conversion
// kernels of smk_download_field write there); after them, room for the same
// count of floats: the FP32 output.
constexpr size_t kNoAllocStageBytes = kNoAllocMaxValues * sizeof(double);
constexpr size_t kNoAllocScratchBytes = kNoAllocStageBytes + kNoAllocMaxValues * sizeof(float);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_addr_mu;
unsigned char *g_addr[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];

// FP64 staging -> the requested output type; alpha_off is added to every
// fourth value (the BF16 smoke temperature offset) in double, as the
// allocating download does on the host before its float conversion.
__global__ void k_stage_to_out(const double *src, void *dst, int out_f32, size_t n, double alpha_off) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = src[i];
        if (alpha_off != 0.0 && (i & 3) == 3) v = v + alpha_off;
        if (out_f32) static_cast<float *>(dst)[i] = (float)v;
        else static_cast<double *>(dst)[i] = v;
    }
}
}  // namespace

hipError_t noalloc_scratch(int device, unsigned char **ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_addr_mu);
    if (!g_addr[device]) {
        // the caller has made `device` current: this is its copy's address
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_addr[device] = static_cast<unsigned char *>(p);
    }
    *ptr = g_addr[device];
    return hipSuccess;
}

std::mutex &noalloc_lock(int device) { return g_call_mu[device]; }

size_t noalloc_stage_bytes() { return kNoAllocStageBytes; }

void launch_stage_to_out(hipStream_t s, const double *src, void *dst, int out_f32, size_t n, double alpha_off) {
    size_t b = (n + 255) / 256;
    unsigned blocks = (unsigned)(b > 65535 * 4 ? 65535 * 4 : (b == 0 ? 1 : b));
    k_stage_to_out<<<blocks, 256, 0, s>>>(src, dst, out_f32, n, alpha_off);
}

}  // namespace smk
