// tcr_search_noalloc: the search of tcr_search_launch + tcr_search_finish
// with its workspace in a file-scope device array instead of a
// caller-provided buffer, for callers that cannot allocate device memory.
//
// `scratch` is one array per GPU the library runs on (the code object's
// device data), held until the process ends.  In host code its name is only
// a stand-in for the symbol, never a device address: the address each device
// holds is looked up once per device with hipGetSymbolAddress and kept in
// g_scratch_ptr; only that pointer is handed to kernels.  Every call on a GPU
// shares that GPU's array and leaves its contents behind; calls that overlap
// on one GPU (two host threads, or two streams) race on it.
#include <mutex>

#include "tcr_internal.h"

namespace tcr {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (looked up on
// the first call per device).
static int scratch_address(void **out) {
    int dev = 0;
    TCR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal beyond " + std::to_string(kMaxDevices));
        return TCR_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return TCR_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TCR_OK;
}

}  // namespace tcr

using namespace tcr;

extern "C" {

int tcr_search_noalloc(const tcr_cloud *clouds, int32_t n_clouds, const float *qpos, const float *qsup,
                       const float *rpos, const float *rsup, int64_t *offsets, void *row, void *col,
                       int32_t index_bytes, int64_t capacity, int64_t *total_pairs, void *stream) {
    PairOut out;
    if (!pair_out(row, col, index_bytes, out)) return TCR_ERR_ARG;
    if (capacity < 0) {
        set_error("capacity must be >= 0");
        return TCR_ERR_ARG;
    }
    const size_t need = tcr_search_workspace_size(clouds, n_clouds);
    if (need == 0) return TCR_ERR_ARG;  // the batch was rejected; the message is set
    if (need > kScratchBytes) {
        set_error("tcr_search_noalloc: this search needs " + std::to_string(need) +
                  " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                  " bytes of its device scratch array; use tcr_search_launch with a workspace");
        return TCR_ERR_SPACE;
    }
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc) return rc;
    return search_complete(clouds, n_clouds, qpos, qsup, rpos, rsup, ws, kScratchBytes, offsets, out, capacity,
                           total_pairs, (hipStream_t)stream);
}

}  // extern "C"
