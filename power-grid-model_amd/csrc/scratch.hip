// scratch.hip -- the fixed device region of the allocation-free entry point.
//
// PGM_amd_block_lu_solve_noalloc keeps everything it needs on the device (structure, input blocks, factors,
// forward vectors, permutations, per-system state) in this file-scope array instead of hipMalloc'd memory.
// Every GPU the library runs on holds its own copy of the array until the process ends. Host code never
// uses the array's name as an address: the device address of each GPU's copy comes from
// hipGetSymbolAddress, called once per device with that device current, and is kept per device.
#include "kernels.hpp"

#include <cstdio>
#include <mutex>

namespace pgm_amd {

constexpr size_t kScratchBytes = size_t(64) << 20;

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_addr[kMaxDevices] = {};
}  // namespace

int scratch_region(void** ptr, size_t* bytes) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_addr[dev]) {
        void* p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(e));
            return -1;
        }
        scratch_addr[dev] = p;
    }
    *ptr = scratch_addr[dev];
    *bytes = kScratchBytes;
    return 0;
}

}  // namespace pgm_amd
