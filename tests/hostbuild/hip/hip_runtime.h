// Host build of the HIP API subset the library uses (tests/hostbuild): kernels are host functions, one host
// thread per GPU thread of a block; streams synchronous, graphs replayed in capture order.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorUnknown = 999 };
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeRelaxed = 2 };
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipEventRecordExternal 1u

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct EmuStream;
struct EmuEvent;
struct EmuGraph;
typedef EmuStream* hipStream_t;
typedef EmuEvent* hipEvent_t;
typedef EmuGraph* hipGraph_t;
typedef EmuGraph* hipGraphExec_t;
typedef void* hipGraphNode_t;

extern thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;

#define __global__
#define __device__
#define __host__
#define __constant__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...)
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) ((const void*)&(x))

// runtime
hipError_t hipMalloc(void** p, size_t n);
template <class T> hipError_t hipMalloc(T** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags = 0);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, void const* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, void const* s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemcpy2D(void* d, size_t dp, void const* s, size_t sp, size_t w, size_t h, hipMemcpyKind k);
hipError_t hipMemset(void* d, int v, size_t n);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventRecordWithFlags(hipEvent_t e, hipStream_t s, unsigned flags);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* n, char* log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipGetDeviceCount(int* n);
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
hipError_t hipGetSymbolAddress(void** p, const void* sym);

// kernel launch: the arguments are converted to the kernel's parameter types and captured by value
void emu_launch(dim3 grid, dim3 block, hipStream_t st, std::function<void()> body);
bool emu_use_checks();
bool emu_is_device(void const* p);
void emu_bad_pointer(char const* what, void const* p);
// a kernel argument that is a pointer must address device memory; descriptor structs (classes of 64 bytes or
// more) are scanned for words that look like user-space pointers outside device memory
template <class T> void emu_check_arg(T const& v) {
    if constexpr (std::is_pointer_v<T>) {
        if (v == nullptr || !emu_use_checks()) return;
        if (!emu_is_device(v)) emu_bad_pointer("kernel pointer argument", v);
        using E = std::remove_cv_t<std::remove_pointer_t<T>>;
        if constexpr (std::is_class_v<E> && sizeof(E) >= 64) {
            auto const* w = reinterpret_cast<unsigned long long const*>(v);
            for (size_t i = 0; i < sizeof(E) / 8; ++i) {
                unsigned long long const x = w[i];
                if (x >= 0x500000000000ull && x < 0x800000000000ull && (x & 3) == 0 && (x & 0xffffffffull) >= 0x10000ull &&
                    !emu_is_device(reinterpret_cast<void const*>(x)))
                    emu_bad_pointer("descriptor word", reinterpret_cast<void const*>(x));
            }
        }
    } else if constexpr (std::is_class_v<T> && sizeof(T) >= 64) {
        if (!emu_use_checks()) return;
        auto const* w = reinterpret_cast<unsigned long long const*>(&v);
        for (size_t i = 0; i < sizeof(T) / 8; ++i) {
            unsigned long long const x = w[i];
            if (x >= 0x500000000000ull && x < 0x800000000000ull && (x & 3) == 0 && (x & 0xffffffffull) >= 0x10000ull &&
                !emu_is_device(reinterpret_cast<void const*>(x)))
                emu_bad_pointer("struct argument word", reinterpret_cast<void const*>(x));
        }
    }
}
bool emu_use_checks();
template <class... P, class... A>
void emu_launch_kernel(void (*k)(P...), dim3 grid, dim3 block, size_t, hipStream_t st, A&&... a) {
    (emu_check_arg(static_cast<std::decay_t<P>>(a)), ...);
    emu_launch(grid, block, st, [k, args = std::make_tuple(static_cast<std::decay_t<P>>(a)...)]() { std::apply(k, args); });
}
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) emu_launch_kernel(K, dim3(G), dim3(B), S, ST, ##__VA_ARGS__)

// device builtins
void __syncthreads();
int __shfl_xor(int v, int mask);
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }  // wave-uniform arguments only
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long v) { double r; std::memcpy(&r, &v, 8); return r; }
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline T emu_atomic_max(T* p, T v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (o < v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return o;
}
inline int atomicMax(int* p, int v) { return emu_atomic_max(p, v); }
inline unsigned atomicMax(unsigned* p, unsigned v) { return emu_atomic_max(p, v); }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) { return emu_atomic_max(p, v); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicOr(unsigned* p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
using std::isfinite;
using std::isinf;
using std::isnan;
struct alignas(16) uint4 {
    unsigned x, y, z, w;
};
