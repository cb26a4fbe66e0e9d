// Host runtime of the host build (tests/hostbuild): the HIP API subset the library uses, with kernel launches run
// by host threads, synchronous streams, graphs replayed in capture order, and optional checks that device and
// host memory are not confused (EMU_MMAP=1).
//
// EMU_DEFER=1: streams are queues that run only when something waits on them (a stream or event
// synchronisation, a wait of another stream on one of their events, hipFree, hipHostFree,
// hipDeviceSynchronize). Work on two streams then runs in the latest order their events allow, so a read of
// the other stream's results without a wait, a write of data the other stream has still to read, and a host
// read of device results before a synchronisation show up as wrong results (or, through EMU_MMAP and the
// sanitizers, as faults). A graph's nodes of the capturing stream go to the stream it is launched on, those of
// other streams to queues of their own, joined at the end of the launch. Host-to-device copies from pageable
// memory take the source when they are issued, as HIP does.
#include "hip/hip_runtime.h"

#include <barrier>
#include <execinfo.h>
#include <map>
#include <sys/mman.h>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <vector>

thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;

// queued work of a stream: 0 run f, 1 event record (nothing to run), 2 wait for stream ws to reach wseq
struct EmuOp {
    int kind;
    std::function<void()> f;
    EmuStream* ws;
    uint64_t wseq;
};
struct EmuStream {
    int id;
    std::deque<EmuOp> q;
    uint64_t issued = 0, done = 0;
};
// the record that a wait refers to: stream s once it has done seq operations
struct EmuEvent {
    int id;
    EmuStream* s = nullptr;
    uint64_t seq = 0;
};
struct EmuNode {
    int kind;
    std::function<void()> f;
    hipStream_t s;
    hipEvent_t ev;
};
struct EmuGraph {
    std::vector<EmuNode> nodes;
    hipStream_t origin = nullptr;
    std::map<hipStream_t, EmuStream*> branch;  // per launched graph: queues of the other streams' nodes
};

namespace {
int fill_byte() {
    static int v = [] {
        char const* e = std::getenv("EMU_FILL");
        return e ? int(std::strtol(e, nullptr, 0)) : 0xA5;
    }();
    return v;
}
std::mutex cap_mu;
EmuGraph* capturing = nullptr;
std::thread::id capture_thread;
bool defer() {
    static bool v = [] {
        char const* e = std::getenv("EMU_DEFER");
        return e && std::atoi(e) != 0;
    }();
    return v;
}
std::recursive_mutex q_mu;
std::vector<EmuStream*> all_streams;
void flush(EmuStream* s, uint64_t upto);
void exec_op(EmuOp& op) {
    if (op.kind == 0) op.f();
    if (op.kind == 2 && op.ws) flush(op.ws, op.wseq);
}
// run stream s until it has done `upto` operations
void flush(EmuStream* s, uint64_t upto) {
    std::lock_guard<std::recursive_mutex> l(q_mu);
    while (s->done < upto && !s->q.empty()) {
        EmuOp op = std::move(s->q.front());
        s->q.pop_front();
        exec_op(op);
        ++s->done;
    }
}
void flush_all() {
    std::lock_guard<std::recursive_mutex> l(q_mu);
    for (bool more = true; more;) {
        more = false;
        for (size_t i = 0; i < all_streams.size(); ++i) {
            EmuStream* s = all_streams[i];
            if (s->done < s->issued) {
                flush(s, s->issued);
                more = true;
            }
        }
    }
}
// issue an operation on stream s (not under capture): queued with EMU_DEFER, else run at once
void issue(hipStream_t s, int kind, std::function<void()> f, hipEvent_t ev) {
    std::lock_guard<std::recursive_mutex> l(q_mu);
    EmuOp op{kind, std::move(f), nullptr, 0};
    if (kind == 1 && ev) {
        ev->s = s;
        ev->seq = s ? s->issued + 1 : 0;
    }
    if (kind == 2 && ev) {
        op.ws = ev->s;
        op.wseq = ev->seq;
    }
    if (!s) {
        exec_op(op);
        return;
    }
    ++s->issued;
    if (defer()) {
        s->q.push_back(std::move(op));
        return;
    }
    exec_op(op);
    ++s->done;
}
// issue, or add to the graph being captured by this thread
void submit(hipStream_t s, int kind, std::function<void()> f, hipEvent_t ev = nullptr) {
    {
        std::lock_guard<std::mutex> l(cap_mu);
        if (capturing && std::this_thread::get_id() == capture_thread) {
            capturing->nodes.push_back(EmuNode{kind, std::move(f), s, ev});
            return;
        }
    }
    issue(s, kind, std::move(f), ev);
}
// live objects of the HIP API, and their peaks (EMU_COUNT=1: reported at exit)
struct Live {
    long now = 0, peak = 0;
    void up() { peak = std::max(peak, ++now); }
    void down() { --now; }
};
Live live_stream, live_event, live_exec, live_dev, live_pinned;
struct LiveReport {
    ~LiveReport() {
        if (!std::getenv("EMU_COUNT")) return;
        std::fprintf(stderr, "[emu] live at exit (peak): streams %ld (%ld) events %ld (%ld) graph execs %ld (%ld) "
                     "device allocations %ld (%ld) pinned allocations %ld (%ld)\n", live_stream.now, live_stream.peak,
                     live_event.now, live_event.peak, live_exec.now, live_exec.peak, live_dev.now, live_dev.peak,
                     live_pinned.now, live_pinned.peak);
    }
} live_report;
bool capturing_here() {
    std::lock_guard<std::mutex> l(cap_mu);
    return capturing && std::this_thread::get_id() == capture_thread;
}

// a team of host threads that runs one block at a time
struct Team {
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    unsigned long long gen = 0;
    int active = 0, remaining = 0;
    std::function<void()> const* fn = nullptr;
    dim3 bidx, bdim, gdim;
    std::barrier<>* bar = nullptr;
    bool stop = false;
    void worker(int i) {
        unsigned long long seen = 0;
        while (true) {
            std::unique_lock<std::mutex> l(mu);
            cv.wait(l, [&] { return stop || gen != seen; });
            if (stop) return;
            seen = gen;
            if (i >= active) continue;
            l.unlock();
            unsigned const nx = bdim.x, ny = bdim.y;
            threadIdx = dim3(unsigned(i) % nx, (unsigned(i) / nx) % ny, unsigned(i) / (nx * ny));
            blockIdx = bidx;
            blockDim = bdim;
            gridDim = gdim;
            (*fn)();
            bar->arrive_and_drop();
            l.lock();
            if (--remaining == 0) done_cv.notify_all();
        }
    }
    void ensure(int n) {
        while (int(th.size()) < n) {
            int const i = int(th.size());
            th.emplace_back([this, i] { worker(i); });
        }
    }
    ~Team() {
        {
            std::lock_guard<std::mutex> l(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : th) t.join();
    }
};
Team& team() {
    static Team t;
    return t;
}
std::barrier<>* cur_bar = nullptr;
std::mutex launch_mu;

// __shfl_xor exchange buffer of the running block
std::vector<int> shfl_buf;

void run_grid(dim3 grid, dim3 block, std::function<void()> const& body) {
    std::lock_guard<std::mutex> lg(launch_mu);
    int const nt = int(block.x * block.y * block.z);
    Team& T = team();
    T.ensure(nt);
    shfl_buf.assign(size_t(nt), 0);
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                std::barrier<> bar(nt);
                cur_bar = &bar;
                std::unique_lock<std::mutex> l(T.mu);
                T.fn = &body;
                T.bidx = dim3(bx, by, bz);
                T.bdim = block;
                T.gdim = grid;
                T.bar = &bar;
                T.active = nt;
                T.remaining = nt;
                ++T.gen;
                T.cv.notify_all();
                T.done_cv.wait(l, [&] { return T.remaining == 0; });
            }
}
}  // namespace

void emu_bad_pointer(char const* what, void const* p) {
    std::fprintf(stderr, "[emu] %s %p is not device memory\n", what, p);
    void* bt[32];
    int const n = backtrace(bt, 32);
    backtrace_symbols_fd(bt, n, 2);
    std::abort();
}
void __syncthreads() { cur_bar->arrive_and_wait(); }
int __shfl_xor(int v, int mask) {
    int const i = int(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    shfl_buf[size_t(i)] = v;
    cur_bar->arrive_and_wait();
    int const j = (i & ~63) | ((i & 63) ^ mask);
    int const r = shfl_buf[size_t(j)];
    cur_bar->arrive_and_wait();
    return r;
}

void emu_launch(dim3 grid, dim3 block, hipStream_t st, std::function<void()> body) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) {
        std::fprintf(stderr, "[emu] launch with an empty grid\n");
        std::abort();
    }
    if (block.x * block.y * block.z > 1024) {
        std::fprintf(stderr, "[emu] block too large\n");
        std::abort();
    }
    submit(st, 0, [grid, block, body = std::move(body)] { run_grid(grid, block, body); });
}

// EMU_MMAP=1: device allocations from a separate address region (page-padded, guard page after), so that a
// host pointer used as a device pointer is recognised
namespace {
std::mutex alloc_mu;
std::map<uintptr_t, size_t> dev_allocs;  // start -> size (device allocations)
std::map<uintptr_t, size_t> host_allocs;  // pinned host allocations
bool use_mmap() {
    static bool v = std::getenv("EMU_MMAP") != nullptr;
    return v;
}
uintptr_t next_region = 0x600000000000ull;
}  // namespace
bool emu_is_device(void const* p) {
    uintptr_t const a = reinterpret_cast<uintptr_t>(p);
    std::lock_guard<std::mutex> l(alloc_mu);
    auto it = dev_allocs.upper_bound(a);
    if (it == dev_allocs.begin()) return false;
    --it;
    return a >= it->first && a <= it->first + it->second;
}
static void* dev_alloc_raw(size_t n) {
    if (!use_mmap()) return std::malloc(n ? n : 1);
    size_t const pg = 4096, len = (n + pg - 1) / pg * pg + pg;
    std::lock_guard<std::mutex> l(alloc_mu);
    void* m = mmap(reinterpret_cast<void*>(next_region), len, PROT_READ | PROT_WRITE,
                   MAP_PRIVATE | MAP_ANONYMOUS | MAP_FIXED_NOREPLACE, -1, 0);
    if (m == MAP_FAILED) return nullptr;
    next_region += len + pg;
    mprotect(static_cast<char*>(m) + len - pg, pg, PROT_NONE);
    return m;
}
bool emu_use_checks() { return use_mmap(); }
hipError_t hipMalloc(void** p, size_t n) {
    void* m = dev_alloc_raw(n);
    if (!m) return hipErrorOutOfMemory;
    std::memset(m, fill_byte(), n);
    {
        std::lock_guard<std::mutex> l(alloc_mu);
        dev_allocs[reinterpret_cast<uintptr_t>(m)] = n;
        live_dev.up();
    }
    *p = m;
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    if (!p) return hipSuccess;
    flush_all();
    size_t n = 0;
    {
        std::lock_guard<std::mutex> l(alloc_mu);
        auto it = dev_allocs.find(reinterpret_cast<uintptr_t>(p));
        if (it == dev_allocs.end()) {
            std::fprintf(stderr, "[emu] hipFree of an unknown pointer %p\n", p);
            std::abort();
        }
        n = it->second;
        dev_allocs.erase(it);
        live_dev.down();
    }
    if (use_mmap()) {
        size_t const pg = 4096, len = (n + pg - 1) / pg * pg + pg;
        munmap(p, len);  // later use faults
    } else {
        std::free(p);
    }
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) {
    *p = std::malloc(n ? n : 1);
    std::lock_guard<std::mutex> l(alloc_mu);
    host_allocs[reinterpret_cast<uintptr_t>(*p)] = n;
    live_pinned.up();
    return hipSuccess;
}
hipError_t hipHostFree(void* p) {
    flush_all();
    {
        std::lock_guard<std::mutex> l(alloc_mu);
        if (host_allocs.erase(reinterpret_cast<uintptr_t>(p))) live_pinned.down();
    }
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, void const* s, size_t n, hipMemcpyKind k) {
    if (use_mmap() && n) {
        bool const dd = emu_is_device(d), sd = emu_is_device(s);
        bool const ok = (k == hipMemcpyHostToDevice && dd && !sd) || (k == hipMemcpyDeviceToHost && !dd && sd) ||
                        (k == hipMemcpyDeviceToDevice && dd && sd) || k == hipMemcpyDefault ||
                        (k == hipMemcpyHostToHost && !dd && !sd);
        if (!ok) {
            std::fprintf(stderr, "[emu] hipMemcpy kind %d with dst %s src %s\n", int(k), dd ? "device" : "host",
                         sd ? "device" : "host");
            std::abort();
        }
    }
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
static bool emu_is_pinned(void const* p) {
    uintptr_t const a = reinterpret_cast<uintptr_t>(p);
    std::lock_guard<std::mutex> l(alloc_mu);
    auto it = host_allocs.upper_bound(a);
    if (it == host_allocs.begin()) return false;
    --it;
    return a >= it->first && a < it->first + std::max<size_t>(it->second, 1);
}
hipError_t hipMemcpyAsync(void* d, void const* s, size_t n, hipMemcpyKind k, hipStream_t st) {
    if (defer() && !capturing_here() && n && !emu_is_device(s) && !emu_is_pinned(s)) {
        // a pageable source is taken when the copy is issued
        auto src = std::make_shared<std::vector<char>>(static_cast<char const*>(s), static_cast<char const*>(s) + n);
        submit(st, 0, [=] { hipMemcpy(d, src->data(), n, k == hipMemcpyDefault ? k : hipMemcpyHostToDevice); });
        return hipSuccess;
    }
    submit(st, 0, [=] { hipMemcpy(d, s, n, k); });
    return hipSuccess;
}
hipError_t hipMemcpy2D(void* d, size_t dp, void const* s, size_t sp, size_t w, size_t h, hipMemcpyKind) {
    for (size_t r = 0; r < h; ++r) std::memmove(static_cast<char*>(d) + r * dp, static_cast<char const*>(s) + r * sp, w);
    return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) {
    if (use_mmap() && n && !emu_is_device(d)) {
        std::fprintf(stderr, "[emu] hipMemset of host memory\n");
        std::abort();
    }
    if (n) std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) {
    if (capturing_here()) std::fprintf(stderr, "[emu] memset recorded into a graph\n");
    submit(st, 0, [=] { hipMemset(d, v, n); });
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    std::lock_guard<std::recursive_mutex> l(q_mu);
    *s = new EmuStream{1};
    all_streams.push_back(*s);
    live_stream.up();
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) {
    // the object stays, since waits may refer to it; its queued work runs now, or (EMU_DEFER=2) at the next
    // device-wide synchronisation, as a destroyed stream's work does on the device
    static bool const late = [] {
        char const* e = std::getenv("EMU_DEFER");
        return e && std::atoi(e) >= 2;
    }();
    if (s && !late) flush(s, s->issued);
    live_stream.down();
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    if (capturing_here()) {
        std::fprintf(stderr, "[emu] synchronize during capture\n");
        return hipErrorUnknown;
    }
    if (s)
        flush(s, s->issued);
    else
        flush_all();
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    submit(s, 2, nullptr, e);
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new EmuEvent{1};
    live_event.up();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t) {
    // kept: nodes of graphs may still name it
    live_event.down();
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    submit(s, 1, nullptr, e);
    return hipSuccess;
}
hipError_t hipEventRecordWithFlags(hipEvent_t e, hipStream_t s, unsigned) { return hipEventRecord(e, s); }
hipError_t hipEventSynchronize(hipEvent_t e) {
    if (e && e->s) flush(e->s, e->seq);
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
    *ms = 0.001f;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    std::lock_guard<std::mutex> l(cap_mu);
    if (capturing) return hipErrorUnknown;
    capturing = new EmuGraph;
    capturing->origin = s;
    capture_thread = std::this_thread::get_id();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    std::lock_guard<std::mutex> l(cap_mu);
    *g = capturing;
    capturing = nullptr;
    return *g ? hipSuccess : hipErrorUnknown;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new EmuGraph(*g);
    live_exec.up();
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> l(q_mu);
    for (auto const& n : e->nodes) {
        hipStream_t t = s;
        if (defer() && n.s != e->origin) {
            auto it = e->branch.find(n.s);
            if (it == e->branch.end()) {
                hipStream_t b;
                hipStreamCreateWithFlags(&b, 0);
                it = e->branch.emplace(n.s, b).first;
            }
            t = it->second;
        }
        issue(t, n.kind, n.f, n.ev);
    }
    // the launch is complete once every branch is
    for (auto const& [k, b] : e->branch) {
        EmuEvent end{0, b, b->issued};
        issue(s, 2, nullptr, &end);
    }
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    flush_all();
    live_exec.down();
    for (auto const& kv : e->branch) {
        (void)kv;
        live_stream.down();
    }
    delete e;
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipPeekAtLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emulated error"; }
hipError_t hipDeviceSynchronize() {
    flush_all();
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    {
        std::lock_guard<std::mutex> l(alloc_mu);
        dev_allocs[reinterpret_cast<uintptr_t>(sym)] = size_t(64) << 20;  // the device symbol's region
    }
    return hipSuccess;
}
