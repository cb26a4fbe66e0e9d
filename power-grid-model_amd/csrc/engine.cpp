// engine.cpp -- device-resident batch: grouping of scenarios by topology, upload, the Newton-Raphson
// iteration loop over level-scheduled sweeps, output records and per-scenario error reporting.
//
// Batch semantics follow the reference (job_dispatch.hpp:35-68, 123-224): every scenario is the base
// model with that scenario's update applied; a scenario whose update changes a branch or source status
// gets its own topology (topology.hpp via build_topology here); errors are collected per scenario and
// reported as "Error in batch #i: <message>". Single calculations report a regular error.
#include "engine.hpp"

#include "../../include/power_grid_model_amd.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <numeric>
#include <cstdlib>
#include <limits>
#include <sstream>
#include <thread>
#include <tuple>

namespace pgm_amd {

#define PGM_HIP_CHECK(expr)                                                                                      \
    do {                                                                                                         \
        hipError_t const e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                                    \
            throw PowerGridError(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_) + "\n");    \
    } while (0)

ChunkPool::~ChunkPool() {
    for (auto const& c : free_) (void)hipFree(c.first);
}
std::pair<void*, size_t> ChunkPool::get(size_t bytes) {
    size_t best = free_.size();
    for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].second >= bytes && (best == free_.size() || free_[i].second < free_[best].second)) best = i;
    if (best == free_.size()) return {nullptr, 0};
    auto const c = free_[best];
    free_.erase(free_.begin() + long(best));
    return c;
}

UploadCache::~UploadCache() {
    for (auto const& [k, p] : dev_) (void)hipFree(p);
}
void const* UploadCache::get(void const* host, size_t bytes) {
    auto const key = std::make_pair(host, bytes);
    auto it = dev_.find(key);
    if (it != dev_.end()) return it->second;
    void* p = nullptr;
    PGM_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 1)));
    hipError_t const e = hipMemcpy(p, host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p);
        throw PowerGridError(std::string("HIP error in upload: ") + hipGetErrorString(e) + "\n");
    }
    dev_.emplace(key, p);
    return p;
}

// device-to-host copy of a large result into pageable caller memory: T host threads, each with its own
// pinned staging buffer and stream, take the chunks in turn (DMA into the staging buffer, then a host
// copy out of it), so the DMA of one chunk overlaps the host copies of the others. The staging buffers
// and streams are made on first use and kept for the process.
namespace {
constexpr int kStageThreads = 8;
constexpr size_t kStageBytes = size_t(32) << 20;
struct Staging {
    std::mutex mu;
    char* buf[kStageThreads] = {};
    hipStream_t st[kStageThreads] = {};
    bool ready = false;
};
Staging& staging() {
    static Staging s;
    return s;
}
}  // namespace

void copy_to_host(void* dst, void const* src, size_t bytes) {
    if (bytes < 4 * kStageBytes) {
        PGM_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return;
    }
    Staging& S = staging();
    std::lock_guard<std::mutex> lock(S.mu);
    if (!S.ready) {
        for (int t = 0; t < kStageThreads; ++t) {
            PGM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&S.buf[t]), kStageBytes));
            PGM_HIP_CHECK(hipStreamCreateWithFlags(&S.st[t], hipStreamNonBlocking));
        }
        S.ready = true;
    }
    size_t const n_chunks = (bytes + kStageBytes - 1) / kStageBytes;
    std::vector<hipError_t> errs(kStageThreads, hipSuccess);
    auto work = [&](int t) {
        for (size_t c = size_t(t); c < n_chunks; c += kStageThreads) {
            size_t const off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
            hipError_t e = hipMemcpyAsync(S.buf[t], static_cast<char const*>(src) + off, len, hipMemcpyDeviceToHost,
                                          S.st[t]);
            if (e == hipSuccess) e = hipStreamSynchronize(S.st[t]);
            if (e != hipSuccess) {
                errs[size_t(t)] = e;
                return;
            }
            std::memcpy(static_cast<char*>(dst) + off, S.buf[t], len);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < kStageThreads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (auto e : errs)
        if (e != hipSuccess) throw PowerGridError(std::string("HIP error in device-to-host copy: ") + hipGetErrorString(e) + "\n");
}

void const* UploadCache::find(void const* host, size_t bytes) const {
    auto it = dev_.find(std::make_pair(host, bytes));
    return it == dev_.end() ? nullptr : it->second;
}

Arena::~Arena() {
    for (auto const& c : chunks_) {
        if (pool_)
            pool_->put(c.first, c.second);
        else
            (void)hipFree(c.first);
    }
}
void* Arena::take(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    if (bytes == 0) return nullptr;
    if (bytes > left_) {
        size_t chunk = std::max<size_t>(bytes, size_t(256) << 20);
        void* p = nullptr;
        if (pool_) std::tie(p, chunk) = pool_->get(chunk);
        if (!p) {
            chunk = std::max<size_t>(bytes, size_t(256) << 20);
            PGM_HIP_CHECK(hipMalloc(&p, chunk));
        }
        chunks_.emplace_back(p, chunk);
        cur_ = static_cast<char*>(p);
        left_ = chunk;
        reserved_ += chunk;
    }
    void* r = cur_;
    cur_ += bytes;
    left_ -= bytes;
    return r;
}

void ensure_device() {
    int n = 0;
    hipError_t const e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        throw PowerGridError("No HIP device available: this library runs the power flow on an MI355X (gfx950) GPU\n");
}

std::string format_double(double x) {
    char buf[64];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, x);
        if (std::strtod(buf, nullptr) == x) break;
    }
    return buf;
}

namespace {

char const* const SparseMatrixErrorMessage =
    "Sparse matrix error, possibly singular matrix!\n"
    "If you get this error from state estimation, it might mean the system is not fully observable, i.e. not "
    "enough measurements.\n"
    "It might also mean that you are running into a corner case where PGM cannot resolve yet.\n"
    "See https://github.com/PowerGridModel/power-grid-model/issues/864.";


template <class T, class A> T* dev_alloc(A& a, size_t count) { return static_cast<T*>(a.take(count * sizeof(T))); }
// blocking copy into the arena (the host vectors are often temporaries); nullptr for an empty vector
template <class T, class A> T* put(A& a, std::vector<T> const& h) {
    if (h.empty()) return nullptr;
    T* d = dev_alloc<T>(a, h.size());
    PGM_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}
std::vector<double> complex_to_doubles(std::vector<cd> const& v) {
    std::vector<double> r(v.size() * 2);
    for (size_t i = 0; i < v.size(); ++i) {
        r[2 * i] = v[i].real();
        r[2 * i + 1] = v[i].imag();
    }
    return r;
}

bool set_status(IntS& tgt, IntS v) {
    if (v == na_IntS) return false;
    bool const nv = v != 0;
    bool const ch = (tgt != 0) != nv;
    tgt = nv;
    return ch;
}
bool set_double(double& tgt, double v) {
    if (std::isnan(v)) return false;
    tgt = v;
    return true;
}
}  // namespace

// component update semantics: branch.hpp:162-183, transformer.hpp:104-123, shunt.hpp:38-48,
// source.hpp:66-78, load_gen.hpp:96-110 (NaN / na attributes keep the current value)
bool apply_update_record(Model const& m, ComponentState& st, int comp, Idx i, void const* rec, bool& structural) {
    bool changed = false;
    switch (comp) {
    case C_LINE: {
        BranchUpdate u;
        std::memcpy(&u, rec, sizeof u);
        bool const a = set_status(st.line_fs[i], u.from_status);
        bool const b = set_status(st.line_ts[i], u.to_status);
        changed = a || b;
        structural |= changed;
        break;
    }
    case C_TRANSFORMER: {
        TransformerUpdate u;
        std::memcpy(&u, rec, sizeof u);
        bool const a = set_status(st.tr_fs[i], u.from_status);
        bool const b = set_status(st.tr_ts[i], u.to_status);
        changed = a || b;
        if (u.tap_pos != na_IntS && u.tap_pos != st.tr_tap[i]) {
            st.tr_tap[i] = tap_limit(m.tstat[i], u.tap_pos);
            changed = true;
        }
        structural |= changed;
        break;
    }
    case C_SHUNT: {
        ShuntUpdate u;
        std::memcpy(&u, rec, sizeof u);
        changed = set_status(st.sh_status[i], u.status);
        changed = set_double(st.sh_g1[i], u.g1) || changed;
        changed = set_double(st.sh_b1[i], u.b1) || changed;
        changed = set_double(st.sh_g0[i], u.g0) || changed;
        changed = set_double(st.sh_b0[i], u.b0) || changed;
        structural |= changed;
        break;
    }
    case C_SOURCE: {
        SourceUpdate u;
        std::memcpy(&u, rec, sizeof u);
        structural |= set_status(st.src_status[i], u.status);
        set_double(st.src_uref[i], u.u_ref);
        set_double(st.src_uang[i], u.u_ref_angle);
        set_double(st.src_sk[i], u.sk);
        set_double(st.src_rx[i], u.rx_ratio);
        set_double(st.src_z01[i], u.z01_ratio);
        break;
    }
    case C_SYM_GEN:
    case C_SYM_LOAD: {
        SymLoadGenUpdate u;
        std::memcpy(&u, rec, sizeof u);
        Idx const l = m.lg_offset[comp == C_SYM_GEN ? 0 : 2] + i;
        set_status(st.lg_status[l], u.status);
        set_double(st.lg_p[l][0], u.p_specified);
        set_double(st.lg_q[l][0], u.q_specified);
        break;
    }
    case C_ASYM_GEN:
    case C_ASYM_LOAD: {
        AsymLoadGenUpdate u;
        std::memcpy(&u, rec, sizeof u);
        Idx const l = m.lg_offset[comp == C_ASYM_GEN ? 1 : 3] + i;
        set_status(st.lg_status[l], u.status);
        for (int ph = 0; ph < 3; ++ph) {
            set_double(st.lg_p[l][ph], u.p_specified[ph]);
            set_double(st.lg_q[l][ph], u.q_specified[ph]);
        }
        break;
    }
    default:
        throw PowerGridError(std::string("component ") + comp_names[comp] + " cannot be updated");
    }
    return changed;
}

// ------------------------------------------------------------------------------------------------
// one group: one topology (and admittance set), a subset of the scenarios
// ------------------------------------------------------------------------------------------------
struct Group {
    ComponentState st;
    Topology topo;
    Plan plan;
    std::vector<int> scenarios;
    int S = 0, S_pad = 0;
    int n_load = 0, n_src = 0, n_shunt = 0;
    long long pair_off = 0, lane_off = 0;
    DevGroup dg;
    DevOutput dout;
    NodeRec node_rec{};
    BranchRec branch_rec[4]{};  // lines, transformers, links, generic branches
    ApplRec appl_rec[N_COMP]{};
    int8_t* load_status = nullptr;
    std::vector<double> level_entries, level_pivots;  // LU entries of the rows / pivots per level
    // union group: scenarios of several small groups that differ only in branch side statuses. The
    // structure is the one of the union state (a side is closed when it is closed in any member); each
    // member keeps its own statuses for its lanes [lane_begin, lane_end)
    bool is_union = false;
    bool packed = false;  // at most 32 scenarios: packed kernels, several tasks per wave
    struct Member {
        int lane_begin = 0, lane_end = 0;
        std::vector<IntS> line_fs, line_ts, tr_fs, tr_ts;
    };
    std::vector<Member> members;
    // branches open on both sides in the union state that some member closes (coupled into the
    // union's math models; the off-pattern entries they add are corrected per lane)
    std::vector<char> couple;
    std::vector<int> xlanes;  // lanes with extra entries
};

namespace {

// groups with fewer scenarios than this whose branches are only opened against the base state are
// solved as one union group on the base structure (contingency-style batches: one line out per scenario)
constexpr int kUnionBelow = 32;
// groups of at most this many scenarios run on the packed kernels (64 / S tasks of a level per wave)
constexpr int kPackedMax = 32;
// target number of parallel tasks of a packed group's schedule (PGM_AMD_PACKED_TARGET overrides)
int packed_target() {
    static int const t = [] {
        char const* v = std::getenv("PGM_AMD_PACKED_TARGET");
        int const x = v ? std::atoi(v) : 0;
        return x > 0 ? x : 256;
    }();
    return t;
}

// waves a plain group's schedule aims at per level (tasks x 64-scenario blocks); PGM_AMD_PLAIN_WAVES
// overrides
int plain_waves() {
    static int const w = [] {
        char const* v = std::getenv("PGM_AMD_PLAIN_WAVES");
        int const x = v ? std::atoi(v) : 0;
        // about two rounds of resident waves over the chip at level 0: shorter tasks shorten the
        // dependent pivot chains more than the extra levels cost (year_20k: 71.1 ms at 4096, 66.0 ms at 12288)
        return x > 0 ? x : 12288;
    }();
    return w;
}

// largest number of tasks per level of a union group's schedule; PGM_AMD_UNION_TASKS overrides. Union
// groups run on the second stream beside the plain groups. year_20k: the same step time at 256, 1024 and
// 2048; above 256 the plain forward sweep runs faster (1.30 against 1.39 ms per launch) only because the
// union group's sweeps, then longer in total than the plain ones, overlap it less
int union_tasks() {
    static int const t = [] {
        char const* v = std::getenv("PGM_AMD_UNION_TASKS");
        int const x = v ? std::atoi(v) : 0;
        return x > 0 ? x : 256;
    }();
    return t;
}

// the state of a topology group: every component's state but the load_gens' (each scenario's loads are
// applied from the base and its own update rows, never from its group's state), so that a batch of many
// groups does not copy the load arrays once per group
ComponentState structural_state(ComponentState const& st) {
    ComponentState r;
    r.line_fs = st.line_fs;
    r.line_ts = st.line_ts;
    r.tr_fs = st.tr_fs;
    r.tr_ts = st.tr_ts;
    r.tr_tap = st.tr_tap;
    r.sh_status = st.sh_status;
    r.sh_g1 = st.sh_g1;
    r.sh_b1 = st.sh_b1;
    r.sh_g0 = st.sh_g0;
    r.sh_b0 = st.sh_b0;
    r.src_status = st.src_status;
    r.src_uref = st.src_uref;
    r.src_uang = st.src_uang;
    r.src_sk = st.src_sk;
    r.src_rx = st.src_rx;
    r.src_z01 = st.src_z01;
    r.reg_status = st.reg_status;
    r.reg_u_set = st.reg_u_set;
    r.reg_u_band = st.reg_u_band;
    r.reg_ldc_r = st.reg_ldc_r;
    r.reg_ldc_x = st.reg_ldc_x;
    return r;
}

// every branch side of st is open where the base state has it open
bool opens_only(Model const& m, ComponentState const& st) {
    auto le = [](std::vector<IntS> const& a, std::vector<IntS> const& b) {
        for (size_t i = 0; i < a.size(); ++i)
            if (a[i] != 0 && b[i] == 0) return false;
        return true;
    };
    return le(st.line_fs, m.base.line_fs) && le(st.line_ts, m.base.line_ts) && le(st.tr_fs, m.base.tr_fs) &&
           le(st.tr_ts, m.base.tr_ts);
}

// branches with a side closed in st that the base state has open (-1 when more than limit)
int closed_against_base(Model const& m, ComponentState const& st, int limit, std::vector<Idx>* list = nullptr) {
    int cnt = 0;
    Idx const nl = m.n_line();
    for (Idx c = 0; c < m.n_branch(); ++c) {
        bool const f = c < nl ? st.line_fs[c] && !m.base.line_fs[c] : st.tr_fs[c - nl] && !m.base.tr_fs[c - nl];
        bool const t = c < nl ? st.line_ts[c] && !m.base.line_ts[c] : st.tr_ts[c - nl] && !m.base.tr_ts[c - nl];
        if (!f && !t) continue;
        if (++cnt > limit) return -1;
        if (list) list->push_back(c);
    }
    return cnt;
}

std::string structural_key(Model const& m, ComponentState const& st, bool branch_status = true) {
    // the components whose state differs from the base define the group
    std::ostringstream k;
    auto const& b = m.base;
    if (branch_status)
        for (Idx i = 0; i < m.n_line(); ++i)
            if (st.line_fs[i] != b.line_fs[i] || st.line_ts[i] != b.line_ts[i])
                k << 'l' << i << ':' << int(st.line_fs[i]) << int(st.line_ts[i]) << ';';
    for (Idx i = 0; i < Idx(m.trafo.size()); ++i) {
        bool const sd = branch_status && (st.tr_fs[i] != b.tr_fs[i] || st.tr_ts[i] != b.tr_ts[i]);
        if (sd || st.tr_tap[i] != b.tr_tap[i])
            k << 't' << i << ':' << (branch_status ? int(st.tr_fs[i]) : 0) << (branch_status ? int(st.tr_ts[i]) : 0)
              << ',' << int(st.tr_tap[i]) << ';';
    }
    for (Idx i = 0; i < Idx(m.shunt.size()); ++i) {
        double const a[4] = {st.sh_g1[i], st.sh_b1[i], st.sh_g0[i], st.sh_b0[i]};
        double const c[4] = {b.sh_g1[i], b.sh_b1[i], b.sh_g0[i], b.sh_b0[i]};
        if (st.sh_status[i] != b.sh_status[i] || std::memcmp(a, c, sizeof a) != 0) {
            k << 's' << i << ':' << int(st.sh_status[i]);
            for (double x : a) k << ',' << format_double(x);
            k << ';';
        }
    }
    for (Idx i = 0; i < Idx(m.source.size()); ++i)
        if (st.src_status[i] != b.src_status[i]) k << 'c' << i << ':' << int(st.src_status[i]) << ';';
    return k.str();
}

// structural_key of a state that differs from the base at most in the listed components (sorted here)
std::string touched_key(Model const& m, ComponentState const& st, std::vector<std::pair<int, Idx>>& touched) {
    std::sort(touched.begin(), touched.end(), [](auto const& a, auto const& b) {
        // key order: lines, transformers, shunts, sources
        auto rank = [](int c) { return c == C_LINE ? 0 : c == C_TRANSFORMER ? 1 : c == C_SHUNT ? 2 : 3; };
        return std::make_pair(rank(a.first), a.second) < std::make_pair(rank(b.first), b.second);
    });
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    std::string k;
    auto const& b = m.base;
    auto num = [&k](long long v) { k += std::to_string(v); };
    for (auto const& [comp, i] : touched) {
        if (comp == C_LINE) {
            if (st.line_fs[i] == b.line_fs[i] && st.line_ts[i] == b.line_ts[i]) continue;
            k += 'l';
            num(i);
            k += ':';
            num(st.line_fs[i]);
            num(st.line_ts[i]);
            k += ';';
        } else if (comp == C_TRANSFORMER) {
            if (st.tr_fs[i] == b.tr_fs[i] && st.tr_ts[i] == b.tr_ts[i] && st.tr_tap[i] == b.tr_tap[i]) continue;
            k += 't';
            num(i);
            k += ':';
            num(st.tr_fs[i]);
            num(st.tr_ts[i]);
            k += ',';
            num(st.tr_tap[i]);
            k += ';';
        } else if (comp == C_SHUNT) {
            double const a[4] = {st.sh_g1[i], st.sh_b1[i], st.sh_g0[i], st.sh_b0[i]};
            double const c[4] = {b.sh_g1[i], b.sh_b1[i], b.sh_g0[i], b.sh_b0[i]};
            if (st.sh_status[i] == b.sh_status[i] && std::memcmp(a, c, sizeof a) == 0) continue;
            k += 's';
            num(i);
            k += ':';
            num(st.sh_status[i]);
            for (double x : a) k += ',' + format_double(x);
            k += ';';
        } else {
            if (st.src_status[i] == b.src_status[i]) continue;
            k += 'c';
            num(i);
            k += ':';
            num(st.src_status[i]);
            k += ';';
        }
    }
    return k;
}

// structure and per-scenario device arrays of one group (records maps included)
void upload_group(Batch& b, Group& g) {
    Model const& m = b.model;
    Arena& A = b.arena;
    Plan const& p = g.plan;
    int const P = p.P, NB = 2 * P;
    size_t const sp = size_t(g.S_pad);
    size_t const n = size_t(p.n), E = size_t(p.E);
    DevGroup& d = g.dg;
    d.P = P;
    d.n = int(n);
    d.E = int(E);
    d.n_mm = int(p.n_mm);
    d.S = g.S;
    d.S_pad = g.S_pad;
    d.fp32 = b.fp32 ? 1 : 0;
    d.row_ptr = put(A, p.row_ptr);
    d.row_end = put(A, p.row_end);
    d.row_stride = put(A, p.row_stride);
    d.col = put(A, p.col);
    d.diag = put(A, p.diag);
    d.transpose = put(A, p.transpose);
    d.fill = put(A, p.fill);
    d.ydata = put(A, complex_to_doubles(p.ydata));
    d.schur_ptr = put(A, p.schur_ptr);
    std::vector<int> sch(p.schur.size() * 3);
    for (size_t i = 0; i < p.schur.size(); ++i)
        for (int k = 0; k < 3; ++k) sch[3 * i + k] = p.schur[i][k];
    d.schur = put(A, sch);
    d.task_ptr = put(A, p.task_ptr);
    d.task_pivot = put(A, p.task_pivot);
    d.task_mm = put(A, p.task_mm);
    d.level_task = put(A, p.level_task);
    d.keep = put(A, p.keep);
    d.phase = put(A, p.phase);
    d.bus_mm = put(A, p.bus_mm);
    d.mm_src_ptr = put(A, p.mm_src_ptr);
    d.mm_src = put(A, p.mm_src);
    {
        std::vector<int> sb(p.src_comp.size());
        for (Idx bus = 0; bus < p.n; ++bus)
            for (int q = p.src_ptr[bus]; q < p.src_ptr[bus + 1]; ++q) sb[q] = int(bus);
        d.src_bus = put(A, sb);
    }
    d.load_ptr = put(A, p.load_ptr);
    d.load_type = put(A, p.load_type);
    d.src_ptr = put(A, p.src_ptr);
    g.n_load = int(p.load_comp.size());
    g.n_src = int(p.src_comp.size());
    g.n_shunt = int(p.shunt_comp.size());
    size_t const fsz = b.fp32 ? sizeof(float) : sizeof(double);
    d.u = dev_alloc<double>(A, n * P * 2 * sp);
    d.xv = dev_alloc<double>(A, n * P * 2 * sp);
    d.y = dev_alloc<double>(A, n * NB * sp);
    d.lu = A.take(E * NB * NB * sp * fsz);
    d.perm = dev_alloc<uint32_t>(A, n * sp);
    d.done = b.pairs.done + g.pair_off;
    d.err = b.pairs.err + g.pair_off;
    d.iters = b.pairs.iters + g.pair_off;
    d.devmax = b.pairs.devmax + g.pair_off;
    d.lastdev = b.pairs.lastdev + g.pair_off;
    if (b.pairs.pert) {
        d.pert = b.pairs.pert + g.pair_off;
        d.rmax = b.pairs.rmax + g.pair_off;
        d.berr = b.pairs.berr + g.pair_off;
        d.refine_all = b.refine_steps ? 1 : 0;
    }
    for (size_t l = 0; l + 1 < p.level_ptr.size(); ++l) {
        double ent = 0, piv = 0;
        for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q) {
            int const tk = p.level_task[q];
            for (int i = p.task_ptr[tk]; i < p.task_ptr[tk + 1]; ++i) {
                int const pv = p.task_pivot[i];
                ent += (p.row_end[pv] - p.row_ptr[pv]) / (p.row_stride.empty() ? 1 : p.row_stride[pv]);
                piv += 1;
            }
        }
        g.level_entries.push_back(ent);
        g.level_pivots.push_back(piv);
    }

    DevOutput& o = g.dout;
    o.P = P;
    o.S = g.S;
    o.S_pad = g.S_pad;
    std::vector<int> si(g.scenarios.begin(), g.scenarios.end());
    si.resize(g.S_pad, 0);
    o.scenario_index = put(A, si);
    o.linear = b.linear ? 1 : 0;
    o.ok = b.d_ok + g.lane_off;
    o.n = int(n);
    o.row_ptr = d.row_ptr;
    o.row_end = d.row_end;
    o.row_stride = d.row_stride;
    o.bus_mm = d.bus_mm;
    o.col = d.col;
    o.fill = d.fill;
    o.ydata = d.ydata;
    o.u = d.u;
    o.n_load = g.n_load;
    o.load_type = d.load_type;
    o.n_src = g.n_src;
    o.src_ptr = d.src_ptr;
    o.load_ptr = d.load_ptr;
    o.shunt_ptr = put(A, p.shunt_ptr);
    o.src_si = dev_alloc<double>(A, size_t(g.n_src) * 2 * P * 2 * sp);
    o.bus_inj = dev_alloc<double>(A, n * P * 2 * sp);
    o.n_shunt = g.n_shunt;
    o.shunt_y = put(A, complex_to_doubles(p.shunt_y));
    o.shunt_si = dev_alloc<double>(A, size_t(g.n_shunt) * 2 * P * 2 * sp);

    // record maps (the plan holds every math model of the topology, in order)
    Topology const& t = g.topo;
    {
        std::vector<int> ids, bus;
        std::vector<double> ur;
        for (Idx i = 0; i < m.n_node(); ++i) {
            ids.push_back(m.node[i].id);
            ur.push_back(m.node[i].u_rated);
            Idx const gg = t.node_group[i];
            bus.push_back(gg < 0 ? -1 : p.bus_label[p.mm_bus_begin[gg] + int(t.node_pos[i])]);
        }
        g.node_rec = NodeRec{int(m.n_node()), put(A, ids), put(A, ur), put(A, bus)};
    }
    {
        std::vector<int> mm_branch_begin(t.models.size() + 1, 0);
        for (size_t k = 0; k < t.models.size(); ++k)
            mm_branch_begin[k + 1] = mm_branch_begin[k] + int(t.models[k].branch.size());
        int* br_from = put(A, p.branch_from);
        int* br_to = put(A, p.branch_to);
        double* br_y = put(A, complex_to_doubles(p.branch_y));
        // kind 0: lines, 1: transformers, 2: links, 3: generic branches (2 and 3 are line slots behind the lines)
        Idx const n_real_line = Idx(m.line.size());
        for (int kind = 0; kind < 4; ++kind) {
            Idx const first = kind == 0 ? 0 : kind == 1 ? m.n_line() : kind == 2 ? n_real_line : generic_slot0(m);
            Idx const cnt = kind == 0   ? n_real_line
                            : kind == 1 ? Idx(m.trafo.size())
                            : kind == 2 ? Idx(m.link.size())
                                        : Idx(m.gbranch.size());
            std::vector<int> ids, math;
            std::vector<double> bif, bit, rating;
            for (Idx k = 0; k < cnt; ++k) {
                Idx const bidx = first + k;
                ids.push_back(kind == 0 ? m.line[k].id
                              : kind == 1 ? m.trafo[k].id
                              : kind == 2 ? m.link[k].id
                                          : m.gbranch[k].id);
                Idx const gg = t.branch_group[bidx];
                math.push_back(gg < 0 ? -1 : mm_branch_begin[gg] + int(t.branch_pos[bidx]));
                if (kind != 1) {
                    bif.push_back(m.lstat[bidx].base_i);
                    bit.push_back(m.lstat[bidx].base_i_to);
                    rating.push_back(kind == 3 ? m.lstat[bidx].sn : m.lstat[bidx].i_n);
                } else {
                    bif.push_back(m.tstat[k].base_i_from);
                    bit.push_back(m.tstat[k].base_i_to);
                    rating.push_back(m.tstat[k].sn);
                }
            }
            g.branch_rec[kind] = BranchRec{int(cnt),    put(A, ids), put(A, math), br_from,        br_to,
                                           br_y,        nullptr,     nullptr,      put(A, bif),    put(A, bit),
                                           put(A, rating), kind == 0 || kind == 2 ? 1 : 0};
        }
    }
    {
        std::vector<int> lg_math(m.n_lg(), -1), src_math(m.source.size(), -1), sh_math(m.shunt.size(), -1);
        for (size_t k = 0; k < p.load_comp.size(); ++k) lg_math[p.load_comp[k]] = int(k);
        for (size_t k = 0; k < p.src_comp.size(); ++k) src_math[p.src_comp[k]] = int(k);
        for (size_t k = 0; k < p.shunt_comp.size(); ++k) sh_math[p.shunt_comp[k]] = int(k);
        auto make = [&](int comp, Idx cnt, auto id_of, auto math_of, auto node_of, auto stat_of, double dir,
                        int8_t const* status_lanes, double const* si) {
            std::vector<int> ids, math, bus;
            std::vector<double> bi;
            std::vector<int8_t> stt;
            for (Idx k = 0; k < cnt; ++k) {
                ids.push_back(id_of(k));
                math.push_back(math_of(k));
                Idx const node = node_of(k), gg = t.node_group[node];
                bus.push_back(gg < 0 ? -1 : p.bus_label[p.mm_bus_begin[gg] + int(t.node_pos[node])]);
                bi.push_back(kBaseP3 / m.node[node].u_rated / kSqrt3);
                stt.push_back(stat_of(k));
            }
            g.appl_rec[comp] = ApplRec{int(cnt), put(A, ids),  put(A, math), put(A, bus), put(A, bi), dir,
                                       status_lanes, put(A, stt), si};
        };
        make(
            C_SHUNT, Idx(m.shunt.size()), [&](Idx k) { return m.shunt[k].id; }, [&](Idx k) { return sh_math[k]; },
            [&](Idx k) { return m.sh_node[k]; }, [&](Idx k) { return int8_t(g.st.sh_status[k] ? 1 : 0); }, -1.0,
            nullptr, o.shunt_si);
        make(
            C_SOURCE, Idx(m.source.size()), [&](Idx k) { return m.source[k].id; },
            [&](Idx k) { return src_math[k]; }, [&](Idx k) { return m.src_node[k]; },
            [&](Idx k) { return int8_t(g.st.src_status[k] ? 1 : 0); }, 1.0, nullptr, o.src_si);
        int const lg_comps[4] = {C_SYM_GEN, C_ASYM_GEN, C_SYM_LOAD, C_ASYM_LOAD};
        for (int kk = 0; kk < 4; ++kk) {
            Idx const off = m.lg_offset[kk];
            make(
                lg_comps[kk], m.lg_count[kk], [&](Idx k) { return m.lg_id[off + k]; },
                [&](Idx k) { return lg_math[off + k]; }, [&](Idx k) { return m.lg_node[off + k]; },
                [&](Idx) { return int8_t(0); }, kk < 2 ? 1.0 : -1.0, g.load_status + size_t(off) * sp, nullptr);
        }
    }
}

// timings of the steps inside one phase of batch creation on stderr when PGM_AMD_TRACE is set
// (diagnostics only)
struct StepTrace {
    char const* phase;
    bool on = std::getenv("PGM_AMD_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(char const* step) {
        if (!on) return;
        auto const now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[pgm_amd]     %s: %-26s %8.2f ms\n", phase, step,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

// per-lane data of a union group: side statuses of every math branch, buses without a source, the
// admittances of every status combination and the Y values of the entries a member's switched
// branches touch, summed in the same order as the shared ones (branches in math order, then shunts)
void upload_union_lanes(Batch& b, Group& g) {
    StepTrace tr{"union lanes"};
    Model const& m = b.model;
    Arena& A = b.arena;
    Plan const& p = g.plan;
    int const PP = p.P * p.P, Q = 2 * PP;
    size_t const sp = size_t(g.S_pad);
    Idx const n = p.n, nbr = Idx(p.branch_comp.size()), nl = m.n_line();
    auto side = [&](std::vector<IntS> const& lf, std::vector<IntS> const& lt, std::vector<IntS> const& tf,
                    std::vector<IntS> const& tt, Idx c) {
        return (c < nl ? (lf[c] != 0) : (tf[c - nl] != 0)) | ((c < nl ? (lt[c] != 0) : (tt[c - nl] != 0)) << 1);
    };
    // admittances of the four status combinations of every math branch (branches split over host threads,
    // each with its own working state)
    std::vector<double> yvar(size_t(nbr) * 16 * PP * 2);
    unsigned const n_ythr =
        unsigned(std::max<Idx>(1, std::min<Idx>({Idx(16), Idx(std::thread::hardware_concurrency()), nbr / 512 + 1})));
    auto yvar_work = [&](unsigned th) {
        ComponentState work = g.st;
        for (Idx mb = th; mb < nbr; mb += n_ythr) {
            Idx const c = p.branch_comp[mb];
            for (int v = 0; v < 4; ++v) {
                IntS const fs = v & 1, ts = (v >> 1) & 1;
                if (c < nl) {
                    work.line_fs[c] = fs;
                    work.line_ts[c] = ts;
                } else {
                    work.tr_fs[c - nl] = fs;
                    work.tr_ts[c - nl] = ts;
                }
                auto const y = branch_param(m, work, c, b.symmetric);
                for (int w = 0; w < 4; ++w)
                    for (int q = 0; q < PP; ++q) {
                        size_t const o = ((size_t(mb) * 4 + v) * 4 * PP + size_t(w) * PP + q) * 2;
                        yvar[o] = y.v[w].m[q].real();
                        yvar[o + 1] = y.v[w].m[q].imag();
                    }
            }
            if (c < nl) {
                work.line_fs[c] = g.st.line_fs[c];
                work.line_ts[c] = g.st.line_ts[c];
            } else {
                work.tr_fs[c - nl] = g.st.tr_fs[c - nl];
                work.tr_ts[c - nl] = g.st.tr_ts[c - nl];
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned t = 1; t < n_ythr; ++t) th.emplace_back(yvar_work, t);
        yvar_work(0);
        for (auto& x : th) x.join();
    }
    tr.mark("branch admittances");
    // buses of the nodes of every math branch, whatever its side statuses in the union state
    Topology const& tp = g.topo;
    std::vector<int> nfrom(nbr, -1), nto(nbr, -1);
    auto node_bus = [&](Idx node) {
        Idx const gg = tp.node_group[node];
        return gg < 0 ? -1 : p.bus_label[p.mm_bus_begin[gg] + int(tp.node_pos[node])];
    };
    for (Idx mb = 0; mb < nbr; ++mb) {
        Idx const c = p.branch_comp[mb];
        nfrom[mb] = node_bus(c < nl ? m.line_f[c] : m.tr_f[c - nl]);
        nto[mb] = node_bus(c < nl ? m.line_t[c] : m.tr_t[c - nl]);
    }
    // math branches incident to each bus, in math order
    std::vector<std::vector<int>> inc(n);
    for (Idx mb = 0; mb < nbr; ++mb) {
        int const f = nfrom[mb], t = nto[mb];
        if (f >= 0) inc[f].push_back(int(mb));
        if (t >= 0 && t != f) inc[t].push_back(int(mb));
    }
    auto find_entry = [&](int r, int c) {
        auto const b0 = p.col.begin() + p.row_ptr[r], b1 = p.col.begin() + p.row_ptr[r + 1];
        auto const it = std::lower_bound(b0, b1, c);
        return it == b1 || *it != c ? -1 : int(it - p.col.begin());
    };
    // extra (off-pattern) entries of the lanes that close branches against the union structure
    bool const extras = !g.couple.empty();
    std::vector<int> xcnt(extras ? sp : 0, 0), xrow(extras ? sp * kExtraMax : 0, 0), xcol(xrow.size(), 0);
    std::vector<int> xm(xcnt.size(), 0), xb(xrow.size(), 0);
    std::vector<double> xy(xrow.size() * 2, 0.0);
    std::vector<int8_t> xbus(extras ? size_t(n) : 0, 0);
    // buses without a source and branch side statuses, one row per member (its lanes share them); the
    // device expands them to the lane-interleaved layout
    // (not value-initialised: every member's thread writes its whole rows, and the first touch of the pages is
    // spread over those threads)
    std::unique_ptr<uint8_t[]> dead_m(new uint8_t[std::max<size_t>(g.members.size() * size_t(n), 1)]);
    std::unique_ptr<uint8_t[]> lstat_m(new uint8_t[std::max<size_t>(g.members.size() * size_t(nbr), 1)]);
    std::vector<int> patch_lane, patch_entry;
    std::vector<double> patch_val;
    // members are independent: several host threads each take members in turn and write the lanes of
    // their own members; patches and extra entries are merged afterwards in member order
    struct MemberOut {
        std::vector<int> pe;
        std::vector<double> pv;
        std::vector<std::pair<std::pair<int, int>, cd>> extras;
        std::vector<int> mbus;
    };
    size_t const n_mem = g.members.size();
    std::vector<MemberOut> mouts(n_mem);
    tr.mark("incidence + buffers");
    unsigned const n_thr =
        unsigned(std::max<size_t>(1, std::min<size_t>({size_t(16), size_t(std::thread::hardware_concurrency()), n_mem})));
    std::vector<std::string> thr_err(n_thr);
    auto member_work = [&](unsigned t) {
        try {
            std::vector<uint8_t> bits(nbr), reach(n);
            std::vector<int> stack;
            std::vector<cd> val(PP);
            std::vector<std::pair<int, int>> touched;
            for (size_t mi = t; mi < n_mem; mi += n_thr) {
                auto const& mem = g.members[mi];
                MemberOut& mo = mouts[mi];
                for (Idx mb = 0; mb < nbr; ++mb)
                    bits[mb] = uint8_t(side(mem.line_fs, mem.line_ts, mem.tr_fs, mem.tr_ts, p.branch_comp[mb]));
                // buses reached from a source over branches closed on both sides
                std::fill(reach.begin(), reach.end(), 0);
                stack.clear();
                for (Idx r = 0; r < n; ++r)
                    if (p.src_ptr[r + 1] > p.src_ptr[r]) {
                        reach[r] = 1;
                        stack.push_back(int(r));
                    }
                while (!stack.empty()) {
                    int const r = stack.back();
                    stack.pop_back();
                    for (int mb : inc[r]) {
                        if (bits[mb] != 3) continue;
                        int const f = nfrom[mb], tt = nto[mb];
                        if (f < 0 || tt < 0) continue;
                        int const o = f == r ? tt : f;
                        if (!reach[o]) {
                            reach[o] = 1;
                            stack.push_back(o);
                        }
                    }
                }
                // entries touched by branches whose statuses differ from the union state
                touched.clear();
                for (Idx mb = 0; mb < nbr; ++mb) {
                    Idx const c = p.branch_comp[mb];
                    if (bits[mb] == side(g.st.line_fs, g.st.line_ts, g.st.tr_fs, g.st.tr_ts, c)) continue;
                    int const f = nfrom[mb], tt = nto[mb];
                    if (f >= 0) touched.emplace_back(f, f);
                    if (tt >= 0) touched.emplace_back(tt, tt);
                    if (f >= 0 && tt >= 0 && f != tt) {
                        touched.emplace_back(f, tt);
                        touched.emplace_back(tt, f);
                    }
                }
                std::sort(touched.begin(), touched.end());
                touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
                for (auto const& [r, c] : touched) {
                    std::fill(val.begin(), val.end(), cd{});
                    for (int mb : inc[r]) {
                        int const f = nfrom[mb], tt = nto[mb];
                        int w = -1;
                        if (r == c) {
                            if (f == r) w = 0;
                            else if (tt == r) w = 3;
                        } else if (f == r && tt == c) {
                            w = 1;
                        } else if (tt == r && f == c) {
                            w = 2;
                        }
                        if (w < 0) continue;
                        double const* y = yvar.data() + ((size_t(mb) * 4 + bits[mb]) * 4 * PP + size_t(w) * PP) * 2;
                        for (int q = 0; q < PP; ++q) val[q] += cd(y[2 * q], y[2 * q + 1]);
                    }
                    if (r == c)
                        for (int q = p.shunt_ptr[r]; q < p.shunt_ptr[r + 1]; ++q)
                            for (int k = 0; k < PP; ++k) val[k] += p.shunt_y[size_t(q) * PP + k];
                    int const e = find_entry(r, c);
                    if (e < 0) {
                        // off the pattern: a branch closed in this member that the union structure has
                        // open; between live buses it becomes an extra entry of the member's lanes
                        if (!extras || PP != 1) throw PowerGridError("union group: LU pattern entry missing\n");
                        if (reach[r] && reach[c] && val[0] != cd{}) mo.extras.push_back({{r, c}, val[0]});
                        continue;
                    }
                    mo.pe.push_back(e);
                    for (int k = 0; k < PP; ++k) {
                        mo.pv.push_back(val[k].real());
                        mo.pv.push_back(val[k].imag());
                    }
                }
                if (int(mo.extras.size()) > kExtraMax)
                    throw PowerGridError("union group: too many extra entries\n");
                for (auto const& [rc, v] : mo.extras) {
                    mo.mbus.push_back(rc.first);
                    mo.mbus.push_back(rc.second);
                }
                std::sort(mo.mbus.begin(), mo.mbus.end());
                mo.mbus.erase(std::unique(mo.mbus.begin(), mo.mbus.end()), mo.mbus.end());
                uint8_t* const dm = dead_m.get() + mi * size_t(n);
                for (Idx r = 0; r < n; ++r) dm[r] = reach[r] ? 0 : 1;
                std::copy(bits.begin(), bits.end(), lstat_m.get() + mi * size_t(nbr));
            }
        } catch (std::exception const& e) {
            thr_err[t] = e.what();
        }
    };
    {
        std::vector<std::thread> th;
        for (unsigned t = 1; t < n_thr; ++t) th.emplace_back(member_work, t);
        member_work(0);
        for (auto& x : th) x.join();
        for (auto const& e : thr_err)
            if (!e.empty()) throw PowerGridError(e);
    }
    tr.mark("members");
    for (size_t mi = 0; mi < n_mem; ++mi) {
        auto const& mem = g.members[mi];
        MemberOut const& mo = mouts[mi];
        size_t const ne = mo.pe.size();
        for (size_t k = 0; k < ne; ++k)
            for (int lane = mem.lane_begin; lane < mem.lane_end; ++lane) {
                patch_lane.push_back(lane);
                patch_entry.push_back(mo.pe[k]);
                patch_val.insert(patch_val.end(), mo.pv.begin() + long(k) * 2 * PP,
                                 mo.pv.begin() + long(k + 1) * 2 * PP);
            }
        for (auto const& [rc, v] : mo.extras) xbus[rc.first] = 1;
        if (mo.extras.empty()) continue;
        for (int lane = mem.lane_begin; lane < mem.lane_end; ++lane) {
            xcnt[lane] = int(mo.extras.size());
            for (size_t e = 0; e < mo.extras.size(); ++e) {
                xrow[size_t(lane) * kExtraMax + e] = mo.extras[e].first.first;
                xcol[size_t(lane) * kExtraMax + e] = mo.extras[e].first.second;
                xy[(size_t(lane) * kExtraMax + e) * 2] = mo.extras[e].second.real();
                xy[(size_t(lane) * kExtraMax + e) * 2 + 1] = mo.extras[e].second.imag();
            }
            xm[lane] = int(mo.mbus.size());
            for (size_t k = 0; k < mo.mbus.size(); ++k) xb[size_t(lane) * kExtraMax + k] = mo.mbus[k];
            g.xlanes.push_back(lane);
        }
    }
    tr.mark("patch lists");
    double* ylane = dev_alloc<double>(A, size_t(p.E) * Q * sp);
    int* d_pl = put(A, patch_lane);
    int* d_pe = put(A, patch_entry);
    double* d_pv = put(A, patch_val);
    launch_expand_ylane(g.dg.ydata, int(p.E), PP, int(sp), ylane, int(patch_lane.size()), d_pl, d_pe, d_pv,
                        b.stream);
    PGM_HIP_CHECK(hipGetLastError());
    PGM_HIP_CHECK(hipStreamSynchronize(b.stream));
    tr.mark("lane admittances");
    std::vector<int> lane_member(sp, -1);
    for (size_t mi = 0; mi < n_mem; ++mi)
        for (int lane = g.members[mi].lane_begin; lane < g.members[mi].lane_end; ++lane) lane_member[size_t(lane)] = int(mi);
    int* d_lane_member = put(A, lane_member);
    uint8_t* d_dead = dev_alloc<uint8_t>(A, size_t(n) * sp);
    uint8_t* d_lstat = dev_alloc<uint8_t>(A, size_t(nbr) * sp);
    auto put_rows = [&](uint8_t const* h, size_t bytes) {
        uint8_t* d = dev_alloc<uint8_t>(A, bytes);
        PGM_HIP_CHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
        return d;
    };
    launch_expand_member_rows(put_rows(dead_m.get(), n_mem * size_t(n)), d_lane_member, int(n), int(sp), d_dead,
                              b.stream);
    launch_expand_member_rows(put_rows(lstat_m.get(), n_mem * size_t(nbr)), d_lane_member, int(nbr), int(sp), d_lstat,
                              b.stream);
    PGM_HIP_CHECK(hipGetLastError());
    double* d_yvar = put(A, yvar);
    g.dg.ylane = ylane;
    g.dg.dead = d_dead;
    g.dout.ylane = ylane;
    g.dout.dead = d_dead;
    int* d_nfrom = put(A, nfrom);
    int* d_nto = put(A, nto);
    tr.mark("member rows");
    for (int k = 0; k < 4; ++k) {
        g.branch_rec[k].lane_status = d_lstat;
        g.branch_rec[k].branch_y_var = d_yvar;
        // sides are masked per lane by the side statuses
        g.branch_rec[k].from_bus = d_nfrom;
        g.branch_rec[k].to_bus = d_nto;
    }
    if (g.xlanes.empty()) return;
    // chains from every bus of an extra entry to the root of its elimination tree (radial structure:
    // the parent of a row is the column of its one U entry)
    std::vector<int> ch_off(n, -1), ch_len(n, 0), ch_node, ch_ue;
    int ch_max = 0;
    for (int lane : g.xlanes)
        for (int k = 0; k < xm[lane]; ++k) {
            int const c = xb[size_t(lane) * kExtraMax + k];
            if (ch_off[c] >= 0) continue;
            ch_off[c] = int(ch_node.size());
            int r = c, prev_ue = -1;
            while (true) {
                ch_node.push_back(r);
                ch_ue.push_back(prev_ue);
                if (p.row_end[r] - p.diag[r] < 2) break;
                prev_ue = p.diag[r] + 1;
                r = p.col[prev_ue];
            }
            ch_len[c] = int(ch_node.size()) - ch_off[c];
            ch_max = std::max(ch_max, ch_len[c]);
        }
    DevGroup& d = g.dg;
    d.xbus = put(A, xbus);
    d.xcnt = put(A, xcnt);
    d.xrow = put(A, xrow);
    d.xcol = put(A, xcol);
    d.xy = put(A, xy);
    d.xblk = dev_alloc<double>(A, sp * kExtraMax * 4);
    d.xm = put(A, xm);
    d.xb = put(A, xb);
    d.ch_off = put(A, ch_off);
    d.ch_len = put(A, ch_len);
    d.ch_node = put(A, ch_node);
    d.ch_ue = put(A, ch_ue);
    d.ch_max = ch_max;
    d.xst = dev_alloc<double>(A, g.xlanes.size() * kExtraMax * size_t(ch_max) * 8);
}

// per-scenario injections, source voltages and admittances of one group
// device_loads: the load_gen updates (every source in device_srcs) are applied on the device; the host
// then only handles the other update rows (sources)
template <class RowFn>
void fill_scenario_data(Batch& b, Group& g, bool has_update, RowFn&& for_each_row, bool device_loads,
                        std::vector<LoadUpdateSource> const& device_srcs, LoadTable table) {
    Model const& m = b.model;
    int const P = b.P;
    bool const sym = b.symmetric;
    Plan const& p = g.plan;
    size_t const sp = size_t(g.S_pad);
    std::vector<double> s_inj(device_loads ? 0 : size_t(g.n_load) * P * 2 * sp, 0.0);
    std::vector<int8_t> lstat(device_loads ? 0 : size_t(m.n_lg()) * sp, 0);
    std::vector<double> u_src(size_t(g.n_src) * 2 * sp, 0.0), y_src(size_t(g.n_src) * 4 * sp, 0.0);
    std::vector<int> lg_math(m.n_lg(), -1), src_math(m.source.size(), -1);
    for (size_t k = 0; k < p.load_comp.size(); ++k) lg_math[p.load_comp[k]] = int(k);
    for (size_t k = 0; k < p.src_comp.size(); ++k) src_math[p.src_comp[k]] = int(k);
    auto put_load = [&](ComponentState const& st, Idx l, int lane) {
        if (device_loads) return;
        lstat[size_t(l) * sp + lane] = st.lg_status[l] ? 1 : 0;
        int const k = lg_math[l];
        if (k < 0) return;
        auto const s3 = load_power(m, st, l, sym);
        for (int ph = 0; ph < P; ++ph) {
            s_inj[((size_t(k) * P + ph) * 2) * sp + lane] = s3[ph].real();
            s_inj[((size_t(k) * P + ph) * 2 + 1) * sp + lane] = s3[ph].imag();
        }
    };
    auto put_src = [&](ComponentState const& st, Idx sidx, int lane) {
        int const k = src_math[sidx];
        if (k < 0) return;
        cd y1, y0;
        source_param(st, sidx, y1, y0);
        cd const ur = st.src_uref[sidx] * std::exp(cd(0.0, st.src_uang[sidx]));
        u_src[(size_t(k) * 2) * sp + lane] = ur.real();
        u_src[(size_t(k) * 2 + 1) * sp + lane] = ur.imag();
        y_src[(size_t(k) * 4) * sp + lane] = y1.real();
        y_src[(size_t(k) * 4 + 1) * sp + lane] = y1.imag();
        y_src[(size_t(k) * 4 + 2) * sp + lane] = y0.real();
        y_src[(size_t(k) * 4 + 3) * sp + lane] = y0.imag();
    };
    // base values in every lane
    for (Idx l = 0; l < (device_loads ? Idx(0) : m.n_lg()); ++l) {
        put_load(m.base, l, 0);
        std::fill_n(lstat.begin() + size_t(l) * sp + 1, g.S - 1, lstat[size_t(l) * sp]);
        int const k = lg_math[l];
        if (k < 0) continue;
        for (int ph = 0; ph < P; ++ph)
            for (int c = 0; c < 2; ++c) {
                size_t const row = ((size_t(k) * P + ph) * 2 + c) * sp;
                std::fill_n(s_inj.begin() + row + 1, g.S - 1, s_inj[row]);
            }
    }
    for (Idx si = 0; si < Idx(m.source.size()); ++si)
        for (int lane = 0; lane < g.S; ++lane) put_src(g.st, si, lane);
    if (has_update) {
        // lanes are independent: several host threads, each with its own working state
        unsigned const hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        unsigned const nth = std::max(1u, std::min<unsigned>(hw, unsigned((g.S + 63) / 64)));
        std::vector<std::exception_ptr> errs(nth);
        auto worker = [&](unsigned t) {
        try {
        ComponentState work = m.base;
        std::vector<std::pair<int, Idx>> touched;
        std::vector<char> scratch;
        for (int lane = int(t); lane < g.S; lane += int(nth)) {
            Idx const s = g.scenarios[lane];
            touched.clear();
            bool structural = false;
            for_each_row(s, scratch, [&](int comp, Idx idx, void const* rec) {
                if (comp == C_LINE || comp == C_TRANSFORMER || comp == C_SHUNT) return;
                apply_update_record(m, work, comp, idx, rec, structural);
                touched.emplace_back(comp, idx);
            });
            for (auto const& [comp, idx] : touched) {
                if (comp == C_SOURCE) {
                    put_src(work, idx, lane);
                } else {
                    int const kk = comp == C_SYM_GEN ? 0 : comp == C_ASYM_GEN ? 1 : comp == C_SYM_LOAD ? 2 : 3;
                    put_load(work, m.lg_offset[kk] + idx, lane);
                }
            }
            for (auto const& [comp, idx] : touched) {  // restore for the next lane
                if (comp == C_SOURCE) {
                    work.src_status[idx] = m.base.src_status[idx];
                    work.src_uref[idx] = m.base.src_uref[idx];
                    work.src_uang[idx] = m.base.src_uang[idx];
                    work.src_sk[idx] = m.base.src_sk[idx];
                    work.src_rx[idx] = m.base.src_rx[idx];
                    work.src_z01[idx] = m.base.src_z01[idx];
                } else {
                    int const kk = comp == C_SYM_GEN ? 0 : comp == C_ASYM_GEN ? 1 : comp == C_SYM_LOAD ? 2 : 3;
                    Idx const l = m.lg_offset[kk] + idx;
                    work.lg_status[l] = m.base.lg_status[l];
                    work.lg_p[l] = m.base.lg_p[l];
                    work.lg_q[l] = m.base.lg_q[l];
                }
            }
        }
        } catch (...) {
            errs[t] = std::current_exception();
        }
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nth; ++t) th.emplace_back(worker, t);
        worker(0);
        for (auto& x : th) x.join();
        for (auto const& e : errs)
            if (e) std::rethrow_exception(e);
    }
    Arena& A = b.arena;
    if (device_loads) {
        // base injections of the group's math loads and base statuses, broadcast to every lane, then
        // each update source scattered into its lanes
        std::vector<double> base_s(size_t(g.n_load) * P * 2, 0.0);
        std::vector<int8_t> base_st(size_t(m.n_lg()), 0);
        for (Idx l = 0; l < m.n_lg(); ++l) {
            base_st[l] = m.base.lg_status[l] ? 1 : 0;
            int const k = lg_math[l];
            if (k < 0) continue;
            auto const s3 = load_power(m, m.base, l, sym);
            for (int ph = 0; ph < P; ++ph) {
                base_s[(size_t(k) * P + ph) * 2] = s3[ph].real();
                base_s[(size_t(k) * P + ph) * 2 + 1] = s3[ph].imag();
            }
        }
        double* d_sinj = dev_alloc<double>(A, size_t(g.n_load) * P * 2 * sp);
        double const* d_base_s = put(A, base_s);
        int8_t const* d_base_st = put(A, base_st);
        launch_load_broadcast(d_sinj, g.load_status, d_base_s, d_base_st, g.n_load, int(m.n_lg()), P, g.S, g.S_pad,
                              b.stream);
        table.math = put(A, lg_math);
        for (auto const& src : device_srcs)
            launch_load_scatter(d_sinj, g.load_status, g.S, g.S_pad, g.dout.scenario_index, src, table, b.stream);
        PGM_HIP_CHECK(hipGetLastError());
        g.dg.s_inj = d_sinj;
    } else {
        g.dg.s_inj = put(A, s_inj);
        PGM_HIP_CHECK(hipMemcpy(g.load_status, lstat.data(), lstat.size(), hipMemcpyHostToDevice));
    }
    g.dg.u_src = put(A, u_src);
    g.dg.y_src = put(A, y_src);
    g.dout.s_inj = g.dg.s_inj;
    g.dout.u_src = g.dg.u_src;
    g.dout.y_src = g.dg.y_src;
}

struct UpdateReader {
    ComponentBuffer const* buf = nullptr;
    int comp = 0;
    bool all_na_ids = false;
    // every scenario lists the same ids in the same order: the component indices of scenario 0's rows
    // serve every scenario (no id lookups per row)
    bool same_ids = false;
    std::vector<Idx> idx0;
    // load_gen updates applied on the device (every scenario lists the same components once each)
    bool device = false;
    LoadUpdateSource dsrc;
};

}  // namespace

void Batch::release() {
    // no work of this batch may still run once its chunks go back to the model's pool (an error thrown
    // between a launch and its synchronisation leaves work queued)
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream_u && stream_u != stream) (void)hipStreamSynchronize(stream_u);
    destroy_graphs();
    for (int k = 0; k < 2; ++k)
        if (ev_iter[k]) (void)hipEventDestroy(ev_iter[k]);
    if (h_act) (void)hipHostFree(h_act);
    groups.clear();
    for (auto e : event_pool) (void)hipEventDestroy(e);
    if (h_active) (void)hipHostFree(h_active);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_lead) (void)hipEventDestroy(ev_lead);
    if (stream_u && stream_u != stream) (void)hipStreamDestroy(stream_u);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {
// phase timings of batch creation on stderr when PGM_AMD_TRACE is set (diagnostics only)
struct PhaseTrace {
    bool on = std::getenv("PGM_AMD_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    double part[4] = {0, 0, 0, 0};
    void mark(char const* phase) {
        if (!on) return;
        auto const now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[pgm_amd] create_batch %-22s %8.1f ms\n", phase,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
    // time since the last mark or add, accumulated into part k
    void add(int k) {
        if (!on) return;
        auto const now = std::chrono::steady_clock::now();
        part[k] += std::chrono::duration<double, std::milli>(now - t).count();
        t = now;
    }
    void report_parts(std::initializer_list<char const*> names) {
        if (!on) return;
        int k = 0;
        for (char const* n : names) std::fprintf(stderr, "[pgm_amd] create_batch   %-20s %8.1f ms\n", n, part[k++]);
    }
};
}  // namespace

std::unique_ptr<Batch> create_batch(Model const& model, PGM_Options const& opt, PGM_ConstDataset const* update,
                                    BatchContext const& ctx) {
    UploadCache* const uploads = ctx.uploads;
    UploadCache const* const resident = ctx.resident;
    PhaseTrace trace;
    if (opt.calculation_type != PGM_power_flow)
        throw PowerGridError("The calculation type is not supported by this build: only power flow is implemented\n");
    if (opt.calculation_method != PGM_default_method && opt.calculation_method != PGM_newton_raphson &&
        opt.calculation_method != PGM_linear && opt.calculation_method != PGM_iterative_current &&
        opt.calculation_method != PGM_linear_current)
        throw PowerGridError("The calculation method is invalid for this calculation!");
    if (update && update->next) throw PowerGridError("Cartesian product batches are not supported by this build\n");
    // a power flow needs every load_gen's type (a state estimation does not use it)
    for (Idx l = 0; l < model.n_lg(); ++l)
        if (model.lg_type[size_t(l)] < 0 || model.lg_type[size_t(l)] > 2)
            throw PowerGridError("invalid load_gen type for id " + std::to_string(model.lg_id[size_t(l)]));
    ensure_device();
    auto bp = std::make_unique<Batch>();
    Batch& b = *bp;
    if (ctx.tuning) b.tuning = *ctx.tuning;
    b.arena.set_pool(ctx.pool);
    b.model = model;
    Model const& m = b.model;
    b.symmetric = opt.symmetric != 0;
    b.P = b.symmetric ? 1 : 3;
    b.err_tol = opt.err_tol;
    b.max_iter = opt.max_iter;
    b.linear = opt.calculation_method == PGM_linear;
    b.current = opt.calculation_method == PGM_iterative_current || opt.calculation_method == PGM_linear_current;
    if (opt.calculation_method == PGM_linear_current) {
        // one iteration of iterative current (math_solver.hpp:151-156)
        b.err_tol = std::numeric_limits<double>::infinity();
        b.max_iter = 1;
    }
    // a single linear solve has no correction step to absorb FP32 factor error: FP64 factors
    // links (y_link = 1e8 (1 + 1j) pu against branch admittances of order 1-100 pu) leave an FP32 factor no
    // significant digits for the rest of the grid: models with links are factored in FP64
    b.fp32 = opt.factor_precision == 1 && !b.linear && m.link.empty();
    b.refine_steps = b.fp32 && opt.refine_steps != 0 && !b.current;
    b.is_batch = update != nullptr;
    b.n_scenarios = update ? update->batch_size : 1;
    PGM_HIP_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    // packed and union groups overlap the plain groups on a second stream, unless PGM_AMD_STREAMS=1
    // asks for one stream (callers that capture or serialise the work)
    if (char const* v = std::getenv("PGM_AMD_STREAMS"); v && std::atoi(v) == 1)
        b.stream_u = b.stream;
    else
        PGM_HIP_CHECK(hipStreamCreateWithFlags(&b.stream_u, hipStreamNonBlocking));
    PGM_HIP_CHECK(hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming));
    PGM_HIP_CHECK(hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming));
    PGM_HIP_CHECK(hipEventCreateWithFlags(&b.ev_lead, hipEventDisableTiming));
    PGM_HIP_CHECK(hipEventCreate(&b.ev_start));
    PGM_HIP_CHECK(hipEventCreate(&b.ev_stop));
    PGM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b.h_active), 2 * sizeof(int)));
    Idx const S = b.n_scenarios;
    bool const sym = b.symmetric;

    // ---- update readers (update.hpp:60-200: rows by id, or by position when every id is NA) ----
    std::vector<UpdateReader> readers;
    if (update) {
        for (auto const& buf : update->buffers) {
            int comp = -1;
            for (int c = 0; c < N_COMP; ++c)
                if (buf.meta->name == comp_names[c]) comp = c;
            if (comp < 0 || comp == C_NODE) throw PowerGridError("update of " + buf.meta->name + " not supported");
            UpdateReader r;
            r.buf = &buf;
            r.comp = comp;
            // elements per scenario of a uniform buffer: a dense one, or an indptr one whose scenarios all hold
            // the same number of rows (dataset.hpp:334-360 is_uniform / uniform_elements_per_scenario); -1
            // otherwise
            Idx eps = buf.elements_per_scenario;
            Idx row0 = 0;  // first row of scenario 0
            if (eps < 0 && buf.indptr && update->batch_size > 0) {
                Idx const k = buf.indptr[1] - buf.indptr[0];
                bool uniform = true;
                for (Idx s = 1; s < update->batch_size && uniform; ++s)
                    uniform = buf.indptr[s + 1] - buf.indptr[s] == k;
                eps = uniform ? k : -1;
                row0 = buf.indptr[0];
            }
            // one pass over the id column (several host threads): NA ids, and whether every scenario
            // repeats scenario 0's ids
            char const* idcol = nullptr;
            size_t idstride = 0;
            if (buf.data) {
                idcol = static_cast<char const*>(buf.data);
                idstride = buf.meta->size;
            } else {
                for (auto const& [a, col] : buf.columns)
                    if (a->name == "id") {
                        idcol = static_cast<char const*>(col);
                        idstride = sizeof(ID);
                    }
            }
            Idx const total = buf.total_elements;
            auto id_at = [&](Idx i) {
                ID id = na_IntID;
                if (idcol) std::memcpy(&id, idcol + size_t(i) * idstride, sizeof id);
                return id;
            };
            unsigned const nth =
                unsigned(std::max<Idx>(1, std::min<Idx>(16, std::min<Idx>(Idx(std::thread::hardware_concurrency()),
                                                                            total / (1 << 20) + 1))));
            std::vector<char> t_na(nth, 0), t_num(nth, 0), t_diff(nth, 0);
            auto scan = [&](unsigned t) {
                Idx const lo = total * Idx(t) / Idx(nth), hi = total * Idx(t + 1) / Idx(nth);
                bool na = false, num = false, diff = false;
                for (Idx i = lo; i < hi; ++i) {
                    ID const id = id_at(i);
                    na = na || id == na_IntID;
                    num = num || id != na_IntID;
                    if (eps > 0 && i >= row0 + eps && !diff) diff = id != id_at(row0 + (i - row0) % eps);
                }
                t_na[t] = na;
                t_num[t] = num;
                t_diff[t] = diff;
            };
            // a resident buffer is scanned where it lies, in HBM
            char const* d_idcol = nullptr;
            if (idcol && resident && total > 0)
                d_idcol = static_cast<char const*>(
                    buf.data ? resident->find(buf.data, size_t(total) * buf.meta->size)
                             : resident->find(idcol, size_t(total) * sizeof(ID)));
            if (!idcol) {
                t_na[0] = total > 0;
            } else if (d_idcol) {
                int* d_flags = static_cast<int*>(b.arena.take(sizeof(int)));
                launch_set_int(d_flags, 0, b.stream);
                launch_id_scan(d_idcol, (long long)(idstride), (long long)(total), (long long)(row0),
                               (long long)(eps), d_flags, b.stream);
                int fl = 0;
                PGM_HIP_CHECK(hipMemcpyAsync(&fl, d_flags, sizeof(int), hipMemcpyDeviceToHost, b.stream));
                PGM_HIP_CHECK(hipStreamSynchronize(b.stream));
                t_na[0] = (fl & 1) != 0;
                t_num[0] = (fl & 2) != 0;
                t_diff[0] = (fl & 4) != 0;
            } else if (nth == 1) {
                scan(0);
            } else {
                std::vector<std::thread> th;
                for (unsigned t = 1; t < nth; ++t) th.emplace_back(scan, t);
                scan(0);
                for (auto& x : th) x.join();
            }
            bool const any_na = std::any_of(t_na.begin(), t_na.end(), [](char c) { return c != 0; });
            bool const any_num = std::any_of(t_num.begin(), t_num.end(), [](char c) { return c != 0; });
            bool const ids_differ = std::any_of(t_diff.begin(), t_diff.end(), [](char c) { return c != 0; });
            if (total > 0 && any_na && any_num)
                throw PowerGridError("IDs contain both numbers and NANs for component " + buf.meta->name +
                                     " in update data!");
            r.all_na_ids = any_na && !any_num && total > 0;
            if (r.all_na_ids && (eps < 0 || eps != m.count(comp)))
                throw PowerGridError("Update data without IDs for component " + buf.meta->name +
                                     " has a different number of elements per scenario then input data!");
            if (!r.all_na_ids && eps > 0 && update->batch_size > 0 && !ids_differ) {
                // every scenario lists the same ids: they are resolved once, before any scenario runs, and
                // an unknown id fails the whole call, not each scenario (job_adapter.hpp:113-124 with
                // main_core/update.hpp:236-255: the sequence of an independent component)
                for (Idx i = 0; i < eps; ++i) r.idx0.push_back(find_component(m, comp, id_at(row0 + i)));
                r.same_ids = true;
            }
            readers.push_back(std::move(r));
        }
    }
    // host_only: leave out readers applied on the device and readers whose rows need no per-scenario
    // id check and carry nothing structural
    auto rows_of = [&](Idx s, std::vector<char>& scratch, bool host_only, auto&& fn) {
        for (auto const& r : readers) {
            if (host_only && r.device) continue;
            scratch.resize(r.buf->meta->size);
            Idx const b0 = r.buf->begin(s);
            for (Idx i = b0; i < r.buf->end(s); ++i) {
                r.buf->read_record(i, scratch.data());
                Idx idx;
                if (r.all_na_ids)
                    idx = i - b0;
                else if (r.same_ids)
                    idx = r.idx0[size_t(i - b0)];
                else {
                    ID id;
                    std::memcpy(&id, scratch.data(), 4);
                    idx = find_component(m, r.comp, id);
                }
                int comp = r.comp;
                line_slot(m, comp, idx);
                fn(comp, idx, scratch.data());
            }
        }
    };
    auto for_each_row = [&](Idx s, std::vector<char>& scratch, auto&& fn) { rows_of(s, scratch, false, fn); };
    auto for_each_host_row = [&](Idx s, std::vector<char>& scratch, auto&& fn) { rows_of(s, scratch, true, fn); };
    trace.mark("update readers");
    // load_gen updates on the device: every load_gen reader lists the same components in every scenario,
    // each once
    bool device_loads = false;
    std::vector<LoadUpdateSource> device_srcs;
    LoadTable table;
    if (update && update->batch_size > 0) {
        bool all_ok = false, any = false;
        for (auto& r : readers) {
            bool const lg = r.comp == C_SYM_GEN || r.comp == C_ASYM_GEN || r.comp == C_SYM_LOAD || r.comp == C_ASYM_LOAD;
            if (!lg) continue;
            Idx const eps = r.buf->elements_per_scenario;
            bool ok = eps > 0 && (r.same_ids || r.all_na_ids);
            if (ok) {
                std::vector<Idx> idx(static_cast<size_t>(eps));
                for (Idx e = 0; e < eps; ++e) idx[e] = r.all_na_ids ? e : r.idx0[size_t(e)];
                std::vector<Idx> sorted = idx;
                std::sort(sorted.begin(), sorted.end());
                ok = std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end();
            }
            all_ok = any ? (all_ok && ok) : ok;
            any = true;
            r.device = ok;
        }
        device_loads = any && all_ok;
        if (!device_loads)
            for (auto& r : readers) r.device = false;
    }
    if (device_loads) {
        Idx const nlg = m.n_lg();
        std::vector<int8_t> st(nlg), gen(nlg), asym(nlg);
        std::vector<double> pp(size_t(nlg) * 3), qq(size_t(nlg) * 3);
        for (Idx l = 0; l < nlg; ++l) {
            st[l] = m.base.lg_status[l] ? 1 : 0;
            gen[l] = m.lg_gen[l] ? 1 : 0;
            asym[l] = m.lg_asym[l] ? 1 : 0;
            for (int ph = 0; ph < 3; ++ph) {
                pp[size_t(l) * 3 + ph] = m.base.lg_p[l][ph];
                qq[size_t(l) * 3 + ph] = m.base.lg_q[l][ph];
            }
        }
        table.n_lg = int(nlg);
        table.P = b.P;
        table.symmetric = sym ? 1 : 0;
        table.status = put(b.arena, st);
        table.p = put(b.arena, pp);
        table.q = put(b.arena, qq);
        table.gen = put(b.arena, gen);
        table.asym = put(b.arena, asym);
        for (auto& r : readers) {
            if (!r.device) continue;
            ComponentBuffer const& buf = *r.buf;
            Idx const eps = buf.elements_per_scenario;
            int const type = r.comp == C_SYM_GEN ? 0 : r.comp == C_ASYM_GEN ? 1 : r.comp == C_SYM_LOAD ? 2 : 3;
            std::vector<int> comp(static_cast<size_t>(eps));
            for (Idx e = 0; e < eps; ++e) comp[e] = int(m.lg_offset[type] + (r.all_na_ids ? e : r.idx0[size_t(e)]));
            LoadUpdateSource& d = r.dsrc;
            d.n_elem = int(eps);
            d.comp = put(b.arena, comp);
            d.asym = (type == 1 || type == 3) ? 1 : 0;
            // row-based: the whole row buffer is uploaded once and the fields address into it
            // device copy of a caller buffer, shared through `uploads` when given
            auto upload = [&](void const* host, size_t bytes) -> char const* {
                if (resident)
                    if (void const* d = resident->find(host, bytes)) return static_cast<char const*>(d);
                if (uploads) return static_cast<char const*>(uploads->get(host, bytes));
                char* d = static_cast<char*>(b.arena.take(bytes));
                PGM_HIP_CHECK(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
                return d;
            };
            char const* drows = nullptr;
            if (buf.data) drows = upload(buf.data, size_t(buf.total_elements) * buf.meta->size);
            auto field = [&](char const* name, char const*& ptr, long long& stride) {
                PGM_MetaAttribute const* a = buf.meta->find(name);
                if (!a) return;
                if (drows) {
                    ptr = drows + a->offset;
                    stride = (long long)(buf.meta->size);
                    return;
                }
                for (auto const& [ca, col] : buf.columns)
                    if (ca == a) {
                        ptr = upload(col, size_t(buf.total_elements) * a->size);
                        stride = (long long)(a->size);
                    }
            };
            field("status", d.status, d.status_stride);
            field("p_specified", d.p, d.p_stride);
            field("q_specified", d.q, d.q_stride);
            device_srcs.push_back(d);
        }
    }
    trace.mark("device load upload");
    // rows whose ids need no per-scenario check (resolved once, or positional)
    bool const ids_resolved =
        std::all_of(readers.begin(), readers.end(), [](UpdateReader const& r) { return r.all_na_ids || r.same_ids; });
    bool const has_structural_update = std::any_of(readers.begin(), readers.end(), [](UpdateReader const& r) {
        return r.comp == C_LINE || r.comp == C_LINK || r.comp == C_GENERIC_BRANCH || r.comp == C_TRANSFORMER ||
               r.comp == C_SHUNT || r.comp == C_SOURCE;
    });

    // ---- group scenarios by their structural state ----
    std::map<std::string, int> key_to_group;
    std::vector<int> scn_group(S, 0);
    std::vector<std::string> scn_error(S);
    auto group_of = [&](std::string const& key, ComponentState const& st) {
        auto it = key_to_group.find(key);
        if (it != key_to_group.end()) return it->second;
        int const gi = int(b.groups.size());
        key_to_group.emplace(key, gi);
        auto g = std::make_unique<Group>();
        g->st = structural_state(st);
        b.groups.push_back(std::move(g));
        return gi;
    };
    std::vector<char> scratch;
    // one working state: each scenario's structural rows are applied to it and undone afterwards, and its
    // key lists only the components those rows touched (the others equal the base)
    ComponentState work = m.base;
    std::vector<std::pair<int, Idx>> touched;
    auto restore = [&] {
        ComponentState const& bs = m.base;
        for (auto const& [comp, i] : touched) {
            if (comp == C_LINE) {
                work.line_fs[i] = bs.line_fs[i];
                work.line_ts[i] = bs.line_ts[i];
            } else if (comp == C_TRANSFORMER) {
                work.tr_fs[i] = bs.tr_fs[i];
                work.tr_ts[i] = bs.tr_ts[i];
                work.tr_tap[i] = bs.tr_tap[i];
            } else if (comp == C_SHUNT) {
                work.sh_status[i] = bs.sh_status[i];
                work.sh_g1[i] = bs.sh_g1[i];
                work.sh_b1[i] = bs.sh_b1[i];
                work.sh_g0[i] = bs.sh_g0[i];
                work.sh_b0[i] = bs.sh_b0[i];
            } else {
                work.src_status[i] = bs.src_status[i];
                work.src_uref[i] = bs.src_uref[i];
                work.src_uang[i] = bs.src_uang[i];
                work.src_sk[i] = bs.src_sk[i];
                work.src_rx[i] = bs.src_rx[i];
                work.src_z01[i] = bs.src_z01[i];
            }
        }
        touched.clear();
    };
    for (Idx s = 0; s < S; ++s) {
        try {
            if (!has_structural_update) {
                // still resolve every id, so that an unknown id fails its scenario
                if (!ids_resolved) for_each_row(s, scratch, [](int, Idx, void const*) {});
                scn_group[s] = group_of(std::string{}, m.base);
            } else {
                bool structural = false;
                for_each_host_row(s, scratch, [&](int comp, Idx idx, void const* rec) {
                    if (comp == C_LINE || comp == C_TRANSFORMER || comp == C_SHUNT || comp == C_SOURCE) {
                        touched.emplace_back(comp, idx);
                        apply_update_record(m, work, comp, idx, rec, structural);
                    }
                });
                scn_group[s] = group_of(structural ? touched_key(m, work, touched) : std::string{},
                                        structural ? work : m.base);
                restore();
            }
        } catch (PowerGridError const& e) {
            restore();
            if (!update) throw;
            scn_error[s] = e.what();
            scn_group[s] = -1;
            continue;
        }
        b.groups[scn_group[s]]->scenarios.push_back(int(s));
    }

    trace.mark("grouping");
    // ---- small groups that differ from the base state only in branch statuses: one union group on
    // the base structure per state of the other components. Members that only open branches run on the
    // shared structure with per-lane admittances. Members that also close branches the base has open
    // (symmetric Newton-Raphson or linear method, radial base structure, at most kExtraMax / 2 such
    // branches, no source at their buses) add off-pattern entries that a low-rank correction takes
    // into account ----
    if (opt.adjoint_layout == 0) {
        StepTrace utr{"union groups"};
        bool const closing_ok = sym && !b.current;
        std::map<std::string, std::vector<int>> by_key;
        // per group: only opens branches of the base state; the branches it closes (when at most
        // kExtraMax / 2, else none listed)
        std::vector<char> opens;
        std::vector<std::vector<Idx>> closed;
        {
            // candidates and their keys, classified on several host threads (each a scan of every branch)
            size_t const ng = b.groups.size();
            std::vector<std::string> keys(ng);
            std::vector<char> cand(ng, 0);
            opens.assign(ng, 0);
            closed.assign(ng, {});
            unsigned const nth = unsigned(std::max<size_t>(
                1, std::min<size_t>({size_t(16), size_t(std::thread::hardware_concurrency()), ng / 8 + 1})));
            auto classify = [&](unsigned t) {
                for (size_t gi = t; gi < ng; gi += nth) {
                    Group const& g = *b.groups[gi];
                    if (g.scenarios.size() >= size_t(kUnionBelow)) continue;
                    opens[gi] = opens_only(m, g.st);
                    // over the limit the scan stops with a partial list: no candidate then
                    if (!opens[gi] && closing_ok && closed_against_base(m, g.st, kExtraMax / 2, &closed[gi]) <= 0)
                        closed[gi].clear();
                    if (opens[gi] || !closed[gi].empty()) {
                        cand[gi] = 1;
                        keys[gi] = structural_key(m, g.st, false);
                    }
                }
            };
            std::vector<std::thread> th;
            for (unsigned t = 1; t < nth; ++t) th.emplace_back(classify, t);
            classify(0);
            for (auto& x : th) x.join();
            for (size_t gi = 0; gi < ng; ++gi)
                if (cand[gi]) by_key[keys[gi]].push_back(int(gi));
        }
        utr.mark("classify");
        std::vector<char> merged(b.groups.size(), 0);
        std::vector<std::unique_ptr<Group>> unions;
        for (auto& [key, cand] : by_key) {
            if (cand.size() < 2) continue;
            ComponentState ust = b.groups[cand[0]]->st;
            ust.line_fs = m.base.line_fs;
            ust.line_ts = m.base.line_ts;
            ust.tr_fs = m.base.tr_fs;
            ust.tr_ts = m.base.tr_ts;
            std::vector<int> list;
            std::vector<char> couple(size_t(m.n_branch()), 0);
            std::vector<int> closing;
            for (int gi : cand) {
                if (opens[size_t(gi)])
                    list.push_back(gi);
                else
                    closing.push_back(gi);
            }
            if (!closing.empty()) {
                // the union structure with every closed branch coupled: the closing members' branches
                // must join two buses of one math model without a source, and the structure must be a
                // forest (one U entry per row: the correction walks single-parent chains)
                std::vector<char> all(size_t(m.n_branch()), 0);
                for (int gi : closing)
                    for (Idx c : closed[size_t(gi)]) all[c] = 1;
                // the coupled structure's plan, kept by the model like the groups' plans
                std::string ckey;
                if (ctx.plans) {
                    ckey = structural_key(m, ust) + "|closing|" + (sym ? "s|" : "a|");
                    for (size_t c = 0; c < all.size(); ++c)
                        if (all[c]) ckey += std::to_string(c) + ",";
                }
                std::shared_ptr<PlanEntry const> cp = ctx.plans ? ctx.plans->find(ckey) : nullptr;
                if (!cp) {
                    PlanEntry e;
                    e.topo = build_topology(m, ust, &all);
                    std::vector<Idx> models(e.topo.models.size());
                    std::iota(models.begin(), models.end(), Idx{0});
                    e.plan = build_plan(m, ust, e.topo, models, sym, 16);
                    cp = std::make_shared<PlanEntry const>(std::move(e));
                    if (ctx.plans) ctx.plans->put(ckey, cp);
                }
                Topology const& t = cp->topo;
                Plan const& pl = cp->plan;
                std::vector<char> src_node(size_t(m.n_node()), 0);
                for (Idx k = 0; k < Idx(m.source.size()); ++k) src_node[m.src_node[k]] = 1;
                utr.mark("closing structure");
                bool radial = true;
                for (Idx r = 0; r < pl.n && radial; ++r) radial = pl.row_end[r] - pl.diag[r] <= 2;
                Idx const nl = m.n_line();
                for (int gi : closing) {
                    if (!radial) break;
                    std::vector<Idx> const& cl = closed[size_t(gi)];
                    bool ok = !cl.empty();
                    Idx const g0 = ok ? t.node_group[cl[0] < nl ? m.line_f[cl[0]] : m.tr_f[cl[0] - nl]] : -1;
                    for (Idx c : cl) {
                        Idx const f = c < nl ? m.line_f[c] : m.tr_f[c - nl];
                        Idx const to = c < nl ? m.line_t[c] : m.tr_t[c - nl];
                        // one math model per member (its lanes iterate and converge together)
                        ok = ok && g0 >= 0 && t.node_group[f] == g0 && t.node_group[to] == g0 && !src_node[f] &&
                             !src_node[to] && t.branch_group[c] >= 0;
                    }
                    if (!ok) continue;
                    list.push_back(gi);
                    for (Idx c : cl)
                        if (t.branch_pos[c] >= 0 && all[c]) couple[c] = 1;
                }
            }
            if (list.size() < 2) continue;
            auto u = std::make_unique<Group>();
            u->is_union = true;
            u->st = std::move(ust);
            if (std::any_of(couple.begin(), couple.end(), [](char c) { return c != 0; })) u->couple = std::move(couple);
            for (int gi : list) {
                Group& g = *b.groups[gi];
                Group::Member mem;
                mem.lane_begin = int(u->scenarios.size());
                mem.lane_end = mem.lane_begin + int(g.scenarios.size());
                mem.line_fs = g.st.line_fs;
                mem.line_ts = g.st.line_ts;
                mem.tr_fs = g.st.tr_fs;
                mem.tr_ts = g.st.tr_ts;
                u->members.push_back(std::move(mem));
                u->scenarios.insert(u->scenarios.end(), g.scenarios.begin(), g.scenarios.end());
                merged[gi] = 1;
            }
            unions.push_back(std::move(u));
        }
        utr.mark("members");
        std::vector<std::unique_ptr<Group>> kept;
        for (size_t gi = 0; gi < b.groups.size(); ++gi)
            if (!merged[gi]) kept.push_back(std::move(b.groups[gi]));
        for (auto& u : unions) kept.push_back(std::move(u));
        b.groups = std::move(kept);
        utr.mark("regroup");
    }

    trace.mark("union groups");
    // ---- per group: topology and plan, built in parallel on the host ----
    {
        size_t const G = b.groups.size();
        unsigned const hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::string> errs(G);
        auto work = [&](size_t t) {
            for (size_t gi = t; gi < G; gi += hw) {
                try {
                    Group& g = *b.groups[gi];
                    int const nblk = int((g.scenarios.size() + 63) / 64);
                    g.packed = !g.is_union && g.scenarios.size() <= size_t(kPackedMax) && opt.adjoint_layout == 0;
                    int const pw = b.tuning.plain_waves > 0 ? b.tuning.plain_waves : plain_waves();
                    int const ut = b.tuning.union_tasks > 0 ? b.tuning.union_tasks : union_tasks();
                    int const pt = b.tuning.packed_target > 0 ? b.tuning.packed_target : packed_target();
                    // a plain group of few 64-scenario blocks takes more, shorter tasks per level: its
                    // levels would otherwise hold too few waves to hide the latency of the pivot chains
                    // (feeders_64, 2 blocks: critical path 281 pivots at 256 tasks, 105 at 4096)
                    int const target = g.packed     ? pt
                                       : g.is_union ? std::clamp(pw / std::max(nblk, 1), 4, ut)
                                                    : std::clamp(pw / std::max(nblk, 1), 4, 4096);
                    int const ilv = g.packed ? kWaveLanes / int(g.scenarios.size()) : 0;
                    // the state's plan from an earlier batch of the model, or built now
                    std::string key;
                    if (ctx.plans) {
                        key = structural_key(m, g.st) + "|" + (sym ? "s" : "a") + std::to_string(target) + "/" +
                              std::to_string(ilv) + "|";
                        for (size_t c = 0; c < g.couple.size(); ++c)
                            if (g.couple[c]) key += std::to_string(c) + ",";
                    }
                    std::shared_ptr<PlanEntry const> hit = ctx.plans ? ctx.plans->find(key) : nullptr;
                    if (hit) {
                        g.topo = hit->topo;
                        g.plan = hit->plan;
                    } else {
                        g.topo = build_topology(m, g.st, g.couple.empty() ? nullptr : &g.couple);
                        std::vector<Idx> all(g.topo.models.size());
                        std::iota(all.begin(), all.end(), Idx{0});
                        g.plan = build_plan(m, g.st, g.topo, all, sym, target);
                        if (g.packed) {
                            // lanes of a packed wave run different tasks: interleave them in memory
                            interleave_plan(g.plan, ilv);
                            check_plan(g.plan);
                        }
                        if (ctx.plans) ctx.plans->put(key, std::make_shared<PlanEntry const>(PlanEntry{g.topo, g.plan}));
                    }
                } catch (std::exception const& e) {
                    errs[gi] = e.what();
                }
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < hw; ++t) th.emplace_back(work, t);
        for (auto& x : th) x.join();
        for (auto const& e : errs)
            if (!e.empty()) throw PowerGridError(e);
    }

    trace.mark("plans");
    // ---- batch-wide iteration state and lanes ----
    long long pairs = 0, lanes = 0;
    for (auto& gp : b.groups) {
        Group& g = *gp;
        g.S = int(g.scenarios.size());
        g.S_pad = g.S > 32 ? (g.S + 63) / 64 * 64 : g.S;
        g.pair_off = pairs;
        g.lane_off = lanes;
        pairs += (long long)(g.plan.n_mm) * g.S_pad;
        lanes += g.S_pad;
        b.real_pairs += (long long)(g.plan.n_mm) * g.S;
        b.total_lanes += g.S;
    }
    b.pairs.n = pairs;
    b.pairs.done = dev_alloc<int>(b.arena, size_t(pairs));
    b.pairs.err = dev_alloc<int>(b.arena, size_t(pairs));
    b.pairs.iters = dev_alloc<int>(b.arena, size_t(pairs));
    b.pairs.devmax = dev_alloc<unsigned long long>(b.arena, size_t(pairs));
    b.pairs.lastdev = dev_alloc<double>(b.arena, size_t(pairs));
    if (b.fp32) {
        b.pairs.pert = dev_alloc<int>(b.arena, size_t(pairs));
        b.pairs.rmax = dev_alloc<unsigned long long>(b.arena, size_t(pairs));
        b.pairs.berr = dev_alloc<unsigned long long>(b.arena, size_t(pairs));
    }
    // pairs solved by the linear method: every pair for calculation_method = linear, else the math
    // models whose load_gens are all const_y (math_solver.hpp:43-64)
    {
        std::vector<uint8_t> lin(size_t(pairs), b.linear ? 1 : 0);
        bool any = b.linear;
        if (!b.linear)
            for (auto& gp : b.groups) {
                Plan const& p = gp->plan;
                std::vector<char> all_y(size_t(p.n_mm), 1);
                for (Idx bus = 0; bus < p.n; ++bus)
                    for (int l = p.load_ptr[bus]; l < p.load_ptr[bus + 1]; ++l)
                        if (p.load_type[l] != 1) all_y[p.bus_mm[bus]] = 0;
                for (Idx k = 0; k < p.n_mm; ++k)
                    if (all_y[k]) {
                        any = true;
                        std::fill_n(lin.begin() + gp->pair_off + k * gp->S_pad, gp->S_pad, uint8_t(1));
                    }
            }
        if (any) b.d_pair_linear = put(b.arena, lin);
    }
    // [0] pairs still iterating after a check, [1] pairs whose step waits for a refinement
    b.d_active = dev_alloc<int>(b.arena, 2);
    b.d_gmax = dev_alloc<unsigned long long>(b.arena, 1);
    b.d_ok = dev_alloc<int>(b.arena, size_t(std::max<long long>(lanes, 1)));
    int8_t* lstat_all = dev_alloc<int8_t>(b.arena, size_t(std::max<long long>(lanes * m.n_lg(), 1)));
    {
        long long off = 0;
        for (auto& gp : b.groups) {
            gp->load_status = lstat_all + off;
            off += (long long)(gp->S_pad) * m.n_lg();
        }
    }
    trace.mark("lane state");
    for (auto& gp : b.groups) {
        upload_group(b, *gp);
        trace.add(0);
        if (gp->is_union) upload_union_lanes(b, *gp);
        trace.add(1);
        fill_scenario_data(b, *gp, update != nullptr, for_each_host_row, device_loads, device_srcs, table);
        trace.add(2);
    }
    trace.report_parts({"upload_group", "union lanes", "scenario data"});

    trace.mark("device tables");
    // ---- device tables and work lists ----
    size_t const G = b.groups.size();
    {
        std::vector<DevGroup> dg;
        std::vector<DevOutput> dout;
        std::vector<NodeRec> nr;
        std::vector<BranchRec> br[4];
        std::vector<ApplRec> ar[N_COMP];
        for (auto& gp : b.groups) {
            dg.push_back(gp->dg);
            dout.push_back(gp->dout);
            nr.push_back(gp->node_rec);
            for (int k = 0; k < 4; ++k) br[k].push_back(gp->branch_rec[k]);
            for (int c = 0; c < N_COMP; ++c) ar[c].push_back(gp->appl_rec[c]);
        }
        b.d_groups = put(b.arena, dg);
        std::vector<ExtraWork> xw;
        for (size_t gi = 0; gi < b.groups.size(); ++gi)
            for (size_t k = 0; k < b.groups[gi]->xlanes.size(); ++k)
                xw.push_back(ExtraWork{int(gi), b.groups[gi]->xlanes[k], int(k)});
        b.n_xwork = int(xw.size());
        if (!xw.empty()) b.d_xwork = put(b.arena, xw);
        b.d_outputs = put(b.arena, dout);
        b.d_node_recs = put(b.arena, nr);
        for (int k = 0; k < 4; ++k) b.d_branch_recs[k] = put(b.arena, br[k]);
        for (int c = 0; c < N_COMP; ++c)
            if (c == C_SHUNT || c == C_SOURCE || c >= C_SYM_GEN) b.d_appl_recs[c] = put(b.arena, ar[c]);
        int max_levels = 0, max_levels_u = 0, max_levels_p = 0;
        for (auto& gp : b.groups) {
            int& ml = gp->is_union ? max_levels_u : gp->packed ? max_levels_p : max_levels;
            ml = std::max(ml, int(gp->plan.level_ptr.size()) - 1);
        }
        double const fsz = b.fp32 ? 4.0 : 8.0;
        // union and packed groups run on their own stream with their own kernels
        for (int l = 0; l < max_levels_u; ++l) {
            std::vector<WorkItem> items;
            for (size_t gi = 0; gi < G; ++gi) {
                Group const& g = *b.groups[gi];
                Plan const& p = g.plan;
                if (!g.is_union || l + 1 >= int(p.level_ptr.size()) || g.S == 0) continue;
                int const nblk = (g.S + 63) / 64;
                for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q)
                    for (int blk = 0; blk < nblk; ++blk) items.push_back(WorkItem{int(gi), p.level_task[q], blk});
            }
            b.d_items_u.push_back(put(b.arena, items));
            b.n_items_u.push_back(int(items.size()));
        }
        for (int l = 0; l < max_levels_p; ++l) {
            std::vector<WorkItem> items;
            for (size_t gi = 0; gi < G; ++gi) {
                Group const& g = *b.groups[gi];
                Plan const& p = g.plan;
                if (!g.packed || l + 1 >= int(p.level_ptr.size()) || g.S == 0) continue;
                int const per_wave = kWaveLanes / g.S;
                for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; q += per_wave)
                    items.push_back(WorkItem{int(gi), q, std::min(per_wave, p.level_ptr[l + 1] - q)});
            }
            b.d_items_p.push_back(put(b.arena, items));
            b.n_items_p.push_back(int(items.size()));
        }
        for (int l = 0; l < max_levels; ++l) {
            std::vector<WorkItem> items;
            double bytes = 0;
            for (size_t gi = 0; gi < G; ++gi) {
                Group const& g = *b.groups[gi];
                Plan const& p = g.plan;
                if (g.is_union || g.packed || l + 1 >= int(p.level_ptr.size()) || g.S == 0) continue;
                int const nblk = (g.S + 63) / 64;
                for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q)
                    for (int blk = 0; blk < nblk; ++blk) items.push_back(WorkItem{int(gi), p.level_task[q], blk});
                // factor written once and read once by the forward substitution (2 E b), voltages read
                // and the forward vector written (32 P bytes per bus)
                bytes += (2.0 * g.level_entries[l] * fsz * 4 * b.P * b.P + 32.0 * b.P * g.level_pivots[l]) * g.S;
            }
            b.d_items.push_back(put(b.arena, items));
            b.n_items.push_back(int(items.size()));
            b.level_bytes_full.push_back(bytes);
        }
        std::vector<long long> bus_off{0};
        for (auto& gp : b.groups) bus_off.push_back(bus_off.back() + (long long)(gp->plan.n) * gp->S_pad);
        b.bus_total = bus_off.back();
        b.d_bus_off = put(b.arena, bus_off);
        for (int c = 0; c < N_COMP; ++c) {
            size_t const rs = c == C_NODE ? (b.symmetric ? sizeof(NodeOutput<1>) : sizeof(NodeOutput<3>))
                              : (c == C_LINE || c == C_TRANSFORMER || c == C_LINK || c == C_GENERIC_BRANCH)
                                  ? (b.symmetric ? sizeof(BranchOutput<1>) : sizeof(BranchOutput<3>))
                                  : (b.symmetric ? sizeof(ApplianceOutput<1>) : sizeof(ApplianceOutput<3>));
            long long const slots = 64LL * record_tile(rs);
            std::vector<long long> off{0};
            for (auto& gp : b.groups) {
                // tiles of W lanes x slots / W components, W = min(S_pad, 64)
                long long const cnt = m.count(c);
                long long const w = std::min<long long>(gp->S_pad, 64), tn = slots / w;
                long long const tiles = cnt == 0 ? 0 : (gp->S_pad + w - 1) / w * ((cnt + tn - 1) / tn);
                off.push_back(off.back() + tiles);
            }
            b.rec_total[c] = off.back();
            b.d_rec_off[c] = put(b.arena, off);
        }
    }

    trace.mark("tables + work lists");
    // ---- output record buffers ----
    for (int c = 0; c < N_COMP; ++c) {
        size_t const rs = c == C_NODE ? (sym ? sizeof(NodeOutput<1>) : sizeof(NodeOutput<3>))
                          : (c == C_LINE || c == C_TRANSFORMER || c == C_LINK || c == C_GENERIC_BRANCH)
                              ? (sym ? sizeof(BranchOutput<1>) : sizeof(BranchOutput<3>))
                              : (sym ? sizeof(ApplianceOutput<1>) : sizeof(ApplianceOutput<3>));
        b.rec_size[c] = rs;  // the record buffer itself is taken by the first run that writes records
    }
    b.status.assign(S, 0);
    b.iterations.assign(S, 0);
    b.messages.assign(S, "");
    for (Idx s = 0; s < S; ++s)
        if (scn_group[s] < 0) {
            b.status[s] = 3;
            b.messages[s] = scn_error[s];
        }
    PGM_HIP_CHECK(hipDeviceSynchronize());
    trace.mark("outputs + sync");
    return bp;
}

void Batch::run() {
    hipStream_t const st = stream;
    fwd_launches = 0;
    fwd_bytes = 0.0;
    std::vector<hipEvent_t> fwd_events;
    double fwd_graph_ms = 0.0;  // forward-level durations of the graph launches (their own events)
    Idx fwd_graph_timed = 0;
    // event timing of the forward launches (stats); PGM_AMD_EVENTS=0 leaves it out, for tools that
    // intercept the queue and take only kernel dispatches
    bool const timing = [] {
        char const* v = std::getenv("PGM_AMD_EVENTS");
        return !(v && std::atoi(v) == 0);
    }();
    auto event_from_pool = [&](size_t i) {
        while (event_pool.size() <= i) {
            hipEvent_t e;
            PGM_HIP_CHECK(hipEventCreate(&e));
            event_pool.push_back(e);
        }
        return event_pool[i];
    };
    int const L = int(n_items.size()), LU = int(n_items_u.size()), LP = int(n_items_p.size());
    hipStream_t const su = stream_u;
    // the union groups' sweeps run on a second stream, overlapping the main groups' sweeps; both
    // join before each convergence check
    auto fork = [&] {
        if (LU + LP == 0 || su == st) return;
        PGM_HIP_CHECK(hipEventRecord(ev_fork, st));
        PGM_HIP_CHECK(hipStreamWaitEvent(su, ev_fork, 0));
    };
    auto join = [&] {
        if (LU + LP == 0 || su == st) return;
        PGM_HIP_CHECK(hipEventRecord(ev_join, su));
        PGM_HIP_CHECK(hipStreamWaitEvent(st, ev_join, 0));
    };
    // the second stream's first forward level runs before the main stream starts its sweeps: the
    // second stream's long upper-level chains then overlap the main stream's wide levels
    bool const lead = su != st && LU + LP > 0;
    auto union_sweeps = [&](int mode) {
        // packed and union groups interleaved level by level on the second stream
        for (int l = 0; l < std::max(LU, LP); ++l) {
            if (l < LP) launch_forward(d_groups, d_items_p[l], n_items_p[l], P, fp32, mode, su, KIND_PACKED);
            if (l < LU) launch_forward(d_groups, d_items_u[l], n_items_u[l], P, fp32, mode, su, KIND_UNION);
            if (l == 0 && lead) {
                PGM_HIP_CHECK(hipEventRecord(ev_lead, su));
                PGM_HIP_CHECK(hipStreamWaitEvent(st, ev_lead, 0));
            }
        }
        // lanes that close branches against their union structure: Woodbury correction of the
        // forward vectors before the backward sweep
        if (n_xwork > 0 && (mode == MODE_INIT || mode == MODE_NR)) launch_lowrank(d_groups, d_xwork, n_xwork, fp32, su);
        for (int l = std::max(LU, LP) - 1; l >= 0; --l) {
            if (l < LP) launch_backward(d_groups, d_items_p[l], n_items_p[l], P, fp32, mode, su, KIND_PACKED);
            if (l < LU) launch_backward(d_groups, d_items_u[l], n_items_u[l], P, fp32, mode, su, KIND_UNION);
        }
    };
    if (timing) PGM_HIP_CHECK(hipEventRecord(ev_start, st));
    launch_init_state(pairs, st);
    fork();
    // the linear system: Newton-Raphson's initial guess, and the answer of the linear method and of
    // the all-const_y math models; iterative current starts flat instead
    if (!current || d_pair_linear) {
        union_sweeps(MODE_INIT);
        for (int l = 0; l < L; ++l) launch_forward(d_groups, d_items[l], n_items[l], P, fp32, MODE_INIT, st);
        for (int l = L - 1; l >= 0; --l) launch_backward(d_groups, d_items[l], n_items[l], P, fp32, MODE_INIT, st);
    }
    double active_frac = 1.0;
    if (d_pair_linear || current) {
        join();
        // the linear solve is the answer of these pairs: no further iteration
        if (d_pair_linear) launch_retire_linear(pairs, d_pair_linear, st);
        if (current) launch_ic_start(P, d_groups, d_bus_off, int(groups.size()), bus_total, st);
        fork();
    }
    // Newton-Raphson iterations as graph launches (PGM_AMD_GRAPH=0: one launch per level, as below)
    bool const graph_wanted = !linear && !current && max_iter > 0 && [] {
        char const* v = std::getenv("PGM_AMD_GRAPH");
        return !(v && std::atoi(v) == 0);
    }();
    // one iteration on the streams: the body that is captured
    auto nr_iteration = [&](int parity, bool events) {
        launch_begin_iteration(d_iter, d_active, st);
        fork();
        union_sweeps(MODE_NR);
        for (int l = 0; l < L; ++l) {
            if (events) PGM_HIP_CHECK(hipEventRecordWithFlags(nr_ev[parity][2 * l], st, hipEventRecordExternal));
            launch_forward(d_groups, d_items[l], n_items[l], P, fp32, MODE_NR, st);
            if (events) PGM_HIP_CHECK(hipEventRecordWithFlags(nr_ev[parity][2 * l + 1], st, hipEventRecordExternal));
        }
        for (int l = L - 1; l >= 0; --l) launch_backward(d_groups, d_items[l], n_items[l], P, fp32, MODE_NR, st);
        join();
        launch_converge(pairs, 0, int(max_iter), err_tol, d_active, st, d_iter);
    };
    if (graph_wanted && !graph_tried) {
        graph_tried = true;
        if (!d_iter) d_iter = static_cast<int*>(arena.take(sizeof(int)));
        if (!h_act) PGM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_act), 4 * sizeof(int)));
        for (int k = 0; k < 2; ++k)
            if (!ev_iter[k]) PGM_HIP_CHECK(hipEventCreateWithFlags(&ev_iter[k], hipEventDisableTiming));
        // the prologue's work on the second stream joins the main stream before the capture starts
        join();
        PGM_HIP_CHECK(hipStreamSynchronize(st));
        graph_ok = true;
        for (int k = 0; k < 2 && graph_ok; ++k) {
            nr_ev[k].clear();
            if (timing) {
                for (int e = 0; e < 2 * L; ++e) {
                    hipEvent_t ev;
                    PGM_HIP_CHECK(hipEventCreate(&ev));
                    nr_ev[k].push_back(ev);
                }
            }
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                graph_ok = false;
                break;
            }
            try {
                nr_iteration(k, timing);
            } catch (PowerGridError const&) {
                graph_ok = false;
            }
            hipError_t const ec = hipStreamEndCapture(st, &graph);
            graph_ok = graph_ok && ec == hipSuccess && graph &&
                       hipGraphInstantiate(&nr_exec[k], graph, nullptr, nullptr, 0) == hipSuccess;
            if (graph) (void)hipGraphDestroy(graph);
        }
        (void)hipGetLastError();
        if (!graph_ok) destroy_graphs();
    }
    graph_launches = 0;
    refinements = 0;
    refine_passes = 0;
    if (graph_wanted && graph_ok) {
        // the prologue's second-stream work joins before the first launch; every launch forks and joins
        join();
        launch_set_int(d_iter, 0, st);
        int launched = 0, done_it = 0, limit = int(max_iter), rejoined = 0;
        auto launch_next = [&] {
            int const k = launched % 2;
            PGM_HIP_CHECK(hipGraphLaunch(nr_exec[k], st));
            PGM_HIP_CHECK(hipMemcpyAsync(&h_act[2 * k], d_active, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
            PGM_HIP_CHECK(hipEventRecord(ev_iter[k], st));
            ++launched;
        };
        launch_next();
        // one iteration ahead: the next iteration is queued before the host waits for the present one, so
        // the device never waits for the host; the iteration queued after the last one finds every lane
        // done and returns at once
        while (true) {
            if (launched < limit) launch_next();
            int const k = done_it % 2;
            PGM_HIP_CHECK(hipEventSynchronize(ev_iter[k]));
            ++done_it;
            if (timing) {
                for (int l = 0; l < L; ++l) {
                    float e = 0;
                    if (hipEventElapsedTime(&e, nr_ev[k][2 * l], nr_ev[k][2 * l + 1]) != hipSuccess) continue;
                    fwd_graph_ms += e;
                    fwd_bytes += level_bytes_full[l] * active_frac;
                    ++fwd_graph_timed;
                }
            }
            fwd_launches += L;
            int act = h_act[2 * k] + rejoined;
            rejoined = 0;
            if (h_act[2 * k + 1] > 0) {
                // steps from perturbed pivots: the iteration queued after this one runs without those pairs;
                // once it is done their steps are refined and applied, and they rejoin one iteration behind
                PGM_HIP_CHECK(hipStreamSynchronize(st));
                rejoined = refine_held_steps();
                act += rejoined;
                ++limit;
            }
            active_frac = real_pairs > 0 ? double(act) / double(real_pairs) : 0.0;
            if (act == 0 || done_it >= limit) break;
        }
        graph_launches = launched;
    }
    int extra_iterations = 0;  // refinements: the pairs that waited are one iteration behind
    for (int it = 1; it <= ((linear || (graph_wanted && graph_ok)) ? 0 : int(max_iter) + extra_iterations); ++it) {
        int const mode = current ? (it == 1 ? MODE_ICF : MODE_RESID) : MODE_NR;
        union_sweeps(mode);
        for (int l = 0; l < L; ++l) {
            if (timing) {
                fwd_events.push_back(event_from_pool(fwd_events.size()));
                PGM_HIP_CHECK(hipEventRecord(fwd_events.back(), st));
            }
            launch_forward(d_groups, d_items[l], n_items[l], P, fp32, mode, st);
            if (timing) {
                fwd_events.push_back(event_from_pool(fwd_events.size()));
                PGM_HIP_CHECK(hipEventRecord(fwd_events.back(), st));
            }
            if (mode == MODE_NR) fwd_bytes += level_bytes_full[l] * active_frac;
            ++fwd_launches;
        }
        for (int l = L - 1; l >= 0; --l) launch_backward(d_groups, d_items[l], n_items[l], P, fp32, mode, st);
        join();
        h_active[0] = h_active[1] = 0;
        PGM_HIP_CHECK(hipMemcpyAsync(d_active, h_active, 2 * sizeof(int), hipMemcpyHostToDevice, st));
        launch_converge(pairs, it, int(max_iter), err_tol, d_active, st);
        fork();
        PGM_HIP_CHECK(hipMemcpyAsync(h_active, d_active, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        PGM_HIP_CHECK(hipStreamSynchronize(st));
        int act = h_active[0];
        if (h_active[1] > 0) {
            act += refine_held_steps();
            ++extra_iterations;
        }
        // lanes still iterating, for the byte count of the next iteration's launches
        active_frac = real_pairs > 0 ? double(act) / double(real_pairs) : 0.0;
        if (act == 0) break;
    }
    PGM_HIP_CHECK(hipGetLastError());
    // per-scenario status
    size_t const np = size_t(pairs.n);
    std::vector<int> done(np), err(np), its(np);
    std::vector<double> ldev(np);
    if (np > 0) {
        PGM_HIP_CHECK(hipMemcpyAsync(err.data(), pairs.err, np * 4, hipMemcpyDeviceToHost, st));
        PGM_HIP_CHECK(hipMemcpyAsync(its.data(), pairs.iters, np * 4, hipMemcpyDeviceToHost, st));
        PGM_HIP_CHECK(hipMemcpyAsync(ldev.data(), pairs.lastdev, np * 8, hipMemcpyDeviceToHost, st));
        PGM_HIP_CHECK(hipStreamSynchronize(st));
    }
    long long lanes = 0;
    for (auto& gp : groups) lanes += gp->S_pad;
    std::vector<int> ok(size_t(std::max<long long>(lanes, 1)), 0);
    for (auto& gp : groups) {
        Group& g = *gp;
        size_t const sp = size_t(g.S_pad);
        size_t const nmm = size_t(g.plan.n_mm);
        for (int lane = 0; lane < g.S; ++lane) {
            Idx const s = g.scenarios[lane];
            int stt = 0;
            Idx it_max = 0;
            std::string msg;
            for (size_t k = 0; k < nmm; ++k) {
                size_t const i = size_t(g.pair_off) + k * sp + lane;
                if (stt == 0 && err[i] == ERR_SPARSE) {
                    stt = 1;
                    msg = SparseMatrixErrorMessage;
                } else if (stt == 0 && err[i] == ERR_DIVERGE) {
                    stt = 2;
                    msg = "Iteration failed to converge after " + std::to_string(max_iter) +
                          " iterations! Max deviation: " + format_double(ldev[i]) +
                          ", error tolerance: " + format_double(err_tol) + ".\n";
                }
                it_max = std::max<Idx>(it_max, its[i]);
            }
            status[s] = stt;
            iterations[s] = stt == 0 ? it_max : 0;
            messages[s] = msg;
            ok[size_t(g.lane_off) + lane] = stt == 0 ? 1 : 0;
        }
    }
    PGM_HIP_CHECK(hipMemcpy(d_ok, ok.data(), ok.size() * 4, hipMemcpyHostToDevice));
    int const G = int(groups.size());
    bool node_pq = want[C_NODE] && cols[C_NODE].empty();
    for (auto const& c : cols[C_NODE]) node_pq = node_pq || c.attr->name == "p" || c.attr->name == "q";
    bool const bus_sums = node_pq || want[C_SOURCE] || want[C_SHUNT];
    auto column_set = [&](int c) {
        ColumnSet cs;
        for (auto const& col : cols[c]) {
            cs.off[cs.n] = int(col.attr->offset);
            cs.size[cs.n] = int(col.attr->size);
            cs.dst[cs.n] = static_cast<char*>(col.dev);
            ++cs.n;
        }
        cs.bus_sums = bus_sums ? 1 : 0;
        return cs;
    };
    // source and shunt flows and the appliance sum of every bus: needed by source and shunt records and
    // by the node's p and q (load_gen records form their own power from the injection and the voltage)
    for (int c = 0; c < N_COMP; ++c)
        if (want[c] && cols[c].empty() && !out[c] && model.count(c) > 0)
            out[c] = arena.take(size_t(n_scenarios) * size_t(model.count(c)) * rec_size[c]);
    if (bus_sums) launch_bus_results(P, d_outputs, d_bus_off, G, bus_total, st);
    if (want[C_NODE])
        launch_node_records(P, d_outputs, d_node_recs, d_rec_off[C_NODE], G, rec_total[C_NODE], out[C_NODE],
                            column_set(C_NODE), st);
    if (want[C_LINE])
        launch_branch_records(P, d_outputs, d_branch_recs[0], d_rec_off[C_LINE], G, rec_total[C_LINE], out[C_LINE],
                              column_set(C_LINE), st);
    if (want[C_TRANSFORMER])
        launch_branch_records(P, d_outputs, d_branch_recs[1], d_rec_off[C_TRANSFORMER], G, rec_total[C_TRANSFORMER],
                              out[C_TRANSFORMER], column_set(C_TRANSFORMER), st);
    if (want[C_LINK])
        launch_branch_records(P, d_outputs, d_branch_recs[2], d_rec_off[C_LINK], G, rec_total[C_LINK], out[C_LINK],
                              column_set(C_LINK), st);
    if (want[C_GENERIC_BRANCH])
        launch_branch_records(P, d_outputs, d_branch_recs[3], d_rec_off[C_GENERIC_BRANCH], G,
                              rec_total[C_GENERIC_BRANCH], out[C_GENERIC_BRANCH], column_set(C_GENERIC_BRANCH), st);
    for (int c : {int(C_SHUNT), int(C_SOURCE), int(C_SYM_GEN), int(C_ASYM_GEN), int(C_SYM_LOAD), int(C_ASYM_LOAD)})
        if (want[c])
            launch_appliance_records(P, d_outputs, d_appl_recs[c], d_rec_off[c], G, rec_total[c], out[c],
                                     column_set(c), st);
    PGM_HIP_CHECK(hipGetLastError());
    float ms = 0;
    if (timing) {
        PGM_HIP_CHECK(hipEventRecord(ev_stop, st));
        PGM_HIP_CHECK(hipEventSynchronize(ev_stop));
        PGM_HIP_CHECK(hipEventElapsedTime(&ms, ev_start, ev_stop));
    } else {
        PGM_HIP_CHECK(hipStreamSynchronize(st));
    }
    kernel_ms = ms;
    double fwd_ms = 0.0;
    for (size_t i = 0; i + 1 < fwd_events.size(); i += 2) {
        float e = 0;
        PGM_HIP_CHECK(hipEventElapsedTime(&e, fwd_events[i], fwd_events[i + 1]));
        fwd_ms += e;
    }
    fwd_ms += fwd_graph_ms;
    fwd_timed = Idx(fwd_events.size() / 2) + fwd_graph_timed;
    fwd_launch_ms = fwd_timed > 0 ? fwd_ms / double(fwd_timed) : 0.0;
    fwd_total_ms = fwd_ms;
    ran = true;
}

int Batch::refine_held_steps() {
    hipStream_t const st = stream;
    // the accumulated steps and the residual denominators of the plain and packed groups (union groups
    // perturb no pivots), made at the first refinement of the batch
    if (!refine_ready) {
        for (auto& gp : groups) {
            Group& g = *gp;
            if (g.is_union || g.dg.xacc) continue;
            size_t const cnt = size_t(g.plan.n) * 2 * size_t(P) * size_t(g.S_pad);
            g.dg.xacc = dev_alloc<double>(arena, cnt);
            g.dg.rden = dev_alloc<double>(arena, cnt);
        }
        std::vector<DevGroup> dg;
        for (auto& gp : groups) dg.push_back(gp->dg);
        PGM_HIP_CHECK(hipMemcpy(d_groups, dg.data(), dg.size() * sizeof(DevGroup), hipMemcpyHostToDevice));
        refine_ready = true;
    }
    int const G = int(groups.size());
    int const L = int(n_items.size()), LP = int(n_items_p.size());
    launch_refine_start(pairs, d_groups, d_bus_off, G, bus_total, P, st);
    for (int pass = 0;; ++pass) {
        PGM_HIP_CHECK(hipMemsetAsync(d_gmax, 0, sizeof(unsigned long long), st));
        launch_refine_residual(pairs, d_groups, d_bus_off, G, bus_total, P, d_gmax, st);
        unsigned long long bits = 0;
        PGM_HIP_CHECK(hipMemcpyAsync(&bits, d_gmax, sizeof bits, hipMemcpyDeviceToHost, st));
        PGM_HIP_CHECK(hipStreamSynchronize(st));
        double berr = 0.0;
        std::memcpy(&berr, &bits, sizeof berr);
        ++refine_passes;
        if (berr <= (refine_steps ? kRefineTolSteps : kRefineTol) || pass == kRefineMaxPasses) break;
        // the correction solved with the stored (perturbed) factors, added to the accumulated step
        for (int l = 0; l < L; ++l) launch_forward(d_groups, d_items[l], n_items[l], P, fp32, MODE_REFINE, st);
        for (int l = 0; l < LP; ++l)
            launch_forward(d_groups, d_items_p[l], n_items_p[l], P, fp32, MODE_REFINE, st, KIND_PACKED);
        for (int l = LP - 1; l >= 0; --l)
            launch_backward(d_groups, d_items_p[l], n_items_p[l], P, fp32, MODE_REFINE, st, KIND_PACKED);
        for (int l = L - 1; l >= 0; --l) launch_backward(d_groups, d_items[l], n_items[l], P, fp32, MODE_REFINE, st);
    }
    h_active[0] = h_active[1] = 0;
    PGM_HIP_CHECK(hipMemcpyAsync(d_active, h_active, 2 * sizeof(int), hipMemcpyHostToDevice, st));
    launch_refine_finish(pairs, d_groups, d_bus_off, G, bus_total, P, int(max_iter), err_tol, d_active, st);
    PGM_HIP_CHECK(hipMemcpyAsync(h_active, d_active, sizeof(int), hipMemcpyDeviceToHost, st));
    PGM_HIP_CHECK(hipStreamSynchronize(st));
    PGM_HIP_CHECK(hipGetLastError());
    ++refinements;
    return h_active[0];
}

void Batch::destroy_graphs() {
    for (int k = 0; k < 2; ++k) {
        if (nr_exec[k]) (void)hipGraphExecDestroy(nr_exec[k]);
        nr_exec[k] = nullptr;
        for (auto e : nr_ev[k]) (void)hipEventDestroy(e);
        nr_ev[k].clear();
    }
}

void Batch::adjoint(double* grad_p, double* grad_q, double* grad_r, double* grad_x) {
    if (!ran) throw PowerGridError("the batch has not been run\n");
    if (!symmetric || linear || current)
        throw PowerGridError("the adjoint is implemented for symmetric Newton-Raphson batches\n");
    if (!model.link.empty() || !model.gbranch.empty())
        throw PowerGridError("the adjoint is implemented for grids without links and generic branches\n");
    for (auto const& gp : groups)
        if (gp->is_union || gp->packed)
            throw PowerGridError("the adjoint needs plain topology groups: build the batch with the adjoint layout "
                                 "(PGM_amd_set_adjoint_layout), or with more than 32 scenarios in every topology\n");
    Model const& m = model;
    Idx const nlg = m.n_lg(), nl = m.n_line();
    size_t const n_grad = size_t(n_scenarios) * size_t(2 * nlg + 2 * nl);
    if (!d_adj) {
        // per group: line incidences, node counts, branch admittances and their derivatives by r1, x1,
        // load and line maps (host, once)
        std::vector<DevAdjoint> adj;
        std::vector<long long> off{0};
        d_grad = dev_alloc<double>(arena, std::max<size_t>(n_grad, 1));
        for (auto const& gp : groups) {
            Group const& g = *gp;
            Plan const& p = g.plan;
            Topology const& t = g.topo;
            Idx const n = p.n, nbr = Idx(p.branch_comp.size());
            DevAdjoint a;
            a.n_lg = int(nlg);
            a.n_line = int(nl);
            std::vector<int> ptr(n + 1, 0), inc;
            std::vector<std::vector<int>> lists(n);
            for (Idx mb = 0; mb < nbr; ++mb) {
                if (p.branch_comp[mb] >= nl) continue;
                int const f = p.branch_from[mb], to = p.branch_to[mb];
                if (f >= 0) lists[f].push_back(int(mb));
                if (to >= 0 && to != f) lists[to].push_back(int(mb));
            }
            for (Idx r = 0; r < n; ++r) {
                ptr[r + 1] = ptr[r] + int(lists[r].size());
                inc.insert(inc.end(), lists[r].begin(), lists[r].end());
            }
            std::vector<int> node_w(n, 0);
            for (Idx i = 0; i < m.n_node(); ++i) {
                Idx const gg = t.node_group[i];
                if (gg >= 0) ++node_w[p.bus_label[p.mm_bus_begin[gg] + int(t.node_pos[i])]];
            }
            std::vector<double> dr(size_t(nbr) * 8, 0.0), dx(size_t(nbr) * 8, 0.0);
            for (Idx mb = 0; mb < nbr; ++mb) {
                Idx const c = p.branch_comp[mb];
                if (c >= nl) continue;
                LineStatic const& ls = m.lstat[c];
                cd const ys = ls.y1s, ysh = ls.y1sh;
                double const u1 = m.node[m.line_f[c]].u_rated;
                double const base_y = (kBaseP3 / u1 / kSqrt3) / (u1 / kSqrt3);
                // y_series = 1 / ((r1 + j x1) base_y): d/dr1 = -ys^2 base_y, d/dx1 = -j ys^2 base_y
                cd const d_r = -ys * ys * base_y, d_x = cd(0.0, -1.0) * ys * ys * base_y;
                bool const fs = g.st.line_fs[c], ts = g.st.line_ts[c];
                std::array<cd, 4> dy{};  // d(ff, ft, tf, tt) / d y_series (line pi model, tap 1)
                if (fs && ts) {
                    dy = {cd(1.0), cd(-1.0), cd(-1.0), cd(1.0)};
                } else if ((fs || ts) && std::abs(ysh) >= kNumTol) {
                    cd const den = 1.0 / ys + 2.0 / ysh;
                    cd const dsh = (1.0 / (ys * ys)) / (den * den);
                    dy[fs ? 0 : 3] = dsh;
                }
                for (int w = 0; w < 4; ++w) {
                    cd const a_r = dy[w] * d_r, a_x = dy[w] * d_x;
                    dr[size_t(mb) * 8 + 2 * w] = a_r.real();
                    dr[size_t(mb) * 8 + 2 * w + 1] = a_r.imag();
                    dx[size_t(mb) * 8 + 2 * w] = a_x.real();
                    dx[size_t(mb) * 8 + 2 * w + 1] = a_x.imag();
                }
            }
            std::vector<int> lg_math(nlg, -1), load_bus(p.load_comp.size(), -1), line_math(nl, -1);
            for (size_t k = 0; k < p.load_comp.size(); ++k) lg_math[p.load_comp[k]] = int(k);
            for (Idx bus = 0; bus < n; ++bus)
                for (int l = p.load_ptr[bus]; l < p.load_ptr[bus + 1]; ++l) load_bus[l] = int(bus);
            std::vector<int> mm_branch_begin(t.models.size() + 1, 0);
            for (size_t k = 0; k < t.models.size(); ++k)
                mm_branch_begin[k + 1] = mm_branch_begin[k] + int(t.models[k].branch.size());
            for (Idx c = 0; c < nl; ++c)
                if (t.branch_group[c] >= 0) line_math[c] = mm_branch_begin[t.branch_group[c]] + int(t.branch_pos[c]);
            std::vector<int8_t> gen(nlg);
            for (Idx l = 0; l < nlg; ++l) gen[l] = m.lg_gen[l] ? 1 : 0;
            a.bus_line_ptr = put(arena, ptr);
            a.bus_line = put(arena, inc);
            a.node_w = put(arena, node_w);
            a.br_from = put(arena, p.branch_from);
            a.br_to = put(arena, p.branch_to);
            a.br_y = put(arena, complex_to_doubles(p.branch_y));
            a.br_dr = put(arena, dr);
            a.br_dx = put(arena, dx);
            a.lg_math = put(arena, lg_math);
            a.load_bus = put(arena, load_bus);
            a.line_math = put(arena, line_math);
            a.lg_gen = put(arena, gen);
            a.load_status = g.load_status;
            a.grad_p = d_grad;
            a.grad_q = d_grad + size_t(n_scenarios) * nlg;
            a.grad_r = d_grad + size_t(n_scenarios) * 2 * nlg;
            a.grad_x = d_grad + size_t(n_scenarios) * (2 * nlg + nl);
            adj.push_back(a);
            off.push_back(off.back() + (long long)(nlg + nl) * g.S_pad);
        }
        d_adj = put(arena, adj);
        d_adj_off = put(arena, off);
        adj_total = off.back();
        // the adjoint always factorises in FP64: a mixed-precision batch gets FP64 factor storage of its
        // own for it (the gradients then carry no FP32 solve error)
        if (fp32) {
            std::vector<DevGroup> dg;
            for (auto const& gp : groups) {
                DevGroup d = gp->dg;
                d.lu = arena.take(size_t(d.E) * 4 * size_t(d.S_pad) * sizeof(double));
                d.fp32 = 0;
                dg.push_back(d);
            }
            d_groups_adj = put(arena, dg);
        } else {
            d_groups_adj = d_groups;
        }
    }
    hipStream_t const st = stream;
    if (n_grad > 0) PGM_HIP_CHECK(hipMemsetAsync(d_grad, 0, n_grad * sizeof(double), st));
    int const L = int(n_items.size());
    DevGroup const* const dga = d_groups_adj;
    // the Jacobian at the converged state, every block kept
    for (int l = 0; l < L; ++l) launch_forward(dga, d_items[l], n_items[l], P, false, MODE_NRA, st);
    launch_adjoint_rhs(dga, d_adj, d_bus_off, int(groups.size()), bus_total, st);
    for (int l = 0; l < L; ++l) launch_adjoint_sweep(dga, d_items[l], n_items[l], false, true, st);
    for (int l = L - 1; l >= 0; --l) launch_adjoint_sweep(dga, d_items[l], n_items[l], false, false, st);
    launch_adjoint_grad(dga, d_outputs, d_adj, d_adj_off, int(groups.size()), adj_total, st);
    PGM_HIP_CHECK(hipGetLastError());
    PGM_HIP_CHECK(hipStreamSynchronize(st));
    size_t const a = size_t(n_scenarios) * nlg, b = size_t(n_scenarios) * nl;
    if (a > 0) {
        PGM_HIP_CHECK(hipMemcpy(grad_p, d_grad, a * 8, hipMemcpyDeviceToHost));
        PGM_HIP_CHECK(hipMemcpy(grad_q, d_grad + a, a * 8, hipMemcpyDeviceToHost));
    }
    if (b > 0) {
        PGM_HIP_CHECK(hipMemcpy(grad_r, d_grad + 2 * a, b * 8, hipMemcpyDeviceToHost));
        PGM_HIP_CHECK(hipMemcpy(grad_x, d_grad + 2 * a + b, b * 8, hipMemcpyDeviceToHost));
    }
}

namespace {
void check_output_dataset(PGM_MutableDataset const& output, bool symmetric, Idx n_scenarios) {
    bool const want_asym = output.meta->name == "asym_output";
    if (output.meta->name != "sym_output" && output.meta->name != "asym_output")
        throw PowerGridError("output dataset must be sym_output or asym_output\n");
    if (want_asym == symmetric) throw PowerGridError("output dataset symmetry does not match the calculation\n");
    if (output.batch_size != n_scenarios)
        throw PowerGridError("output batch size does not match the number of scenarios\n");
}
int comp_slot(std::string const& name) {
    for (int c = 0; c < N_COMP; ++c)
        if (name == comp_names[c]) return c;
    return -1;
}
}  // namespace

Batch::~Batch() {
    for (void* p : col_mem) (void)hipFree(p);
    release();
}

void Batch::set_output(PGM_MutableDataset const& output) {
    check_output_dataset(output, symmetric, n_scenarios);
    for (void* p : col_mem) PGM_HIP_CHECK(hipFree(p));
    col_mem.clear();
    for (int c = 0; c < N_COMP; ++c) {
        want[c] = false;
        cols[c].clear();
    }
    for (auto const& buf : output.buffers) {
        int const comp = comp_slot(buf.meta->name);
        if (comp < 0) continue;
        want[comp] = true;
        if (!buf.columnar()) continue;
        if (buf.columns.size() > size_t(kMaxColumns)) throw PowerGridError("too many output columns\n");
        Idx const cnt = model.count(comp);
        for (auto const& [attr, host] : buf.columns) {
            void* d = nullptr;
            size_t const bytes = std::max<size_t>(size_t(n_scenarios) * size_t(cnt) * attr->size, 1);
            PGM_HIP_CHECK(hipMalloc(&d, bytes));
            col_mem.push_back(d);
            cols[comp].push_back(DevColumn{attr, d});
        }
    }
    ran = false;
}

void Batch::fetch(PGM_MutableDataset const& output, PGM_Handle* handle) {
    if (!ran) throw PowerGridError("the batch has not been run\n");
    check_output_dataset(output, symmetric, n_scenarios);
    std::vector<char> host;
    bool const all_ok = std::all_of(status.begin(), status.end(), [](int x) { return x == 0; });
    for (auto const& buf : output.buffers) {
        int const comp = comp_slot(buf.meta->name);
        if (comp < 0) continue;
        Idx const cnt = model.count(comp);
        if (cnt == 0) continue;
        if (!want[comp])
            throw PowerGridError("component " + buf.meta->name + " was not selected for output of this batch\n");
        if (!cols[comp].empty()) {
            // device columns: each requested column must have been produced
            if (!buf.columnar())
                throw PowerGridError("component " + buf.meta->name + " was selected as columns for this batch\n");
            for (auto const& [attr, hcol] : buf.columns) {
                void const* dcol = nullptr;
                for (auto const& dc : cols[comp])
                    if (dc.attr == attr) dcol = dc.dev;
                if (!dcol)
                    throw PowerGridError("attribute " + attr->name + " of " + buf.meta->name +
                                         " was not selected for output of this batch\n");
                size_t const as = attr->size;
                if (all_ok && buf.elements_per_scenario == cnt) {
                    copy_to_host(hcol, dcol, size_t(n_scenarios) * cnt * as);
                    continue;
                }
                host.resize(size_t(n_scenarios) * cnt * as);
                copy_to_host(host.data(), dcol, host.size());
                for (Idx s = 0; s < n_scenarios; ++s) {
                    if (status[s] != 0) continue;
                    Idx const b0 = buf.begin(s), b1 = buf.end(s);
                    if (b1 - b0 != cnt)
                        throw PowerGridError("output buffer of " + buf.meta->name +
                                             " must hold every element per scenario\n");
                    std::memcpy(static_cast<char*>(hcol) + size_t(b0) * as, host.data() + size_t(s) * cnt * as,
                                size_t(cnt) * as);
                }
            }
            continue;
        }
        size_t const rs = rec_size[comp];
        bool const direct = !buf.columnar() && buf.elements_per_scenario == cnt;
        if (direct && all_ok) {
            copy_to_host(buf.data, out[comp], size_t(n_scenarios) * cnt * rs);
            continue;
        }
        host.resize(size_t(n_scenarios) * cnt * rs);
        copy_to_host(host.data(), out[comp], host.size());
        for (Idx s = 0; s < n_scenarios; ++s) {
            if (status[s] != 0) continue;
            Idx const b0 = buf.begin(s), b1 = buf.end(s);
            if (b1 - b0 != cnt)
                throw PowerGridError("output buffer of " + buf.meta->name + " must hold every element per scenario\n");
            for (Idx i = 0; i < cnt; ++i) buf.write_record(b0 + i, host.data() + (size_t(s) * cnt + i) * rs);
        }
    }
    // errors (job_dispatch.hpp:208-224, model.cpp:205-220)
    if (handle) {
        std::vector<Idx> failed;
        std::string combined;
        std::vector<std::string> msgs;
        for (Idx s = 0; s < n_scenarios; ++s)
            if (status[s] != 0) {
                failed.push_back(s);
                msgs.push_back(messages[s]);
                combined += "Error in batch #" + std::to_string(s) + ": " + messages[s] + "\n";
            }
        if (!failed.empty()) {
            if (is_batch) {
                handle->err_code = PGM_batch_error;
                handle->err_msg = combined;
                handle->failed_scenarios = failed;
                handle->batch_errs = msgs;
            } else {
                handle->err_code = PGM_regular_error;
                handle->err_msg = messages[0];
            }
        }
    }
}

std::shared_ptr<PlanEntry const> PlanCache::find(std::string const& key) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = map_.find(key);
    return it == map_.end() ? nullptr : it->second;
}
void PlanCache::put(std::string const& key, std::shared_ptr<PlanEntry const> e) {
    std::lock_guard<std::mutex> lock(mu_);
    if (map_.size() >= 512) map_.clear();  // bounded: a long run of distinct states starts over
    map_[key] = std::move(e);
}
void PlanCache::clear() {
    std::lock_guard<std::mutex> lock(mu_);
    map_.clear();
}

BatchContext model_context(PGM_PowerGridModel const& model, PGM_Options const& opt, PGM_ConstDataset const* update) {
    BatchContext ctx;
    ctx.pool = model.pool;
    ctx.resident = model.resident.get();
    ctx.plans = model.plans.get();
    auto const tu = model.tuned.find(tune_key(opt, update));
    if (tu != model.tuned.end()) ctx.tuning = &tu->second;
    return ctx;
}

std::string tune_key(PGM_Options const& opt, PGM_ConstDataset const* update) {
    return std::to_string(update ? update->batch_size : 1) + (opt.symmetric ? "/sym" : "/asym") +
           (opt.factor_precision == 1 ? "/fp32" : "/fp64") + "/m" + std::to_string(int(opt.calculation_method)) +
           (opt.adjoint_layout ? "/adjoint" : "") + (opt.refine_steps ? "/refine" : "");
}

Tuning autotune(PGM_PowerGridModel& model, PGM_Options const& opt, PGM_ConstDataset const* update,
                std::vector<TuneTrial>& trials) {
    auto measure = [&](Tuning const& t) {
        BatchContext ctx = model_context(model, opt, update);
        ctx.tuning = &t;
        auto b = create_batch(model.model, opt, update, ctx);
        b->run();
        b->run();
        double const ms = b->kernel_ms;
        bool const has_union = std::any_of(b->groups.begin(), b->groups.end(), [](auto const& g) { return g->is_union; });
        bool const has_packed = std::any_of(b->groups.begin(), b->groups.end(), [](auto const& g) { return g->packed; });
        trials.push_back({t, ms});
        return std::make_tuple(ms, has_union, has_packed);
    };
    Tuning best{plain_waves(), union_tasks(), packed_target()};
    auto [best_ms, has_union, has_packed] = measure(best);
    auto try_values = [&](int Tuning::*field, std::initializer_list<int> values) {
        for (int v : values) {
            if (v == best.*field) continue;
            Tuning t = best;
            t.*field = v;
            double const ms = std::get<0>(measure(t));
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    };
    try_values(&Tuning::plain_waves, {3072, 6144, 12288, 24576, 49152});
    if (has_union) try_values(&Tuning::union_tasks, {64, 256, 1024});
    if (has_packed) try_values(&Tuning::packed_target, {64, 256, 1024});
    model.tuned[tune_key(opt, update)] = best;
    return best;
}

size_t make_resident(Model const& model, PGM_ConstDataset const& update, UploadCache& cache) {
    (void)model;
    ensure_device();
    size_t bytes = 0;
    for (auto const& buf : update.buffers) {
        std::string const& name = buf.meta->name;
        if (name != comp_names[C_SYM_GEN] && name != comp_names[C_ASYM_GEN] && name != comp_names[C_SYM_LOAD] &&
            name != comp_names[C_ASYM_LOAD])
            continue;
        // the same host ranges the batch's device load updates read (create_batch)
        if (buf.data) {
            size_t const n = size_t(buf.total_elements) * buf.meta->size;
            if (!cache.find(buf.data, n)) bytes += n;
            cache.get(buf.data, n);
            continue;
        }
        for (auto const& [a, col] : buf.columns) {
            if (a->name != "id" && a->name != "status" && a->name != "p_specified" && a->name != "q_specified")
                continue;
            size_t const n = size_t(buf.total_elements) * a->size;
            if (!cache.find(col, n)) bytes += n;
            cache.get(col, n);
        }
    }
    return bytes;
}

void batch_plan_sizes(Batch const& b, double& buses, double& entries) {
    double sb = 0, se = 0, ss = 0;
    for (auto const& g : b.groups) {
        sb += double(g->plan.n) * g->S;
        se += double(g->plan.E) * g->S;
        ss += g->S;
    }
    buses = ss > 0 ? sb / ss : 0;
    entries = ss > 0 ? se / ss : 0;
}

namespace {
// bump allocation inside a fixed device region (the allocation-free entry point's scratch); a request
// beyond the region throws before any kernel runs
struct ScratchArena {
    char* base;
    size_t cap;
    size_t used = 0;
    void* take(size_t bytes) {
        bytes = (bytes + 255) / 256 * 256;
        if (bytes == 0) return nullptr;
        if (used + bytes > cap)
            throw ScratchTooSmall("the block LU batch needs more than the " + std::to_string(cap >> 20) +
                                  " MiB device scratch of the allocation-free entry point\n");
        void* p = base + used;
        used += bytes;
        return p;
    }
};

template <class Ar>
int block_lu_impl(Ar& A, hipStream_t st, Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices,
                  Idx const* diag_lu, Idx batch, double const* data, double const* rhs, double* x, int32_t* status,
                  bool fp32);
}  // namespace

int device_block_lu(Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices, Idx const* diag_lu, Idx batch,
                    double const* data, double const* rhs, double* x, int32_t* status, bool fp32) {
    if (nb != 2 && nb != 6) return -1;
    ensure_device();
    hipStream_t st;
    PGM_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{st};
    Arena A;
    return block_lu_impl(A, st, nb, n, row_indptr, col_indices, diag_lu, batch, data, rhs, x, status, fp32);
}

int device_block_lu_noalloc(Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices, Idx const* diag_lu,
                            Idx batch, double const* data, double const* rhs, double* x, int32_t* status, bool fp32) {
    if (nb != 2 && nb != 6) return -1;
    ensure_device();
    void* base = nullptr;
    size_t cap = 0;
    if (scratch_region(&base, &cap) != 0) throw PowerGridError("the device scratch could not be located\n");
    ScratchArena A{static_cast<char*>(base), cap};
    // the default stream: no stream is created either
    return block_lu_impl(A, hipStream_t(nullptr), nb, n, row_indptr, col_indices, diag_lu, batch, data, rhs, x,
                         status, fp32);
}

namespace {
template <class Ar>
int block_lu_impl(Ar& A, hipStream_t st, Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices,
                  Idx const* diag_lu, Idx batch, double const* data, double const* rhs, double* x, int32_t* status,
                  bool fp32) {
    int const P = int(nb / 2), NB = int(nb), NB2 = NB * NB;
    Plan p;
    p.P = P;
    p.n = n;
    p.n_mm = 1;
    p.row_ptr.assign(row_indptr, row_indptr + n + 1);
    p.E = p.row_ptr[n];
    p.col.assign(col_indices, col_indices + p.E);
    p.diag.assign(diag_lu, diag_lu + n);
    p.fill.assign(p.E, 0);
    p.bus_mm.assign(n, 0);
    p.transpose.resize(p.E);
    for (Idx r = 0; r < n; ++r)
        for (int k = p.row_ptr[r]; k < p.row_ptr[r + 1]; ++k) {
            int const c = p.col[k];
            auto const b0 = p.col.begin() + p.row_ptr[c], b1 = p.col.begin() + p.row_ptr[c + 1];
            auto const it = std::lower_bound(b0, b1, int(r));
            if (it == b1 || *it != int(r)) return -1;  // the pattern must be symmetric
            p.transpose[k] = int(it - p.col.begin());
        }
    p.row_end.assign(p.row_ptr.begin() + 1, p.row_ptr.end());
    build_schedule(p, 16);
    int const S = int(batch), S_pad = S > 32 ? (S + 63) / 64 * 64 : S;
    size_t const sp = size_t(S_pad);
    DevGroup d;
    d.P = P;
    d.n = int(n);
    d.E = int(p.E);
    d.n_mm = 1;
    d.S = S;
    d.S_pad = S_pad;
    d.fp32 = fp32 ? 1 : 0;
    d.row_ptr = put(A, p.row_ptr);
    d.row_end = put(A, p.row_end);
    d.col = put(A, p.col);
    d.diag = put(A, p.diag);
    d.transpose = put(A, p.transpose);
    d.fill = put(A, p.fill);
    d.schur_ptr = put(A, p.schur_ptr);
    std::vector<int> sch(p.schur.size() * 3);
    for (size_t i = 0; i < p.schur.size(); ++i)
        for (int k = 0; k < 3; ++k) sch[3 * i + k] = p.schur[i][k];
    d.schur = put(A, sch);
    d.task_ptr = put(A, p.task_ptr);
    d.task_pivot = put(A, p.task_pivot);
    d.task_mm = put(A, p.task_mm);
    d.keep = put(A, p.keep);
    std::vector<double> a(size_t(p.E) * NB2 * sp, 0.0), bv(size_t(n) * NB * sp, 0.0);
    for (int s = 0; s < S; ++s) {
        for (Idx e = 0; e < p.E; ++e)
            for (int k = 0; k < NB2; ++k) a[(size_t(e) * NB2 + k) * sp + s] = data[(size_t(s) * p.E + e) * NB2 + k];
        for (Idx r = 0; r < n; ++r)
            for (int k = 0; k < NB; ++k) bv[(size_t(r) * NB + k) * sp + s] = rhs[(size_t(s) * n + r) * NB + k];
    }
    d.raw_a = put(A, a);
    d.raw_b = put(A, bv);
    d.y = dev_alloc<double>(A, size_t(n) * NB * sp);
    d.lu = A.take(size_t(p.E) * NB2 * sp * (fp32 ? 4 : 8));
    d.perm = dev_alloc<uint32_t>(A, size_t(n) * sp);
    PairState ps;
    ps.n = (long long)sp;
    ps.done = dev_alloc<int>(A, sp);
    ps.err = dev_alloc<int>(A, sp);
    ps.iters = dev_alloc<int>(A, sp);
    ps.devmax = dev_alloc<unsigned long long>(A, sp);
    ps.lastdev = dev_alloc<double>(A, sp);
    d.done = ps.done;
    d.err = ps.err;
    d.iters = ps.iters;
    d.devmax = ps.devmax;
    d.lastdev = ps.lastdev;
    DevGroup* dg = put(A, std::vector<DevGroup>{d});
    int const nblk = (S + 63) / 64;
    std::vector<WorkItem*> items;
    std::vector<int> counts;
    for (size_t l = 0; l + 1 < p.level_ptr.size(); ++l) {
        std::vector<WorkItem> it;
        for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q)
            for (int blk = 0; blk < nblk; ++blk) it.push_back(WorkItem{0, p.level_task[q], blk});
        items.push_back(put(A, it));
        counts.push_back(int(it.size()));
    }
    launch_init_state(ps, st);
    int const L = int(items.size());
    for (int l = 0; l < L; ++l) launch_forward(dg, items[l], counts[l], P, fp32, MODE_RAW, st);
    for (int l = L - 1; l >= 0; --l) launch_backward(dg, items[l], counts[l], P, fp32, MODE_RAW, st);
    PGM_HIP_CHECK(hipGetLastError());
    std::vector<double> yh(bv.size());
    std::vector<int> eh(sp);
    PGM_HIP_CHECK(hipMemcpyAsync(yh.data(), d.y, yh.size() * 8, hipMemcpyDeviceToHost, st));
    PGM_HIP_CHECK(hipMemcpyAsync(eh.data(), ps.err, eh.size() * 4, hipMemcpyDeviceToHost, st));
    PGM_HIP_CHECK(hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s) {
        status[s] = eh[s] == ERR_SPARSE ? 1 : 0;
        for (Idx r = 0; r < n; ++r)
            for (int k = 0; k < NB; ++k) x[(size_t(s) * n + r) * NB + k] = yh[(size_t(r) * NB + k) * sp + s];
    }
    return 0;
}
}  // namespace

}  // namespace pgm_amd

PGM_PowerGridModel::PGM_PowerGridModel()
    : pool(std::make_shared<pgm_amd::ChunkPool>()),
      resident(std::make_unique<pgm_amd::UploadCache>()),
      plans(std::make_shared<pgm_amd::PlanCache>()) {}
PGM_PowerGridModel::PGM_PowerGridModel(PGM_PowerGridModel const& other)
    : model(other.model),
      pool(std::make_shared<pgm_amd::ChunkPool>()),
      resident(std::make_unique<pgm_amd::UploadCache>()),
      plans(std::make_shared<pgm_amd::PlanCache>()) {}
PGM_PowerGridModel::~PGM_PowerGridModel() = default;
