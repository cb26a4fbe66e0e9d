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
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <sstream>

namespace pgm_amd {

#define PGM_HIP_CHECK(expr)                                                                                      \
    do {                                                                                                         \
        hipError_t const e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                                    \
            throw PowerGridError(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_) + "\n");    \
    } while (0)

void DeviceBuffer::alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    PGM_HIP_CHECK(hipMalloc(&ptr_, bytes));
    bytes_ = bytes;
}
void DeviceBuffer::release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
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

void ensure_device() {
    int n = 0;
    hipError_t const e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        throw PowerGridError("No HIP device available: this library runs the power flow on an MI355X (gfx950) GPU\n");
}

// blocking copy: the host vectors are often temporaries
template <class T> void upload(DeviceBuffer& d, std::vector<T> const& h, hipStream_t st) {
    (void)st;
    d.alloc(h.size() * sizeof(T));
    if (!h.empty()) PGM_HIP_CHECK(hipMemcpy(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
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
    // device: structure
    DeviceBuffer row_ptr, col, diag, transpose, fill, ydata, schur_ptr, schur, task_ptr, task_pivot, task_mm,
        level_task, load_ptr, load_type, src_ptr, shunt_ptr, load_bus, src_bus, shunt_bus, shunt_y;
    // device: per scenario
    DeviceBuffer u, xv, y, lu, perm, s_inj, u_src, y_src, done, err, iters, devmax, lastdev, active;
    DeviceBuffer scenario_index, ok, load_status, load_si, src_si, shunt_si, bus_inj;
    // device: record maps
    DeviceBuffer node_id, node_urated, node_bus;
    DeviceBuffer br_id[2], br_math[2], br_bif[2], br_bit[2], br_rating[2], br_from, br_to, br_y;
    DeviceBuffer ap_id[N_COMP], ap_math[N_COMP], ap_base_i[N_COMP], ap_status_const[N_COMP];
    DevGroup dg;
    DevOutput dout;
    std::vector<int> level_ptr;
    int* h_active = nullptr;  // pinned
    Group() = default;
    Group(Group const&) = delete;
    Group& operator=(Group const&) = delete;
    ~Group() {
        if (h_active) (void)hipHostFree(h_active);
    }
};

namespace {

std::string structural_key(Model const& m, ComponentState const& st) {
    // the components whose state differs from the base define the group
    std::ostringstream k;
    auto const& b = m.base;
    for (Idx i = 0; i < m.n_line(); ++i)
        if (st.line_fs[i] != b.line_fs[i] || st.line_ts[i] != b.line_ts[i])
            k << 'l' << i << ':' << int(st.line_fs[i]) << int(st.line_ts[i]) << ';';
    for (Idx i = 0; i < Idx(m.trafo.size()); ++i)
        if (st.tr_fs[i] != b.tr_fs[i] || st.tr_ts[i] != b.tr_ts[i] || st.tr_tap[i] != b.tr_tap[i])
            k << 't' << i << ':' << int(st.tr_fs[i]) << int(st.tr_ts[i]) << ',' << int(st.tr_tap[i]) << ';';
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

// the records of component `comp` of scenario s, as (index within the component, record pointer)
struct UpdateReader {
    ComponentBuffer const* buf = nullptr;
    std::vector<char> tmp;
    int comp = 0;
    bool all_na_ids = false;
};

}  // namespace

Batch::~Batch() {
    groups.clear();
    for (auto e : event_pool) (void)hipEventDestroy(e);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (ev_fwd0) (void)hipEventDestroy(ev_fwd0);
    if (ev_fwd1) (void)hipEventDestroy(ev_fwd1);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

void upload_group(Batch& b, Group& g, Model const& m) {
    Plan const& p = g.plan;
    hipStream_t const st = b.stream;
    int const P = p.P, NB = 2 * P;
    g.S = int(g.scenarios.size());
    g.S_pad = (g.S + 63) / 64 * 64;
    size_t const sp = size_t(g.S_pad);
    upload(g.row_ptr, p.row_ptr, st);
    upload(g.col, p.col, st);
    upload(g.diag, p.diag, st);
    upload(g.transpose, p.transpose, st);
    upload(g.fill, p.fill, st);
    std::vector<double> yd(p.ydata.size() * 2);
    for (size_t i = 0; i < p.ydata.size(); ++i) {
        yd[2 * i] = p.ydata[i].real();
        yd[2 * i + 1] = p.ydata[i].imag();
    }
    upload(g.ydata, yd, st);
    upload(g.schur_ptr, p.schur_ptr, st);
    std::vector<int> sch(p.schur.size() * 3);
    for (size_t i = 0; i < p.schur.size(); ++i)
        for (int k = 0; k < 3; ++k) sch[3 * i + k] = p.schur[i][k];
    upload(g.schur, sch, st);
    upload(g.task_ptr, p.task_ptr, st);
    upload(g.task_pivot, p.task_pivot, st);
    upload(g.task_mm, p.task_mm, st);
    upload(g.level_task, p.level_task, st);
    g.level_ptr = p.level_ptr;
    upload(g.load_ptr, p.load_ptr, st);
    upload(g.load_type, p.load_type, st);
    upload(g.src_ptr, p.src_ptr, st);
    upload(g.shunt_ptr, p.shunt_ptr, st);
    g.n_load = int(p.load_comp.size());
    g.n_src = int(p.src_comp.size());
    g.n_shunt = int(p.shunt_comp.size());
    std::vector<double> shy(p.shunt_y.size() * 2);
    for (size_t i = 0; i < p.shunt_y.size(); ++i) {
        shy[2 * i] = p.shunt_y[i].real();
        shy[2 * i + 1] = p.shunt_y[i].imag();
    }
    upload(g.shunt_y, shy, st);

    size_t const n = size_t(p.n), E = size_t(p.E);
    size_t const fsz = b.fp32 ? sizeof(float) : sizeof(double);
    g.u.alloc(n * P * 2 * sp * 8);
    g.xv.alloc(n * P * 2 * sp * 8);
    g.y.alloc(n * NB * sp * 8);
    g.lu.alloc(E * NB * NB * sp * fsz);
    g.perm.alloc(n * sp * 4);
    g.done.alloc(size_t(p.n_mm) * sp * 4);
    g.err.alloc(size_t(p.n_mm) * sp * 4);
    g.iters.alloc(size_t(p.n_mm) * sp * 4);
    g.devmax.alloc(size_t(p.n_mm) * sp * 8);
    g.lastdev.alloc(size_t(p.n_mm) * sp * 8);
    g.active.alloc(4);
    g.load_si.alloc(size_t(g.n_load) * 2 * P * 2 * sp * 8);
    g.src_si.alloc(size_t(g.n_src) * 2 * P * 2 * sp * 8);
    g.shunt_si.alloc(size_t(g.n_shunt) * 2 * P * 2 * sp * 8);
    g.bus_inj.alloc(n * P * 2 * sp * 8);
    PGM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&g.h_active), sizeof(int)));
    std::vector<int> si(g.scenarios.begin(), g.scenarios.end());
    si.resize(g.S_pad, 0);
    upload(g.scenario_index, si, st);
    g.ok.alloc(sp * 4);

    DevGroup& d = g.dg;
    d.P = P;
    d.n = int(n);
    d.E = int(E);
    d.n_mm = int(p.n_mm);
    d.S = g.S;
    d.S_pad = g.S_pad;
    d.fp32 = b.fp32 ? 1 : 0;
    d.row_ptr = g.row_ptr.as<int>();
    d.col = g.col.as<int>();
    d.diag = g.diag.as<int>();
    d.transpose = g.transpose.as<int>();
    d.fill = g.fill.as<int8_t>();
    d.ydata = g.ydata.as<double>();
    d.schur_ptr = g.schur_ptr.as<int>();
    d.schur = g.schur.as<int>();
    d.task_ptr = g.task_ptr.as<int>();
    d.task_pivot = g.task_pivot.as<int>();
    d.task_mm = g.task_mm.as<int>();
    d.load_ptr = g.load_ptr.as<int>();
    d.load_type = g.load_type.as<int8_t>();
    d.src_ptr = g.src_ptr.as<int>();
    d.u = g.u.as<double>();
    d.xv = g.xv.as<double>();
    d.y = g.y.as<double>();
    d.lu = g.lu.get();
    d.perm = g.perm.as<uint32_t>();
    d.s_inj = g.s_inj.as<double>();
    d.u_src = g.u_src.as<double>();
    d.y_src = g.y_src.as<double>();
    d.done = g.done.as<int>();
    d.err = g.err.as<int>();
    d.iters = g.iters.as<int>();
    d.devmax = g.devmax.as<unsigned long long>();
    d.lastdev = g.lastdev.as<double>();

    DevOutput& o = g.dout;
    o.P = P;
    o.S = g.S;
    o.S_pad = g.S_pad;
    o.scenario_index = g.scenario_index.as<int>();
    o.ok = g.ok.as<int>();
    o.n = int(n);
    o.row_ptr = d.row_ptr;
    o.col = d.col;
    o.fill = d.fill;
    o.ydata = d.ydata;
    o.u = d.u;
    o.n_load = g.n_load;
    o.load_type = d.load_type;
    o.s_inj = d.s_inj;
    o.load_si = g.load_si.as<double>();
    o.n_src = g.n_src;
    o.src_ptr = d.src_ptr;
    o.load_ptr = d.load_ptr;
    o.shunt_ptr = g.shunt_ptr.as<int>();
    o.u_src = d.u_src;
    o.y_src = d.y_src;
    o.src_si = g.src_si.as<double>();
    o.bus_inj = g.bus_inj.as<double>();
    o.n_shunt = g.n_shunt;
    o.shunt_y = g.shunt_y.as<double>();
    o.shunt_si = g.shunt_si.as<double>();

    // record maps
    Topology const& t = g.topo;
    std::vector<int> mm_rank(t.models.size(), -1);  // math model -> position in the plan
    // the plan holds all models of the topology in order
    for (size_t k = 0; k < t.models.size(); ++k) mm_rank[k] = int(k);
    {
        std::vector<int> ids, bus;
        std::vector<double> ur;
        for (Idx i = 0; i < m.n_node(); ++i) {
            ids.push_back(m.node[i].id);
            ur.push_back(m.node[i].u_rated);
            Idx const gg = t.node_group[i];
            bus.push_back(gg < 0 ? -1 : p.mm_bus_begin[mm_rank[gg]] + int(t.node_pos[i]));
        }
        upload(g.node_id, ids, st);
        upload(g.node_urated, ur, st);
        upload(g.node_bus, bus, st);
    }
    {
        // math branch index in plan order: models concatenated
        std::vector<int> mm_branch_begin(t.models.size() + 1, 0);
        for (size_t k = 0; k < t.models.size(); ++k)
            mm_branch_begin[k + 1] = mm_branch_begin[k] + int(t.models[k].branch.size());
        for (int kind = 0; kind < 2; ++kind) {
            Idx const first = kind == 0 ? 0 : m.n_line();
            Idx const cnt = kind == 0 ? m.n_line() : Idx(m.trafo.size());
            std::vector<int> ids, math;
            std::vector<double> bif, bit, rating;
            for (Idx k = 0; k < cnt; ++k) {
                Idx const bidx = first + k;
                ids.push_back(kind == 0 ? m.line[k].id : m.trafo[k].id);
                Idx const gg = t.branch_group[bidx];
                math.push_back(gg < 0 ? -1 : mm_branch_begin[mm_rank[gg]] + int(t.branch_pos[bidx]));
                if (kind == 0) {
                    bif.push_back(m.lstat[k].base_i);
                    bit.push_back(m.lstat[k].base_i);
                    rating.push_back(m.lstat[k].i_n);
                } else {
                    bif.push_back(m.tstat[k].base_i_from);
                    bit.push_back(m.tstat[k].base_i_to);
                    rating.push_back(m.tstat[k].sn);
                }
            }
            upload(g.br_id[kind], ids, st);
            upload(g.br_math[kind], math, st);
            upload(g.br_bif[kind], bif, st);
            upload(g.br_bit[kind], bit, st);
            upload(g.br_rating[kind], rating, st);
        }
        upload(g.br_from, p.branch_from, st);
        upload(g.br_to, p.branch_to, st);
        std::vector<double> by(p.branch_y.size() * 2);
        for (size_t i = 0; i < p.branch_y.size(); ++i) {
            by[2 * i] = p.branch_y[i].real();
            by[2 * i + 1] = p.branch_y[i].imag();
        }
        upload(g.br_y, by, st);
    }
    {
        // appliance maps: index into the plan's load / source / shunt arrays
        std::vector<int> lg_math(m.n_lg(), -1), src_math(m.source.size(), -1), sh_math(m.shunt.size(), -1);
        for (size_t k = 0; k < p.load_comp.size(); ++k) lg_math[p.load_comp[k]] = int(k);
        for (size_t k = 0; k < p.src_comp.size(); ++k) src_math[p.src_comp[k]] = int(k);
        for (size_t k = 0; k < p.shunt_comp.size(); ++k) sh_math[p.shunt_comp[k]] = int(k);
        auto set = [&](int comp, std::vector<int> const& ids, std::vector<int> const& math,
                       std::vector<double> const& base_i, std::vector<int8_t> const& status) {
            upload(g.ap_id[comp], ids, st);
            upload(g.ap_math[comp], math, st);
            upload(g.ap_base_i[comp], base_i, st);
            upload(g.ap_status_const[comp], status, st);
        };
        {
            std::vector<int> ids, math;
            std::vector<double> bi;
            std::vector<int8_t> stt;
            for (size_t k = 0; k < m.shunt.size(); ++k) {
                ids.push_back(m.shunt[k].id);
                math.push_back(sh_math[k]);
                bi.push_back(kBaseP3 / m.node[m.sh_node[k]].u_rated / kSqrt3);
                stt.push_back(g.st.sh_status[k] ? 1 : 0);
            }
            set(C_SHUNT, ids, math, bi, stt);
        }
        {
            std::vector<int> ids, math;
            std::vector<double> bi;
            std::vector<int8_t> stt;
            for (size_t k = 0; k < m.source.size(); ++k) {
                ids.push_back(m.source[k].id);
                math.push_back(src_math[k]);
                bi.push_back(kBaseP3 / m.node[m.src_node[k]].u_rated / kSqrt3);
                stt.push_back(g.st.src_status[k] ? 1 : 0);
            }
            set(C_SOURCE, ids, math, bi, stt);
        }
        int const lg_comps[4] = {C_SYM_GEN, C_ASYM_GEN, C_SYM_LOAD, C_ASYM_LOAD};
        for (int kk = 0; kk < 4; ++kk) {
            std::vector<int> ids, math;
            std::vector<double> bi;
            std::vector<int8_t> stt;
            for (Idx k = 0; k < m.lg_count[kk]; ++k) {
                Idx const l = m.lg_offset[kk] + k;
                ids.push_back(m.lg_id[l]);
                math.push_back(lg_math[l]);
                bi.push_back(kBaseP3 / m.node[m.lg_node[l]].u_rated / kSqrt3);
                stt.push_back(0);
            }
            set(lg_comps[kk], ids, math, bi, stt);
        }
    }
}

}  // namespace

std::unique_ptr<Batch> create_batch(Model const& model, PGM_Options const& opt, PGM_ConstDataset const* update) {
    if (opt.calculation_type != PGM_power_flow)
        throw PowerGridError("The calculation type is not supported by this build: only power flow is implemented\n");
    if (opt.calculation_method != PGM_default_method && opt.calculation_method != PGM_newton_raphson)
        throw PowerGridError("The calculation method is invalid for this calculation!");
    if (update && update->next) throw PowerGridError("Cartesian product batches are not supported by this build\n");
    ensure_device();
    auto b = std::make_unique<Batch>();
    b->model = model;
    Model const& m = b->model;
    b->symmetric = opt.symmetric != 0;
    b->P = b->symmetric ? 1 : 3;
    b->err_tol = opt.err_tol;
    b->max_iter = opt.max_iter;
    b->fp32 = opt.factor_precision == 1;
    b->is_batch = update != nullptr;
    b->n_scenarios = update ? update->batch_size : 1;
    PGM_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    PGM_HIP_CHECK(hipEventCreate(&b->ev_start));
    PGM_HIP_CHECK(hipEventCreate(&b->ev_stop));
    PGM_HIP_CHECK(hipEventCreate(&b->ev_fwd0));
    PGM_HIP_CHECK(hipEventCreate(&b->ev_fwd1));
    Idx const S = b->n_scenarios;
    bool const sym = b->symmetric;
    int const P = b->P;

    // ---- update readers ----
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
            r.tmp.resize(buf.meta->size);
            // ids: all NA means positional update (update.hpp:60-90)
            bool all_na = true, any_na = false;
            for (Idx i = 0; i < buf.total_elements; ++i) {
                buf.read_record(i, r.tmp.data());
                ID id;
                std::memcpy(&id, r.tmp.data(), 4);
                all_na = all_na && id == na_IntID;
                any_na = any_na || id == na_IntID;
            }
            if (buf.total_elements > 0 && any_na && !all_na)
                throw PowerGridError("IDs contain both numbers and NANs for component " + buf.meta->name +
                                     " in update data!");
            r.all_na_ids = all_na && buf.total_elements > 0;
            if (r.all_na_ids && (buf.elements_per_scenario < 0 || buf.elements_per_scenario != m.count(comp)))
                throw PowerGridError("Update data without IDs for component " + buf.meta->name +
                                     " has a different number of elements per scenario then input data!");
            readers.push_back(std::move(r));
        }
    }
    auto for_each_row = [&](Idx s, auto&& fn) {
        for (auto& r : readers) {
            for (Idx i = r.buf->begin(s); i < r.buf->end(s); ++i) {
                r.buf->read_record(i, r.tmp.data());
                ID id;
                std::memcpy(&id, r.tmp.data(), 4);
                Idx const idx = r.all_na_ids ? i - r.buf->begin(s) : find_component(m, r.comp, id);
                fn(r.comp, idx, r.tmp.data());
            }
        }
    };

    // ---- group scenarios by structural state ----
    std::map<std::string, int> key_to_group;
    std::vector<int> scn_group(S), scn_lane(S);
    std::vector<std::string> scn_error(S);
    for (Idx s = 0; s < S; ++s) {
        ComponentState st = m.base;
        bool structural = false;
        try {
            for_each_row(s, [&](int comp, Idx idx, void const* rec) {
                if (comp == C_LINE || comp == C_TRANSFORMER || comp == C_SHUNT || comp == C_SOURCE)
                    apply_update_record(m, st, comp, idx, rec, structural);
            });
        } catch (PowerGridError const& e) {
            if (!update) throw;
            scn_error[s] = e.what();
            scn_group[s] = -1;
            continue;
        }
        std::string const key = structural ? structural_key(m, st) : std::string{};
        auto it = key_to_group.find(key);
        int gi;
        if (it == key_to_group.end()) {
            gi = int(b->groups.size());
            key_to_group.emplace(key, gi);
            auto g = std::make_unique<Group>();
            g->st = st;
            b->groups.push_back(std::move(g));
        } else {
            gi = it->second;
        }
        scn_group[s] = gi;
        scn_lane[s] = int(b->groups[gi]->scenarios.size());
        b->groups[gi]->scenarios.push_back(int(s));
    }
    if (!update && !scn_error[0].empty()) throw PowerGridError(scn_error[0]);

    // ---- per group: topology, plan, upload ----
    for (auto& gp : b->groups) {
        Group& g = *gp;
        g.topo = build_topology(m, g.st);
        std::vector<Idx> all(g.topo.models.size());
        std::iota(all.begin(), all.end(), Idx{0});
        int const nblk = int((g.scenarios.size() + 63) / 64);
        int const target = std::clamp(4096 / std::max(nblk, 1), 4, 256);
        g.plan = build_plan(m, g.st, g.topo, all, sym, target);
        upload_group(*b, g, m);
    }

    // ---- per-scenario injections, source voltages and admittances ----
    for (auto& gp : b->groups) {
        Group& g = *gp;
        Plan const& p = g.plan;
        size_t const sp = size_t(g.S_pad);
        std::vector<double> s_inj(size_t(g.n_load) * P * 2 * sp, 0.0);
        std::vector<int8_t> lstat(size_t(m.n_lg()) * sp, 0);
        std::vector<double> u_src(size_t(g.n_src) * 2 * sp, 0.0), y_src(size_t(g.n_src) * 4 * sp, 0.0);
        std::vector<int> lg_math(m.n_lg(), -1), src_math(m.source.size(), -1);
        for (size_t k = 0; k < p.load_comp.size(); ++k) lg_math[p.load_comp[k]] = int(k);
        for (size_t k = 0; k < p.src_comp.size(); ++k) src_math[p.src_comp[k]] = int(k);
        // base values broadcast to every lane
        auto put_load = [&](ComponentState const& st, Idx l, int lane) {
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
        for (int lane = 0; lane < g.S; ++lane) {
            Idx const s = g.scenarios[lane];
            (void)s;
        }
        // fill base for every lane, then apply each scenario's own load / source rows
        for (Idx l = 0; l < m.n_lg(); ++l) {
            put_load(m.base, l, 0);
            for (int lane = 1; lane < g.S; ++lane) {
                lstat[size_t(l) * sp + lane] = lstat[size_t(l) * sp];
                int const k = lg_math[l];
                if (k < 0) continue;
                for (int ph = 0; ph < P; ++ph)
                    for (int c = 0; c < 2; ++c)
                        s_inj[((size_t(k) * P + ph) * 2 + c) * sp + lane] = s_inj[((size_t(k) * P + ph) * 2 + c) * sp];
            }
        }
        for (Idx si = 0; si < Idx(m.source.size()); ++si)
            for (int lane = 0; lane < g.S; ++lane) put_src(g.st, si, lane);
        if (update) {
            ComponentState work = m.base;
            for (int lane = 0; lane < g.S; ++lane) {
                Idx const s = g.scenarios[lane];
                std::vector<std::pair<int, Idx>> touched;
                bool structural = false;
                for_each_row(s, [&](int comp, Idx idx, void const* rec) {
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
                // restore the touched entries for the next lane
                for (auto const& [comp, idx] : touched) {
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
        }
        upload(g.s_inj, s_inj, b->stream);
        upload(g.u_src, u_src, b->stream);
        upload(g.y_src, y_src, b->stream);
        upload(g.load_status, lstat, b->stream);
        g.dg.s_inj = g.s_inj.as<double>();
        g.dg.u_src = g.u_src.as<double>();
        g.dg.y_src = g.y_src.as<double>();
        g.dout.s_inj = g.dg.s_inj;
        g.dout.u_src = g.dg.u_src;
        g.dout.y_src = g.dg.y_src;
    }

    // ---- output record buffers ----
    for (int c = 0; c < N_COMP; ++c) {
        Idx const cnt = m.count(c);
        size_t const rs = c == C_NODE ? (sym ? sizeof(NodeOutput<1>) : sizeof(NodeOutput<3>))
                          : (c == C_LINE || c == C_TRANSFORMER)
                              ? (sym ? sizeof(BranchOutput<1>) : sizeof(BranchOutput<3>))
                              : (sym ? sizeof(ApplianceOutput<1>) : sizeof(ApplianceOutput<3>));
        b->rec_size[c] = rs;
        if (cnt > 0) b->out[c].alloc(size_t(S) * size_t(cnt) * rs);
    }
    b->status.assign(S, 0);
    b->iterations.assign(S, 0);
    b->messages.assign(S, "");
    for (Idx s = 0; s < S; ++s)
        if (scn_group[s] < 0) {
            b->status[s] = 3;
            b->messages[s] = scn_error[s];
        }
    PGM_HIP_CHECK(hipStreamSynchronize(b->stream));
    return b;
}

void Batch::run() {
    hipStream_t const st = stream;
    PGM_HIP_CHECK(hipEventRecord(ev_start, st));
    fwd_launches = 0;
    double fwd_ms = 0.0;
    std::vector<hipEvent_t> fwd_events;
    auto event_from_pool = [&](size_t i) {
        while (event_pool.size() <= i) {
            hipEvent_t e;
            PGM_HIP_CHECK(hipEventCreate(&e));
            event_pool.push_back(e);
        }
        return event_pool[i];
    };
    for (auto& gp : groups) {
        Group& g = *gp;
        if (g.S == 0) continue;
        DevGroup const& d = g.dg;
        int const L = int(g.level_ptr.size()) - 1;
        int const* lt = g.level_task.as<int>();
        auto level = [&](int l) { return std::make_pair(lt + g.level_ptr[l], g.level_ptr[l + 1] - g.level_ptr[l]); };
        launch_init_state(d, st);
        for (int l = 0; l < L; ++l) launch_forward(d, MODE_INIT, level(l).first, level(l).second, st);
        for (int l = L - 1; l >= 0; --l) launch_backward(d, MODE_INIT, level(l).first, level(l).second, st);
        for (int it = 1; it <= int(max_iter); ++it) {
            for (int l = 0; l < L; ++l) {
                bool const timed = fwd_events.size() < 2 * kMaxTimedLaunches;
                if (timed) {
                    fwd_events.push_back(event_from_pool(fwd_events.size()));
                    PGM_HIP_CHECK(hipEventRecord(fwd_events.back(), st));
                }
                launch_forward(d, MODE_NR, level(l).first, level(l).second, st);
                if (timed) {
                    fwd_events.push_back(event_from_pool(fwd_events.size()));
                    PGM_HIP_CHECK(hipEventRecord(fwd_events.back(), st));
                }
                ++fwd_launches;
            }
            for (int l = L - 1; l >= 0; --l) launch_backward(d, MODE_NR, level(l).first, level(l).second, st);
            *g.h_active = 0;
            PGM_HIP_CHECK(hipMemcpyAsync(g.active.get(), g.h_active, sizeof(int), hipMemcpyHostToDevice, st));
            launch_converge(d, it, int(max_iter), err_tol, g.active.as<int>(), st);
            PGM_HIP_CHECK(hipMemcpyAsync(g.h_active, g.active.get(), sizeof(int), hipMemcpyDeviceToHost, st));
            PGM_HIP_CHECK(hipStreamSynchronize(st));
            if (*g.h_active == 0) break;
        }
        PGM_HIP_CHECK(hipGetLastError());
    }
    // per-scenario status
    for (auto& gp : groups) {
        Group& g = *gp;
        if (g.S == 0) continue;
        size_t const sp = size_t(g.S_pad);
        size_t const nmm = size_t(g.dg.n_mm);
        std::vector<int> done(nmm * sp), err(nmm * sp), its(nmm * sp);
        std::vector<double> ldev(nmm * sp);
        if (nmm > 0) {
            PGM_HIP_CHECK(hipMemcpyAsync(done.data(), g.done.get(), done.size() * 4, hipMemcpyDeviceToHost, st));
            PGM_HIP_CHECK(hipMemcpyAsync(err.data(), g.err.get(), err.size() * 4, hipMemcpyDeviceToHost, st));
            PGM_HIP_CHECK(hipMemcpyAsync(its.data(), g.iters.get(), its.size() * 4, hipMemcpyDeviceToHost, st));
            PGM_HIP_CHECK(hipMemcpyAsync(ldev.data(), g.lastdev.get(), ldev.size() * 8, hipMemcpyDeviceToHost, st));
            PGM_HIP_CHECK(hipStreamSynchronize(st));
        }
        std::vector<int> ok(sp, 0);
        for (int lane = 0; lane < g.S; ++lane) {
            Idx const s = g.scenarios[lane];
            int stt = 0;
            Idx it_max = 0;
            std::string msg;
            for (size_t k = 0; k < nmm; ++k) {
                size_t const i = k * sp + lane;
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
            ok[lane] = stt == 0 ? 1 : 0;
        }
        PGM_HIP_CHECK(hipMemcpy(g.ok.get(), ok.data(), sp * 4, hipMemcpyHostToDevice));
        // results and records
        launch_bus_results(g.dout, st);
        NodeRec nr{int(model.n_node()), g.node_id.as<int>(), g.node_urated.as<double>(), g.node_bus.as<int>()};
        launch_node_records(g.dout, nr, out[C_NODE].get(), st);
        for (int kind = 0; kind < 2; ++kind) {
            int const comp = kind == 0 ? C_LINE : C_TRANSFORMER;
            BranchRec br{int(model.count(comp)), g.br_id[kind].as<int>(), g.br_math[kind].as<int>(),
                         g.br_from.as<int>(), g.br_to.as<int>(), g.br_y.as<double>(), g.br_bif[kind].as<double>(),
                         g.br_bit[kind].as<double>(), g.br_rating[kind].as<double>(), kind == 0 ? 1 : 0};
            launch_branch_records(g.dout, br, out[comp].get(), st);
        }
        {
            ApplRec ar{int(model.count(C_SHUNT)), g.ap_id[C_SHUNT].as<int>(), g.ap_math[C_SHUNT].as<int>(),
                       g.ap_base_i[C_SHUNT].as<double>(), -1.0, nullptr, g.ap_status_const[C_SHUNT].as<int8_t>(),
                       g.shunt_si.as<double>()};
            launch_appliance_records(g.dout, ar, out[C_SHUNT].get(), st);
            ApplRec as{int(model.count(C_SOURCE)), g.ap_id[C_SOURCE].as<int>(), g.ap_math[C_SOURCE].as<int>(),
                       g.ap_base_i[C_SOURCE].as<double>(), 1.0, nullptr, g.ap_status_const[C_SOURCE].as<int8_t>(),
                       g.src_si.as<double>()};
            launch_appliance_records(g.dout, as, out[C_SOURCE].get(), st);
        }
        int const lg_comps[4] = {C_SYM_GEN, C_ASYM_GEN, C_SYM_LOAD, C_ASYM_LOAD};
        for (int kk = 0; kk < 4; ++kk) {
            int const comp = lg_comps[kk];
            ApplRec al{int(model.count(comp)),
                       g.ap_id[comp].as<int>(),
                       g.ap_math[comp].as<int>(),
                       g.ap_base_i[comp].as<double>(),
                       (comp == C_SYM_GEN || comp == C_ASYM_GEN) ? 1.0 : -1.0,
                       g.load_status.as<int8_t>() + size_t(model.lg_offset[kk]) * g.S_pad,
                       nullptr,
                       g.load_si.as<double>()};
            launch_appliance_records(g.dout, al, out[comp].get(), st);
        }
        PGM_HIP_CHECK(hipGetLastError());
    }
    PGM_HIP_CHECK(hipEventRecord(ev_stop, st));
    PGM_HIP_CHECK(hipEventSynchronize(ev_stop));
    float ms = 0;
    PGM_HIP_CHECK(hipEventElapsedTime(&ms, ev_start, ev_stop));
    kernel_ms = ms;
    for (size_t i = 0; i + 1 < fwd_events.size(); i += 2) {
        float e = 0;
        PGM_HIP_CHECK(hipEventElapsedTime(&e, fwd_events[i], fwd_events[i + 1]));
        fwd_ms += e;
    }
    fwd_timed = Idx(fwd_events.size() / 2);
    fwd_launch_ms = fwd_timed > 0 ? fwd_ms / double(fwd_timed) : 0.0;
    ran = true;
}

void Batch::fetch(PGM_MutableDataset const& output, PGM_Handle* handle) {
    if (!ran) throw PowerGridError("the batch has not been run\n");
    bool const want_asym = output.meta->name == "asym_output";
    if (output.meta->name != "sym_output" && output.meta->name != "asym_output")
        throw PowerGridError("output dataset must be sym_output or asym_output\n");
    if (want_asym == symmetric) throw PowerGridError("output dataset symmetry does not match the calculation\n");
    if (output.batch_size != n_scenarios)
        throw PowerGridError("output batch size does not match the number of scenarios\n");
    std::vector<char> host;
    for (auto const& buf : output.buffers) {
        int comp = -1;
        for (int c = 0; c < N_COMP; ++c)
            if (buf.meta->name == comp_names[c]) comp = c;
        if (comp < 0) continue;
        Idx const cnt = model.count(comp);
        if (cnt == 0) continue;
        size_t const rs = rec_size[comp];
        bool const direct = !buf.columnar() && buf.elements_per_scenario == cnt;
        bool const all_ok = std::all_of(status.begin(), status.end(), [](int x) { return x == 0; });
        if (direct && all_ok) {
            PGM_HIP_CHECK(hipMemcpy(buf.data, out[comp].get(), size_t(n_scenarios) * cnt * rs, hipMemcpyDeviceToHost));
            continue;
        }
        host.resize(size_t(n_scenarios) * cnt * rs);
        PGM_HIP_CHECK(hipMemcpy(host.data(), out[comp].get(), host.size(), hipMemcpyDeviceToHost));
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

int device_block_lu(Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices, Idx const* diag_lu, Idx batch,
                    double const* data, double const* rhs, double* x, int32_t* status, bool fp32) {
    if (nb != 2 && nb != 6) return -1;
    ensure_device();
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
            if (it == b1 || *it != int(r)) return -1;  // pattern must be symmetric
            p.transpose[k] = int(it - p.col.begin());
        }
    build_schedule(p, 16);
    hipStream_t st;
    PGM_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    int const S = int(batch), S_pad = (S + 63) / 64 * 64;
    size_t const sp = size_t(S_pad);
    DeviceBuffer row_ptr, col, diag, transpose, fill, schur_ptr, schur, task_ptr, task_pivot, task_mm, level_task;
    upload(row_ptr, p.row_ptr, st);
    upload(col, p.col, st);
    upload(diag, p.diag, st);
    upload(transpose, p.transpose, st);
    upload(fill, p.fill, st);
    upload(schur_ptr, p.schur_ptr, st);
    std::vector<int> sch(p.schur.size() * 3);
    for (size_t i = 0; i < p.schur.size(); ++i)
        for (int k = 0; k < 3; ++k) sch[3 * i + k] = p.schur[i][k];
    upload(schur, sch, st);
    upload(task_ptr, p.task_ptr, st);
    upload(task_pivot, p.task_pivot, st);
    upload(task_mm, p.task_mm, st);
    upload(level_task, p.level_task, st);
    std::vector<double> a(size_t(p.E) * NB2 * sp, 0.0), bv(size_t(n) * NB * sp, 0.0);
    for (int s = 0; s < S; ++s) {
        for (Idx e = 0; e < p.E; ++e)
            for (int k = 0; k < NB2; ++k) a[(size_t(e) * NB2 + k) * sp + s] = data[(size_t(s) * p.E + e) * NB2 + k];
        for (Idx r = 0; r < n; ++r)
            for (int k = 0; k < NB; ++k) bv[(size_t(r) * NB + k) * sp + s] = rhs[(size_t(s) * n + r) * NB + k];
    }
    DeviceBuffer da, db, y, lu, perm, done, err, iters, devmax, lastdev;
    upload(da, a, st);
    upload(db, bv, st);
    y.alloc(size_t(n) * NB * sp * 8);
    lu.alloc(size_t(p.E) * NB2 * sp * (fp32 ? 4 : 8));
    perm.alloc(size_t(n) * sp * 4);
    done.alloc(sp * 4);
    err.alloc(sp * 4);
    iters.alloc(sp * 4);
    devmax.alloc(sp * 8);
    lastdev.alloc(sp * 8);
    DevGroup d;
    d.P = P;
    d.n = int(n);
    d.E = int(p.E);
    d.n_mm = 1;
    d.S = S;
    d.S_pad = S_pad;
    d.fp32 = fp32 ? 1 : 0;
    d.row_ptr = row_ptr.as<int>();
    d.col = col.as<int>();
    d.diag = diag.as<int>();
    d.transpose = transpose.as<int>();
    d.fill = fill.as<int8_t>();
    d.schur_ptr = schur_ptr.as<int>();
    d.schur = schur.as<int>();
    d.task_ptr = task_ptr.as<int>();
    d.task_pivot = task_pivot.as<int>();
    d.task_mm = task_mm.as<int>();
    d.y = y.as<double>();
    d.lu = lu.get();
    d.perm = perm.as<uint32_t>();
    d.done = done.as<int>();
    d.err = err.as<int>();
    d.iters = iters.as<int>();
    d.devmax = devmax.as<unsigned long long>();
    d.lastdev = lastdev.as<double>();
    d.raw_a = da.as<double>();
    d.raw_b = db.as<double>();
    launch_init_state(d, st);
    int const L = int(p.level_ptr.size()) - 1;
    int const* lt = level_task.as<int>();
    for (int l = 0; l < L; ++l) launch_forward(d, MODE_RAW, lt + p.level_ptr[l], p.level_ptr[l + 1] - p.level_ptr[l], st);
    for (int l = L - 1; l >= 0; --l)
        launch_backward(d, MODE_RAW, lt + p.level_ptr[l], p.level_ptr[l + 1] - p.level_ptr[l], st);
    PGM_HIP_CHECK(hipGetLastError());
    std::vector<double> yh(bv.size());
    std::vector<int> eh(sp);
    PGM_HIP_CHECK(hipMemcpyAsync(yh.data(), y.get(), yh.size() * 8, hipMemcpyDeviceToHost, st));
    PGM_HIP_CHECK(hipMemcpyAsync(eh.data(), err.get(), eh.size() * 4, hipMemcpyDeviceToHost, st));
    PGM_HIP_CHECK(hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s) {
        status[s] = eh[s] == ERR_SPARSE ? 1 : 0;
        for (Idx r = 0; r < n; ++r)
            for (int k = 0; k < NB; ++k) x[(size_t(s) * n + r) * NB + k] = yh[(size_t(r) * NB + k) * sp + s];
    }
    return 0;
}

}  // namespace pgm_amd
