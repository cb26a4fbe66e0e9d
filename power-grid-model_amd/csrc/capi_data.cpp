// capi_data.cpp -- handle, options, metadata, buffers and datasets of the C ABI.
//
// Mirrors the reference C API behaviour (power_grid_model_c/power_grid_model_c/src/{handle,options,
// meta_data,buffer,dataset}.cpp): errors are reported through the handle, never thrown across the ABI.
#include "dataset.hpp"

#include "../../include/power_grid_model_amd.h"

#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <limits>
#include <new>

using namespace pgm_amd;

namespace pgm_amd {

namespace {
constexpr double kNan = std::numeric_limits<double>::quiet_NaN();

// offsetof through a null pointer is not constant-evaluated here; use a real object instead
template <class T> struct Probe {
    static T const& obj() {
        static T t{};
        return t;
    }
};
template <class T, class M> PGM_MetaAttribute mk(char const* name, M T::*member, int ctype) {
    T const& t = Probe<T>::obj();
    auto const off = size_t(reinterpret_cast<char const*>(&(t.*member)) - reinterpret_cast<char const*>(&t));
    size_t size = ctype == PGM_int32 ? 4 : ctype == PGM_int8 ? 1 : ctype == PGM_double ? 8 : 24;
    return PGM_MetaAttribute{name, ctype, off, size};
}
template <class T> PGM_MetaComponent comp(char const* name, std::vector<PGM_MetaAttribute> attrs) {
    return PGM_MetaComponent{name, sizeof(T), alignof(T), std::move(attrs)};
}

template <int P> std::vector<PGM_MetaComponent> output_components() {
    int const v = P == 1 ? PGM_double : PGM_double3;
    using N = NodeOutput<P>;
    using B = BranchOutput<P>;
    using A = ApplianceOutput<P>;
    auto node = comp<N>("node", {mk("id", &N::id, PGM_int32), mk("energized", &N::energized, PGM_int8),
                                 mk("u_pu", &N::u_pu, v), mk("u", &N::u, v), mk("u_angle", &N::u_angle, v),
                                 mk("p", &N::p, v), mk("q", &N::q, v)});
    auto branch = [&](char const* n) {
        return comp<B>(n, {mk("id", &B::id, PGM_int32), mk("energized", &B::energized, PGM_int8),
                           mk("loading", &B::loading, PGM_double), mk("p_from", &B::p_from, v),
                           mk("q_from", &B::q_from, v), mk("i_from", &B::i_from, v), mk("s_from", &B::s_from, v),
                           mk("p_to", &B::p_to, v), mk("q_to", &B::q_to, v), mk("i_to", &B::i_to, v),
                           mk("s_to", &B::s_to, v)});
    };
    auto appl = [&](char const* n) {
        return comp<A>(n, {mk("id", &A::id, PGM_int32), mk("energized", &A::energized, PGM_int8),
                           mk("p", &A::p, v), mk("q", &A::q, v), mk("i", &A::i, v), mk("s", &A::s, v),
                           mk("pf", &A::pf, v)});
    };
    using R = TapRegulatorOutput;
    auto reg = comp<R>(kTapRegulatorName, {mk("id", &R::id, PGM_int32), mk("energized", &R::energized, PGM_int8),
                                           mk("tap_pos", &R::tap_pos, PGM_int8)});
    using F = FaultOutput;
    auto fault = comp<F>(kFaultName, {mk("id", &F::id, PGM_int32), mk("energized", &F::energized, PGM_int8)});
    using VS = VoltageSensorOutput<P>;
    auto vs = comp<VS>(kVoltageSensorName, {mk("id", &VS::id, PGM_int32), mk("energized", &VS::energized, PGM_int8),
                                            mk("u_residual", &VS::u_residual, v),
                                            mk("u_angle_residual", &VS::u_angle_residual, v)});
    using PS = PowerSensorOutput<P>;
    auto ps = comp<PS>(kPowerSensorName, {mk("id", &PS::id, PGM_int32), mk("energized", &PS::energized, PGM_int8),
                                          mk("p_residual", &PS::p_residual, v), mk("q_residual", &PS::q_residual, v)});
    return {node,           branch("line"),  branch("link"),    branch("generic_branch"), branch("transformer"), appl("shunt"),
            appl("source"), appl("sym_gen"), appl("asym_gen"), appl("sym_load"),     appl("asym_load"),
            reg,            fault,           vs,                ps};
}

std::vector<PGM_MetaDataset> build_meta() {
    std::vector<PGM_MetaDataset> ds;
    using NI = NodeInput;
    using LI = LineInput;
    using LkI = LinkInput;
    using GBI = GenericBranchInput;
    using FI = FaultInput;
    using TI = TransformerInput;
    using SI = ShuntInput;
    using SrcI = SourceInput;
    using LG = SymLoadGenInput;
    using ALG = AsymLoadGenInput;
    using RI = TapRegulatorInput;
    using VSI = SymVoltageSensorInput;
    using PSI = SymPowerSensorInput;
    auto lg_in = [&](char const* n) {
        return comp<LG>(n, {mk("id", &LG::id, PGM_int32), mk("node", &LG::node, PGM_int32),
                            mk("status", &LG::status, PGM_int8), mk("type", &LG::type, PGM_int8),
                            mk("p_specified", &LG::p_specified, PGM_double),
                            mk("q_specified", &LG::q_specified, PGM_double)});
    };
    auto alg_in = [&](char const* n) {
        return comp<ALG>(n, {mk("id", &ALG::id, PGM_int32), mk("node", &ALG::node, PGM_int32),
                             mk("status", &ALG::status, PGM_int8), mk("type", &ALG::type, PGM_int8),
                             mk("p_specified", &ALG::p_specified, PGM_double3),
                             mk("q_specified", &ALG::q_specified, PGM_double3)});
    };
    ds.push_back(PGM_MetaDataset{
        "input",
        {comp<NI>("node", {mk("id", &NI::id, PGM_int32), mk("u_rated", &NI::u_rated, PGM_double)}),
         comp<LI>("line", {mk("id", &LI::id, PGM_int32), mk("from_node", &LI::from_node, PGM_int32),
                           mk("to_node", &LI::to_node, PGM_int32), mk("from_status", &LI::from_status, PGM_int8),
                           mk("to_status", &LI::to_status, PGM_int8), mk("r1", &LI::r1, PGM_double),
                           mk("x1", &LI::x1, PGM_double), mk("c1", &LI::c1, PGM_double),
                           mk("tan1", &LI::tan1, PGM_double), mk("r0", &LI::r0, PGM_double),
                           mk("x0", &LI::x0, PGM_double), mk("c0", &LI::c0, PGM_double),
                           mk("tan0", &LI::tan0, PGM_double), mk("i_n", &LI::i_n, PGM_double)}),
         comp<LkI>("link", {mk("id", &LkI::id, PGM_int32), mk("from_node", &LkI::from_node, PGM_int32),
                            mk("to_node", &LkI::to_node, PGM_int32), mk("from_status", &LkI::from_status, PGM_int8),
                            mk("to_status", &LkI::to_status, PGM_int8)}),
         comp<GBI>("generic_branch",
                   {mk("id", &GBI::id, PGM_int32), mk("from_node", &GBI::from_node, PGM_int32),
                    mk("to_node", &GBI::to_node, PGM_int32), mk("from_status", &GBI::from_status, PGM_int8),
                    mk("to_status", &GBI::to_status, PGM_int8), mk("r1", &GBI::r1, PGM_double),
                    mk("x1", &GBI::x1, PGM_double), mk("g1", &GBI::g1, PGM_double), mk("b1", &GBI::b1, PGM_double),
                    mk("k", &GBI::k, PGM_double), mk("theta", &GBI::theta, PGM_double),
                    mk("sn", &GBI::sn, PGM_double)}),
         comp<TI>("transformer",
                  {mk("id", &TI::id, PGM_int32),
                   mk("from_node", &TI::from_node, PGM_int32),
                   mk("to_node", &TI::to_node, PGM_int32),
                   mk("from_status", &TI::from_status, PGM_int8),
                   mk("to_status", &TI::to_status, PGM_int8),
                   mk("u1", &TI::u1, PGM_double),
                   mk("u2", &TI::u2, PGM_double),
                   mk("sn", &TI::sn, PGM_double),
                   mk("uk", &TI::uk, PGM_double),
                   mk("pk", &TI::pk, PGM_double),
                   mk("i0", &TI::i0, PGM_double),
                   mk("p0", &TI::p0, PGM_double),
                   mk("i0_zero_sequence", &TI::i0_zero_sequence, PGM_double),
                   mk("p0_zero_sequence", &TI::p0_zero_sequence, PGM_double),
                   mk("winding_from", &TI::winding_from, PGM_int8),
                   mk("winding_to", &TI::winding_to, PGM_int8),
                   mk("clock", &TI::clock, PGM_int8),
                   mk("tap_side", &TI::tap_side, PGM_int8),
                   mk("tap_pos", &TI::tap_pos, PGM_int8),
                   mk("tap_min", &TI::tap_min, PGM_int8),
                   mk("tap_max", &TI::tap_max, PGM_int8),
                   mk("tap_nom", &TI::tap_nom, PGM_int8),
                   mk("tap_size", &TI::tap_size, PGM_double),
                   mk("uk_min", &TI::uk_min, PGM_double),
                   mk("uk_max", &TI::uk_max, PGM_double),
                   mk("pk_min", &TI::pk_min, PGM_double),
                   mk("pk_max", &TI::pk_max, PGM_double),
                   mk("r_grounding_from", &TI::r_grounding_from, PGM_double),
                   mk("x_grounding_from", &TI::x_grounding_from, PGM_double),
                   mk("r_grounding_to", &TI::r_grounding_to, PGM_double),
                   mk("x_grounding_to", &TI::x_grounding_to, PGM_double)}),
         comp<SI>("shunt", {mk("id", &SI::id, PGM_int32), mk("node", &SI::node, PGM_int32),
                            mk("status", &SI::status, PGM_int8), mk("g1", &SI::g1, PGM_double),
                            mk("b1", &SI::b1, PGM_double), mk("g0", &SI::g0, PGM_double),
                            mk("b0", &SI::b0, PGM_double)}),
         comp<SrcI>("source", {mk("id", &SrcI::id, PGM_int32), mk("node", &SrcI::node, PGM_int32),
                               mk("status", &SrcI::status, PGM_int8), mk("u_ref", &SrcI::u_ref, PGM_double),
                               mk("u_ref_angle", &SrcI::u_ref_angle, PGM_double), mk("sk", &SrcI::sk, PGM_double),
                               mk("rx_ratio", &SrcI::rx_ratio, PGM_double),
                               mk("z01_ratio", &SrcI::z01_ratio, PGM_double)}),
         lg_in("sym_gen"), alg_in("asym_gen"), lg_in("sym_load"), alg_in("asym_load"),
         comp<FI>(kFaultName, {mk("id", &FI::id, PGM_int32), mk("status", &FI::status, PGM_int8),
                               mk("fault_type", &FI::fault_type, PGM_int8), mk("fault_phase", &FI::fault_phase, PGM_int8),
                               mk("fault_object", &FI::fault_object, PGM_int32), mk("r_f", &FI::r_f, PGM_double),
                               mk("x_f", &FI::x_f, PGM_double)}),
         comp<VSI>(kVoltageSensorName,
                   {mk("id", &VSI::id, PGM_int32), mk("measured_object", &VSI::measured_object, PGM_int32),
                    mk("u_sigma", &VSI::u_sigma, PGM_double), mk("u_measured", &VSI::u_measured, PGM_double),
                    mk("u_angle_measured", &VSI::u_angle_measured, PGM_double)}),
         comp<PSI>(kPowerSensorName,
                   {mk("id", &PSI::id, PGM_int32), mk("measured_object", &PSI::measured_object, PGM_int32),
                    mk("measured_terminal_type", &PSI::measured_terminal_type, PGM_int8),
                    mk("power_sigma", &PSI::power_sigma, PGM_double), mk("p_measured", &PSI::p_measured, PGM_double),
                    mk("q_measured", &PSI::q_measured, PGM_double), mk("p_sigma", &PSI::p_sigma, PGM_double),
                    mk("q_sigma", &PSI::q_sigma, PGM_double)}),
         comp<RI>(kTapRegulatorName,
                  {mk("id", &RI::id, PGM_int32), mk("regulated_object", &RI::regulated_object, PGM_int32),
                   mk("status", &RI::status, PGM_int8), mk("control_side", &RI::control_side, PGM_int8),
                   mk("u_set", &RI::u_set, PGM_double), mk("u_band", &RI::u_band, PGM_double),
                   mk("line_drop_compensation_r", &RI::line_drop_compensation_r, PGM_double),
                   mk("line_drop_compensation_x", &RI::line_drop_compensation_x, PGM_double)})}});
    using BU = BranchUpdate;
    using TU = TransformerUpdate;
    using SU = ShuntUpdate;
    using SrcU = SourceUpdate;
    using LGU = SymLoadGenUpdate;
    using ALGU = AsymLoadGenUpdate;
    using RU = TapRegulatorUpdate;
    using VSU = SymVoltageSensorUpdate;
    using PSU = SymPowerSensorUpdate;
    auto lg_up = [&](char const* n) {
        return comp<LGU>(n, {mk("id", &LGU::id, PGM_int32), mk("status", &LGU::status, PGM_int8),
                             mk("p_specified", &LGU::p_specified, PGM_double),
                             mk("q_specified", &LGU::q_specified, PGM_double)});
    };
    auto alg_up = [&](char const* n) {
        return comp<ALGU>(n, {mk("id", &ALGU::id, PGM_int32), mk("status", &ALGU::status, PGM_int8),
                              mk("p_specified", &ALGU::p_specified, PGM_double3),
                              mk("q_specified", &ALGU::q_specified, PGM_double3)});
    };
    ds.push_back(PGM_MetaDataset{
        "update",
        {comp<BU>("line", {mk("id", &BU::id, PGM_int32), mk("from_status", &BU::from_status, PGM_int8),
                           mk("to_status", &BU::to_status, PGM_int8)}),
         comp<BU>("link", {mk("id", &BU::id, PGM_int32), mk("from_status", &BU::from_status, PGM_int8),
                           mk("to_status", &BU::to_status, PGM_int8)}),
         comp<BU>("generic_branch", {mk("id", &BU::id, PGM_int32), mk("from_status", &BU::from_status, PGM_int8),
                                     mk("to_status", &BU::to_status, PGM_int8)}),
         comp<TU>("transformer", {mk("id", &TU::id, PGM_int32), mk("from_status", &TU::from_status, PGM_int8),
                                  mk("to_status", &TU::to_status, PGM_int8), mk("tap_pos", &TU::tap_pos, PGM_int8)}),
         comp<SU>("shunt", {mk("id", &SU::id, PGM_int32), mk("status", &SU::status, PGM_int8),
                            mk("g1", &SU::g1, PGM_double), mk("b1", &SU::b1, PGM_double),
                            mk("g0", &SU::g0, PGM_double), mk("b0", &SU::b0, PGM_double)}),
         comp<SrcU>("source", {mk("id", &SrcU::id, PGM_int32), mk("status", &SrcU::status, PGM_int8),
                               mk("u_ref", &SrcU::u_ref, PGM_double),
                               mk("u_ref_angle", &SrcU::u_ref_angle, PGM_double), mk("sk", &SrcU::sk, PGM_double),
                               mk("rx_ratio", &SrcU::rx_ratio, PGM_double),
                               mk("z01_ratio", &SrcU::z01_ratio, PGM_double)}),
         lg_up("sym_gen"), alg_up("asym_gen"), lg_up("sym_load"), alg_up("asym_load"),
         comp<VSU>(kVoltageSensorName,
                   {mk("id", &VSU::id, PGM_int32), mk("u_sigma", &VSU::u_sigma, PGM_double),
                    mk("u_measured", &VSU::u_measured, PGM_double),
                    mk("u_angle_measured", &VSU::u_angle_measured, PGM_double)}),
         comp<PSU>(kPowerSensorName,
                   {mk("id", &PSU::id, PGM_int32), mk("power_sigma", &PSU::power_sigma, PGM_double),
                    mk("p_measured", &PSU::p_measured, PGM_double), mk("q_measured", &PSU::q_measured, PGM_double),
                    mk("p_sigma", &PSU::p_sigma, PGM_double), mk("q_sigma", &PSU::q_sigma, PGM_double)}),
         comp<RU>(kTapRegulatorName,
                  {mk("id", &RU::id, PGM_int32), mk("status", &RU::status, PGM_int8),
                   mk("u_set", &RU::u_set, PGM_double), mk("u_band", &RU::u_band, PGM_double),
                   mk("line_drop_compensation_r", &RU::line_drop_compensation_r, PGM_double),
                   mk("line_drop_compensation_x", &RU::line_drop_compensation_x, PGM_double)})}});
    ds.push_back(PGM_MetaDataset{"sym_output", output_components<1>()});
    ds.push_back(PGM_MetaDataset{"asym_output", output_components<3>()});
    return ds;
}
}  // namespace

std::vector<PGM_MetaDataset> const& meta_datasets() {
    static std::vector<PGM_MetaDataset> const m = build_meta();
    return m;
}
PGM_MetaDataset const& meta_dataset(std::string const& name) {
    for (auto const& d : meta_datasets())
        if (d.name == name) return d;
    throw DatasetError("unknown dataset " + name);
}

void set_null(PGM_MetaAttribute const& a, void* record) {
    char* p = static_cast<char*>(record) + a.offset;
    switch (a.ctype) {
    case PGM_int32: {
        int32_t const v = na_IntID;
        std::memcpy(p, &v, 4);
        break;
    }
    case PGM_int8: {
        int8_t const v = na_IntS;
        std::memcpy(p, &v, 1);
        break;
    }
    case PGM_double:
        std::memcpy(p, &kNan, 8);
        break;
    default:
        for (int k = 0; k < 3; ++k) std::memcpy(p + 8 * k, &kNan, 8);
    }
}
void set_null_records(PGM_MetaComponent const& c, void* ptr, Idx n) {
    char* base = static_cast<char*>(ptr);
    for (Idx i = 0; i < n; ++i) {
        std::memset(base + i * c.size, 0, c.size);
        for (auto const& a : c.attributes) set_null(a, base + i * c.size);
    }
}

void ComponentBuffer::read_record(Idx i, void* dst) const {
    if (data) {
        std::memcpy(dst, static_cast<char const*>(data) + i * meta->size, meta->size);
        return;
    }
    set_null_records(*meta, dst, 1);
    for (auto const& [a, col] : columns)
        std::memcpy(static_cast<char*>(dst) + a->offset, static_cast<char const*>(col) + i * a->size, a->size);
}
void ComponentBuffer::write_record(Idx i, void const* src) const {
    if (data) {
        std::memcpy(static_cast<char*>(data) + i * meta->size, src, meta->size);
        return;
    }
    for (auto const& [a, col] : columns)
        std::memcpy(static_cast<char*>(col) + i * a->size, static_cast<char const*>(src) + a->offset, a->size);
}

}  // namespace pgm_amd

namespace {
template <class DS> void fill_info(DS const& d) {
    auto& in = d.info.d;
    in.name = d.meta->name;
    in.is_batch = d.is_batch;
    in.batch_size = d.batch_size;
    in.comp_names.clear();
    in.eps.clear();
    in.total.clear();
    in.attr_indications.clear();
    for (auto const& b : d.buffers) {
        in.comp_names.push_back(b.meta->name);
        in.eps.push_back(b.elements_per_scenario);
        in.total.push_back(b.total_elements);
        std::vector<std::string> ai;
        for (auto const& c : b.columns) ai.push_back(c.first->name);
        in.attr_indications.push_back(ai);
    }
}

void handle_exception(PGM_Handle* h) {
    if (!h) return;
    try {
        throw;
    } catch (std::exception const& e) {
        h->err_code = PGM_regular_error;
        h->err_msg = e.what();
    } catch (...) {
        h->err_code = PGM_regular_error;
        h->err_msg = "Unknown error!\n";
    }
}

template <class F> auto call_with_catch(PGM_Handle* h, F&& f, decltype(f()) fallback) -> decltype(f()) {
    if (h) h->clear();
    try {
        return f();
    } catch (...) {
        handle_exception(h);
        return fallback;
    }
}
template <class F> void call_with_catch_void(PGM_Handle* h, F&& f) {
    if (h) h->clear();
    try {
        f();
    } catch (...) {
        handle_exception(h);
    }
}

template <class DS> void add_buffer(DS* ds, char const* component, Idx eps, Idx total, Idx const* indptr, void* data) {
    auto const* c = ds->meta->find(component);
    if (!c) throw DatasetError(std::string("unknown component ") + component + " in dataset " + ds->meta->name);
    for (auto const& b : ds->buffers)
        if (b.meta == c) throw DatasetError(std::string("component ") + component + " added twice");
    if (!ds->is_batch && eps < 0) throw DatasetError("a single dataset must be uniform");
    if (eps < 0 && !indptr) throw DatasetError("sparse buffer needs indptr");
    if (eps >= 0 && eps * ds->batch_size != total)
        throw DatasetError("total elements do not match elements per scenario times batch size");
    if (eps < 0 && (indptr[0] != 0 || indptr[ds->batch_size] != total))
        throw DatasetError("indptr does not match total elements");
    ComponentBuffer b;
    b.meta = c;
    b.elements_per_scenario = eps;
    b.total_elements = total;
    b.indptr = eps < 0 ? indptr : nullptr;
    b.data = data;
    ds->buffers.push_back(b);
    ds->refresh_info();
}
template <class DS> void add_attr(DS* ds, char const* component, char const* attribute, void* data) {
    for (auto& b : ds->buffers) {
        if (b.meta->name != component) continue;
        if (b.data) throw DatasetError("cannot add attribute buffer to a row-based component buffer");
        auto const* a = b.meta->find(attribute);
        if (!a) throw DatasetError(std::string("unknown attribute ") + attribute);
        for (auto const& c : b.columns)
            if (c.first == a) throw DatasetError(std::string("attribute ") + attribute + " added twice");
        b.columns.emplace_back(a, data);
        ds->refresh_info();
        return;
    }
    throw DatasetError(std::string("component ") + component + " has no buffer");
}
}  // namespace

void PGM_ConstDataset::refresh_info() const { fill_info(*this); }
void PGM_MutableDataset::refresh_info() const { fill_info(*this); }

// ------------------------------------------------------------------------------------------------
extern "C" {

PGM_Handle* PGM_create_handle(void) noexcept { return new (std::nothrow) PGM_Handle{}; }
void PGM_destroy_handle(PGM_Handle* handle) noexcept { delete handle; }
PGM_Idx PGM_error_code(PGM_Handle const* handle) noexcept { return handle ? handle->err_code : 0; }
char const* PGM_error_message(PGM_Handle const* handle) noexcept { return handle ? handle->err_msg.c_str() : ""; }
PGM_Idx PGM_n_failed_scenarios(PGM_Handle const* handle) noexcept {
    return handle ? PGM_Idx(handle->failed_scenarios.size()) : 0;
}
PGM_Idx const* PGM_failed_scenarios(PGM_Handle const* handle) noexcept {
    return handle ? handle->failed_scenarios.data() : nullptr;
}
char const** PGM_batch_errors(PGM_Handle const* handle) noexcept {
    if (!handle) return nullptr;
    handle->batch_errs_c.clear();
    for (auto const& s : handle->batch_errs) handle->batch_errs_c.push_back(s.c_str());
    return handle->batch_errs_c.data();
}
void PGM_clear_error(PGM_Handle* handle) noexcept {
    if (handle) handle->clear();
}
char const* PGM_version(void) noexcept { return "1.13.amd-mi355x"; }

PGM_Options* PGM_create_options(PGM_Handle* handle) noexcept {
    return call_with_catch(handle, [] { return new PGM_Options{}; }, static_cast<PGM_Options*>(nullptr));
}
void PGM_destroy_options(PGM_Options* opt) noexcept { delete opt; }
void PGM_set_calculation_type(PGM_Handle*, PGM_Options* opt, PGM_Idx type) noexcept { opt->calculation_type = type; }
void PGM_set_calculation_method(PGM_Handle*, PGM_Options* opt, PGM_Idx method) noexcept {
    opt->calculation_method = method;
}
void PGM_set_symmetric(PGM_Handle*, PGM_Options* opt, PGM_Idx sym) noexcept { opt->symmetric = sym; }
void PGM_set_err_tol(PGM_Handle*, PGM_Options* opt, double err_tol) noexcept { opt->err_tol = err_tol; }
void PGM_set_max_iter(PGM_Handle*, PGM_Options* opt, PGM_Idx max_iter) noexcept { opt->max_iter = max_iter; }
void PGM_set_threading(PGM_Handle*, PGM_Options* opt, PGM_Idx threading) noexcept { opt->threading = threading; }
void PGM_set_short_circuit_voltage_scaling(PGM_Handle*, PGM_Options* opt, PGM_Idx v) noexcept {
    opt->short_circuit_voltage_scaling = v;
}
void PGM_set_tap_changing_strategy(PGM_Handle*, PGM_Options* opt, PGM_Idx v) noexcept {
    opt->tap_changing_strategy = v;
}
void PGM_set_experimental_features(PGM_Handle*, PGM_Options* opt, PGM_Idx v) noexcept {
    opt->experimental_features = v;
}
void PGM_amd_set_factor_precision(PGM_Handle* handle, PGM_Options* opt, PGM_Idx precision) noexcept {
    call_with_catch_void(handle, [&] {
        if (precision != 0 && precision != 1) throw PowerGridError("factor precision must be 0 (FP64) or 1 (FP32)");
        opt->factor_precision = precision;
    });
}

void PGM_amd_set_adjoint_layout(PGM_Handle* handle, PGM_Options* opt, PGM_Idx value) noexcept {
    call_with_catch_void(handle, [&] {
        if (value != 0 && value != 1) throw PowerGridError("adjoint layout must be 0 or 1");
        opt->adjoint_layout = value;
    });
}

void PGM_amd_set_step_refinement(PGM_Handle* handle, PGM_Options* opt, PGM_Idx value) noexcept {
    call_with_catch_void(handle, [&] {
        if (value != 0 && value != 1) throw PowerGridError("step refinement must be 0 or 1");
        opt->refine_steps = value;
    });
}

PGM_Idx PGM_meta_n_datasets(PGM_Handle*) noexcept { return PGM_Idx(meta_datasets().size()); }
PGM_MetaDataset const* PGM_meta_get_dataset_by_idx(PGM_Handle* handle, PGM_Idx idx) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_MetaDataset const* {
            if (idx < 0 || idx >= PGM_Idx(meta_datasets().size())) throw PowerGridError("dataset index out of range");
            return &meta_datasets()[idx];
        },
        static_cast<PGM_MetaDataset const*>(nullptr));
}
PGM_MetaDataset const* PGM_meta_get_dataset_by_name(PGM_Handle* handle, char const* dataset) noexcept {
    return call_with_catch(
        handle, [&]() -> PGM_MetaDataset const* { return &meta_dataset(dataset); },
        static_cast<PGM_MetaDataset const*>(nullptr));
}
char const* PGM_meta_dataset_name(PGM_Handle*, PGM_MetaDataset const* dataset) noexcept {
    return dataset->name.c_str();
}
PGM_Idx PGM_meta_n_components(PGM_Handle*, PGM_MetaDataset const* dataset) noexcept {
    return PGM_Idx(dataset->components.size());
}
PGM_MetaComponent const* PGM_meta_get_component_by_idx(PGM_Handle* handle, PGM_MetaDataset const* dataset,
                                                       PGM_Idx idx) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_MetaComponent const* {
            if (idx < 0 || idx >= PGM_Idx(dataset->components.size()))
                throw PowerGridError("component index out of range");
            return &dataset->components[idx];
        },
        static_cast<PGM_MetaComponent const*>(nullptr));
}
PGM_MetaComponent const* PGM_meta_get_component_by_name(PGM_Handle* handle, char const* dataset,
                                                        char const* component) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_MetaComponent const* {
            auto const* c = meta_dataset(dataset).find(component);
            if (!c) throw DatasetError(std::string("unknown component ") + component);
            return c;
        },
        static_cast<PGM_MetaComponent const*>(nullptr));
}
char const* PGM_meta_component_name(PGM_Handle*, PGM_MetaComponent const* c) noexcept { return c->name.c_str(); }
size_t PGM_meta_component_size(PGM_Handle*, PGM_MetaComponent const* c) noexcept { return c->size; }
size_t PGM_meta_component_alignment(PGM_Handle*, PGM_MetaComponent const* c) noexcept { return c->alignment; }
PGM_Idx PGM_meta_n_attributes(PGM_Handle*, PGM_MetaComponent const* c) noexcept {
    return PGM_Idx(c->attributes.size());
}
PGM_MetaAttribute const* PGM_meta_get_attribute_by_idx(PGM_Handle* handle, PGM_MetaComponent const* c,
                                                       PGM_Idx idx) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_MetaAttribute const* {
            if (idx < 0 || idx >= PGM_Idx(c->attributes.size())) throw PowerGridError("attribute index out of range");
            return &c->attributes[idx];
        },
        static_cast<PGM_MetaAttribute const*>(nullptr));
}
PGM_MetaAttribute const* PGM_meta_get_attribute_by_name(PGM_Handle* handle, char const* dataset,
                                                        char const* component, char const* attribute) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_MetaAttribute const* {
            auto const* c = meta_dataset(dataset).find(component);
            if (!c) throw DatasetError(std::string("unknown component ") + component);
            auto const* a = c->find(attribute);
            if (!a) throw DatasetError(std::string("unknown attribute ") + attribute);
            return a;
        },
        static_cast<PGM_MetaAttribute const*>(nullptr));
}
char const* PGM_meta_attribute_name(PGM_Handle*, PGM_MetaAttribute const* a) noexcept { return a->name.c_str(); }
PGM_Idx PGM_meta_attribute_ctype(PGM_Handle*, PGM_MetaAttribute const* a) noexcept { return a->ctype; }
size_t PGM_meta_attribute_offset(PGM_Handle*, PGM_MetaAttribute const* a) noexcept { return a->offset; }
int PGM_is_little_endian(PGM_Handle*) noexcept { return 1; }

void* PGM_create_buffer(PGM_Handle* handle, PGM_MetaComponent const* component, PGM_Idx size) noexcept {
    return call_with_catch(
        handle,
        [&]() -> void* {
            size_t const bytes = std::max<size_t>(1, size_t(size) * component->size);
            size_t const al = std::max<size_t>(component->alignment, 64);
            void* p = std::aligned_alloc(al, (bytes + al - 1) / al * al);
            if (!p) throw std::bad_alloc();
            return p;
        },
        static_cast<void*>(nullptr));
}
void PGM_destroy_buffer(void* ptr) noexcept { std::free(ptr); }
void PGM_buffer_set_nan(PGM_Handle* handle, PGM_MetaComponent const* component, void* ptr, PGM_Idx offset,
                        PGM_Idx size) noexcept {
    call_with_catch_void(handle, [&] {
        set_null_records(*component, static_cast<char*>(ptr) + offset * component->size, size);
    });
}
void PGM_buffer_set_value(PGM_Handle* handle, PGM_MetaAttribute const* attribute, void* buffer_ptr,
                          void const* src_ptr, PGM_Idx buffer_offset, PGM_Idx size, PGM_Idx src_stride) noexcept {
    call_with_catch_void(handle, [&] {
        PGM_Idx const stride = src_stride < 0 ? PGM_Idx(attribute->size) : src_stride;
        // the attribute's component size is not known from the attribute alone: find it
        size_t comp_size = 0;
        for (auto const& d : meta_datasets())
            for (auto const& c : d.components)
                for (auto const& a : c.attributes)
                    if (&a == attribute) comp_size = c.size;
        for (PGM_Idx i = 0; i < size; ++i)
            std::memcpy(static_cast<char*>(buffer_ptr) + (buffer_offset + i) * comp_size + attribute->offset,
                        static_cast<char const*>(src_ptr) + i * stride, attribute->size);
    });
}
void PGM_buffer_get_value(PGM_Handle* handle, PGM_MetaAttribute const* attribute, void const* buffer_ptr,
                          void* dest_ptr, PGM_Idx buffer_offset, PGM_Idx size, PGM_Idx dest_stride) noexcept {
    call_with_catch_void(handle, [&] {
        PGM_Idx const stride = dest_stride < 0 ? PGM_Idx(attribute->size) : dest_stride;
        size_t comp_size = 0;
        for (auto const& d : meta_datasets())
            for (auto const& c : d.components)
                for (auto const& a : c.attributes)
                    if (&a == attribute) comp_size = c.size;
        for (PGM_Idx i = 0; i < size; ++i)
            std::memcpy(static_cast<char*>(dest_ptr) + i * stride,
                        static_cast<char const*>(buffer_ptr) + (buffer_offset + i) * comp_size + attribute->offset,
                        attribute->size);
    });
}

char const* PGM_dataset_info_name(PGM_Handle*, PGM_DatasetInfo const* info) noexcept { return info->d.name.c_str(); }
PGM_Idx PGM_dataset_info_is_batch(PGM_Handle*, PGM_DatasetInfo const* info) noexcept { return info->d.is_batch; }
PGM_Idx PGM_dataset_info_batch_size(PGM_Handle*, PGM_DatasetInfo const* info) noexcept { return info->d.batch_size; }
PGM_Idx PGM_dataset_info_n_components(PGM_Handle*, PGM_DatasetInfo const* info) noexcept {
    return PGM_Idx(info->d.comp_names.size());
}
char const* PGM_dataset_info_component_name(PGM_Handle*, PGM_DatasetInfo const* info, PGM_Idx i) noexcept {
    return info->d.comp_names[i].c_str();
}
PGM_Idx PGM_dataset_info_elements_per_scenario(PGM_Handle*, PGM_DatasetInfo const* info, PGM_Idx i) noexcept {
    return info->d.eps[i];
}
PGM_Idx PGM_dataset_info_total_elements(PGM_Handle*, PGM_DatasetInfo const* info, PGM_Idx i) noexcept {
    return info->d.total[i];
}
PGM_Idx PGM_dataset_info_has_attribute_indications(PGM_Handle*, PGM_DatasetInfo const* info, PGM_Idx i) noexcept {
    return !info->d.attr_indications[i].empty();
}
PGM_Idx PGM_dataset_info_n_attribute_indications(PGM_Handle*, PGM_DatasetInfo const* info, PGM_Idx i) noexcept {
    return PGM_Idx(info->d.attr_indications[i].size());
}
char const* PGM_dataset_info_attribute_name(PGM_Handle*, PGM_DatasetInfo const* info, PGM_Idx i,
                                            PGM_Idx j) noexcept {
    return info->d.attr_indications[i][j].c_str();
}

PGM_ConstDataset* PGM_create_dataset_const(PGM_Handle* handle, char const* dataset, PGM_Idx is_batch,
                                           PGM_Idx batch_size) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_ConstDataset* {
            if (!is_batch && batch_size != 1) throw DatasetError("a single dataset has batch size 1");
            if (batch_size < 0) throw DatasetError("negative batch size");
            auto* d = new PGM_ConstDataset{};
            d->meta = &meta_dataset(dataset);
            d->is_batch = is_batch != 0;
            d->batch_size = batch_size;
            d->refresh_info();
            return d;
        },
        static_cast<PGM_ConstDataset*>(nullptr));
}
PGM_ConstDataset* PGM_create_dataset_const_from_mutable(PGM_Handle* handle,
                                                        PGM_MutableDataset const* m) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_ConstDataset* {
            auto* d = new PGM_ConstDataset{};
            d->meta = m->meta;
            d->is_batch = m->is_batch;
            d->batch_size = m->batch_size;
            d->buffers = m->buffers;
            d->refresh_info();
            return d;
        },
        static_cast<PGM_ConstDataset*>(nullptr));
}
void PGM_destroy_dataset_const(PGM_ConstDataset* dataset) noexcept { delete dataset; }
void PGM_dataset_const_add_buffer(PGM_Handle* handle, PGM_ConstDataset* dataset, char const* component,
                                  PGM_Idx eps, PGM_Idx total, PGM_Idx const* indptr, void const* data) noexcept {
    call_with_catch_void(handle,
                         [&] { add_buffer(dataset, component, eps, total, indptr, const_cast<void*>(data)); });
}
void PGM_dataset_const_add_attribute_buffer(PGM_Handle* handle, PGM_ConstDataset* dataset, char const* component,
                                            char const* attribute, void const* data) noexcept {
    call_with_catch_void(handle, [&] { add_attr(dataset, component, attribute, const_cast<void*>(data)); });
}
PGM_DatasetInfo const* PGM_dataset_const_get_info(PGM_Handle*, PGM_ConstDataset const* dataset) noexcept {
    dataset->refresh_info();
    return &dataset->info;
}
void PGM_dataset_const_set_next_cartesian_product_dimension(PGM_Handle* handle, PGM_ConstDataset* dataset,
                                                            PGM_ConstDataset const* next) noexcept {
    call_with_catch_void(handle, [&] {
        for (auto const* p = next; p; p = p->next)
            if (p == dataset) throw DatasetError("cyclic cartesian product dimension");
        dataset->next = next;
    });
}
PGM_MutableDataset* PGM_create_dataset_mutable(PGM_Handle* handle, char const* dataset, PGM_Idx is_batch,
                                               PGM_Idx batch_size) noexcept {
    return call_with_catch(
        handle,
        [&]() -> PGM_MutableDataset* {
            if (!is_batch && batch_size != 1) throw DatasetError("a single dataset has batch size 1");
            auto* d = new PGM_MutableDataset{};
            d->meta = &meta_dataset(dataset);
            d->is_batch = is_batch != 0;
            d->batch_size = batch_size;
            d->refresh_info();
            return d;
        },
        static_cast<PGM_MutableDataset*>(nullptr));
}
void PGM_destroy_dataset_mutable(PGM_MutableDataset* dataset) noexcept { delete dataset; }
void PGM_dataset_mutable_add_buffer(PGM_Handle* handle, PGM_MutableDataset* dataset, char const* component,
                                    PGM_Idx eps, PGM_Idx total, PGM_Idx const* indptr, void* data) noexcept {
    call_with_catch_void(handle, [&] { add_buffer(dataset, component, eps, total, indptr, data); });
}
void PGM_dataset_mutable_add_attribute_buffer(PGM_Handle* handle, PGM_MutableDataset* dataset,
                                              char const* component, char const* attribute, void* data) noexcept {
    call_with_catch_void(handle, [&] { add_attr(dataset, component, attribute, data); });
}
PGM_DatasetInfo const* PGM_dataset_mutable_get_info(PGM_Handle*, PGM_MutableDataset const* dataset) noexcept {
    dataset->refresh_info();
    return &dataset->info;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// writable datasets and (de)serialization: not provided by this build (dataset.h:241-269,
// serialization.h:29-114); the reference's Python layer binds them at import, so they exist and fail
// with a regular error
// ------------------------------------------------------------------------------------------------
namespace {
void unsupported(PGM_Handle* h, char const* what) {
    if (!h) return;
    h->clear();
    h->err_code = PGM_regular_error;
    h->err_msg = std::string(what) + " is not supported by this build (power flow only)\n";
}
}  // namespace

PGM_DatasetInfo const* PGM_dataset_writable_get_info(PGM_Handle* handle, PGM_WritableDataset const*) noexcept {
    unsupported(handle, "Writable datasets (deserialization)");
    return nullptr;
}
void PGM_dataset_writable_set_buffer(PGM_Handle* handle, PGM_WritableDataset*, char const*, PGM_Idx*, void*) noexcept {
    unsupported(handle, "Writable datasets (deserialization)");
}
void PGM_dataset_writable_set_attribute_buffer(PGM_Handle* handle, PGM_WritableDataset*, char const*, char const*,
                                               void*) noexcept {
    unsupported(handle, "Writable datasets (deserialization)");
}
PGM_Deserializer* PGM_create_deserializer_from_binary_buffer(PGM_Handle* handle, char const*, PGM_Idx,
                                                             PGM_Idx) noexcept {
    unsupported(handle, "Deserialization");
    return nullptr;
}
PGM_Deserializer* PGM_create_deserializer_from_null_terminated_string(PGM_Handle* handle, char const*,
                                                                      PGM_Idx) noexcept {
    unsupported(handle, "Deserialization");
    return nullptr;
}
PGM_WritableDataset* PGM_deserializer_get_dataset(PGM_Handle* handle, PGM_Deserializer*) noexcept {
    unsupported(handle, "Deserialization");
    return nullptr;
}
void PGM_deserializer_parse_to_buffer(PGM_Handle* handle, PGM_Deserializer*) noexcept {
    unsupported(handle, "Deserialization");
}
void PGM_destroy_deserializer(PGM_Deserializer*) noexcept {}
PGM_Serializer* PGM_create_serializer(PGM_Handle* handle, PGM_ConstDataset const*, PGM_Idx) noexcept {
    unsupported(handle, "Serialization");
    return nullptr;
}
void PGM_serializer_get_to_binary_buffer(PGM_Handle* handle, PGM_Serializer*, PGM_Idx, char const** data,
                                         PGM_Idx* size) noexcept {
    if (data) *data = nullptr;
    if (size) *size = 0;
    unsupported(handle, "Serialization");
}
char const* PGM_serializer_get_to_zero_terminated_string(PGM_Handle* handle, PGM_Serializer*, PGM_Idx,
                                                         PGM_Idx) noexcept {
    unsupported(handle, "Serialization");
    return nullptr;
}
void PGM_destroy_serializer(PGM_Serializer*) noexcept {}
