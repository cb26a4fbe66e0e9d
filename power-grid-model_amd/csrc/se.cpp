// se.cpp -- batched Newton-Raphson state estimation (calculation_type state_estimation, method
// newton_raphson, symmetric).
//
// The host reads every scenario's sensors and component states, groups the scenarios by topology (the
// scenarios of a group share one bus ordering and Y-bus pattern), aggregates each scenario's measurements
// per bus, branch side and shunt (measured_values.hpp) and checks their observability (observability.hpp).
// The device assembles and solves the gain system of every scenario and iterates to the error tolerance
// (se_kernels.hip, 4x4-block sweeps of kernels.hip). The results (node voltages and injections, branch and
// appliance flows, sensor residuals) are formed on the host from the estimated voltages.
//
// Reference (power_grid_model_c/power_grid_model/include/power_grid_model/):
//   sensors and calculation parameters   component/{voltage_sensor,power_sensor}.hpp
//   state-estimation input per math model main_core/calculation_input_preparation.hpp:191-258
//   measurement aggregation               math_solver/measured_values.hpp:55-651
//   observability                         math_solver/observability.hpp:60-240, 644-698
//   solver loop                           math_solver/newton_raphson_se_solver.hpp:163-211
//   results                               math_solver/common_solver_functions.hpp:449-456,
//                                         math_solver/y_bus.hpp:482-546, main_core/output.hpp:189-280
#include "se.hpp"

#include "../../include/power_grid_model_amd.h"
#include "engine.hpp"
#include "kernels.hpp"
#include "topology.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace pgm_amd {

namespace {
// target number of parallel tasks per level of a group's schedule (PGM_AMD_SE_TASKS overrides). The sweeps of
// the 4x4 gain blocks are latency-bound on their pivot chains: se_meshed_1000 (1024 scenarios) takes the same
// 48 ms of iterations at every target from 64 to 4096
int se_tasks(int blocks) {
    (void)blocks;
    static int const t = [] {
        char const* v = std::getenv("PGM_AMD_SE_TASKS");
        int const x = v ? std::atoi(v) : 0;
        return x > 0 ? x : 256;
    }();
    return t;
}

// phase timings of a state-estimation call on stderr when PGM_AMD_TRACE is set (diagnostics only)
struct SeTrace {
    bool on = std::getenv("PGM_AMD_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(char const* what) {
        if (!on) return;
        auto const now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[pgm_amd se] %-28s %8.2f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};
}  // namespace

#define SE_HIP_CHECK(expr)                                                                                       \
    do {                                                                                                         \
        hipError_t const e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                                    \
            throw PowerGridError(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_) + "\n");    \
    } while (0)

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr int kRefineSolves = 6;  // sparse_lu_solver.hpp:286, 524-528: one solve and five refinements
constexpr double kRefineConverge = 1e-13;

struct NotObservable : PowerGridError {
    explicit NotObservable(std::string const& msg)
        : PowerGridError("Not enough measurements available for state estimation.\n" + msg + "\n") {}
};

// measurements (statistics.hpp): a complex voltage with one variance (imaginary NaN: magnitude only), a
// power with separate variances of its real and imaginary parts
struct VoltMeas {
    double re = 0.0, im = 0.0, var = kInf;
};
struct PowMeas {
    double p = 0.0, q = 0.0, vp = kInf, vq = kInf;
};

// statistics::combine over uniform real variables: inverse-variance weighted mean; infinite variance when
// the accumulated inverse variance is not a normal number
void combine_real(double& value, double& var, std::vector<std::pair<double, double>> const& xs) {
    double acc = 0.0, wsum = 0.0;
    for (auto const& [v, s2] : xs) {
        acc += 1.0 / s2;
        wsum += v / s2;
    }
    if (!std::isnormal(acc)) {
        value = wsum;
        var = kInf;
        return;
    }
    value = wsum / acc;
    var = 1.0 / acc;
}

// measured_values.hpp:276-305: magnitude-only combination when any sensor lacks an angle
VoltMeas combine_voltage(std::vector<VoltMeas> const& xs) {
    VoltMeas r;
    bool any_nan = false;
    for (auto const& x : xs) any_nan = any_nan || std::isnan(x.im);
    if (any_nan) {
        std::vector<std::pair<double, double>> mags;
        for (auto const& x : xs) mags.push_back({std::isnan(x.im) ? x.re : std::hypot(x.re, x.im), x.var});
        combine_real(r.re, r.var, mags);
        r.im = kNaN;
        return r;
    }
    double acc = 0.0, wre = 0.0, wim = 0.0;
    for (auto const& x : xs) {
        acc += 1.0 / x.var;
        wre += x.re / x.var;
        wim += x.im / x.var;
    }
    if (!std::isnormal(acc)) {
        r.re = wre;
        r.im = wim;
        r.var = kInf;
        return r;
    }
    r.re = wre / acc;
    r.im = wim / acc;
    r.var = 1.0 / acc;
    return r;
}

// statistics::combine over decomposed complex variables (statistics.hpp:348-366); the real and imaginary parts
// combine as combine_real does, in one pass without temporaries
template <class Range> PowMeas combine_power(Range const& xs) {
    double accp = 0.0, wp = 0.0, accq = 0.0, wq = 0.0;
    for (auto const& x : xs) {
        accp += 1.0 / x.vp;
        wp += x.p / x.vp;
        accq += 1.0 / x.vq;
        wq += x.q / x.vq;
    }
    PowMeas r;
    if (std::isnormal(accp)) {
        r.p = wp / accp;
        r.vp = 1.0 / accp;
    } else {
        r.p = wp;
        r.vp = kInf;
    }
    if (std::isnormal(accq)) {
        r.q = wq / accq;
        r.vq = 1.0 / accq;
    } else {
        r.q = wq;
        r.vq = kInf;
    }
    if (!(std::isnormal(r.vp) && std::isnormal(r.vq))) r.vp = r.vq = kInf;
    return r;
}
PowMeas combine_power(std::initializer_list<PowMeas> xs) { return combine_power<std::initializer_list<PowMeas>>(xs); }

// sensor calculation parameters in per unit (voltage_sensor.hpp:126-135, power_sensor.hpp:110-152)
VoltMeas voltage_param(Model const& m, ComponentState const& st, Idx i) {
    double const ur = m.node[m.vs_node[i]].u_rated;
    double const u = st.vs_u[i] / ur, sigma = st.vs_sigma[i] / ur, ang = st.vs_angle[i];
    VoltMeas v;
    v.var = sigma * sigma;
    if (!std::isnan(ang)) {
        v.re = u * std::cos(ang);
        v.im = u * std::sin(ang);
    } else {
        v.re = u;
        v.im = kNaN;
    }
    return v;
}
double power_direction(int term) { return term == T_LOAD || term == T_SHUNT ? -1.0 : 1.0; }
PowMeas power_param(Model const& m, ComponentState const& st, Idx i) {
    double const dir = power_direction(m.ps_term[i]);
    double const inv = 1.0 / 1e6;
    double const ps = st.ps_psig[i] * inv, qs = st.ps_qsig[i] * inv, ss = st.ps_sigma[i] * inv;
    PowMeas r;
    if (std::isnormal(ps) && std::isnormal(qs)) {
        r.vp = ps * ps;
        r.vq = qs * qs;
    } else {
        double const var = std::isnan(ps) ? ss * ss / 2 : kInf;
        r.vp = r.vq = var;
    }
    r.p = dir * st.ps_p[i] * inv;
    r.q = dir * st.ps_q[i] * inv;
    return r;
}

// one scenario's state: the base state with the scenario's update rows applied
struct ScenarioReader {
    Model const& m;
    PGM_ConstDataset const* update;
    std::vector<char> rec;
    // apply scenario s's rows to st; returns the structural key of the rows (components that change the
    // topology or the admittances, with their new values)
    std::string apply(Idx s, ComponentState& st) {
        std::string key;
        if (!update) return key;
        for (auto const& buf : update->buffers) {
            std::string const& name = buf.meta->name;
            int comp = -1;
            for (int c = 0; c < N_COMP; ++c)
                if (name == comp_names[c]) comp = c;
            if (name == kVoltageSensorName) comp = C_VOLTAGE_SENSOR;
            if (name == kPowerSensorName) comp = C_POWER_SENSOR;
            if (comp < 0)
                throw PowerGridError("component " + name + " cannot be updated by a state estimation of this build\n");
            rec.resize(buf.meta->size);
            Idx const b0 = buf.begin(s), b1 = buf.end(s);
            for (Idx r = b0; r < b1; ++r) {
                buf.read_record(r, rec.data());
                ID id;
                std::memcpy(&id, rec.data(), 4);
                Idx idx = id == na_IntID ? r - b0 : find_component(m, comp, id);
                if (comp == C_VOLTAGE_SENSOR) {
                    SymVoltageSensorUpdate u;
                    std::memcpy(&u, rec.data(), sizeof u);
                    apply_voltage_sensor_update(st, idx, u);
                    continue;
                }
                if (comp == C_POWER_SENSOR) {
                    SymPowerSensorUpdate u;
                    std::memcpy(&u, rec.data(), sizeof u);
                    apply_power_sensor_update(st, idx, u);
                    continue;
                }
                int slot = comp;
                line_slot(m, slot, idx);
                bool structural = false;
                apply_update_record(m, st, slot, idx, rec.data(), structural);
                if (slot == C_LINE || slot == C_TRANSFORMER || slot == C_SHUNT || slot == C_SOURCE) {
                    key += std::to_string(slot) + ":" + std::to_string(idx) + ";";
                }
            }
        }
        return key;
    }
};

// the structural part of a state in a key (only the components the scenario's rows touched differ from the
// base state)
std::string state_key(Model const& m, ComponentState const& st, std::string const& touched) {
    if (touched.empty()) return {};
    std::string k;
    char buf[64];
    size_t pos = 0;
    while (pos < touched.size()) {
        size_t const c = touched.find(':', pos), e = touched.find(';', pos);
        int const slot = std::stoi(touched.substr(pos, c - pos));
        Idx const i = std::stoll(touched.substr(c + 1, e - c - 1));
        pos = e + 1;
        if (slot == C_LINE) {
            std::snprintf(buf, sizeof buf, "l%lld:%d%d;", (long long)i, int(st.line_fs[i]), int(st.line_ts[i]));
        } else if (slot == C_TRANSFORMER) {
            std::snprintf(buf, sizeof buf, "t%lld:%d%d%d;", (long long)i, int(st.tr_fs[i]), int(st.tr_ts[i]),
                          int(st.tr_tap[i]));
        } else if (slot == C_SHUNT) {
            std::snprintf(buf, sizeof buf, "s%lld:%d,%a,%a,%a,%a;", (long long)i, int(st.sh_status[i]), st.sh_g1[i],
                          st.sh_b1[i], st.sh_g0[i], st.sh_b0[i]);
        } else {
            std::snprintf(buf, sizeof buf, "c%lld:%d;", (long long)i, int(st.src_status[i]));
        }
        k += buf;
    }
    // the same component may appear more than once in the rows: a sorted, unique list
    std::vector<std::string> parts;
    size_t p0 = 0;
    while (p0 < k.size()) {
        size_t const e = k.find(';', p0);
        parts.push_back(k.substr(p0, e - p0 + 1));
        p0 = e + 1;
    }
    std::sort(parts.begin(), parts.end());
    parts.erase(std::unique(parts.begin(), parts.end()), parts.end());
    std::string r;
    for (auto const& s : parts) r += s;
    return r;
}

// a topology group of scenarios
struct SeGroup {
    ComponentState st;  // structural state (loads and sensors per scenario)
    Topology topo;
    Plan plan;
    std::vector<Idx> scenarios;
    int S = 0, S_pad = 0;
    // sensors to plan objects
    std::vector<int> vs_bus;    // per voltage sensor: plan bus or -1
    std::vector<int> ps_math;   // per power sensor: plan object (bus, branch, shunt, load_gen, source) or -1
    std::vector<int> branch_of; // per component branch slot: plan branch or -1
    std::vector<int> mm_branch_begin;
    // Y-bus elements of every LU entry (kind in the low 2 bits: 0 shunt, 1 / 2 branch seen from its from / to
    // side), branches in math order, then shunts
    std::vector<int> el_ptr, el;
};

// per lane results needed after the solve: the measured values of the appliances
struct LaneAppliances {
    std::vector<int8_t> lg_state, src_state;  // 0 disconnected, 1 unmeasured, 2 measured
    std::vector<PowMeas> lg_meas, src_meas;
    std::vector<PowMeas> bus_appl;            // sum of measured appliances per bus
    std::vector<int> bus_unmeasured;          // appliances connected without a measurement per bus
};

struct HostLanes {
    size_t sp = 0;
    std::vector<int8_t> v_flag, inj_flag, bf_flag, bt_flag, sh_flag, has_angle, use_pert;
    std::vector<double> v_val, inj_val, bf_val, bt_val, sh_val, mean_shift;
    std::vector<int> vref;
    std::vector<int> done;
    std::vector<std::string> lane_error;
    std::vector<LaneAppliances> appl;
};

// measured values and observability of every math model of one lane (measured_values.hpp,
// observability.hpp); fills the lane's columns of h
void measure_lane(Model const& m, SeGroup const& g, ComponentState const& st, int lane, HostLanes& h,
                  std::vector<std::vector<int>> const& vs_at_bus,
                  std::vector<std::vector<std::vector<int>>> const& ps_at) {
    Plan const& p = g.plan;
    size_t const sp = h.sp;
    int const n = int(p.n);
    int const nbr = int(p.branch_comp.size());
    auto& la = h.appl[size_t(lane)];
    // per-thread scratch for the sensors of one object (no allocation per object)
    static thread_local std::vector<PowMeas> obj_xs, bus_xs;
    static thread_local std::vector<VoltMeas> volt_xs;
    la.lg_state.assign(p.load_comp.size(), 0);
    la.src_state.assign(p.src_comp.size(), 0);
    la.lg_meas.assign(p.load_comp.size(), PowMeas{});
    la.src_meas.assign(p.src_comp.size(), PowMeas{});
    la.bus_appl.assign(size_t(n), PowMeas{0.0, 0.0, 0.0, 0.0});
    la.bus_unmeasured.assign(size_t(n), 0);
    // sensor lists per object, in sensor order: ps_at[kind][object] with kind 0 branch from, 1 branch to,
    // 2 shunt, 3 load_gen, 4 source, 5 node
    auto lane_of = [&](size_t item) { return item * sp + size_t(lane); };
    for (int k = 0; k < int(p.n_mm); ++k) {
        int const b0 = p.mm_bus_begin[k], b1 = p.mm_bus_begin[k + 1];
        size_t const pair = size_t(k) * sp + size_t(lane);
        try {
            // voltages
            double angle_cum = 0.0;
            int n_volt = 0, n_angle = 0, first_volt = -1;
            std::vector<double> main_vars;
            std::vector<VoltMeas> vm(static_cast<size_t>(b1 - b0));
            for (int b = b0; b < b1; ++b) {
                h.v_flag[lane_of(size_t(b))] = 0;
                if (vs_at_bus[size_t(b)].empty()) continue;
                volt_xs.clear();
                for (int i : vs_at_bus[size_t(b)]) volt_xs.push_back(voltage_param(m, st, i));
                VoltMeas const a = combine_voltage(volt_xs);
                if (std::isinf(a.var)) continue;
                vm[size_t(b - b0)] = a;
                ++n_volt;
                if (first_volt < 0) first_volt = b;
                bool const ang = !std::isnan(a.im);
                h.v_flag[lane_of(size_t(b))] = ang ? 2 : 1;
                if (ang) {
                    ++n_angle;
                    std::complex<double> const z = std::complex<double>(a.re, a.im) * std::exp(std::complex<double>(0.0, -p.phase[size_t(b)]));
                    angle_cum += std::arg(z);
                }
            }
            h.mean_shift[pair] = n_angle > 0 ? angle_cum / double(n_angle) : 0.0;
            h.has_angle[pair] = n_angle > 0 ? 1 : 0;
            // appliances and injections, bus by bus
            std::vector<PowMeas> inj(static_cast<size_t>(b1 - b0));
            std::vector<int8_t> inj_ok(static_cast<size_t>(b1 - b0), 0);
            std::vector<PowMeas> shm(p.shunt_comp.size());
            std::vector<int8_t> sh_ok(p.shunt_comp.size(), 0);
            auto one_object = [&](int kind, int obj, bool connected, PowMeas& out) -> int {
                // process_one_object: 0 disconnected, 1 unmeasured, 2 measured
                if (!connected) return 0;
                auto const& sensors = ps_at[size_t(kind)][size_t(obj)];
                if (sensors.empty()) return 1;
                obj_xs.clear();
                for (int i : sensors) obj_xs.push_back(power_param(m, st, i));
                PowMeas const c = combine_power(obj_xs);
                if (std::isinf(c.vp) || std::isinf(c.vq)) return 1;
                out = c;
                return 2;
            };
            for (int b = b0; b < b1; ++b) {
                for (int q = p.shunt_ptr[size_t(b)]; q < p.shunt_ptr[size_t(b) + 1]; ++q)
                    sh_ok[size_t(q)] = one_object(2, q, st.sh_status[size_t(p.shunt_comp[size_t(q)])] != 0, shm[size_t(q)]) == 2;
                for (int l = p.load_ptr[size_t(b)]; l < p.load_ptr[size_t(b) + 1]; ++l)
                    la.lg_state[size_t(l)] =
                        int8_t(one_object(3, l, st.lg_status[size_t(p.load_comp[size_t(l)])] != 0, la.lg_meas[size_t(l)]));
                for (int q = p.src_ptr[size_t(b)]; q < p.src_ptr[size_t(b) + 1]; ++q)
                    la.src_state[size_t(q)] =
                        int8_t(one_object(4, q, st.src_status[size_t(p.src_comp[size_t(q)])] != 0, la.src_meas[size_t(q)]));
                // combine_appliances_to_injection_measurements
                int n_unmeasured = 0;
                PowMeas appl{0.0, 0.0, 0.0, 0.0};
                auto add = [&](int8_t state, PowMeas const& x) {
                    if (state == 1) {
                        ++n_unmeasured;
                    } else if (state == 2) {
                        appl.p += x.p;
                        appl.q += x.q;
                        appl.vp += x.vp;
                        appl.vq += x.vq;
                    }
                };
                for (int l = p.load_ptr[size_t(b)]; l < p.load_ptr[size_t(b) + 1]; ++l)
                    add(la.lg_state[size_t(l)], la.lg_meas[size_t(l)]);
                for (int q = p.src_ptr[size_t(b)]; q < p.src_ptr[size_t(b) + 1]; ++q)
                    add(la.src_state[size_t(q)], la.src_meas[size_t(q)]);
                la.bus_appl[size_t(b)] = appl;
                la.bus_unmeasured[size_t(b)] = n_unmeasured;
                bus_xs.clear();
                for (int i : ps_at[5][size_t(b)]) bus_xs.push_back(power_param(m, st, i));
                PowMeas const direct = combine_power(bus_xs);
                bool const uncertain = std::isinf(direct.vp) || std::isinf(direct.vq);
                int const lb = b - b0;
                if (n_unmeasured > 0) {
                    if (!uncertain) {
                        inj[size_t(lb)] = direct;
                        inj_ok[size_t(lb)] = 1;
                    }
                } else if (uncertain || appl.vp == 0.0 || appl.vq == 0.0) {
                    inj[size_t(lb)] = appl;
                    inj_ok[size_t(lb)] = 1;
                } else {
                    inj[size_t(lb)] = combine_power({direct, appl});
                    inj_ok[size_t(lb)] = 1;
                }
            }
            // branch sides
            std::vector<PowMeas> bfm(static_cast<size_t>(nbr)), btm(static_cast<size_t>(nbr));
            std::vector<int8_t> bf_ok(static_cast<size_t>(nbr), 0), bt_ok(size_t(nbr), 0);
            int const br0 = g.mm_branch_begin[size_t(k)], br1 = g.mm_branch_begin[size_t(k) + 1];
            for (int b = br0; b < br1; ++b) {
                bf_ok[size_t(b)] = one_object(0, b, p.branch_from[size_t(b)] >= 0, bfm[size_t(b)]) == 2;
                bt_ok[size_t(b)] = one_object(1, b, p.branch_to[size_t(b)] >= 0, btm[size_t(b)]) == 2;
            }
            // normalize_variance: the smallest non-zero variance of the main values becomes one
            double min_var = kInf;
            auto take = [&](double v) {
                if (v != 0.0) min_var = std::min(min_var, v);
            };
            for (int b = b0; b < b1; ++b)
                if (h.v_flag[lane_of(size_t(b))]) take(vm[size_t(b - b0)].var);
            for (int b = b0; b < b1; ++b)
                for (int q = p.shunt_ptr[size_t(b)]; q < p.shunt_ptr[size_t(b) + 1]; ++q)
                    if (sh_ok[size_t(q)]) take(shm[size_t(q)].vp + shm[size_t(q)].vq);
            for (int lb = 0; lb < b1 - b0; ++lb)
                if (inj_ok[size_t(lb)]) take(inj[size_t(lb)].vp + inj[size_t(lb)].vq);
            for (int b = br0; b < br1; ++b) {
                if (bf_ok[size_t(b)]) take(bfm[size_t(b)].vp + bfm[size_t(b)].vq);
                if (bt_ok[size_t(b)]) take(btm[size_t(b)].vp + btm[size_t(b)].vq);
            }
            double const inv = 1.0 / min_var;
            auto put_pow = [&](std::vector<double>& val, size_t item, PowMeas const& x) {
                val[(item * 4) * sp + size_t(lane)] = x.p;
                val[(item * 4 + 1) * sp + size_t(lane)] = x.q;
                val[(item * 4 + 2) * sp + size_t(lane)] = x.vp * inv;
                val[(item * 4 + 3) * sp + size_t(lane)] = x.vq * inv;
            };
            for (int b = b0; b < b1; ++b) {
                size_t const it = size_t(b);
                if (h.v_flag[lane_of(it)]) {
                    VoltMeas const& a = vm[size_t(b - b0)];
                    h.v_val[(it * 3) * sp + size_t(lane)] = a.re;
                    h.v_val[(it * 3 + 1) * sp + size_t(lane)] = a.im;
                    h.v_val[(it * 3 + 2) * sp + size_t(lane)] = a.var * inv;
                }
                h.inj_flag[lane_of(it)] = inj_ok[size_t(b - b0)];
                if (inj_ok[size_t(b - b0)]) put_pow(h.inj_val, it, inj[size_t(b - b0)]);
                for (int q = p.shunt_ptr[it]; q < p.shunt_ptr[it + 1]; ++q) {
                    h.sh_flag[lane_of(size_t(q))] = sh_ok[size_t(q)];
                    if (sh_ok[size_t(q)]) put_pow(h.sh_val, size_t(q), shm[size_t(q)]);
                }
            }
            for (int b = br0; b < br1; ++b) {
                h.bf_flag[lane_of(size_t(b))] = bf_ok[size_t(b)];
                h.bt_flag[lane_of(size_t(b))] = bt_ok[size_t(b)];
                if (bf_ok[size_t(b)]) put_pow(h.bf_val, size_t(b), bfm[size_t(b)]);
                if (bt_ok[size_t(b)]) put_pow(h.bt_val, size_t(b), btm[size_t(b)]);
            }
            // observability (observability.hpp:644-698)
            if (n_volt == 0) throw NotObservable("No voltage sensor found!\n");
            int const nb = b1 - b0;
            std::vector<int8_t> flow(p.col.size(), 0), vph(size_t(nb), 0);
            bool ill = false;
            int total_inj = 0;
            for (int b = b0; b < b1; ++b) {
                bool any = false;
                int const dg = p.diag[size_t(b)];
                if (inj_ok[size_t(b - b0)]) {
                    ++total_inj;
                    flow[size_t(dg)] = 1;
                    any = true;
                }
                for (int e = dg + 1; e < p.row_ptr[size_t(b) + 1]; ++e) {
                    if (p.fill[size_t(e)]) continue;
                    // the entry's branches (both sides connected): one measured branch makes a flow sensor
                    for (int q = g.el_ptr[size_t(e)]; q < g.el_ptr[size_t(e) + 1]; ++q) {
                        int const code = g.el[size_t(q)];
                        if ((code & 3) == 0) continue;
                        int const br = code >> 2;
                        if (bf_ok[size_t(br)] || bt_ok[size_t(br)]) {
                            flow[size_t(e)] = 1;
                            any = true;
                            break;
                        }
                    }
                }
                if (h.v_flag[lane_of(size_t(b))] == 2) {
                    any = true;
                    vph[size_t(b - b0)] = 1;
                }
                if (!any && b != b1 - 1) ill = true;
            }
            int n_flow = 0, n_vph = 0;
            for (int b = b0; b < b1; ++b)
                for (int e = p.row_ptr[size_t(b)]; e < p.row_ptr[size_t(b) + 1]; ++e) n_flow += flow[size_t(e)];
            for (auto v : vph) n_vph += v;
            if (n_vph == 0 && n_flow < nb - 1)
                throw NotObservable("The total number of independent power sensors is not enough to make the grid "
                                    "observable.");
            if (n_vph > 0 && n_flow + n_vph < nb)
                throw NotObservable("The total number of independent power sensors and voltage phasor sensors is not "
                                    "enough to make the grid observable.");
            bool const radial = g.topo.models[size_t(k)].radial;
            bool observable = true;
            if (n_vph > 1 && !radial) {
                observable = false;
            } else if (total_inj > nb - 2) {
                observable = true;
            } else if (radial) {
                // assign_independent_sensors_radial: every branch takes an injection or voltage phasor sensor
                // of its end buses when it has no flow sensor of its own
                for (int b = b0; b < b1 - 1; ++b) {
                    int const cur = p.diag[size_t(b)];
                    int const up_e = cur + 1;
                    int const up_bus = p.col[size_t(up_e)];
                    int const up_diag = p.diag[size_t(up_bus)];
                    int const next_lb = b + 1 - b0;
                    if (flow[size_t(up_e)] == 0) {
                        if (flow[size_t(cur)] == 1) {
                            std::swap(flow[size_t(up_e)], flow[size_t(cur)]);
                        } else if (flow[size_t(up_diag)] == 1) {
                            std::swap(flow[size_t(up_e)], flow[size_t(up_diag)]);
                        } else if (vph[size_t(b - b0)] == 1) {
                            std::swap(flow[size_t(up_e)], vph[size_t(b - b0)]);
                        } else if (vph[size_t(next_lb)] == 1) {
                            std::swap(flow[size_t(up_e)], vph[size_t(next_lb)]);
                        }
                    }
                    flow[size_t(cur)] = 0;
                }
                flow[size_t(p.diag[size_t(b1 - 1)])] = 0;
                int n_ind = 0, n_rem = 0;
                for (int b = b0; b < b1; ++b)
                    for (int e = p.row_ptr[size_t(b)]; e < p.row_ptr[size_t(b) + 1]; ++e) n_ind += flow[size_t(e)];
                for (auto v : vph) n_rem += v;
                if (n_ind < nb - 1 || (n_vph > 0 && n_rem < 1))
                    throw NotObservable("The number of power, current, and voltage phasor sensors appears sufficient, "
                                        "but they are not independent enough. The system is still not observable.\n");
            }
            // a meshed model that passes the counting condition is taken as observable here: the reference's
            // meshed sufficient-condition search is not restated (an unobservable meshed model then fails in
            // the factorisation)
            h.use_pert[pair] = ill && observable ? 1 : 0;
            int const slack = int(p.mm_bus_begin[size_t(k)] + g.topo.models[size_t(k)].slack);
            h.vref[pair] = h.v_flag[lane_of(size_t(slack))] ? slack : first_volt;
        } catch (NotObservable const& e) {
            if (h.lane_error[size_t(lane)].empty()) h.lane_error[size_t(lane)] = e.what();
        }
    }
}

template <class T> T* dput(Arena& A, std::vector<T> const& v) {
    if (v.empty()) return nullptr;
    T* d = static_cast<T*>(A.take(v.size() * sizeof(T)));
    SE_HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}
template <class T> T* dalloc(Arena& A, size_t n) { return static_cast<T*>(A.take(std::max<size_t>(n, 1) * sizeof(T))); }

using cdd = std::complex<double>;

}  // namespace

void run_state_estimation(Model const& m, PGM_Options const& opt, PGM_MutableDataset const& output,
                          PGM_ConstDataset const* update, PGM_Handle* handle, std::shared_ptr<ChunkPool> pool) {
    if (!opt.symmetric)
        throw PowerGridError("asymmetric state estimation is not supported by this build: use symmetric\n");
    if (opt.calculation_method != PGM_newton_raphson)
        throw PowerGridError("The calculation method is not supported by this build for state estimation: only "
                             "newton_raphson is implemented\n");
    ensure_device();
    SeTrace tr;
    Idx const S = update ? update->batch_size : 1;
    ComponentState base = m.base;
    // ---- scenarios grouped by topology ----
    std::map<std::string, int> key_to_group;
    std::vector<std::unique_ptr<SeGroup>> groups;
    std::vector<int> scn_group(static_cast<size_t>(S), -1);
    std::vector<int> scn_lane(static_cast<size_t>(S), -1);
    std::vector<std::string> scn_error(static_cast<size_t>(S));
    ScenarioReader reader{m, update, {}};
    for (Idx s = 0; s < S; ++s) {
        ComponentState st = base;
        try {
            std::string const touched = reader.apply(s, st);
            std::string const key = state_key(m, st, touched);
            auto it = key_to_group.find(key);
            int gi;
            if (it == key_to_group.end()) {
                gi = int(groups.size());
                key_to_group.emplace(key, gi);
                auto g = std::make_unique<SeGroup>();
                g->st = st;
                groups.push_back(std::move(g));
            } else {
                gi = it->second;
            }
            scn_group[size_t(s)] = gi;
            scn_lane[size_t(s)] = int(groups[size_t(gi)]->scenarios.size());
            groups[size_t(gi)]->scenarios.push_back(s);
        } catch (PowerGridError const& e) {
            if (!update) throw;
            scn_error[size_t(s)] = e.what();
        }
    }
    std::vector<int> status(static_cast<size_t>(S), 0);
    std::vector<std::string> messages(static_cast<size_t>(S));
    for (Idx s = 0; s < S; ++s)
        if (!scn_error[size_t(s)].empty()) {
            status[size_t(s)] = 1;
            messages[size_t(s)] = scn_error[size_t(s)];
        }
    tr.mark("grouping");
    hipStream_t st = nullptr;
    SE_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{st};
    double const err_tol = opt.err_tol;
    Idx const max_iter = opt.max_iter;
    for (auto& gp : groups) {
        SeGroup& g = *gp;
        g.topo = build_topology(m, g.st);
        std::vector<Idx> all(g.topo.models.size());
        std::iota(all.begin(), all.end(), Idx{0});
        g.plan = build_plan(m, g.st, g.topo, all, true, se_tasks((int(g.scenarios.size()) + 63) / 64));
        Plan const& p = g.plan;
        g.S = int(g.scenarios.size());
        g.S_pad = (g.S + 63) / 64 * 64;
        size_t const sp = size_t(g.S_pad);
        int const n = int(p.n), E = int(p.E), nmm = int(p.n_mm);
        int const nbr = int(p.branch_comp.size()), nsh = int(p.shunt_comp.size());
        // sensor maps
        g.mm_branch_begin.assign(g.topo.models.size() + 1, 0);
        for (size_t k = 0; k < g.topo.models.size(); ++k)
            g.mm_branch_begin[k + 1] = g.mm_branch_begin[k] + int(g.topo.models[k].branch.size());
        g.branch_of.assign(size_t(m.n_branch()), -1);
        for (Idx b = 0; b < m.n_branch(); ++b) {
            Idx const gg = g.topo.branch_group[size_t(b)];
            if (gg >= 0) g.branch_of[size_t(b)] = g.mm_branch_begin[size_t(gg)] + int(g.topo.branch_pos[size_t(b)]);
        }
        auto node_bus = [&](Idx node) {
            Idx const gg = g.topo.node_group[size_t(node)];
            return gg < 0 ? -1 : int(p.mm_bus_begin[size_t(gg)] + g.topo.node_pos[size_t(node)]);
        };
        std::vector<int> lg_math(size_t(m.n_lg()), -1), src_math(m.source.size(), -1), sh_math(m.shunt.size(), -1);
        for (size_t k = 0; k < p.load_comp.size(); ++k) lg_math[size_t(p.load_comp[k])] = int(k);
        for (size_t k = 0; k < p.src_comp.size(); ++k) src_math[size_t(p.src_comp[k])] = int(k);
        for (size_t k = 0; k < p.shunt_comp.size(); ++k) sh_math[size_t(p.shunt_comp[k])] = int(k);
        std::vector<std::vector<int>> vs_at_bus(static_cast<size_t>(n));
        g.vs_bus.assign(m.vs_id.size(), -1);
        for (size_t i = 0; i < m.vs_id.size(); ++i) {
            int const b = node_bus(m.vs_node[i]);
            g.vs_bus[i] = b;
            if (b >= 0) vs_at_bus[size_t(b)].push_back(int(i));
        }
        std::vector<std::vector<std::vector<int>>> ps_at(6);
        ps_at[0].resize(size_t(nbr));
        ps_at[1].resize(size_t(nbr));
        ps_at[2].resize(size_t(nsh));
        ps_at[3].resize(p.load_comp.size());
        ps_at[4].resize(p.src_comp.size());
        ps_at[5].resize(size_t(n));
        g.ps_math.assign(m.ps_id.size(), -1);
        for (size_t i = 0; i < m.ps_id.size(); ++i) {
            int const t = m.ps_term[i];
            Idx const o = m.ps_obj[i];
            int kind = -1, mo = -1;
            switch (t) {
            case T_BRANCH_FROM: kind = 0; mo = g.branch_of[size_t(o)]; break;
            case T_BRANCH_TO: kind = 1; mo = g.branch_of[size_t(o)]; break;
            case T_SHUNT: kind = 2; mo = sh_math[size_t(o)]; break;
            case T_LOAD:
            case T_GENERATOR: kind = 3; mo = lg_math[size_t(o)]; break;
            case T_SOURCE: kind = 4; mo = src_math[size_t(o)]; break;
            case T_NODE: kind = 5; mo = node_bus(o); break;
            default: break;
            }
            g.ps_math[i] = mo;
            if (kind >= 0 && mo >= 0) ps_at[size_t(kind)][size_t(mo)].push_back(int(i));
        }
        // per-entry Y-bus elements: branches in math order, then shunts
        std::vector<std::vector<int>> el_of(static_cast<size_t>(E));
        auto find_entry = [&](int r, int c) {
            auto const b0 = p.col.begin() + p.row_ptr[size_t(r)], b1 = p.col.begin() + p.row_ptr[size_t(r) + 1];
            auto const it = std::lower_bound(b0, b1, c);
            return it == b1 || *it != c ? -1 : int(it - p.col.begin());
        };
        for (int b = 0; b < nbr; ++b) {
            int const f = p.branch_from[size_t(b)], t = p.branch_to[size_t(b)];
            if (f < 0 || t < 0) continue;
            int const eft = find_entry(f, t), etf = find_entry(t, f);
            if (eft >= 0) el_of[size_t(eft)].push_back((b << 2) | 1);
            if (etf >= 0 && etf != eft) el_of[size_t(etf)].push_back((b << 2) | 2);
            if (f == t && eft >= 0) el_of[size_t(eft)].push_back((b << 2) | 2);
        }
        for (int b = 0; b < n; ++b)
            for (int q = p.shunt_ptr[size_t(b)]; q < p.shunt_ptr[size_t(b) + 1]; ++q)
                el_of[size_t(p.diag[size_t(b)])].push_back(q << 2);
        g.el_ptr.assign(size_t(E) + 1, 0);
        g.el.clear();
        for (int k = 0; k < E; ++k) {
            g.el.insert(g.el.end(), el_of[size_t(k)].begin(), el_of[size_t(k)].end());
            g.el_ptr[size_t(k) + 1] = int(g.el.size());
        }
        std::vector<int> const& el_ptr = g.el_ptr;
        std::vector<int> const& el = g.el;
        tr.mark("structure");
        // ---- measured values of every lane ----
        HostLanes h;
        h.sp = sp;
        h.v_flag.assign(size_t(n) * sp, 0);
        h.v_val.assign(size_t(n) * 3 * sp, 1.0);
        h.inj_flag.assign(size_t(n) * sp, 0);
        h.inj_val.assign(size_t(n) * 4 * sp, 1.0);
        h.bf_flag.assign(size_t(std::max(nbr, 1)) * sp, 0);
        h.bt_flag.assign(size_t(std::max(nbr, 1)) * sp, 0);
        h.bf_val.assign(size_t(std::max(nbr, 1)) * 4 * sp, 1.0);
        h.bt_val.assign(size_t(std::max(nbr, 1)) * 4 * sp, 1.0);
        h.sh_flag.assign(size_t(std::max(nsh, 1)) * sp, 0);
        h.sh_val.assign(size_t(std::max(nsh, 1)) * 4 * sp, 1.0);
        h.has_angle.assign(size_t(nmm) * sp, 0);
        h.use_pert.assign(size_t(nmm) * sp, 0);
        h.mean_shift.assign(size_t(nmm) * sp, 0.0);
        h.vref.assign(size_t(nmm) * sp, 0);
        h.lane_error.assign(sp, {});
        h.appl.assign(sp, {});
        {
            unsigned const nth = unsigned(std::max<size_t>(1, std::min<size_t>({size_t(16), size_t(std::thread::hardware_concurrency()),
                                                                                 size_t(g.S)})));
            std::vector<std::string> errs(nth);
            auto work = [&](unsigned t) {
                try {
                    ScenarioReader rd{m, update, {}};
                    // contiguous runs of 64 lanes per thread: the host arrays interleave the lanes, and threads
                    // writing neighbouring lanes would share cache lines
                    int const runs = (g.S + 63) / 64;
                    for (int r = int(t); r < runs; r += int(nth))
                    for (int lane = r * 64; lane < std::min(g.S, r * 64 + 64); ++lane) {
                        ComponentState stl = base;
                        rd.apply(g.scenarios[size_t(lane)], stl);
                        measure_lane(m, g, stl, lane, h, vs_at_bus, ps_at);
                    }
                } catch (std::exception const& e) {
                    errs[t] = e.what();
                }
            };
            std::vector<std::thread> th;
            for (unsigned t = 1; t < nth; ++t) th.emplace_back(work, t);
            work(0);
            for (auto& x : th) x.join();
            for (auto const& e : errs)
                if (!e.empty()) throw PowerGridError(e);
        }
        tr.mark("measured values");
        bool const trace = std::getenv("PGM_AMD_TRACE") != nullptr;
        if (trace)
            for (int k = 0; k < nmm; ++k)
                for (int lane = 0; lane < std::min(g.S, 2); ++lane) {
                    size_t const i = size_t(k) * sp + size_t(lane);
                    std::fprintf(stderr, "[pgm_amd se] group lane %d model %d: buses %d use_pert %d has_angle %d vref %d error '%s'\n",
                                 lane, k, int(p.mm_bus_begin[size_t(k) + 1] - p.mm_bus_begin[size_t(k)]), int(h.use_pert[i]),
                                 int(h.has_angle[i]), h.vref[i], h.lane_error[size_t(lane)].c_str());
                }
        // ---- device ----
        Arena A;
        A.set_pool(pool);
        // destroyed before the arena: the work queued on the stream ends before its chunks go back to the
        // model's pool (an error thrown between a launch and its synchronisation leaves work queued)
        struct DrainGuard {
            hipStream_t s;
            ~DrainGuard() { (void)hipStreamSynchronize(s); }
        } drain{st};
        SeDev d;
        d.n = n;
        d.n_mm = nmm;
        d.S_pad = (long long)sp;
        d.row_ptr = dput(A, p.row_ptr);
        d.col = dput(A, p.col);
        d.diag = dput(A, p.diag);
        d.transpose = dput(A, p.transpose);
        d.fill = dput(A, p.fill);
        d.bus_mm = dput(A, p.bus_mm);
        d.phase = dput(A, p.phase);
        d.el_ptr = dput(A, el_ptr);
        d.el = el.empty() ? dalloc<int>(A, 1) : dput(A, el);
        std::vector<double> yent(static_cast<size_t>(E) * 2), ybr(size_t(std::max(nbr, 1)) * 8, 0.0), ysh(size_t(std::max(nsh, 1)) * 2, 0.0);
        for (int k = 0; k < E; ++k) {
            yent[size_t(k) * 2] = p.ydata[size_t(k)].real();
            yent[size_t(k) * 2 + 1] = p.ydata[size_t(k)].imag();
        }
        for (int b = 0; b < nbr; ++b)
            for (int w = 0; w < 4; ++w) {
                ybr[size_t(b) * 8 + 2 * w] = p.branch_y[size_t(b) * 4 + w].real();
                ybr[size_t(b) * 8 + 2 * w + 1] = p.branch_y[size_t(b) * 4 + w].imag();
            }
        for (int q = 0; q < nsh; ++q) {
            ysh[size_t(q) * 2] = p.shunt_y[size_t(q)].real();
            ysh[size_t(q) * 2 + 1] = p.shunt_y[size_t(q)].imag();
        }
        d.yent = dput(A, yent);
        d.ybr = dput(A, ybr);
        d.ysh = dput(A, ysh);
        d.v_flag = dput(A, h.v_flag);
        d.v_val = dput(A, h.v_val);
        d.inj_flag = dput(A, h.inj_flag);
        d.inj_val = dput(A, h.inj_val);
        d.bf_flag = dput(A, h.bf_flag);
        d.bf_val = dput(A, h.bf_val);
        d.bt_flag = dput(A, h.bt_flag);
        d.bt_val = dput(A, h.bt_val);
        d.sh_flag = dput(A, h.sh_flag);
        d.sh_val = dput(A, h.sh_val);
        std::vector<int> slack(static_cast<size_t>(nmm));
        for (int k = 0; k < nmm; ++k) slack[size_t(k)] = int(p.mm_bus_begin[size_t(k)] + g.topo.models[size_t(k)].slack);
        d.slack = dput(A, slack);
        d.vref = dput(A, h.vref);
        d.has_angle = dput(A, h.has_angle);
        d.use_pert = dput(A, h.use_pert);
        size_t const npairs = size_t(nmm) * sp;
        // pair state: lanes beyond the batch and lanes whose measurements failed start done
        std::vector<int> done(npairs, ACTIVE), err(npairs, ERR_NONE);
        for (int k = 0; k < nmm; ++k)
            for (size_t lane = 0; lane < sp; ++lane)
                if (int(lane) >= g.S || !h.lane_error[lane].empty()) done[size_t(k) * sp + lane] = FAILED;
        d.done = dput(A, done);
        int* d_err = dput(A, err);
        int* d_iters = dput(A, std::vector<int>(npairs, 0));
        d.devmax = dput(A, std::vector<unsigned long long>(npairs, 0ull));
        double* d_lastdev = dput(A, std::vector<double>(npairs, 0.0));
        d.norm = dalloc<unsigned long long>(A, npairs);
        d.pthr = dalloc<double>(A, npairs);
        d.pert = dalloc<int>(A, npairs);
        d.rmax = dalloc<unsigned long long>(A, npairs);
        d.berr = dalloc<unsigned long long>(A, npairs);
        d.offset = dalloc<double>(A, npairs);
        d.x = dalloc<double>(A, size_t(n) * 4 * sp);
        d.u = dalloc<double>(A, size_t(n) * 2 * sp);
        d.raw_a = dalloc<double>(A, size_t(E) * 16 * sp);
        d.raw_b = dalloc<double>(A, size_t(n) * 4 * sp);
        d.y = dalloc<double>(A, size_t(n) * 4 * sp);
        d.xacc = dalloc<double>(A, size_t(n) * 4 * sp);
        d.rden = dalloc<double>(A, size_t(n) * 4 * sp);
        double* d_shift = dput(A, h.mean_shift);
        // the 4x4 block LU over the plan's schedule
        DevGroup lu;
        lu.P = 2;
        lu.n = n;
        lu.E = E;
        lu.n_mm = nmm;
        lu.S = g.S_pad;
        lu.S_pad = g.S_pad;
        lu.row_ptr = d.row_ptr;
        lu.row_end = dput(A, p.row_end);
        lu.col = d.col;
        lu.diag = d.diag;
        lu.transpose = d.transpose;
        lu.fill = d.fill;
        std::vector<int> sch;
        for (auto const& t3 : p.schur) sch.insert(sch.end(), t3.begin(), t3.end());
        lu.schur_ptr = dput(A, p.schur_ptr);
        lu.schur = sch.empty() ? dalloc<int>(A, 1) : dput(A, sch);
        lu.task_ptr = dput(A, p.task_ptr);
        lu.task_pivot = dput(A, p.task_pivot);
        lu.task_mm = dput(A, p.task_mm);
        lu.keep = dput(A, p.keep);
        lu.raw_a = d.raw_a;
        lu.raw_b = d.raw_b;
        lu.pthr = d.pthr;
        lu.pert = d.pert;
        lu.y = d.y;
        lu.xacc = d.xacc;
        lu.lu = A.take(size_t(E) * 16 * sp * 8);
        lu.perm = dalloc<uint32_t>(A, size_t(n) * sp);
        lu.done = d.done;
        lu.err = d_err;
        lu.iters = d_iters;
        lu.devmax = d.devmax;
        lu.lastdev = d_lastdev;
        DevGroup* d_lu = dput(A, std::vector<DevGroup>{lu});
        int const nblk = int(sp / 64);
        std::vector<WorkItem*> items;
        std::vector<int> counts;
        for (size_t l = 0; l + 1 < p.level_ptr.size(); ++l) {
            std::vector<WorkItem> it;
            for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q)
                for (int blk = 0; blk < nblk; ++blk) it.push_back(WorkItem{0, p.level_task[size_t(q)], blk});
            items.push_back(dput(A, it));
            counts.push_back(int(it.size()));
        }
        int const L = int(items.size());
        PairState ps;
        ps.n = (long long)npairs;
        ps.done = d.done;
        ps.err = d_err;
        ps.iters = d_iters;
        ps.devmax = d.devmax;
        ps.lastdev = d_lastdev;
        int* d_active = dalloc<int>(A, 2);
        bool any_pert = std::any_of(h.use_pert.begin(), h.use_pert.end(), [](int8_t v) { return v != 0; });
        launch_se_init(d, d_shift, st);
        std::vector<int> hpert(npairs), hdone(npairs);
        std::vector<unsigned long long> hberr(npairs);
        tr.mark("device setup");
        for (Idx it = 1; it <= max_iter; ++it) {
            launch_se_pair_reset(d, st);
            launch_se_gain(d, st);
            for (int l = 0; l < L; ++l) launch_gain_sweep(d_lu, items[size_t(l)], counts[size_t(l)], false, true, st);
            for (int l = L - 1; l >= 0; --l) launch_gain_sweep(d_lu, items[size_t(l)], counts[size_t(l)], false, false, st);
            SE_HIP_CHECK(hipGetLastError());
            if (trace) {
                std::vector<int> td(npairs), te(npairs), tp(npairs);
                std::vector<double> tthr(npairs);
                SE_HIP_CHECK(hipMemcpyAsync(td.data(), d.done, npairs * 4, hipMemcpyDeviceToHost, st));
                SE_HIP_CHECK(hipMemcpyAsync(te.data(), d_err, npairs * 4, hipMemcpyDeviceToHost, st));
                SE_HIP_CHECK(hipMemcpyAsync(tp.data(), d.pert, npairs * 4, hipMemcpyDeviceToHost, st));
                SE_HIP_CHECK(hipMemcpyAsync(tthr.data(), d.pthr, npairs * 8, hipMemcpyDeviceToHost, st));
                SE_HIP_CHECK(hipStreamSynchronize(st));
                std::fprintf(stderr, "[pgm_amd se] iteration %lld pair 0: done %d err %d pert %d threshold %.3e\n",
                             (long long)it, td[0], te[0], tp[0], tthr[0]);
            }
            if (any_pert) {
                // pairs whose factorisation perturbed a pivot: iterative refinement against the gain matrix
                SE_HIP_CHECK(hipMemcpyAsync(hpert.data(), d.pert, npairs * 4, hipMemcpyDeviceToHost, st));
                SE_HIP_CHECK(hipMemcpyAsync(hdone.data(), d.done, npairs * 4, hipMemcpyDeviceToHost, st));
                SE_HIP_CHECK(hipStreamSynchronize(st));
                bool refine = false;
                for (size_t i = 0; i < npairs; ++i)
                    if (hpert[i] && hdone[i] == ACTIVE) {
                        hdone[i] = REFINE;
                        refine = true;
                    }
                if (refine) {
                    SE_HIP_CHECK(hipMemcpyAsync(d.done, hdone.data(), npairs * 4, hipMemcpyHostToDevice, st));
                    launch_se_refine_begin(d, st);
                    for (int pass = 2; pass <= kRefineSolves && refine; ++pass) {
                        launch_se_residual(d, st);
                        launch_se_berr(d, st);
                        for (int l = 0; l < L; ++l)
                            launch_gain_sweep(d_lu, items[size_t(l)], counts[size_t(l)], true, true, st);
                        for (int l = L - 1; l >= 0; --l)
                            launch_gain_sweep(d_lu, items[size_t(l)], counts[size_t(l)], true, false, st);
                        SE_HIP_CHECK(hipMemcpyAsync(hberr.data(), d.berr, npairs * 8, hipMemcpyDeviceToHost, st));
                        SE_HIP_CHECK(hipStreamSynchronize(st));
                        if (trace) {
                            double be0;
                            std::memcpy(&be0, &hberr[0], 8);
                            std::fprintf(stderr, "[pgm_amd se]   refinement solve %d pair 0: state %d backward error %.3e\n",
                                         pass, hdone[0], be0);
                        }
                        refine = false;
                        for (size_t i = 0; i < npairs; ++i) {
                            if (hdone[i] != REFINE) continue;
                            double be;
                            std::memcpy(&be, &hberr[i], 8);
                            if (be <= kRefineConverge) {
                                hdone[i] = ACTIVE;
                            } else {
                                refine = true;
                            }
                        }
                        SE_HIP_CHECK(hipMemcpyAsync(d.done, hdone.data(), npairs * 4, hipMemcpyHostToDevice, st));
                    }
                    if (refine) {
                        // the refinement did not reach its bound within the reference's number of solves
                        std::vector<int> herr(npairs);
                        SE_HIP_CHECK(hipMemcpyAsync(herr.data(), d_err, npairs * 4, hipMemcpyDeviceToHost, st));
                        SE_HIP_CHECK(hipStreamSynchronize(st));
                        for (size_t i = 0; i < npairs; ++i)
                            if (hdone[i] == REFINE) {
                                hdone[i] = FAILED;
                                herr[i] = ERR_SPARSE;
                            }
                        SE_HIP_CHECK(hipMemcpyAsync(d.done, hdone.data(), npairs * 4, hipMemcpyHostToDevice, st));
                        SE_HIP_CHECK(hipMemcpyAsync(d_err, herr.data(), npairs * 4, hipMemcpyHostToDevice, st));
                    }
                }
            }
            launch_se_update(d, st);
            int zero[2] = {0, 0};
            SE_HIP_CHECK(hipMemcpyAsync(d_active, zero, sizeof zero, hipMemcpyHostToDevice, st));
            launch_converge(ps, int(it), int(max_iter), err_tol, d_active, st);
            int act[2] = {0, 0};
            SE_HIP_CHECK(hipMemcpyAsync(act, d_active, sizeof act, hipMemcpyDeviceToHost, st));
            SE_HIP_CHECK(hipStreamSynchronize(st));
            SE_HIP_CHECK(hipGetLastError());
            if (act[0] == 0) break;
        }
        tr.mark("iterations");
        // ---- per-lane status and voltages ----
        std::vector<int> herr(npairs), hits(npairs);
        std::vector<double> hldev(npairs), hu(size_t(n) * 2 * sp);
        SE_HIP_CHECK(hipMemcpyAsync(hdone.data(), d.done, npairs * 4, hipMemcpyDeviceToHost, st));
        SE_HIP_CHECK(hipMemcpyAsync(herr.data(), d_err, npairs * 4, hipMemcpyDeviceToHost, st));
        SE_HIP_CHECK(hipMemcpyAsync(hldev.data(), d_lastdev, npairs * 8, hipMemcpyDeviceToHost, st));
        SE_HIP_CHECK(hipMemcpyAsync(hu.data(), d.u, hu.size() * 8, hipMemcpyDeviceToHost, st));
        SE_HIP_CHECK(hipStreamSynchronize(st));
        std::vector<char> lane_ok(static_cast<size_t>(g.S), 0);
        for (int lane = 0; lane < g.S; ++lane) {
            Idx const s = g.scenarios[size_t(lane)];
            std::string msg = h.lane_error[size_t(lane)];
            if (msg.empty())
                for (int k = 0; k < nmm && msg.empty(); ++k) {
                    size_t const i = size_t(k) * sp + size_t(lane);
                    if (herr[i] == ERR_SPARSE) {
                        msg = "Sparse matrix error, possibly singular matrix!\n"
                              "If you get this error from state estimation, it might mean the system is not fully "
                              "observable, i.e. not enough measurements.\n"
                              "It might also mean that you are running into a corner case where PGM cannot resolve "
                              "yet.\n"
                              "See https://github.com/PowerGridModel/power-grid-model/issues/864.";
                    } else if (herr[i] == ERR_DIVERGE) {
                        msg = "Iteration failed to converge after " + std::to_string(max_iter) +
                              " iterations! Max deviation: " + format_double(hldev[i]) +
                              ", error tolerance: " + format_double(err_tol) + ".\n";
                    }
                }
            if (!msg.empty()) {
                status[size_t(s)] = 1;
                messages[size_t(s)] = msg;
            } else {
                lane_ok[size_t(lane)] = 1;
            }
        }
        tr.mark("status");
        // ---- results of every converged lane, written into the output dataset ----
        unsigned const nth = unsigned(std::max<size_t>(1, std::min<size_t>({size_t(16), size_t(std::thread::hardware_concurrency()),
                                                                             size_t(g.S)})));
        std::vector<std::string> errs(nth);
        auto write_lane = [&](int lane, ComponentState const& stl) {
            Idx const s = g.scenarios[size_t(lane)];
            auto U = [&](int b) { return cdd(hu[(size_t(b) * 2) * sp + size_t(lane)], hu[(size_t(b) * 2 + 1) * sp + size_t(lane)]); };
            // bus injections, branch and shunt flows (y_bus.hpp:482-546)
            std::vector<cdd> sbus(static_cast<size_t>(n));
            for (int b = 0; b < n; ++b) {
                cdd i_inj = 0.0;
                for (int k = p.row_ptr[size_t(b)]; k < p.row_ptr[size_t(b) + 1]; ++k) {
                    if (p.fill[size_t(k)]) continue;
                    i_inj += p.ydata[size_t(k)] * U(p.col[size_t(k)]);
                }
                sbus[size_t(b)] = std::conj(i_inj) * U(b);
            }
            std::vector<cdd> sf(static_cast<size_t>(nbr)), st_(static_cast<size_t>(nbr)), if_(static_cast<size_t>(nbr)), it_(static_cast<size_t>(nbr));
            for (int b = 0; b < nbr; ++b) {
                int const f = p.branch_from[size_t(b)], t = p.branch_to[size_t(b)];
                cdd const uf = f >= 0 ? U(f) : cdd(0.0), ut = t >= 0 ? U(t) : cdd(0.0);
                cdd const* y = &p.branch_y[size_t(b) * 4];
                if_[size_t(b)] = y[0] * uf + y[1] * ut;
                it_[size_t(b)] = y[2] * uf + y[3] * ut;
                sf[size_t(b)] = uf * std::conj(if_[size_t(b)]);
                st_[size_t(b)] = ut * std::conj(it_[size_t(b)]);
            }
            std::vector<cdd> ssh(static_cast<size_t>(nsh)), ish(static_cast<size_t>(nsh));
            for (int b = 0; b < n; ++b)
                for (int q = p.shunt_ptr[size_t(b)]; q < p.shunt_ptr[size_t(b) + 1]; ++q) {
                    ish[size_t(q)] = -p.shunt_y[size_t(q)] * U(b);
                    ssh[size_t(q)] = U(b) * std::conj(ish[size_t(q)]);
                }
            // load_gen and source flows from the measurements (measured_values.hpp:124-150, 604-647)
            LaneAppliances const& la = h.appl[size_t(lane)];
            std::vector<cdd> slg(p.load_comp.size(), 0.0), ssrc(p.src_comp.size(), 0.0);
            for (int b = 0; b < n; ++b) {
                cdd const sb = sbus[size_t(b)];
                PowMeas const& ap = la.bus_appl[size_t(b)];
                int const nu = la.bus_unmeasured[size_t(b)];
                if (nu > 0) {
                    cdd const resid = (sb - cdd(ap.p, ap.q)) / double(nu);
                    for (int l = p.load_ptr[size_t(b)]; l < p.load_ptr[size_t(b) + 1]; ++l) {
                        if (la.lg_state[size_t(l)] == 2) slg[size_t(l)] = cdd(la.lg_meas[size_t(l)].p, la.lg_meas[size_t(l)].q);
                        else if (la.lg_state[size_t(l)] == 1) slg[size_t(l)] = resid;
                    }
                    for (int q = p.src_ptr[size_t(b)]; q < p.src_ptr[size_t(b) + 1]; ++q) {
                        if (la.src_state[size_t(q)] == 2) ssrc[size_t(q)] = cdd(la.src_meas[size_t(q)].p, la.src_meas[size_t(q)].q);
                        else if (la.src_state[size_t(q)] == 1) ssrc[size_t(q)] = resid;
                    }
                } else {
                    cdd const delta = cdd(ap.p, ap.q) - sb;
                    cdd const mu = cdd(delta.real() / ap.vp, delta.imag() / ap.vq);
                    auto adj = [&](PowMeas const& x) { return cdd(x.p - x.vp * mu.real(), x.q - x.vq * mu.imag()); };
                    for (int l = p.load_ptr[size_t(b)]; l < p.load_ptr[size_t(b) + 1]; ++l)
                        if (la.lg_state[size_t(l)] == 2) slg[size_t(l)] = adj(la.lg_meas[size_t(l)]);
                    for (int q = p.src_ptr[size_t(b)]; q < p.src_ptr[size_t(b) + 1]; ++q)
                        if (la.src_state[size_t(q)] == 2) ssrc[size_t(q)] = adj(la.src_meas[size_t(q)]);
                }
            }
            // records
            auto write_comp = [&](char const* name, auto const& fill_one) {
                ComponentBuffer const* buf = output.find(name);
                if (!buf) return;
                Idx const b0 = buf->begin(s), b1 = buf->end(s);
                std::vector<char> r(buf->meta->size);
                for (Idx i = b0; i < b1; ++i) {
                    set_null_records(*buf->meta, r.data(), 1);
                    fill_one(i - b0, r.data());
                    buf->write_record(i, r.data());
                }
            };
            double const bp = 1e6;
            write_comp("node", [&](Idx i, char* r) {
                NodeOutput<1> o{};
                o.id = m.node[size_t(i)].id;
                int const b = node_bus(i);
                o.energized = b >= 0 ? 1 : 0;
                if (b >= 0) {
                    cdd const u = U(b);
                    o.u_pu[0] = std::abs(u);
                    o.u[0] = o.u_pu[0] * m.node[size_t(i)].u_rated;
                    o.u_angle[0] = std::arg(u);
                    o.p[0] = bp * sbus[size_t(b)].real();
                    o.q[0] = bp * sbus[size_t(b)].imag();
                } else {
                    o.u_pu[0] = o.u[0] = o.u_angle[0] = o.p[0] = o.q[0] = 0.0;
                }
                std::memcpy(r, &o, sizeof o);
            });
            auto branch_rec = [&](Idx slot, ID id, double bif, double bit, double rating, bool is_line, char* r) {
                BranchOutput<1> o{};
                o.id = id;
                int const b = g.branch_of[size_t(slot)];
                o.energized = b >= 0 ? 1 : 0;
                if (b >= 0) {
                    o.p_from[0] = bp * sf[size_t(b)].real();
                    o.q_from[0] = bp * sf[size_t(b)].imag();
                    o.i_from[0] = bif * std::abs(if_[size_t(b)]);
                    o.s_from[0] = bp * std::abs(sf[size_t(b)]);
                    o.p_to[0] = bp * st_[size_t(b)].real();
                    o.q_to[0] = bp * st_[size_t(b)].imag();
                    o.i_to[0] = bit * std::abs(it_[size_t(b)]);
                    o.s_to[0] = bp * std::abs(st_[size_t(b)]);
                    o.loading = is_line ? std::max(o.i_from[0], o.i_to[0]) / rating : std::max(o.s_from[0], o.s_to[0]) / rating;
                }
                std::memcpy(r, &o, sizeof o);
            };
            write_comp("line", [&](Idx i, char* r) {
                branch_rec(i, m.line[size_t(i)].id, m.lstat[size_t(i)].base_i, m.lstat[size_t(i)].base_i_to, m.lstat[size_t(i)].i_n, true, r);
            });
            write_comp("link", [&](Idx i, char* r) {
                Idx const slot = Idx(m.line.size()) + i;
                branch_rec(slot, m.link[size_t(i)].id, m.lstat[size_t(slot)].base_i, m.lstat[size_t(slot)].base_i_to, m.lstat[size_t(slot)].i_n, true, r);
            });
            write_comp("generic_branch", [&](Idx i, char* r) {
                Idx const slot = generic_slot0(m) + i;
                branch_rec(slot, m.gbranch[size_t(i)].id, m.lstat[size_t(slot)].base_i, m.lstat[size_t(slot)].base_i_to, m.lstat[size_t(slot)].sn, false, r);
            });
            write_comp("transformer", [&](Idx i, char* r) {
                branch_rec(m.n_line() + i, m.trafo[size_t(i)].id, m.tstat[size_t(i)].base_i_from, m.tstat[size_t(i)].base_i_to, m.tstat[size_t(i)].sn, false, r);
            });
            auto appl_rec = [&](ID id, int8_t energized, int bus, cdd sv, cdd iv, double dir, double base_i, char* r) {
                ApplianceOutput<1> o{};
                o.id = id;
                o.energized = bus >= 0 ? energized : 0;
                if (bus >= 0) {
                    o.p[0] = bp * sv.real() * dir;
                    o.q[0] = bp * sv.imag() * dir;
                    o.s[0] = bp * std::abs(sv);
                    o.i[0] = base_i * std::abs(iv);
                    o.pf[0] = o.s[0] < 1e-8 ? 0.0 : o.p[0] / o.s[0];
                }
                std::memcpy(r, &o, sizeof o);
            };
            auto base_i_of = [&](Idx node) { return 1e6 / m.node[size_t(node)].u_rated / std::sqrt(3.0); };
            write_comp("shunt", [&](Idx i, char* r) {
                int const q = sh_math[size_t(i)];
                int const bus = q >= 0 ? node_bus(m.sh_node[size_t(i)]) : -1;
                appl_rec(m.shunt[size_t(i)].id, int8_t(stl.sh_status[size_t(i)] ? 1 : 0), bus, q >= 0 ? ssh[size_t(q)] : cdd(0.0),
                         q >= 0 ? ish[size_t(q)] : cdd(0.0), -1.0, base_i_of(m.sh_node[size_t(i)]), r);
            });
            write_comp("source", [&](Idx i, char* r) {
                int const q = src_math[size_t(i)];
                int const bus = q >= 0 ? node_bus(m.src_node[size_t(i)]) : -1;
                cdd const sv = q >= 0 ? ssrc[size_t(q)] : cdd(0.0);
                cdd const iv = q >= 0 ? std::conj(sv / U(bus)) : cdd(0.0);
                appl_rec(m.source[size_t(i)].id, int8_t(stl.src_status[size_t(i)] ? 1 : 0), bus, sv, iv, 1.0,
                         base_i_of(m.src_node[size_t(i)]), r);
            });
            int const lg_comps[4] = {C_SYM_GEN, C_ASYM_GEN, C_SYM_LOAD, C_ASYM_LOAD};
            for (int kk = 0; kk < 4; ++kk) {
                Idx const off = m.lg_offset[size_t(kk)];
                double const dir = kk < 2 ? 1.0 : -1.0;
                write_comp(comp_names[lg_comps[kk]], [&](Idx i, char* r) {
                    Idx const l = off + i;
                    int const q = lg_math[size_t(l)];
                    int const bus = q >= 0 ? node_bus(m.lg_node[size_t(l)]) : -1;
                    cdd const sv = q >= 0 ? slg[size_t(q)] : cdd(0.0);
                    cdd const iv = q >= 0 ? std::conj(sv / U(bus)) : cdd(0.0);
                    appl_rec(m.lg_id[size_t(l)], int8_t(stl.lg_status[size_t(l)] ? 1 : 0), bus, sv, iv, dir,
                             base_i_of(m.lg_node[size_t(l)]), r);
                });
            }
            write_comp(kVoltageSensorName, [&](Idx i, char* r) {
                VoltageSensorOutput<1> o{};
                o.id = m.vs_id[size_t(i)];
                int const b = g.vs_bus[size_t(i)];
                if (b >= 0) {
                    o.energized = 1;
                    VoltMeas const v = voltage_param(m, stl, i);
                    cdd const u = U(b);
                    double const ur = m.node[size_t(m.vs_node[size_t(i)])].u_rated;
                    if (!std::isnan(v.im)) {
                        o.u_residual[0] = (std::hypot(v.re, v.im) - std::abs(u)) * ur;
                        double const a = std::atan2(v.im, v.re) - std::arg(u);
                        o.u_angle_residual[0] = std::atan2(std::sin(a), std::cos(a));
                    } else {
                        o.u_residual[0] = (v.re - std::abs(u)) * ur;
                        o.u_angle_residual[0] = kNaN;
                    }
                } else {
                    o.energized = 0;
                    o.u_residual[0] = o.u_angle_residual[0] = 0.0;
                }
                std::memcpy(r, &o, sizeof o);
            });
            write_comp(kPowerSensorName, [&](Idx i, char* r) {
                PowerSensorOutput<1> o{};
                o.id = m.ps_id[size_t(i)];
                int const mo = g.ps_math[size_t(i)];
                if (mo >= 0) {
                    o.energized = 1;
                    cdd sc = 0.0;
                    switch (m.ps_term[size_t(i)]) {
                    case T_BRANCH_FROM: sc = sf[size_t(mo)]; break;
                    case T_BRANCH_TO: sc = st_[size_t(mo)]; break;
                    case T_SOURCE: sc = ssrc[size_t(mo)]; break;
                    case T_SHUNT: sc = ssh[size_t(mo)]; break;
                    case T_LOAD:
                    case T_GENERATOR: sc = slg[size_t(mo)]; break;
                    case T_NODE: sc = sbus[size_t(mo)]; break;
                    default: break;
                    }
                    PowMeas const pm = power_param(m, stl, i);
                    double const dir = power_direction(m.ps_term[size_t(i)]);
                    cdd const res = (cdd(pm.p, pm.q) - sc) * dir * bp;
                    o.p_residual[0] = res.real();
                    o.q_residual[0] = res.imag();
                } else {
                    o.energized = 0;
                    o.p_residual[0] = o.q_residual[0] = 0.0;
                }
                std::memcpy(r, &o, sizeof o);
            });
        };
        auto out_work = [&](unsigned t) {
            try {
                ScenarioReader rd{m, update, {}};
                int const runs = (g.S + 63) / 64;
                for (int r = int(t); r < runs; r += int(nth))
                for (int lane = r * 64; lane < std::min(g.S, r * 64 + 64); ++lane) {
                    if (!lane_ok[size_t(lane)]) continue;
                    ComponentState stl = base;
                    rd.apply(g.scenarios[size_t(lane)], stl);
                    write_lane(lane, stl);
                }
            } catch (std::exception const& e) {
                errs[t] = e.what();
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nth; ++t) th.emplace_back(out_work, t);
        out_work(0);
        for (auto& x : th) x.join();
        for (auto const& e : errs)
            if (!e.empty()) throw PowerGridError(e);
    }
    tr.mark("results");
    // ---- errors (job_dispatch.hpp:208-224) ----
    if (handle) {
        std::vector<Idx> failed;
        std::string combined;
        std::vector<std::string> msgs;
        for (Idx s = 0; s < S; ++s)
            if (status[size_t(s)] != 0) {
                failed.push_back(s);
                msgs.push_back(messages[size_t(s)]);
                combined += "Error in batch #" + std::to_string(s) + ": " + messages[size_t(s)] + "\n";
            }
        if (!failed.empty()) {
            if (update) {
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

}  // namespace pgm_amd
