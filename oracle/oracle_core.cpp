// oracle_core.cpp -- CPU FP64 restatement of power-grid-model's Newton-Raphson power flow.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path, and the CPU baseline timed by
// bench.py. The product never links or calls this file.
//
// Each part cites the reference file it restates (paths relative to
// power_grid_model_c/power_grid_model/include/power_grid_model/):
//   per-unit parameters   component/{line,branch,transformer,transformer_utils,source,load_gen,shunt}.hpp
//   topology + ordering   topology.hpp:196-368, sparse_ordering.hpp:20-219
//   Y-bus structure       math_solver/y_bus.hpp:36-293, 342-431
//   NR iteration          math_solver/newton_raphson_pf_solver.hpp:255-349, 462-547, 706-870
//   NR driver             math_solver/iterative_pf_solver.hpp:33-90
//   block sparse LU       math_solver/sparse_lu_solver.hpp:85-200, 346-495, 769-827
//   results               math_solver/y_bus.hpp:482-546, common_solver_functions.hpp:80-160, 383-446
//   outputs               component/{node,branch,appliance}.hpp get_output, main_core/output.hpp,
//                         main_core/topological_node_output.hpp:69-117
//   batch semantics       job_dispatch.hpp:35-172 (threads stride scenarios; per-scenario errors)
//   tap changer           optimizer/tap_position_optimizer.hpp:209-1110, transformer_tap_regulator.hpp
//   other methods         math_solver.hpp:43-64 and 142-156, linear_pf_solver.hpp,
//                         iterative_current_pf_solver.hpp (complex systems solved in their real
//                         [[G, -B], [B, G]] form: equal to the reference's up to rounding)
//
// Pinning (tests/test_oracle_golden.py): minimum degree, topology, Y-bus pattern and block LU against
// the reference's known answers; NR against closed-form grids at 1e-12; every method against the
// reference's validation fixtures that list it. Parity unpinned: per-scenario iteration counts on large
// generated grids and bitwise agreement with Eigen's operation order (no reference fixture holds them).

#include "pgm_oracle.h"

#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

using cd = std::complex<double>;
using Idx = int64_t;
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt3 = 1.7320508075688772935;
constexpr double kDeg30 = kPi / 6.0;
const cd kA{-0.5, kSqrt3 / 2.0};
const cd kA2{-0.5, -kSqrt3 / 2.0};
constexpr double kNumTol = 1e-8;
constexpr double kBaseP3 = 1e6;
constexpr double kBaseP1 = 1e6 / 3.0;
constexpr int8_t kNaS = -128;
constexpr int32_t kNaID = std::numeric_limits<int32_t>::min();
constexpr double kEps = std::numeric_limits<double>::epsilon();
const double kNan = std::numeric_limits<double>::quiet_NaN();

struct SparseMatrixError : std::runtime_error {
    SparseMatrixError()
        : std::runtime_error(
              "Sparse matrix error, possibly singular matrix!\n"
              "If you get this error from state estimation, it might mean the system is not fully observable, i.e. "
              "not enough measurements.\n"
              "It might also mean that you are running into a corner case where PGM cannot resolve yet.\n"
              "See https://github.com/PowerGridModel/power-grid-model/issues/864.") {}
};
struct IterationDiverge : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct OtherError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

std::string fmt_double(double x) {
    // shortest round-trip representation, like std::format("{}", x)
    char buf[64];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, x);
        if (std::strtod(buf, nullptr) == x) break;
    }
    return buf;
}

// ------------------------------------------------------------------------------------------------
// phase-generic complex values and tensors (three_phase_tensor.hpp)
// ------------------------------------------------------------------------------------------------
template <int P> struct Val {
    std::array<cd, P> v{};
};
template <int P> struct Ten {
    std::array<cd, P * P> m{};  // row-major
    cd& operator()(int r, int c) { return m[r * P + c]; }
    cd operator()(int r, int c) const { return m[r * P + c]; }
};
template <int P> Ten<P> ten_sm(cd s, cd mm) {  // Tensor(s, m): s on diagonal, m off diagonal
    Ten<P> t;
    for (int r = 0; r < P; ++r)
        for (int c = 0; c < P; ++c) t(r, c) = r == c ? s : (P == 1 ? s : mm);
    return t;
}
template <int P> Val<P> balanced(cd x) {  // ComplexValue<asym>{x}: a, b=x*a^2, c=x*a
    Val<P> r;
    r.v[0] = x;
    if (P == 3) {
        r.v[1] = x * kA2;
        r.v[2] = x * kA;
    }
    return r;
}
template <int P> Val<P> matvec(Ten<P> const& t, Val<P> const& x) {
    Val<P> r;
    for (int i = 0; i < P; ++i) {
        cd s = 0;
        for (int j = 0; j < P; ++j) s += t(i, j) * x.v[j];
        r.v[i] = s;
    }
    return r;
}
template <int P> Ten<P> operator+(Ten<P> a, Ten<P> const& b) {
    for (int i = 0; i < P * P; ++i) a.m[i] += b.m[i];
    return a;
}
template <int P> Ten<P> operator-(Ten<P> a) {
    for (auto& x : a.m) x = -x;
    return a;
}

struct BranchParamSym {
    cd v[4]{};  // ff, ft, tf, tt
};
template <int P> struct BranchParam {
    Ten<P> v[4];
};

// ------------------------------------------------------------------------------------------------
// component parameters (component/*.hpp)
// ------------------------------------------------------------------------------------------------
// branch.hpp:196-245 calc_param_y_sym
BranchParamSym calc_param_y_sym(cd y_series, cd y_shunt, cd tap_ratio, bool from_status, bool to_status) {
    double const tap = std::abs(tap_ratio);
    BranchParamSym p;
    if (!(from_status && to_status)) {
        if (from_status || to_status) {
            cd branch_shunt;
            if (std::abs(y_shunt) < kNumTol) {
                branch_shunt = 0.0;
            } else {
                branch_shunt = 0.5 * y_shunt + 1.0 / (1.0 / y_series + 2.0 / y_shunt);
            }
            p.v[0] = from_status ? (1.0 / tap / tap) * branch_shunt : 0.0;
            p.v[3] = to_status ? branch_shunt : 0.0;
        }
    } else {
        p.v[3] = y_series + 0.5 * y_shunt;
        p.v[0] = (1.0 / tap / tap) * p.v[3];
        p.v[1] = (-1.0 / std::conj(tap_ratio)) * y_series;
        p.v[2] = (-1.0 / tap_ratio) * y_series;
    }
    return p;
}
BranchParam<3> sym_to_asym(BranchParamSym const& p1, BranchParamSym const& p0) {
    BranchParam<3> r;
    for (int i = 0; i < 4; ++i) r.v[i] = ten_sm<3>((2.0 * p1.v[i] + p0.v[i]) / 3.0, (p0.v[i] - p1.v[i]) / 3.0);
    return r;
}

struct LineParam {
    cd y1s, y1sh, y0s, y0sh;
    double base_i, i_n, base_i_to;
    cd ratio{1.0, 0.0};  // generic branch: k e^(j theta)
    double shift = 0.0;
    double sn = 0.0;     // generic branch: rated power, NaN when not given
    bool generic = false;
};
// line.hpp:26-36
LineParam line_param(OrLineInput const& in, double f, double u) {
    LineParam p;
    p.base_i = kBaseP3 / u / kSqrt3;
    double const base_y = p.base_i / (u / kSqrt3);
    p.y1s = 1.0 / cd(in.r1, in.x1) / base_y;
    p.y1sh = 2.0 * kPi * f * in.c1 / base_y * cd(in.tan1, 1.0);
    p.y0s = 1.0 / cd(in.r0, in.x0) / base_y;
    p.y0sh = 2.0 * kPi * f * in.c0 / base_y * cd(in.tan0, 1.0);
    p.i_n = in.i_n;
    p.base_i_to = p.base_i;
    return p;
}
// link.hpp:19-36 and common.hpp:97-101: y_series = y_link = g_link (1 + 1j) with g_link = 1e6 S at 10 kV in
// per unit, no shunt, tap 1; base currents from the rated voltages of both nodes; loading 0 (i_n infinite)
LineParam link_param(double u1, double u2) {
    LineParam p;
    double const g_link = 1e6 / (kBaseP3 / 10e3 / 10e3);
    p.y1s = p.y0s = cd(g_link, g_link);
    p.y1sh = p.y0sh = cd(0.0, 0.0);
    p.base_i = kBaseP3 / u1 / kSqrt3;
    p.base_i_to = kBaseP3 / u2 / kSqrt3;
    p.i_n = std::numeric_limits<double>::infinity();
    return p;
}
// generic_branch.hpp:24-60: y1 in per unit of the to side, ratio k e^(j theta) with k = 1 and theta = 0 when
// not given (theta modulo 2 pi), loading max s / sn or 0 without sn
LineParam generic_param(OrGenericBranchInput const& in, double u1, double u2) {
    LineParam p;
    p.base_i = kBaseP3 / u1 / kSqrt3;
    p.base_i_to = kBaseP3 / u2 / kSqrt3;
    double const base_y = p.base_i_to / (u2 / kSqrt3);
    p.y1s = p.y0s = 1.0 / cd(in.r1, in.x1) / base_y;
    p.y1sh = p.y0sh = cd(in.g1, in.b1) / base_y;
    double const k = std::isnan(in.k) ? 1.0 : in.k;
    p.shift = std::isnan(in.theta) ? 0.0 : std::fmod(in.theta, 2.0 * kPi);
    p.ratio = k * std::exp(cd(0.0, p.shift));
    p.i_n = kNan;
    p.sn = in.sn;
    p.generic = true;
    return p;
}

struct TrafoStatic {
    double u1, u2, sn, tap_size, uk, pk, i0, p0, i0_zs, p0_zs, uk_min, uk_max, pk_min, pk_max;
    int8_t winding_from, winding_to, clock, tap_side, tap_min, tap_max, tap_nom, tap_direction;
    double base_i_from, base_i_to, nominal_ratio;
    cd z_ground_from, z_ground_to;
};
bool is_nan(double x) { return std::isnan(x); }
// transformer.hpp:32-90 constructor
TrafoStatic trafo_static(OrTransformerInput const& in, double u1r, double u2r, int8_t& tap_pos) {
    TrafoStatic t{};
    t.u1 = in.u1;
    t.u2 = in.u2;
    t.sn = in.sn;
    t.tap_size = in.tap_size;
    t.uk = in.uk;
    t.pk = in.pk;
    t.i0 = in.i0;
    t.p0 = in.p0;
    t.i0_zs = is_nan(in.i0_zero_sequence) ? t.i0 : in.i0_zero_sequence;
    t.p0_zs = is_nan(in.p0_zero_sequence) ? t.p0 + t.pk * (t.i0_zs * t.i0_zs - t.i0 * t.i0) : in.p0_zero_sequence;
    t.winding_from = in.winding_from;
    t.winding_to = in.winding_to;
    t.clock = in.clock;
    t.tap_side = in.tap_side;
    t.tap_min = in.tap_min;
    t.tap_max = in.tap_max;
    t.tap_nom = in.tap_nom == kNaS ? 0 : in.tap_nom;
    t.tap_direction = t.tap_max > t.tap_min ? 1 : -1;
    t.uk_min = is_nan(in.uk_min) ? t.uk : in.uk_min;
    t.uk_max = is_nan(in.uk_max) ? t.uk : in.uk_max;
    t.pk_min = is_nan(in.pk_min) ? t.pk : in.pk_min;
    t.pk_max = is_nan(in.pk_max) ? t.pk : in.pk_max;
    t.base_i_from = kBaseP3 / u1r / kSqrt3;
    t.base_i_to = kBaseP3 / u2r / kSqrt3;
    t.nominal_ratio = u1r / u2r;
    auto zpu = [](double r, double x, double u) {
        r = is_nan(r) ? 0 : r;
        x = is_nan(x) ? 0 : x;
        double const bz = u * u / kBaseP3;
        return cd(r / bz, x / bz);
    };
    t.z_ground_from = zpu(in.r_grounding_from, in.x_grounding_from, u1r);
    t.z_ground_to = zpu(in.r_grounding_to, in.x_grounding_to, u2r);
    if (in.tap_pos == kNaS)
        tap_pos = in.tap_nom == kNaS ? 0 : in.tap_nom;
    else
        tap_pos = in.tap_pos;
    // transformer_utils.hpp is_valid_clock
    bool const even = (t.clock % 2) == 0;
    bool const from_wye = t.winding_from == 0 || t.winding_from == 1;
    bool const to_wye = t.winding_to == 0 || t.winding_to == 1;
    if (even != (from_wye == to_wye))
        throw OtherError("Invalid clock for transformer " + std::to_string(in.id) + ", clock " +
                         std::to_string(int(t.clock)) + "\n");
    t.clock = int8_t(((t.clock % 12) + 12) % 12);
    return t;
}
int8_t tap_limit(TrafoStatic const& t, int8_t tap) {
    tap = std::min<int8_t>(tap, std::max(t.tap_max, t.tap_min));
    tap = std::max<int8_t>(tap, std::min(t.tap_max, t.tap_min));
    return tap;
}
// transformer_utils.hpp tap_adjust_impedance
double tap_adjust(double tap_pos, double tap_min, double tap_max, double tap_nom, double xk, double xk_min,
                  double xk_max) {
    if (tap_pos <= std::max(tap_nom, tap_max) && tap_pos >= std::min(tap_nom, tap_max)) {
        if (tap_max == tap_nom) return xk;
        return xk + (tap_pos - tap_nom) * ((xk_max - xk) / (tap_max - tap_nom));
    }
    if (tap_min == tap_nom) return xk;
    return xk + (tap_pos - tap_nom) * ((xk_min - xk) / (tap_min - tap_nom));
}
struct TrafoParams {
    cd y_series, y_shunt, y0_shunt;
    double k;
};
// transformer.hpp transformer_params
TrafoParams trafo_params(TrafoStatic const& t, int8_t tap_pos) {
    double const base_y_to = t.base_i_to * t.base_i_to / kBaseP1;
    double u1 = t.u1, u2 = t.u2;
    if (t.tap_side == 0)
        u1 += t.tap_direction * (tap_pos - t.tap_nom) * t.tap_size;
    else
        u2 += t.tap_direction * (tap_pos - t.tap_nom) * t.tap_size;
    double const k = (u1 / u2) / t.nominal_ratio;
    double const uk = tap_adjust(tap_pos, t.tap_min, t.tap_max, t.tap_nom, t.uk, t.uk_min, t.uk_max);
    double const pk = tap_adjust(tap_pos, t.tap_min, t.tap_max, t.tap_nom, t.pk, t.pk_min, t.pk_max);
    cd z_series;
    double const uk_sign = uk >= 0 ? 1.0 : -1.0;
    double const z_abs = std::abs(uk) * u2 * u2 / t.sn;
    z_series.real(pk * u2 * u2 / t.sn / t.sn);
    double const zi2 = z_abs * z_abs - z_series.real() * z_series.real();
    z_series.imag(uk_sign * (zi2 > 0.0 ? std::sqrt(zi2) : 0.0));
    cd const y_series = (1.0 / z_series) / base_y_to;
    auto shunt = [&](double i0, double p0) {
        cd y;
        double const yabs = i0 * t.sn / u2 / u2;
        y.real(p0 / u2 / u2);
        double const yi2 = yabs * yabs - y.real() * y.real();
        y.imag(yi2 > 0.0 ? -std::sqrt(yi2) : 0.0);
        return y / base_y_to;
    };
    return {y_series, shunt(t.i0, t.p0), shunt(t.i0_zs, t.p0_zs), k};
}
BranchParamSym trafo_param_sym(TrafoStatic const& t, int8_t tap_pos, bool fs, bool ts) {
    auto const tp = trafo_params(t, tap_pos);
    return calc_param_y_sym(tp.y_series, tp.y_shunt, tp.k * std::exp(cd(0, t.clock * kDeg30)), fs, ts);
}
Ten<3> sym_matrix() {
    Ten<3> m;
    cd const vals[9] = {1.0, 1.0, 1.0, 1.0, kA2, kA, 1.0, kA, kA2};
    for (int i = 0; i < 9; ++i) m.m[i] = vals[i];
    return m;
}
Ten<3> sym_matrix_inv() {
    Ten<3> m;
    cd const vals[9] = {1.0, 1.0, 1.0, 1.0, kA, kA2, 1.0, kA2, kA};
    for (int i = 0; i < 9; ++i) m.m[i] = vals[i] / 3.0;
    return m;
}
Ten<3> matmul(Ten<3> const& a, Ten<3> const& b) {
    Ten<3> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd s = 0;
            for (int k = 0; k < 3; ++k) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}
// transformer.hpp asym_calc_param
BranchParam<3> trafo_param_asym(TrafoStatic const& t, int8_t tap_pos, bool fs, bool ts) {
    auto const tp = trafo_params(t, tap_pos);
    auto const y_series = tp.y_series;
    double const k = tp.k;
    auto const p1 = calc_param_y_sym(y_series, tp.y_shunt, k * std::exp(cd(0, t.clock * kDeg30)), fs, ts);
    auto const p2 = calc_param_y_sym(y_series, tp.y_shunt, k * std::exp(cd(0, -t.clock * kDeg30)), fs, ts);
    BranchParamSym p0{};
    constexpr int8_t wye_n = 1, delta = 2, zigzag_n = 4;
    if (t.winding_from == wye_n && t.winding_to == wye_n) {
        double ps0 = 0.0;
        if (t.clock == 2 || t.clock == 6 || t.clock == 10) ps0 = 6.0 * kDeg30;
        cd const z0s = 1.0 / y_series + 3.0 * (t.z_ground_to + t.z_ground_from / k / k);
        p0 = calc_param_y_sym(1.0 / z0s, tp.y0_shunt, k * std::exp(cd(0, ps0)), fs, ts);
    } else if (t.winding_from == wye_n && fs) {
        cd y0 = tp.y0_shunt;
        if (t.winding_to == delta) y0 += y_series;
        if (y0 != cd(0.0, 0.0)) {
            cd const z0 = 1.0 / y0 + 3.0 * t.z_ground_from / k / k;
            y0 = 1.0 / z0;
            p0.v[0] = y0 / k / k;
        }
    } else if (t.winding_to == wye_n && ts) {
        cd y0 = tp.y0_shunt;
        if (t.winding_from == delta) y0 += y_series;
        if (y0 != cd(0.0, 0.0)) {
            cd const z0 = 1.0 / y0 + 3.0 * t.z_ground_to;
            y0 = 1.0 / z0;
            p0.v[3] = y0;
        }
    }
    if (t.winding_from == zigzag_n && fs) {
        cd const z0s = (1.0 / y_series) * 0.1 + 3.0 * t.z_ground_from / k / k;
        p0.v[0] = (1.0 / z0s) / k / k;
    }
    if (t.winding_to == zigzag_n && ts) {
        cd const z0s = (1.0 / y_series) * 0.1 + 3.0 * t.z_ground_to;
        p0.v[3] = 1.0 / z0s;
    }
    double const low_b = -1e-8 * t.sn / kBaseP3 / t.uk;
    auto avail = [](int8_t a, int8_t b) {
        if (a == wye_n) return b == wye_n || b == delta;
        if (a == zigzag_n) return true;
        return false;
    };
    if (!avail(t.winding_from, t.winding_to) && fs) p0.v[0] += cd(0.0, low_b);
    if (!avail(t.winding_to, t.winding_from) && ts) p0.v[3] += cd(0.0, low_b);
    Ten<3> const sm = sym_matrix(), smi = sym_matrix_inv();
    BranchParam<3> r;
    for (int i = 0; i < 4; ++i) {
        Ten<3> d;
        d(0, 0) = p0.v[i];
        d(1, 1) = p1.v[i];
        d(2, 2) = p2.v[i];
        r.v[i] = matmul(matmul(sm, d), smi);
    }
    return r;
}

// ------------------------------------------------------------------------------------------------
// minimum degree ordering (sparse_ordering.hpp:20-219)
// ------------------------------------------------------------------------------------------------
using AdjMap = std::map<Idx, std::vector<Idx>>;
struct DegreeLookup {
    std::map<Idx, Idx> v2d;
    std::map<Idx, std::set<Idx>> d2v;
    void remove_degree(Idx u, Idx d) {
        auto it = d2v.find(d);
        if (it != d2v.end()) {
            it->second.erase(u);
            if (it->second.empty()) d2v.erase(it);
        }
    }
    void set(Idx u, Idx d) {
        auto it = v2d.find(u);
        if (it != v2d.end()) {
            remove_degree(u, it->second);
            it->second = d;
        } else {
            v2d.emplace(u, d);
        }
        d2v[d].insert(u);
    }
    void erase(Idx u) {
        auto it = v2d.find(u);
        if (it == v2d.end()) return;
        Idx const d = it->second;
        v2d.erase(it);
        remove_degree(u, d);
    }
    Idx min_vertex() const { return *d2v.begin()->second.begin(); }
};
bool in_graph(Idx a, Idx b, AdjMap const& d) {
    auto it = d.find(a);
    return it != d.end() && std::find(it->second.begin(), it->second.end(), b) != it->second.end();
}
std::vector<Idx> remove_vertices_update_degrees(Idx u, AdjMap& d, DegreeLookup& dgd,
                                                std::vector<std::pair<Idx, Idx>>& fills) {
    // check_indistguishable
    std::vector<Idx> nbs = d.at(u);
    std::vector<Idx> lu = nbs;
    lu.push_back(u);
    std::sort(lu.begin(), lu.end());
    std::vector<Idx> rl;
    for (Idx v : nbs) {
        std::vector<Idx> lv = d.at(v);
        lv.push_back(v);
        std::sort(lv.begin(), lv.end());
        if (lu == lv) rl.push_back(v);
    }
    std::vector<Idx> alpha = rl;
    rl.push_back(u);
    for (Idx uu : rl) {
        if (uu != u) nbs.erase(std::remove(nbs.begin(), nbs.end(), uu), nbs.end());
        dgd.erase(uu);
        std::vector<Idx> el;
        std::vector<Idx> const adj_uu = d[uu];
        for (Idx e : adj_uu) {
            auto& a = d[e];
            a.erase(std::remove(a.begin(), a.end(), uu), a.end());
            if (a.empty()) el.push_back(e);
        }
        el.push_back(uu);
        for (Idx e : el) d.erase(e);
    }
    // make_clique(nbs): map keyed by vertex, each with all other members in nbs order
    std::map<Idx, std::vector<Idx>> dd;
    for (size_t i = 0; i < nbs.size(); ++i) {
        std::vector<Idx> sl;
        for (size_t j = 0; j < nbs.size(); ++j)
            if (j != i) sl.push_back(nbs[j]);
        dd[nbs[i]] = std::move(sl);
    }
    for (auto const& [k, adjacent] : dd) {
        for (Idx e : adjacent) {
            if (!in_graph(k, e, d)) {
                d[k].push_back(e);
                d[e].push_back(k);
                fills.emplace_back(k, e);
            }
        }
    }
    for (Idx e : nbs) {
        auto it = d.find(e);
        dgd.set(e, it == d.end() ? 0 : Idx(it->second.size()));
    }
    return alpha;
}
std::pair<std::vector<Idx>, std::vector<std::pair<Idx, Idx>>> minimum_degree_ordering(AdjMap d) {
    AdjMap sym = d;
    for (auto const& [k, adj] : d)
        for (Idx e : adj) sym[e].push_back(k);
    d = std::move(sym);
    for (auto& [k, adj] : d) {
        std::set<Idx> const s(adj.begin(), adj.end());
        adj.assign(s.begin(), s.end());
    }
    DegreeLookup dgd;
    for (auto const& [k, adj] : d) dgd.set(k, Idx(adj.size()));
    Idx const n = Idx(d.size());
    std::vector<Idx> alpha;
    std::vector<std::pair<Idx, Idx>> fills;
    for (Idx k = 0; k < n; ++k) {
        Idx const u = dgd.min_vertex();
        alpha.push_back(u);
        if (d.size() == 2) {
            Idx const from = d.begin()->first;
            Idx const to = d.begin()->second[0];
            alpha.push_back(alpha.back() == from ? to : from);
            return {alpha, fills};
        }
        auto res = remove_vertices_update_degrees(u, d, dgd, fills);
        alpha.insert(alpha.end(), res.begin(), res.end());
        if (d.empty()) return {alpha, fills};
    }
    return {alpha, fills};
}

// ------------------------------------------------------------------------------------------------
// topology (topology.hpp:196-368)
// ------------------------------------------------------------------------------------------------
struct CompBranch {
    Idx from, to;
    bool fs, ts;
    double phase_shift;
};
struct TopoResult {
    std::vector<Idx> node_group, node_pos;
    std::vector<double> phase;  // phase shift of every node against its source (topology.hpp dfs)
    std::vector<std::vector<Idx>> bus_nodes;  // per math model: node index per bus
    std::vector<Idx> slack;
    std::vector<bool> radial;
    std::vector<std::vector<std::array<Idx, 2>>> fill_in;
};
TopoResult build_topology(Idx n_node, std::vector<CompBranch> const& branches, std::vector<Idx> const& source_node,
                          std::vector<bool> const& source_on) {
    // CSR digraph, out edges grouped by source vertex in insertion order (stable histogram sort)
    std::vector<std::array<Idx, 2>> edges;
    std::vector<double> eps;
    for (auto const& b : branches) {
        if (b.fs && b.ts && b.from != b.to) {
            edges.push_back({b.from, b.to});
            eps.push_back(-b.phase_shift);
            edges.push_back({b.to, b.from});
            eps.push_back(b.phase_shift);
        }
    }
    std::vector<Idx> start(n_node + 1, 0);
    for (auto const& e : edges) ++start[e[0] + 1];
    for (Idx i = 0; i < n_node; ++i) start[i + 1] += start[i];
    std::vector<Idx> tgt(edges.size());
    std::vector<double> eph(edges.size());
    {
        std::vector<Idx> pos(start.begin(), start.end() - 1);
        for (size_t k = 0; k < edges.size(); ++k) {
            Idx const p = pos[edges[k][0]]++;
            tgt[p] = edges[k][1];
            eph[p] = eps[k];
        }
    }
    TopoResult res;
    res.node_group.assign(n_node, -1);
    res.node_pos.assign(n_node, -1);
    std::vector<double> phase(n_node, 0.0);
    std::vector<Idx> pred(n_node);
    std::iota(pred.begin(), pred.end(), Idx{0});
    constexpr Idx not_processed = -1, in_cycle = -2;
    std::vector<Idx> node_status(n_node, not_processed);
    std::vector<int8_t> color(n_node, 0);  // 0 white, 1 gray, 2 black
    Idx group = 0;
    for (size_t s = 0; s < source_node.size(); ++s) {
        if (!source_on[s]) continue;
        Idx const root = source_node[s];
        if (res.node_group[root] != -1) continue;
        std::vector<Idx> dfs_node;
        std::vector<std::pair<Idx, Idx>> back_edges;
        // iterative depth-first visit equivalent to the recursive one (boost::depth_first_visit)
        struct Frame {
            Idx u, ei, ee;
        };
        std::vector<Frame> stack;
        color[root] = 1;
        res.node_group[root] = group;
        dfs_node.push_back(root);
        stack.push_back({root, start[root], start[root + 1]});
        while (!stack.empty()) {
            Frame& f = stack.back();
            if (f.ei == f.ee) {
                color[f.u] = 2;
                stack.pop_back();
                continue;
            }
            Idx const e = f.ei++;
            Idx const v = tgt[e];
            if (color[v] == 0) {
                phase[v] = phase[f.u] + eph[e];
                pred[v] = f.u;
                color[v] = 1;
                res.node_group[v] = group;
                dfs_node.push_back(v);
                Idx const u = f.u;
                (void)u;
                stack.push_back({v, start[v], start[v + 1]});
            } else if (color[v] == 1) {
                Idx const u = f.u;
                if (pred[u] != v) back_edges.emplace_back(u, v);
            }
        }
        std::vector<std::array<Idx, 2>> fills;
        bool radial;
        if (back_edges.empty()) {
            std::reverse(dfs_node.begin(), dfs_node.end());
            radial = true;
        } else {
            radial = false;
            // reorder_node
            std::vector<Idx> const copy = dfs_node;
            dfs_node.clear();
            for (auto const& be : back_edges) {
                Idx x = be.first;
                while (node_status[x] != in_cycle) {
                    node_status[x] = in_cycle;
                    x = pred[x];
                }
            }
            for (auto it = copy.rbegin(); it != copy.rend(); ++it)
                if (node_status[*it] == not_processed) dfs_node.push_back(*it);
            std::vector<Idx> cyclic;
            for (Idx x : copy)
                if (node_status[x] == in_cycle) cyclic.push_back(x);
            if (cyclic.size() < 4) {
                for (auto it = cyclic.rbegin(); it != cyclic.rend(); ++it) dfs_node.push_back(*it);
            } else {
                AdjMap nn;
                for (Idx x : cyclic)
                    if (pred[x] != x) nn[x] = {pred[x]};
                for (auto const& [from, to] : back_edges)
                    if (!in_graph(from, to, nn)) nn[from].push_back(to);
                auto [reordered, fl] = minimum_degree_ordering(std::move(nn));
                Idx const n_non_cyclic = Idx(dfs_node.size());
                std::map<Idx, Idx> perm;
                for (Idx i = 0; i < Idx(reordered.size()); ++i) perm[reordered[i]] = n_non_cyclic + i;
                dfs_node.insert(dfs_node.end(), reordered.begin(), reordered.end());
                for (auto [a, b] : fl) fills.push_back({perm[a], perm[b]});
            }
        }
        for (Idx i = 0; i < Idx(dfs_node.size()); ++i) res.node_pos[dfs_node[i]] = i;
        res.bus_nodes.push_back(dfs_node);
        res.slack.push_back(res.node_pos[root]);
        res.radial.push_back(radial);
        res.fill_in.push_back(fills);
        ++group;
    }
    res.phase = phase;
    return res;
}

// ------------------------------------------------------------------------------------------------
// Y-bus structure (y_bus.hpp:36-293)
// ------------------------------------------------------------------------------------------------
struct YElem {
    Idx row, col;
    int type;  // 0 ff, 1 ft, 2 tf, 3 tt, 4 shunt, 5 fill ft, 6 fill tf
    Idx idx;
};
struct YStruct {
    std::vector<Idx> row_indptr, col_indices, bus_entry, row_indptr_lu, col_indices_lu, diag_lu, map_lu_y_bus,
        lu_transpose_entry, y_bus_entry_indptr;
    std::vector<YElem> y_bus_element;
};
YStruct build_ystruct(Idx n_bus, std::vector<std::array<Idx, 2>> const& branch_bus,
                      std::vector<std::vector<Idx>> const& shunts_per_bus,
                      std::vector<std::array<Idx, 2>> const& fill_in) {
    Idx const n_branch = Idx(branch_bus.size());
    Idx const n_fill = Idx(fill_in.size());
    std::vector<YElem> vec;
    auto append = [&](Idx a, Idx b, int t, Idx i) {
        if (a == -1 || b == -1) return;
        vec.push_back({a, b, t, i});
    };
    for (Idx br = 0; br < n_branch; ++br)
        for (int i = 0; i < 4; ++i) append(branch_bus[br][i / 2], branch_bus[br][i % 2], i, br);
    Idx n_shunt = 0;
    for (Idx bus = 0; bus < n_bus; ++bus)
        for (Idx sh : shunts_per_bus[bus]) {
            append(bus, bus, 4, sh);
            ++n_shunt;
        }
    for (Idx f = 0; f < n_fill; ++f) {
        append(fill_in[f][0], fill_in[f][1], 5, f);
        append(fill_in[f][1], fill_in[f][0], 6, f);
    }
    // stable counting sort by column, then by row
    auto csort = [&](bool by_row) {
        std::vector<Idx> counter(n_bus, 0);
        for (auto const& e : vec) ++counter[by_row ? e.row : e.col];
        for (Idx i = 1; i < n_bus; ++i) counter[i] += counter[i - 1];
        std::vector<YElem> tmp(vec.size());
        for (auto it = vec.rbegin(); it != vec.rend(); ++it) tmp[--counter[by_row ? it->row : it->col]] = *it;
        vec.swap(tmp);
    };
    csort(false);
    csort(true);
    YStruct ys;
    ys.row_indptr.assign(n_bus + 1, 0);
    ys.row_indptr_lu.assign(n_bus + 1, 0);
    ys.bus_entry.assign(n_bus, 0);
    ys.diag_lu.assign(n_bus, 0);
    ys.y_bus_entry_indptr.push_back(0);
    std::vector<std::array<Idx, 2>> off_diag(n_branch + n_fill, {0, 0});
    for (auto const& m : vec)
        if (m.type < 5) ys.y_bus_element.push_back(m);
    Idx nnz = 0, row_start = 0, nnz_lu = 0, row_start_lu = 0, fill_counter = 0;
    for (size_t i = 0; i < vec.size();) {
        Idx const row = vec[i].row, col = vec[i].col;
        ys.col_indices_lu.push_back(col);
        if (row > row_start_lu) ys.row_indptr_lu[++row_start_lu] = nnz_lu;
        if (vec[i].type < 5) {
            ys.col_indices.push_back(col);
            ys.map_lu_y_bus.push_back(nnz);
            if (row > row_start) ys.row_indptr[++row_start] = nnz;
            if (row == col) {
                ys.bus_entry[row] = nnz;
                ys.diag_lu[row] = nnz_lu;
            } else {
                off_diag[vec[i].idx][vec[i].type - 1] = nnz_lu;
            }
            while (i < vec.size() && vec[i].row == row && vec[i].col == col) ++i;
            ys.y_bus_entry_indptr.push_back(Idx(i) - fill_counter);
            ++nnz;
            ++nnz_lu;
        } else {
            off_diag[vec[i].idx + n_branch][vec[i].type - 5] = nnz_lu;
            ys.map_lu_y_bus.push_back(-1);
            ++fill_counter;
            ++nnz_lu;
            ++i;
        }
    }
    ys.row_indptr[++row_start] = nnz;
    ys.row_indptr_lu[++row_start_lu] = nnz_lu;
    if (n_branch == 0 && n_shunt == 0) {
        ys.row_indptr = {0, 1};
        ys.col_indices = {0};
        ys.bus_entry = {0};
        ys.y_bus_entry_indptr = {0, 0};
        ys.row_indptr_lu = {0, 1};
        ys.col_indices_lu = {0};
        ys.diag_lu = {0};
        ys.map_lu_y_bus = {0};
        nnz_lu = 1;
    }
    ys.lu_transpose_entry.resize(nnz_lu);
    std::iota(ys.lu_transpose_entry.begin(), ys.lu_transpose_entry.end(), Idx{0});
    if (!(n_branch == 0 && n_shunt == 0)) {
        for (auto const& od : off_diag) {
            ys.lu_transpose_entry[od[0]] = od[1];
            ys.lu_transpose_entry[od[1]] = od[0];
        }
    }
    return ys;
}

// ------------------------------------------------------------------------------------------------
// dense full-pivot block LU and block sparse LU (sparse_lu_solver.hpp:85-200, 346-495, 769-827)
// ------------------------------------------------------------------------------------------------
template <int NB> struct BlockPerm {
    std::array<int8_t, NB> r{}, c{};  // row / column transpositions per pivot step
};
template <int NB> using Blk = std::array<double, NB * NB>;  // column-major
template <int NB> using Vec = std::array<double, NB>;
template <int NB> inline double& at(Blk<NB>& b, int r, int c) { return b[c * NB + r]; }
template <int NB> inline double at(Blk<NB> const& b, int r, int c) { return b[c * NB + r]; }

// pivot perturbation (sparse_lu_solver.hpp:39-48, 85-165): with `use_pert`, a chosen pivot whose magnitude is
// below `pert_thr` takes that magnitude with its own sign (+ for a zero pivot) and `has_pert` is set; `has_pert`
// belongs to the whole factorisation, and once set the pivot test checks only for non-normal pivots
template <int NB>
void factorize_block(Blk<NB>& a, BlockPerm<NB>& perm, double pert_thr = 0.0, bool use_pert = false,
                     bool* has_pert = nullptr) {
    double max_pivot = 0.0;
    bool dummy = false;
    bool& pert = has_pert ? *has_pert : dummy;
    for (int t = 0; t < NB; ++t) {
        double best = -1.0;
        int rb = t, cb = t;
        for (int c = t; c < NB; ++c)
            for (int r = t; r < NB; ++r) {
                double const s = at<NB>(a, r, c) * at<NB>(a, r, c);
                if (s > best) {
                    best = s;
                    rb = r;
                    cb = c;
                }
            }
        if (best == 0.0 && !use_pert) {
            for (int k = t; k < NB; ++k) perm.r[k] = perm.c[k] = int8_t(k);
            break;
        }
        double abs_pivot = std::sqrt(best);
        if (abs_pivot < pert_thr) {
            double& v = at<NB>(a, rb, cb);
            v = (abs_pivot == 0.0 ? 1.0 : v / abs_pivot) * pert_thr;
            abs_pivot = pert_thr;
            pert = true;
        }
        max_pivot = std::max(max_pivot, abs_pivot);
        perm.r[t] = int8_t(rb);
        perm.c[t] = int8_t(cb);
        if (rb != t)
            for (int c = 0; c < NB; ++c) std::swap(at<NB>(a, t, c), at<NB>(a, rb, c));
        if (cb != t)
            for (int r = 0; r < NB; ++r) std::swap(at<NB>(a, r, t), at<NB>(a, r, cb));
        if (t < NB - 1) {
            double const piv = at<NB>(a, t, t);
            for (int r = t + 1; r < NB; ++r) at<NB>(a, r, t) /= piv;
            for (int c = t + 1; c < NB; ++c)
                for (int r = t + 1; r < NB; ++r) at<NB>(a, r, c) -= at<NB>(a, r, t) * at<NB>(a, t, c);
        }
    }
    double const thr = pert ? 0.0 : kEps * max_pivot;
    for (int t = 0; t < NB; ++t) {
        double const d = at<NB>(a, t, t);
        if (std::abs(d) < thr || !std::isnormal(d)) throw SparseMatrixError{};
    }
}
// P * rows of a block (row transpositions applied in pivot order)
template <int NB> void perm_rows(Blk<NB>& m, BlockPerm<NB> const& p) {
    for (int t = 0; t < NB; ++t)
        if (p.r[t] != t)
            for (int c = 0; c < NB; ++c) std::swap(at<NB>(m, t, c), at<NB>(m, p.r[t], c));
}
template <int NB> void perm_cols(Blk<NB>& m, BlockPerm<NB> const& p) {  // m * Q
    for (int t = 0; t < NB; ++t)
        if (p.c[t] != t)
            for (int r = 0; r < NB; ++r) std::swap(at<NB>(m, r, t), at<NB>(m, r, p.c[t]));
}
template <int NB> void perm_vec_p(Vec<NB>& v, BlockPerm<NB> const& p) {
    for (int t = 0; t < NB; ++t)
        if (p.r[t] != t) std::swap(v[t], v[p.r[t]]);
}
template <int NB> void perm_vec_q(Vec<NB>& v, BlockPerm<NB> const& p) {  // Q * v
    for (int t = NB - 1; t >= 0; --t)
        if (p.c[t] != t) std::swap(v[t], v[p.c[t]]);
}
// left, unit lower: per column of m, forward substitution
template <int NB> void left_lower_solve(Blk<NB> const& lu, Blk<NB>& m) {
    for (int r = 0; r < NB; ++r)
        for (int k = 0; k < r; ++k) {
            double const l = at<NB>(lu, r, k);
            for (int c = 0; c < NB; ++c) at<NB>(m, r, c) -= l * at<NB>(m, k, c);
        }
}
// right, upper: m := m * U^-1
template <int NB> void right_upper_solve(Blk<NB> const& lu, Blk<NB>& m) {
    for (int c = 0; c < NB; ++c) {
        for (int k = 0; k < c; ++k) {
            double const u = at<NB>(lu, k, c);
            for (int r = 0; r < NB; ++r) at<NB>(m, r, c) -= u * at<NB>(m, r, k);
        }
        double const d = at<NB>(lu, c, c);
        for (int r = 0; r < NB; ++r) at<NB>(m, r, c) /= d;
    }
}
template <int NB> void vec_lower(Blk<NB> const& lu, Vec<NB>& x) {
    for (int r = 0; r < NB; ++r)
        for (int k = 0; k < r; ++k) x[r] -= at<NB>(lu, r, k) * x[k];
}
template <int NB> void vec_upper(Blk<NB> const& lu, Vec<NB>& x) {
    for (int s = 0; s < NB; ++s) {
        int const r = NB - 1 - s;
        for (int ps = 0; ps < s; ++ps) {
            int const k = NB - 1 - ps;
            x[r] -= at<NB>(lu, r, k) * x[k];
        }
        x[r] /= at<NB>(lu, r, r);
    }
}
template <int NB> void block_mul_sub(Blk<NB>& a, Blk<NB> const& l, Blk<NB> const& u) {  // a -= l*u
    Blk<NB> prod{};
    for (int c = 0; c < NB; ++c)
        for (int r = 0; r < NB; ++r) {
            double s = 0.0;
            for (int k = 0; k < NB; ++k) s += at<NB>(l, r, k) * at<NB>(u, k, c);
            at<NB>(prod, r, c) = s;
        }
    for (int i = 0; i < NB * NB; ++i) a[i] -= prod[i];
}
template <int NB> void matvec_sub(Vec<NB>& x, Blk<NB> const& m, Vec<NB> const& y) {
    Vec<NB> prod{};
    for (int r = 0; r < NB; ++r) {
        double s = 0.0;
        for (int k = 0; k < NB; ++k) s += at<NB>(m, r, k) * y[k];
        prod[r] = s;
    }
    for (int r = 0; r < NB; ++r) x[r] -= prod[r];
}

template <int NB> struct SparseLU {
    std::vector<Idx> const& indptr;
    std::vector<Idx> const& cols;
    std::vector<Idx> const& diag;
    Idx n;
    std::vector<BlockPerm<NB>> perm;
    // pivot perturbation of the whole factorisation (sparse_lu_solver.hpp:346-360)
    double pert_thr = 0.0;
    bool use_pert = false;
    bool has_pert = false;
    SparseLU(std::vector<Idx> const& ip, std::vector<Idx> const& ci, std::vector<Idx> const& dg)
        : indptr(ip), cols(ci), diag(dg), n(Idx(ip.size()) - 1), perm(n) {}

    Idx find_entry(Idx row, Idx col, Idx b, Idx e) const {
        for (Idx k = b; k < e; ++k)
            if (cols[k] == col) return k;
        throw OtherError("sparse pattern entry not found");
    }
    void factorize(std::vector<Blk<NB>>& lu) {
        std::vector<Idx> colpos(indptr.begin(), indptr.end() - 1);
        for (Idx p = 0; p < n; ++p) {
            Idx const pidx = diag[p];
            factorize_block<NB>(lu[pidx], perm[p], pert_thr, use_pert, &has_pert);
            auto const& bp = perm[p];
            Blk<NB> const piv = lu[pidx];
            for (Idx l_idx = indptr[p]; l_idx < pidx; ++l_idx) {
                perm_rows<NB>(lu[l_idx], bp);
                Idx const urow = cols[l_idx];
                Idx const u_idx = colpos[urow];
                perm_cols<NB>(lu[u_idx], bp);
                ++colpos[urow];
            }
            for (Idx u_idx = pidx + 1; u_idx < indptr[p + 1]; ++u_idx) {
                perm_rows<NB>(lu[u_idx], bp);
                left_lower_solve<NB>(piv, lu[u_idx]);
            }
            for (Idx lr = pidx + 1; lr < indptr[p + 1]; ++lr) {
                Idx const l_row = cols[lr];
                Idx const l_idx = colpos[l_row];
                perm_cols<NB>(lu[l_idx], bp);
                right_upper_solve<NB>(piv, lu[l_idx]);
                Idx a_idx = l_idx;
                for (Idx u_idx = pidx + 1; u_idx < indptr[p + 1]; ++u_idx) {
                    a_idx = find_entry(l_row, cols[u_idx], a_idx + 1, indptr[l_row + 1]);
                    block_mul_sub<NB>(lu[a_idx], lu[l_idx], lu[u_idx]);
                }
                ++colpos[l_row];
            }
            ++colpos[p];
        }
    }
    void solve(std::vector<Blk<NB>> const& lu, std::vector<Vec<NB>> const& rhs, std::vector<Vec<NB>>& x) const {
        for (Idx r = 0; r < n; ++r) {
            x[r] = rhs[r];
            perm_vec_p<NB>(x[r], perm[r]);
            for (Idx k = indptr[r]; k < diag[r]; ++k) matvec_sub<NB>(x[r], lu[k], x[cols[k]]);
            vec_lower<NB>(lu[diag[r]], x[r]);
        }
        for (Idx r = n - 1; r >= 0; --r) {
            for (Idx k = indptr[r + 1] - 1; k > diag[r]; --k) matvec_sub<NB>(x[r], lu[k], x[cols[k]]);
            vec_upper<NB>(lu[diag[r]], x[r]);
        }
        for (Idx r = 0; r < n; ++r) perm_vec_q<NB>(x[r], perm[r]);
    }
};

// ------------------------------------------------------------------------------------------------
// math model and Newton-Raphson (newton_raphson_pf_solver.hpp, iterative_pf_solver.hpp)
// ------------------------------------------------------------------------------------------------
template <int P> struct MathModel {
    Idx n_bus = 0;
    std::vector<double> phase_shift;  // per bus
    std::vector<std::array<Idx, 2>> branch_bus;
    std::vector<BranchParam<P>> branch_param;
    std::vector<Idx> shunt_bus;
    std::vector<Ten<P>> shunt_param;
    std::vector<Idx> source_bus;
    std::vector<cd> source_y1, source_y0;
    std::vector<Idx> load_bus;
    std::vector<int8_t> load_type;
    std::vector<std::array<Idx, 2>> fill_in;
};
template <int P> struct MathInput {
    std::vector<Val<P>> s_inj;
    std::vector<cd> u_source;
};
template <int P> struct MathOutput {
    std::vector<Val<P>> u, bus_injection;
    std::vector<Val<P>> br_sf, br_st, br_if, br_it;
    std::vector<Val<P>> sh_s, sh_i, ld_s, ld_i, src_s, src_i;
    Idx iterations = 0;
};
template <int P> Ten<P> y_ref(cd y1, cd y0) {
    if (P == 1) return ten_sm<P>(y1, y1);
    return ten_sm<P>((2.0 * y1 + y0) / 3.0, (y0 - y1) / 3.0);
}

template <int P> struct NRSolver {
    static constexpr int NB = 2 * P;
    MathModel<P> const& mm;
    YStruct ys;
    std::vector<Ten<P>> ydata;  // per Y-bus entry
    std::vector<std::vector<Idx>> loads_per_bus, sources_per_bus;

    explicit NRSolver(MathModel<P> const& m) : mm(m) {
        std::vector<std::vector<Idx>> shunts_per_bus(m.n_bus);
        for (Idx s = 0; s < Idx(m.shunt_bus.size()); ++s) shunts_per_bus[m.shunt_bus[s]].push_back(s);
        ys = build_ystruct(m.n_bus, m.branch_bus, shunts_per_bus, m.fill_in);
        Idx const nnz = ys.row_indptr.back();
        ydata.assign(nnz, Ten<P>{});
        for (Idx e = 0; e < nnz; ++e) {
            Ten<P> acc{};
            for (Idx k = ys.y_bus_entry_indptr[e]; k < ys.y_bus_entry_indptr[e + 1]; ++k) {
                auto const& el = ys.y_bus_element[k];
                if (el.type == 4)
                    acc = acc + m.shunt_param[el.idx];
                else
                    acc = acc + m.branch_param[el.idx].v[el.type];
            }
            ydata[e] = acc;
        }
        loads_per_bus.resize(m.n_bus);
        sources_per_bus.resize(m.n_bus);
        for (Idx l = 0; l < Idx(m.load_bus.size()); ++l) loads_per_bus[m.load_bus[l]].push_back(l);
        for (Idx s = 0; s < Idx(m.source_bus.size()); ++s) sources_per_bus[m.source_bus[s]].push_back(s);
    }

    // calculate_hnml: pf = (ui outer conj(uj)) .* conj(yij); H = L = Im pf, N = -M = Re pf
    static Blk<NB> hnml(Ten<P> const& y, Val<P> const& ui, Val<P> const& uj) {
        Blk<NB> b{};
        for (int r = 0; r < P; ++r)
            for (int c = 0; c < P; ++c) {
                cd const pf = ui.v[r] * std::conj(uj.v[c]) * std::conj(y(r, c));
                at<NB>(b, r, c) = pf.imag();
                at<NB>(b, r, c + P) = pf.real();
                at<NB>(b, r + P, c) = -pf.real();
                at<NB>(b, r + P, c + P) = pf.imag();
            }
        return b;
    }

    // method: 0 Newton-Raphson, 1 linear, 2 iterative current, 3 linear current (math_solver.hpp:43-64,
    // 142-156: a model whose load_gens are all const_y always runs linear). Linear: the initial-guess
    // solve is the answer (linear_pf_solver.hpp), and the results treat every load as const_y
    MathOutput<P> run(MathInput<P> const& in, double err_tol, Idx max_iter, int method = 0) {
        bool const all_const_y =
            std::all_of(mm.load_type.begin(), mm.load_type.end(), [](int8_t t) { return t == 1; });
        bool const linear = method == 1 || all_const_y;
        if (!linear && method == 3) return run_iterative_current(in, std::numeric_limits<double>::infinity(), 1);
        if (!linear && method == 2) return run_iterative_current(in, err_tol, max_iter);
        Idx const n = mm.n_bus;
        Idx const nnz_lu = ys.row_indptr_lu.back();
        std::vector<Blk<NB>> jac(nnz_lu);
        std::vector<Vec<NB>> rhs(n), dx(n);
        SparseLU<NB> lu(ys.row_indptr_lu, ys.col_indices_lu, ys.diag_lu);
        MathOutput<P> out;
        out.u.assign(n, Val<P>{});
        std::vector<std::array<double, P>> theta(n), vmag(n);

        // initial guess: linear solve [[G,-B],[B,G]] [Ur,Ui] = [Ir,Ii]  (:255-303, :706-762, :856-870)
        for (Idx k = 0; k < nnz_lu; ++k) {
            jac[k] = Blk<NB>{};
            Idx const ky = ys.map_lu_y_bus[k];
            if (ky == -1) continue;
            auto const& y = ydata[ky];
            for (int r = 0; r < P; ++r)
                for (int c = 0; c < P; ++c) {
                    at<NB>(jac[k], r, c) = y(r, c).real();
                    at<NB>(jac[k], r, c + P) = -y(r, c).imag();
                    at<NB>(jac[k], r + P, c) = y(r, c).imag();
                    at<NB>(jac[k], r + P, c + P) = y(r, c).real();
                }
        }
        for (Idx b = 0; b < n; ++b) {
            rhs[b] = Vec<NB>{};
            auto& blk = jac[ys.diag_lu[b]];
            for (Idx l : loads_per_bus[b]) {
                for (int ph = 0; ph < P; ++ph) {
                    cd const yl = -std::conj(in.s_inj[l].v[ph]);
                    at<NB>(blk, ph, ph + P) += -yl.imag();
                    at<NB>(blk, ph, ph) += yl.real();
                    at<NB>(blk, ph + P, ph + P) += yl.real();
                    at<NB>(blk, ph + P, ph) += yl.imag();
                }
            }
            for (Idx s : sources_per_bus[b]) {
                Ten<P> const ysrc = y_ref<P>(mm.source_y1[s], mm.source_y0[s]);
                Val<P> const us = balanced<P>(in.u_source[s]);
                for (int r = 0; r < P; ++r)
                    for (int c = 0; c < P; ++c) {
                        at<NB>(blk, r, c + P) -= ysrc(r, c).imag();
                        at<NB>(blk, r, c) += ysrc(r, c).real();
                        at<NB>(blk, r + P, c + P) += ysrc(r, c).real();
                        at<NB>(blk, r + P, c) += ysrc(r, c).imag();
                    }
                Val<P> const irhs = matvec<P>(ysrc, us);
                for (int r = 0; r < P; ++r) {
                    rhs[b][r] += irhs.v[r].real();
                    rhs[b][r + P] += irhs.v[r].imag();
                }
            }
        }
        lu.factorize(jac);
        lu.solve(jac, rhs, dx);
        for (Idx b = 0; b < n; ++b)
            for (int ph = 0; ph < P; ++ph) {
                out.u[b].v[ph] = cd(dx[b][ph], dx[b][ph + P]);
                vmag[b][ph] = std::abs(out.u[b].v[ph]);
                theta[b][ph] = std::arg(out.u[b].v[ph]);
            }
        if (linear) {
            out.iterations = 1;
            calculate_result(in, out, true);
            return out;
        }

        // iterations (iterative_pf_solver.hpp:55-74)
        double max_dev = std::numeric_limits<double>::infinity();
        Idx num_iter = 0;
        while (max_dev > err_tol || num_iter == 0) {
            if (num_iter++ == max_iter)
                throw IterationDiverge("Iteration failed to converge after " + std::to_string(max_iter) +
                                       " iterations! Max deviation: " + fmt_double(max_dev) +
                                       ", error tolerance: " + fmt_double(err_tol) + ".\n");
            // prepare_matrix_and_rhs_from_network_perspective (:473-521)
            for (Idx row = 0; row < n; ++row) {
                rhs[row] = Vec<NB>{};
                for (Idx k = ys.row_indptr_lu[row]; k < ys.row_indptr_lu[row + 1]; ++k) {
                    Idx const ky = ys.map_lu_y_bus[k];
                    if (ky == -1) {
                        jac[k] = Blk<NB>{};
                        continue;
                    }
                    Idx const j = ys.col_indices_lu[k];
                    jac[k] = hnml(ydata[ky], out.u[row], out.u[j]);
                    for (int r = 0; r < P; ++r) {
                        double sn = 0.0, sh = 0.0;
                        for (int c = 0; c < P; ++c) {
                            sn += at<NB>(jac[k], r, c + P);
                            sh += at<NB>(jac[k], r, c);
                        }
                        rhs[row][r] -= sn;
                        rhs[row][r + P] -= sh;
                    }
                }
                auto& d = jac[ys.diag_lu[row]];
                for (int r = 0; r < P; ++r) {
                    at<NB>(d, r, r) += rhs[row][r + P];
                    at<NB>(d, r, r + P) -= rhs[row][r];
                    at<NB>(d, r + P, r) -= rhs[row][r];
                    at<NB>(d, r + P, r + P) -= rhs[row][r + P];
                }
            }
            // add_loads / add_sources (:764-852)
            for (Idx b = 0; b < n; ++b) {
                auto& d = jac[ys.diag_lu[b]];
                for (Idx l : loads_per_bus[b]) {
                    for (int ph = 0; ph < P; ++ph) {
                        cd const s = in.s_inj[l].v[ph];
                        double const v = vmag[b][ph];
                        switch (mm.load_type[l]) {
                        case 0:
                            rhs[b][ph] += s.real();
                            rhs[b][ph + P] += s.imag();
                            break;
                        case 1:
                            rhs[b][ph] += s.real() * v * v;
                            rhs[b][ph + P] += s.imag() * v * v;
                            at<NB>(d, ph, ph + P) += -s.real() * 2.0 * v * v;
                            at<NB>(d, ph + P, ph + P) += -s.imag() * 2.0 * v * v;
                            break;
                        case 2:
                            rhs[b][ph] += s.real() * v;
                            rhs[b][ph + P] += s.imag() * v;
                            at<NB>(d, ph, ph + P) += -s.real() * v;
                            at<NB>(d, ph + P, ph + P) += -s.imag() * v;
                            break;
                        default:
                            throw OtherError("unknown load type");
                        }
                    }
                }
                for (Idx s : sources_per_bus[b]) {
                    Ten<P> const yr = y_ref<P>(mm.source_y1[s], mm.source_y0[s]);
                    Val<P> const uref = balanced<P>(in.u_source[s]);
                    Blk<NB> mmb = hnml(yr, out.u[b], out.u[b]);
                    Blk<NB> const ms = hnml(-yr, out.u[b], uref);
                    std::array<double, P> pc{}, qc{};
                    for (int r = 0; r < P; ++r)
                        for (int c = 0; c < P; ++c) {
                            pc[r] += at<NB>(mmb, r, c + P) + at<NB>(ms, r, c + P);
                            qc[r] += at<NB>(mmb, r, c) + at<NB>(ms, r, c);
                        }
                    for (int r = 0; r < P; ++r) {
                        at<NB>(mmb, r, r) += -qc[r];
                        at<NB>(mmb, r, r + P) += pc[r];
                        at<NB>(mmb, r + P, r) += pc[r];
                        at<NB>(mmb, r + P, r + P) += qc[r];
                        rhs[b][r] -= pc[r];
                        rhs[b][r + P] -= qc[r];
                    }
                    for (int i = 0; i < NB * NB; ++i) d[i] += mmb[i];
                }
            }
            lu.factorize(jac);
            lu.solve(jac, rhs, dx);
            // iterate_unknown (:325-349)
            max_dev = 0.0;
            for (Idx b = 0; b < n; ++b) {
                double dev = 0.0;
                for (int ph = 0; ph < P; ++ph) {
                    theta[b][ph] += dx[b][ph];
                    vmag[b][ph] += vmag[b][ph] * dx[b][ph + P];
                    cd const un = vmag[b][ph] * std::exp(cd(0.0, theta[b][ph]));
                    double const d = std::abs(un - out.u[b].v[ph]);
                    dev = ph == 0 ? d : std::max(dev, d);
                    out.u[b].v[ph] = un;
                }
                max_dev = std::max(dev, max_dev);
            }
        }
        out.iterations = num_iter;
        calculate_result(in, out);
        return out;
    }

    // iterative_current_pf_solver.hpp: (Y + y_ref) factorised once; each iteration solves for the
    // injected currents of the loads (by type) and sources at the last voltages; flat start at the
    // averaged source voltage with the buses' phase shifts. The complex system is solved in its real
    // form [[G, -B], [B, G]], which differs from the reference's complex LU by rounding only.
    MathOutput<P> run_iterative_current(MathInput<P> const& in, double err_tol, Idx max_iter) {
        Idx const n = mm.n_bus;
        Idx const nnz_lu = ys.row_indptr_lu.back();
        std::vector<Blk<NB>> mat(nnz_lu);
        std::vector<Vec<NB>> rhs(n), x(n);
        SparseLU<NB> lu(ys.row_indptr_lu, ys.col_indices_lu, ys.diag_lu);
        for (Idx k = 0; k < nnz_lu; ++k) {
            mat[k] = Blk<NB>{};
            Idx const ky = ys.map_lu_y_bus[k];
            if (ky == -1) continue;
            auto const& y = ydata[ky];
            for (int r = 0; r < P; ++r)
                for (int c = 0; c < P; ++c) {
                    at<NB>(mat[k], r, c) = y(r, c).real();
                    at<NB>(mat[k], r, c + P) = -y(r, c).imag();
                    at<NB>(mat[k], r + P, c) = y(r, c).imag();
                    at<NB>(mat[k], r + P, c + P) = y(r, c).real();
                }
        }
        for (Idx b = 0; b < n; ++b)
            for (Idx s : sources_per_bus[b]) {
                Ten<P> const ysrc = y_ref<P>(mm.source_y1[s], mm.source_y0[s]);
                auto& blk = mat[ys.diag_lu[b]];
                for (int r = 0; r < P; ++r)
                    for (int c = 0; c < P; ++c) {
                        at<NB>(blk, r, c) += ysrc(r, c).real();
                        at<NB>(blk, r, c + P) -= ysrc(r, c).imag();
                        at<NB>(blk, r + P, c) += ysrc(r, c).imag();
                        at<NB>(blk, r + P, c + P) += ysrc(r, c).real();
                    }
            }
        lu.factorize(mat);
        MathOutput<P> out;
        out.u.assign(n, Val<P>{});
        // make_flat_start
        cd sum_u_ref = 0.0;
        for (Idx s = 0; s < Idx(mm.source_bus.size()); ++s)
            sum_u_ref += in.u_source[s] * std::exp(cd(0.0, -mm.phase_shift[mm.source_bus[s]]));
        cd const u_ref = sum_u_ref / double(mm.source_bus.size());
        for (Idx b = 0; b < n; ++b) out.u[b] = balanced<P>(u_ref * std::exp(cd(0.0, mm.phase_shift[b])));
        double max_dev = std::numeric_limits<double>::infinity();
        Idx num_iter = 0;
        while (max_dev > err_tol || num_iter == 0) {
            if (num_iter++ == max_iter)
                throw IterationDiverge("Iteration failed to converge after " + std::to_string(max_iter) +
                                       " iterations! Max deviation: " + fmt_double(max_dev) +
                                       ", error tolerance: " + fmt_double(err_tol) + ".\n");
            for (Idx b = 0; b < n; ++b) {
                Val<P> i_inj{};
                for (Idx l : loads_per_bus[b])
                    for (int ph = 0; ph < P; ++ph) {
                        cd const sv = in.s_inj[l].v[ph];
                        cd const u = out.u[b].v[ph];
                        switch (mm.load_type[l]) {
                        case 0: i_inj.v[ph] += std::conj(sv / u); break;
                        case 1: i_inj.v[ph] += std::conj(sv) * u; break;
                        case 2: i_inj.v[ph] += std::conj(sv * std::abs(u) / u); break;
                        default: throw OtherError("unknown load type");
                        }
                    }
                for (Idx s : sources_per_bus[b]) {
                    Val<P> const iv =
                        matvec<P>(y_ref<P>(mm.source_y1[s], mm.source_y0[s]), balanced<P>(in.u_source[s]));
                    for (int ph = 0; ph < P; ++ph) i_inj.v[ph] += iv.v[ph];
                }
                for (int ph = 0; ph < P; ++ph) {
                    rhs[b][ph] = i_inj.v[ph].real();
                    rhs[b][ph + P] = i_inj.v[ph].imag();
                }
            }
            lu.solve(mat, rhs, x);
            max_dev = 0.0;
            for (Idx b = 0; b < n; ++b) {
                double dev = 0.0;
                for (int ph = 0; ph < P; ++ph) {
                    cd const un(x[b][ph], x[b][ph + P]);
                    dev = std::max(dev, std::abs(un - out.u[b].v[ph]));
                    out.u[b].v[ph] = un;
                }
                max_dev = std::max(max_dev, dev);
            }
        }
        out.iterations = num_iter;
        calculate_result(in, out);
        return out;
    }

    // calculate_pf_result (common_solver_functions.hpp:414-446), y_bus.hpp:482-546
    void calculate_result(MathInput<P> const& in, MathOutput<P>& out, bool const_y = false) const {
        Idx const n = mm.n_bus;
        Idx const nbr = Idx(mm.branch_bus.size());
        out.br_if.resize(nbr);
        out.br_it.resize(nbr);
        out.br_sf.resize(nbr);
        out.br_st.resize(nbr);
        for (Idx b = 0; b < nbr; ++b) {
            auto const [f, t] = mm.branch_bus[b];
            Val<P> uf{}, ut{};
            if (f != -1) uf = out.u[f];
            if (t != -1) ut = out.u[t];
            auto const& pr = mm.branch_param[b];
            Val<P> a = matvec<P>(pr.v[0], uf), bb = matvec<P>(pr.v[1], ut);
            Val<P> c = matvec<P>(pr.v[2], uf), d = matvec<P>(pr.v[3], ut);
            for (int ph = 0; ph < P; ++ph) {
                out.br_if[b].v[ph] = a.v[ph] + bb.v[ph];
                out.br_it[b].v[ph] = c.v[ph] + d.v[ph];
                out.br_sf[b].v[ph] = uf.v[ph] * std::conj(out.br_if[b].v[ph]);
                out.br_st[b].v[ph] = ut.v[ph] * std::conj(out.br_it[b].v[ph]);
            }
        }
        Idx const nsh = Idx(mm.shunt_bus.size());
        out.sh_i.resize(nsh);
        out.sh_s.resize(nsh);
        for (Idx s = 0; s < nsh; ++s) {
            Idx const bus = mm.shunt_bus[s];
            Val<P> const iy = matvec<P>(mm.shunt_param[s], out.u[bus]);
            for (int ph = 0; ph < P; ++ph) {
                out.sh_i[s].v[ph] = -iy.v[ph];
                out.sh_s[s].v[ph] = out.u[bus].v[ph] * std::conj(out.sh_i[s].v[ph]);
            }
        }
        out.bus_injection.resize(n);
        for (Idx b = 0; b < n; ++b) {
            Val<P> acc{};
            for (Idx k = ys.row_indptr[b]; k < ys.row_indptr[b + 1]; ++k) {
                Val<P> const t = matvec<P>(ydata[k], out.u[ys.col_indices[k]]);
                for (int ph = 0; ph < P; ++ph) acc.v[ph] += t.v[ph];
            }
            for (int ph = 0; ph < P; ++ph) out.bus_injection[b].v[ph] = std::conj(acc.v[ph]) * out.u[b].v[ph];
        }
        Idx const nl = Idx(mm.load_bus.size());
        out.ld_s.resize(nl);
        out.ld_i.resize(nl);
        Idx const ns = Idx(mm.source_bus.size());
        out.src_s.resize(ns);
        out.src_i.resize(ns);
        for (Idx b = 0; b < n; ++b) {
            Val<P> iload{};
            for (Idx l : loads_per_bus[b]) {
                for (int ph = 0; ph < P; ++ph) {
                    cd const s = in.s_inj[l].v[ph];
                    cd const u = out.u[b].v[ph];
                    cd sl;
                    switch (const_y ? int8_t(1) : mm.load_type[l]) {
                    case 0: sl = s; break;
                    case 1: sl = s * std::norm(u); break;
                    default: sl = s * std::abs(u); break;
                    }
                    out.ld_s[l].v[ph] = sl;
                    out.ld_i[l].v[ph] = std::conj(sl / u);
                    iload.v[ph] += out.ld_i[l].v[ph];
                }
            }
            auto const& srcs = sources_per_bus[b];
            if (srcs.empty()) continue;
            Val<P> iinj;
            for (int ph = 0; ph < P; ++ph)
                iinj.v[ph] = std::conj(out.bus_injection[b].v[ph] / out.u[b].v[ph]) - iload.v[ph];
            if (srcs.size() == 1) {
                out.src_i[srcs[0]] = iinj;
                for (int ph = 0; ph < P; ++ph)
                    out.src_s[srcs[0]].v[ph] = out.u[b].v[ph] * std::conj(iinj.v[ph]);
            } else if (P == 1) {
                cd y_ref_t = 0, i_ref_t = 0;
                for (Idx s : srcs) y_ref_t += mm.source_y1[s];
                cd const z_ref_t = 1.0 / y_ref_t;
                for (Idx s : srcs) i_ref_t += in.u_source[s] * mm.source_y1[s];
                for (Idx s : srcs) {
                    cd const ratio = mm.source_y1[s] * z_ref_t;
                    cd const lhs = ratio * (in.u_source[s] * y_ref_t - i_ref_t);
                    out.src_i[s].v[0] = lhs + ratio * iinj.v[0];
                    out.src_s[s].v[0] = out.u[b].v[0] * std::conj(out.src_i[s].v[0]);
                }
            } else {
                std::array<cd, 3> yt{};
                cd i_ref_1_t = 0;
                for (Idx s : srcs) {
                    yt[0] += mm.source_y0[s];
                    yt[1] += mm.source_y1[s];
                    yt[2] += mm.source_y1[s];
                    i_ref_1_t += in.u_source[s] * mm.source_y1[s];
                }
                Ten<3> const smi = sym_matrix_inv(), sm = sym_matrix();
                Val<3> it3;
                for (int ph = 0; ph < 3; ++ph) it3.v[ph] = iinj.v[ph];
                Val<3> const i012 = matvec<3>(smi, it3);
                for (Idx s : srcs) {
                    cd const r1 = mm.source_y1[s] / yt[1];
                    cd const lhs = r1 * (in.u_source[s] * yt[1] - i_ref_1_t);
                    Val<3> v;
                    v.v[0] = (mm.source_y0[s] / yt[0]) * i012.v[0];
                    v.v[1] = lhs + r1 * i012.v[1];
                    v.v[2] = (mm.source_y1[s] / yt[2]) * i012.v[2];
                    Val<3> const iabc = matvec<3>(sm, v);
                    for (int ph = 0; ph < P; ++ph) {
                        out.src_i[s].v[ph] = iabc.v[ph];
                        out.src_s[s].v[ph] = out.u[b].v[ph] * std::conj(iabc.v[ph]);
                    }
                }
            }
        }
    }
};

// ------------------------------------------------------------------------------------------------
// component model, batch and outputs
// ------------------------------------------------------------------------------------------------
struct State {
    std::vector<int8_t> line_fs, line_ts;
    std::vector<int8_t> tr_fs, tr_ts, tr_tap;
    std::vector<int8_t> sh_status;
    std::vector<double> sh_g1, sh_b1, sh_g0, sh_b0;
    std::vector<int8_t> src_status;
    std::vector<double> src_uref, src_uang, src_sk, src_rx, src_z01;
    // load_gens in container order: sym_gen, asym_gen, sym_load, asym_load
    std::vector<int8_t> lg_status;
    std::vector<std::array<double, 3>> lg_p, lg_q;
};

struct Model {
    double freq = 50.0;
    std::vector<OrNodeInput> node;
    std::vector<OrLineInput> line;
    std::vector<OrLinkInput> link;
    std::vector<OrGenericBranchInput> gbranch;
    std::vector<OrTransformerInput> trafo;
    std::vector<OrShuntInput> shunt;
    std::vector<OrSourceInput> source;
    std::vector<int32_t> lg_id, lg_node;
    std::vector<int8_t> lg_type;
    std::vector<bool> lg_asym, lg_gen;
    std::vector<Idx> lg_node_idx;
    Idx n_lg_comp[4] = {0, 0, 0, 0};  // counts: sym_gen, asym_gen, sym_load, asym_load
    std::map<int32_t, std::pair<int, Idx>> ids;
    std::map<int32_t, Idx> node_index;
    std::vector<LineParam> lparam;
    std::vector<TrafoStatic> tstatic;
    std::vector<Idx> line_f, line_t, tr_f, tr_t, sh_node, src_node;
    State base;

    // line slots: the lines, then the links (a link is a line of admittance y_link)
    Idx n_lslot() const { return Idx(line_f.size()); }
    Idx n_branch() const { return Idx(line_f.size() + trafo.size()); }
};

Idx node_idx(Model const& m, int32_t id) {
    auto it = m.node_index.find(id);
    if (it == m.node_index.end()) throw OtherError("The id cannot be found: " + std::to_string(id) + "\n");
    return it->second;
}

Model build_model(OrConstDataset const& in, double f) {
    Model m;
    m.freq = f;
    auto reg = [&](int32_t id, int comp, Idx i) {
        if (!m.ids.emplace(id, std::make_pair(comp, i)).second)
            throw OtherError("Conflicting id detected: " + std::to_string(id) + "\n");
    };
    auto const* nodes = static_cast<OrNodeInput const*>(in.data[OR_NODE]);
    for (Idx i = 0; i < in.count[OR_NODE]; ++i) {
        m.node.push_back(nodes[i]);
        reg(nodes[i].id, OR_NODE, i);
        m.node_index[nodes[i].id] = i;
    }
    auto const* lines = static_cast<OrLineInput const*>(in.data[OR_LINE]);
    for (Idx i = 0; i < in.count[OR_LINE]; ++i) {
        auto const& l = lines[i];
        m.line.push_back(l);
        reg(l.id, OR_LINE, i);
        Idx const f_ = node_idx(m, l.from_node), t_ = node_idx(m, l.to_node);
        double const u1 = m.node[f_].u_rated, u2 = m.node[t_].u_rated;
        if (std::abs(u1 - u2) > kNumTol) throw OtherError("Conflicting voltage for line " + std::to_string(l.id));
        m.line_f.push_back(f_);
        m.line_t.push_back(t_);
        m.lparam.push_back(line_param(l, f, u1));
        m.base.line_fs.push_back(l.from_status);
        m.base.line_ts.push_back(l.to_status);
    }
    auto const* links = static_cast<OrLinkInput const*>(in.data[OR_LINK]);
    for (Idx i = 0; i < in.count[OR_LINK]; ++i) {
        auto const& l = links[i];
        m.link.push_back(l);
        reg(l.id, OR_LINK, i);
        Idx const f_ = node_idx(m, l.from_node), t_ = node_idx(m, l.to_node);
        m.line_f.push_back(f_);
        m.line_t.push_back(t_);
        m.lparam.push_back(link_param(m.node[f_].u_rated, m.node[t_].u_rated));
        m.base.line_fs.push_back(l.from_status);
        m.base.line_ts.push_back(l.to_status);
    }
    auto const* gbs = static_cast<OrGenericBranchInput const*>(in.data[OR_GENERIC_BRANCH]);
    for (Idx i = 0; i < in.count[OR_GENERIC_BRANCH]; ++i) {
        auto const& g = gbs[i];
        m.gbranch.push_back(g);
        reg(g.id, OR_GENERIC_BRANCH, i);
        Idx const f_ = node_idx(m, g.from_node), t_ = node_idx(m, g.to_node);
        m.line_f.push_back(f_);
        m.line_t.push_back(t_);
        m.lparam.push_back(generic_param(g, m.node[f_].u_rated, m.node[t_].u_rated));
        m.base.line_fs.push_back(g.from_status);
        m.base.line_ts.push_back(g.to_status);
    }
    auto const* trs = static_cast<OrTransformerInput const*>(in.data[OR_TRANSFORMER]);
    for (Idx i = 0; i < in.count[OR_TRANSFORMER]; ++i) {
        auto const& t = trs[i];
        m.trafo.push_back(t);
        reg(t.id, OR_TRANSFORMER, i);
        Idx const f_ = node_idx(m, t.from_node), t_ = node_idx(m, t.to_node);
        m.tr_f.push_back(f_);
        m.tr_t.push_back(t_);
        int8_t tap;
        m.tstatic.push_back(trafo_static(t, m.node[f_].u_rated, m.node[t_].u_rated, tap));
        m.base.tr_tap.push_back(tap_limit(m.tstatic.back(), tap));
        m.base.tr_fs.push_back(t.from_status);
        m.base.tr_ts.push_back(t.to_status);
    }
    auto const* shs = static_cast<OrShuntInput const*>(in.data[OR_SHUNT]);
    for (Idx i = 0; i < in.count[OR_SHUNT]; ++i) {
        m.shunt.push_back(shs[i]);
        reg(shs[i].id, OR_SHUNT, i);
        m.sh_node.push_back(node_idx(m, shs[i].node));
        m.base.sh_status.push_back(shs[i].status);
        m.base.sh_g1.push_back(shs[i].g1);
        m.base.sh_b1.push_back(shs[i].b1);
        m.base.sh_g0.push_back(shs[i].g0);
        m.base.sh_b0.push_back(shs[i].b0);
    }
    auto const* srcs = static_cast<OrSourceInput const*>(in.data[OR_SOURCE]);
    for (Idx i = 0; i < in.count[OR_SOURCE]; ++i) {
        auto const& s = srcs[i];
        m.source.push_back(s);
        reg(s.id, OR_SOURCE, i);
        m.src_node.push_back(node_idx(m, s.node));
        m.base.src_status.push_back(s.status);
        m.base.src_uref.push_back(s.u_ref);
        m.base.src_uang.push_back(is_nan(s.u_ref_angle) ? 0.0 : s.u_ref_angle);
        m.base.src_sk.push_back(is_nan(s.sk) ? 1e10 : s.sk);
        m.base.src_rx.push_back(is_nan(s.rx_ratio) ? 0.1 : s.rx_ratio);
        m.base.src_z01.push_back(is_nan(s.z01_ratio) ? 1.0 : s.z01_ratio);
    }
    int const lg_comps[4] = {OR_SYM_GEN, OR_ASYM_GEN, OR_SYM_LOAD, OR_ASYM_LOAD};
    for (int k = 0; k < 4; ++k) {
        int const comp = lg_comps[k];
        bool const asym = comp == OR_ASYM_GEN || comp == OR_ASYM_LOAD;
        bool const gen = comp == OR_SYM_GEN || comp == OR_ASYM_GEN;
        m.n_lg_comp[k] = in.count[comp];
        for (Idx i = 0; i < in.count[comp]; ++i) {
            int32_t id, node;
            int8_t status, type;
            std::array<double, 3> p{kNan, kNan, kNan}, q{kNan, kNan, kNan};
            if (asym) {
                auto const& r = static_cast<OrAsymLoadGenInput const*>(in.data[comp])[i];
                id = r.id;
                node = r.node;
                status = r.status;
                type = r.type;
                for (int ph = 0; ph < 3; ++ph) {
                    p[ph] = r.p_specified[ph];
                    q[ph] = r.q_specified[ph];
                }
            } else {
                auto const& r = static_cast<OrSymLoadGenInput const*>(in.data[comp])[i];
                id = r.id;
                node = r.node;
                status = r.status;
                type = r.type;
                p[0] = r.p_specified;
                q[0] = r.q_specified;
            }
            reg(id, comp, i);
            m.lg_id.push_back(id);
            m.lg_node.push_back(node);
            m.lg_node_idx.push_back(node_idx(m, node));
            m.lg_type.push_back(type);
            m.lg_asym.push_back(asym);
            m.lg_gen.push_back(gen);
            m.base.lg_status.push_back(status);
            m.base.lg_p.push_back(p);
            m.base.lg_q.push_back(q);
        }
    }
    return m;
}

Idx lg_offset(Model const& m, int comp) {
    Idx off = 0;
    int const order[4] = {OR_SYM_GEN, OR_ASYM_GEN, OR_SYM_LOAD, OR_ASYM_LOAD};
    for (int k = 0; k < 4; ++k) {
        if (order[k] == comp) return off;
        off += m.n_lg_comp[k];
    }
    return off;
}

// apply one scenario's update rows (component update semantics: NaN / na means "keep")
void apply_update(Model const& m, State& st, OrConstDataset const& upd, Idx scenario, bool& topo_changed,
                  bool& param_changed) {
    auto lookup = [&](int32_t id, int comp, Idx row, Idx n_in_update) -> Idx {
        if (id == kNaID) {
            // optional id: positional when the update holds every element of the component
            (void)n_in_update;
            return row;
        }
        auto it = m.ids.find(id);
        if (it == m.ids.end()) throw OtherError("The id cannot be found: " + std::to_string(id) + "\n");
        if (it->second.first != comp) throw OtherError("Wrong type for object with id " + std::to_string(id) + "\n");
        return it->second.second;
    };
    auto set_s = [](int8_t& tgt, int8_t v) {
        if (v == kNaS) return false;
        bool const ch = (tgt != 0) != (v != 0);
        tgt = v != 0;
        return ch;
    };
    Idx cnt = 0;
    for (int comp : {int(OR_LINE), int(OR_LINK), int(OR_GENERIC_BRANCH)}) {
        cnt = upd.count[comp];
        auto const* lu = static_cast<OrBranchUpdate const*>(upd.data[comp]);
        Idx const slot0 = comp == OR_LINE ? 0 : comp == OR_LINK ? Idx(m.line.size()) : Idx(m.line.size() + m.link.size());
        for (Idx r = 0; r < cnt; ++r) {
            auto const& u = lu[scenario * cnt + r];
            Idx const i = slot0 + lookup(u.id, comp, r, cnt);
            bool ch = set_s(st.line_fs[i], u.from_status);
            ch = set_s(st.line_ts[i], u.to_status) || ch;
            topo_changed |= ch;
            param_changed |= ch;
        }
    }
    cnt = upd.count[OR_TRANSFORMER];
    auto const* tu = static_cast<OrTransformerUpdate const*>(upd.data[OR_TRANSFORMER]);
    for (Idx r = 0; r < cnt; ++r) {
        auto const& u = tu[scenario * cnt + r];
        Idx const i = lookup(u.id, OR_TRANSFORMER, r, cnt);
        bool ch = set_s(st.tr_fs[i], u.from_status);
        ch = set_s(st.tr_ts[i], u.to_status) || ch;
        topo_changed |= ch;
        param_changed |= ch;
        if (u.tap_pos != kNaS && u.tap_pos != st.tr_tap[i]) {
            st.tr_tap[i] = tap_limit(m.tstatic[i], u.tap_pos);
            param_changed = true;
        }
    }
    cnt = upd.count[OR_SHUNT];
    auto const* su = static_cast<OrShuntUpdate const*>(upd.data[OR_SHUNT]);
    for (Idx r = 0; r < cnt; ++r) {
        auto const& u = su[scenario * cnt + r];
        Idx const i = lookup(u.id, OR_SHUNT, r, cnt);
        if (u.status != kNaS) st.sh_status[i] = u.status != 0;
        if (!is_nan(u.g1)) st.sh_g1[i] = u.g1;
        if (!is_nan(u.b1)) st.sh_b1[i] = u.b1;
        if (!is_nan(u.g0)) st.sh_g0[i] = u.g0;
        if (!is_nan(u.b0)) st.sh_b0[i] = u.b0;
        param_changed = true;
    }
    cnt = upd.count[OR_SOURCE];
    auto const* cu = static_cast<OrSourceUpdate const*>(upd.data[OR_SOURCE]);
    for (Idx r = 0; r < cnt; ++r) {
        auto const& u = cu[scenario * cnt + r];
        Idx const i = lookup(u.id, OR_SOURCE, r, cnt);
        bool const ch = set_s(st.src_status[i], u.status);
        topo_changed |= ch;
        if (!is_nan(u.u_ref)) st.src_uref[i] = u.u_ref;
        if (!is_nan(u.u_ref_angle)) st.src_uang[i] = u.u_ref_angle;
        if (!is_nan(u.sk)) st.src_sk[i] = u.sk;
        if (!is_nan(u.rx_ratio)) st.src_rx[i] = u.rx_ratio;
        if (!is_nan(u.z01_ratio)) st.src_z01[i] = u.z01_ratio;
        param_changed = true;
    }
    int const lg_comps[4] = {OR_SYM_GEN, OR_ASYM_GEN, OR_SYM_LOAD, OR_ASYM_LOAD};
    for (int comp : lg_comps) {
        cnt = upd.count[comp];
        bool const asym = comp == OR_ASYM_GEN || comp == OR_ASYM_LOAD;
        Idx const off = lg_offset(m, comp);
        for (Idx r = 0; r < cnt; ++r) {
            int32_t id;
            int8_t status;
            double p[3] = {kNan, kNan, kNan}, q[3] = {kNan, kNan, kNan};
            if (asym) {
                auto const& u = static_cast<OrAsymLoadGenUpdate const*>(upd.data[comp])[scenario * cnt + r];
                id = u.id;
                status = u.status;
                for (int ph = 0; ph < 3; ++ph) {
                    p[ph] = u.p_specified[ph];
                    q[ph] = u.q_specified[ph];
                }
            } else {
                auto const& u = static_cast<OrSymLoadGenUpdate const*>(upd.data[comp])[scenario * cnt + r];
                id = u.id;
                status = u.status;
                p[0] = u.p_specified;
                q[0] = u.q_specified;
            }
            Idx const i = off + lookup(id, comp, r, cnt);
            if (status != kNaS) st.lg_status[i] = status != 0;
            int const nph = asym ? 3 : 1;
            for (int ph = 0; ph < nph; ++ph) {
                if (!is_nan(p[ph])) st.lg_p[i][ph] = p[ph];
                if (!is_nan(q[ph])) st.lg_q[i][ph] = q[ph];
            }
        }
    }
}

struct TopoCache {
    TopoResult topo;
    // per math model: branch list (component branch index), loads, sources, shunts in math order
    std::vector<std::vector<Idx>> mm_branch, mm_load, mm_source, mm_shunt;
    std::vector<std::vector<std::array<Idx, 2>>> mm_branch_bus;
    std::vector<Idx> br_group, br_pos, lg_group, lg_pos, src_group, src_pos, sh_group, sh_pos;
};

TopoCache build_topo_cache(Model const& m, State const& st) {
    TopoCache tc;
    std::vector<CompBranch> br;
    Idx const nl = m.n_lslot();
    for (Idx i = 0; i < nl; ++i)
        br.push_back({m.line_f[i], m.line_t[i], st.line_fs[i] != 0, st.line_ts[i] != 0, m.lparam[i].shift});
    for (Idx i = 0; i < Idx(m.trafo.size()); ++i)
        br.push_back({m.tr_f[i], m.tr_t[i], st.tr_fs[i] != 0, st.tr_ts[i] != 0, m.tstatic[i].clock * kDeg30});
    std::vector<bool> son;
    for (auto s : st.src_status) son.push_back(s != 0);
    tc.topo = build_topology(Idx(m.node.size()), br, m.src_node, son);
    Idx const nmm = Idx(tc.topo.bus_nodes.size());
    tc.mm_branch.resize(nmm);
    tc.mm_branch_bus.resize(nmm);
    tc.mm_load.resize(nmm);
    tc.mm_source.resize(nmm);
    tc.mm_shunt.resize(nmm);
    auto const& ng = tc.topo.node_group;
    auto const& np = tc.topo.node_pos;
    // couple_branch (topology.hpp:370-413)
    tc.br_group.assign(br.size(), -1);
    tc.br_pos.assign(br.size(), -1);
    for (Idx b = 0; b < Idx(br.size()); ++b) {
        Idx g = -1;
        if (br[b].fs && ng[br[b].from] != -1)
            g = ng[br[b].from];
        else if (br[b].ts && ng[br[b].to] != -1)
            g = ng[br[b].to];
        if (g == -1) continue;
        tc.br_group[b] = g;
        tc.br_pos[b] = Idx(tc.mm_branch[g].size());
        tc.mm_branch[g].push_back(b);
        tc.mm_branch_bus[g].push_back({br[b].fs ? np[br[b].from] : -1, br[b].ts ? np[br[b].to] : -1});
    }
    // couple appliances, stable by bus (build_sparse_mapping / build_dense_mapping)
    auto couple = [&](std::vector<Idx> const& node_of, std::vector<bool> const& include,
                      std::vector<std::vector<Idx>>& per_mm, std::vector<Idx>& grp, std::vector<Idx>& pos) {
        Idx const n = Idx(node_of.size());
        grp.assign(n, -1);
        pos.assign(n, -1);
        std::vector<std::vector<std::pair<Idx, Idx>>> lists(nmm);
        for (Idx i = 0; i < n; ++i) {
            if (!include[i]) continue;
            Idx const g = ng[node_of[i]];
            if (g < 0) continue;
            lists[g].push_back({np[node_of[i]], i});
        }
        for (Idx g = 0; g < nmm; ++g) {
            std::stable_sort(lists[g].begin(), lists[g].end(),
                             [](auto const& a, auto const& b) { return a.first < b.first; });
            for (auto const& [bus, i] : lists[g]) {
                grp[i] = g;
                pos[i] = Idx(per_mm[g].size());
                per_mm[g].push_back(i);
            }
        }
    };
    std::vector<bool> all_sh(m.shunt.size(), true), all_lg(m.lg_id.size(), true), src_inc(m.source.size());
    for (size_t i = 0; i < m.source.size(); ++i) src_inc[i] = st.src_status[i] != 0;
    couple(m.sh_node, all_sh, tc.mm_shunt, tc.sh_group, tc.sh_pos);
    couple(m.lg_node_idx, all_lg, tc.mm_load, tc.lg_group, tc.lg_pos);
    couple(m.src_node, src_inc, tc.mm_source, tc.src_group, tc.src_pos);
    return tc;
}

template <int P> MathModel<P> make_math_model(Model const& m, State const& st, TopoCache const& tc, Idx g) {
    MathModel<P> mm;
    mm.n_bus = Idx(tc.topo.bus_nodes[g].size());
    for (Idx node : tc.topo.bus_nodes[g]) mm.phase_shift.push_back(tc.topo.phase[node]);
    mm.branch_bus = tc.mm_branch_bus[g];
    mm.fill_in = tc.topo.fill_in[g];
    Idx const nl = m.n_lslot();
    for (Idx b : tc.mm_branch[g]) {
        BranchParam<P> bp;
        if (b < nl) {
            bool const fs = st.line_fs[b] != 0, ts = st.line_ts[b] != 0;
            auto const& lp = m.lparam[b];
            if (P == 1) {
                auto const p = calc_param_y_sym(lp.y1s, lp.y1sh, lp.ratio, fs, ts);
                for (int i = 0; i < 4; ++i) bp.v[i] = ten_sm<P>(p.v[i], p.v[i]);
            } else {
                if (lp.generic) throw OtherError("The functionality is not implemented. asymmetric generic_branch\n");
                auto const p1 = calc_param_y_sym(lp.y1s, lp.y1sh, 1.0, fs, ts);
                auto const p0 = calc_param_y_sym(lp.y0s, lp.y0sh, 1.0, fs, ts);
                auto const a = sym_to_asym(p1, p0);
                for (int i = 0; i < 4; ++i)
                    for (int k = 0; k < P * P; ++k) bp.v[i].m[k] = a.v[i].m[k];
            }
        } else {
            Idx const t = b - nl;
            bool const fs = st.tr_fs[t] != 0, ts = st.tr_ts[t] != 0;
            if (P == 1) {
                auto const p = trafo_param_sym(m.tstatic[t], st.tr_tap[t], fs, ts);
                for (int i = 0; i < 4; ++i) bp.v[i] = ten_sm<P>(p.v[i], p.v[i]);
            } else {
                auto const a = trafo_param_asym(m.tstatic[t], st.tr_tap[t], fs, ts);
                for (int i = 0; i < 4; ++i)
                    for (int k = 0; k < P * P; ++k) bp.v[i].m[k] = a.v[i].m[k];
            }
        }
        mm.branch_param.push_back(bp);
    }
    for (Idx s : tc.mm_shunt[g]) {
        mm.shunt_bus.push_back(tc.topo.node_pos[m.sh_node[s]]);
        Ten<P> y{};
        if (st.sh_status[s]) {
            double const u = m.node[m.sh_node[s]].u_rated;
            double const base_y = (kBaseP3 / u / kSqrt3) / (u / kSqrt3);
            cd const y1 = cd(st.sh_g1[s], st.sh_b1[s]) / base_y;
            cd const y0 = cd(st.sh_g0[s], st.sh_b0[s]) / base_y;
            y = P == 1 ? ten_sm<P>(y1, y1) : ten_sm<P>((2.0 * y1 + y0) / 3.0, (y0 - y1) / 3.0);
        }
        mm.shunt_param.push_back(y);
    }
    for (Idx s : tc.mm_source[g]) {
        mm.source_bus.push_back(tc.topo.node_pos[m.src_node[s]]);
        // source.hpp:40-48
        double const z_abs = kBaseP3 / st.src_sk[s];
        double const rx = st.src_rx[s];
        double const x1 = z_abs / std::sqrt(rx * rx + 1.0);
        double const r1 = x1 * rx;
        cd const y1 = 1.0 / cd(r1, x1);
        mm.source_y1.push_back(y1);
        mm.source_y0.push_back(y1 / st.src_z01[s]);
    }
    for (Idx l : tc.mm_load[g]) {
        mm.load_bus.push_back(tc.topo.node_pos[m.lg_node_idx[l]]);
        mm.load_type.push_back(m.lg_type[l]);
    }
    return mm;
}

// load_gen.hpp set_power / calc_param: s = direction * (p + jq) / base_power
template <int P> Val<P> load_s(Model const& m, State const& st, Idx l) {
    Val<P> r;
    if (!st.lg_status[l]) return r;
    double const dir = m.lg_gen[l] ? 1.0 : -1.0;
    if (!m.lg_asym[l]) {
        cd const s = dir * cd(st.lg_p[l][0], st.lg_q[l][0]) / kBaseP3;
        if (std::isnan(s.real()) || std::isnan(s.imag())) {
            for (int ph = 0; ph < P; ++ph) r.v[ph] = cd(kNan, kNan);
            return r;
        }
        for (int ph = 0; ph < P; ++ph) r.v[ph] = s;  // piecewise, no rotation
        return r;
    }
    std::array<cd, 3> sp;
    for (int ph = 0; ph < 3; ++ph) sp[ph] = dir * cd(st.lg_p[l][ph], st.lg_q[l][ph]) / kBaseP1;
    if (P == 1)
        r.v[0] = (sp[0] + sp[1] + sp[2]) / 3.0;
    else
        for (int ph = 0; ph < 3; ++ph) r.v[ph] = sp[ph];
    return r;
}

struct OutPtrs {
    OrMutableDataset const* out;
    Idx scenario;
};

template <int P> void write_outputs(Model const& m, State const& st, TopoCache const& tc,
                                    std::vector<MathOutput<P>> const& res, OrMutableDataset const& out, Idx scn) {
    double const base_power = P == 1 ? kBaseP3 : kBaseP1;
    double const u_scale = P == 1 ? 1.0 : 1.0 / kSqrt3;
    Idx const nn = Idx(m.node.size());
    // node injection = sum of appliance s in order Source, SymLoad, SymGen, AsymLoad, AsymGen
    std::vector<Val<P>> node_inj(nn);
    for (Idx s = 0; s < Idx(m.source.size()); ++s) {
        if (tc.src_group[s] < 0) continue;
        auto const& v = res[tc.src_group[s]].src_s[tc.src_pos[s]];
        for (int ph = 0; ph < P; ++ph) node_inj[m.src_node[s]].v[ph] += v.v[ph];
    }
    int const inj_order[4] = {OR_SYM_LOAD, OR_SYM_GEN, OR_ASYM_LOAD, OR_ASYM_GEN};
    for (int comp : inj_order) {
        Idx const off = lg_offset(m, comp);
        Idx const cnt = comp == OR_SYM_GEN ? m.n_lg_comp[0]
                        : comp == OR_ASYM_GEN ? m.n_lg_comp[1]
                        : comp == OR_SYM_LOAD ? m.n_lg_comp[2]
                                              : m.n_lg_comp[3];
        for (Idx k = 0; k < cnt; ++k) {
            Idx const l = off + k;
            if (tc.lg_group[l] < 0) continue;
            auto const& v = res[tc.lg_group[l]].ld_s[tc.lg_pos[l]];
            for (int ph = 0; ph < P; ++ph) node_inj[m.lg_node_idx[l]].v[ph] += v.v[ph];
        }
    }
    auto put = [](double* dst, double v, int ph) { dst[ph] = v; };
    (void)put;
    if (out.data[OR_NODE]) {
        for (Idx i = 0; i < nn; ++i) {
            Idx const g = tc.topo.node_group[i];
            if (P == 1) {
                auto& o = static_cast<OrNodeOutputSym*>(out.data[OR_NODE])[scn * nn + i];
                o = OrNodeOutputSym{};
                o.id = m.node[i].id;
                if (g < 0) continue;
                o.energized = 1;
                cd const u = res[g].u[tc.topo.node_pos[i]].v[0];
                o.u_pu = std::abs(u);
                o.u = u_scale * m.node[i].u_rated * o.u_pu;
                o.u_angle = std::arg(u);
                o.p = base_power * node_inj[i].v[0].real();
                o.q = base_power * node_inj[i].v[0].imag();
            } else {
                auto& o = static_cast<OrNodeOutputAsym*>(out.data[OR_NODE])[scn * nn + i];
                o = OrNodeOutputAsym{};
                o.id = m.node[i].id;
                if (g < 0) continue;
                o.energized = 1;
                for (int ph = 0; ph < 3; ++ph) {
                    cd const u = res[g].u[tc.topo.node_pos[i]].v[ph];
                    o.u_pu[ph] = std::abs(u);
                    o.u[ph] = u_scale * m.node[i].u_rated * o.u_pu[ph];
                    o.u_angle[ph] = std::arg(u);
                    o.p[ph] = base_power * node_inj[i].v[ph].real();
                    o.q[ph] = base_power * node_inj[i].v[ph].imag();
                }
            }
        }
    }
    // branches (branch.hpp get_output)
    Idx const nl = m.n_lslot();
    auto branch_out = [&](int comp, Idx first, Idx count) {
        if (!out.data[comp]) return;
        for (Idx k = 0; k < count; ++k) {
            Idx const b = first + k;
            int32_t const id = comp == OR_LINE            ? m.line[k].id
                               : comp == OR_LINK           ? m.link[k].id
                               : comp == OR_GENERIC_BRANCH ? m.gbranch[k].id
                                                           : m.trafo[k].id;
            double bif, bit;
            if (comp != OR_TRANSFORMER) {
                bif = m.lparam[b].base_i;
                bit = m.lparam[b].base_i_to;
            } else {
                bif = m.tstatic[k].base_i_from;
                bit = m.tstatic[k].base_i_to;
            }
            Idx const g = tc.br_group[b];
            std::array<double, 3> pf{}, qf{}, if_{}, sf{}, pt{}, qt{}, it{}, st_{};
            double loading = 0.0;
            int8_t en = 0;
            if (g >= 0) {
                en = 1;
                Idx const pos = tc.br_pos[b];
                auto const& r = res[g];
                double sum_sf = 0, sum_st = 0, max_i = 0;
                for (int ph = 0; ph < P; ++ph) {
                    pf[ph] = base_power * r.br_sf[pos].v[ph].real();
                    qf[ph] = base_power * r.br_sf[pos].v[ph].imag();
                    if_[ph] = bif * std::abs(r.br_if[pos].v[ph]);
                    sf[ph] = base_power * std::abs(r.br_sf[pos].v[ph]);
                    pt[ph] = base_power * r.br_st[pos].v[ph].real();
                    qt[ph] = base_power * r.br_st[pos].v[ph].imag();
                    it[ph] = bit * std::abs(r.br_it[pos].v[ph]);
                    st_[ph] = base_power * std::abs(r.br_st[pos].v[ph]);
                    sum_sf += sf[ph];
                    sum_st += st_[ph];
                }
                double mi_f = if_[0], mi_t = it[0];
                for (int ph = 1; ph < P; ++ph) {
                    mi_f = std::max(mi_f, if_[ph]);
                    mi_t = std::max(mi_t, it[ph]);
                }
                max_i = std::max(mi_f, mi_t);
                double const max_s = std::max(sum_sf, sum_st);
                if (comp == OR_GENERIC_BRANCH)
                    loading = std::isnan(m.lparam[b].sn) ? 0.0 : max_s / m.lparam[b].sn;
                else
                    loading = comp != OR_TRANSFORMER ? max_i / m.lparam[b].i_n : max_s / m.tstatic[k].sn;
            }
            if (P == 1) {
                auto& o = static_cast<OrBranchOutputSym*>(out.data[comp])[scn * count + k];
                o = OrBranchOutputSym{};
                o.id = id;
                o.energized = en;
                o.loading = loading;
                o.p_from = pf[0];
                o.q_from = qf[0];
                o.i_from = if_[0];
                o.s_from = sf[0];
                o.p_to = pt[0];
                o.q_to = qt[0];
                o.i_to = it[0];
                o.s_to = st_[0];
            } else {
                auto& o = static_cast<OrBranchOutputAsym*>(out.data[comp])[scn * count + k];
                o = OrBranchOutputAsym{};
                o.id = id;
                o.energized = en;
                o.loading = loading;
                for (int ph = 0; ph < 3; ++ph) {
                    o.p_from[ph] = pf[ph];
                    o.q_from[ph] = qf[ph];
                    o.i_from[ph] = if_[ph];
                    o.s_from[ph] = sf[ph];
                    o.p_to[ph] = pt[ph];
                    o.q_to[ph] = qt[ph];
                    o.i_to[ph] = it[ph];
                    o.s_to[ph] = st_[ph];
                }
            }
        }
    };
    branch_out(OR_LINE, 0, Idx(m.line.size()));
    branch_out(OR_LINK, Idx(m.line.size()), Idx(m.link.size()));
    branch_out(OR_GENERIC_BRANCH, Idx(m.line.size() + m.link.size()), Idx(m.gbranch.size()));
    branch_out(OR_TRANSFORMER, nl, Idx(m.trafo.size()));
    // appliances (appliance.hpp get_output)
    auto app_out = [&](int comp, Idx k, Idx count, int32_t id, double base_i, double dir, bool status, Idx g,
                       Val<P> const* s, Val<P> const* i) {
        std::array<double, 3> p{}, q{}, ii{}, ss{}, pf{};
        int8_t en = 0;
        if (g >= 0) {
            en = status ? 1 : 0;
            for (int ph = 0; ph < P; ++ph) {
                p[ph] = base_power * s->v[ph].real() * dir;
                q[ph] = base_power * s->v[ph].imag() * dir;
                ss[ph] = base_power * std::abs(s->v[ph]);
                ii[ph] = base_i * std::abs(i->v[ph]);
                pf[ph] = ss[ph] < kNumTol ? 0.0 : p[ph] / ss[ph];
            }
        }
        if (P == 1) {
            auto& o = static_cast<OrApplianceOutputSym*>(out.data[comp])[scn * count + k];
            o = OrApplianceOutputSym{};
            o.id = id;
            o.energized = en;
            o.p = p[0];
            o.q = q[0];
            o.i = ii[0];
            o.s = ss[0];
            o.pf = pf[0];
        } else {
            auto& o = static_cast<OrApplianceOutputAsym*>(out.data[comp])[scn * count + k];
            o = OrApplianceOutputAsym{};
            o.id = id;
            o.energized = en;
            for (int ph = 0; ph < 3; ++ph) {
                o.p[ph] = p[ph];
                o.q[ph] = q[ph];
                o.i[ph] = ii[ph];
                o.s[ph] = ss[ph];
                o.pf[ph] = pf[ph];
            }
        }
    };
    if (out.data[OR_SHUNT]) {
        Idx const c = Idx(m.shunt.size());
        for (Idx k = 0; k < c; ++k) {
            Idx const g = tc.sh_group[k];
            double const bi = kBaseP3 / m.node[m.sh_node[k]].u_rated / kSqrt3;
            app_out(OR_SHUNT, k, c, m.shunt[k].id, bi, -1.0, st.sh_status[k] != 0, g,
                    g >= 0 ? &res[g].sh_s[tc.sh_pos[k]] : nullptr, g >= 0 ? &res[g].sh_i[tc.sh_pos[k]] : nullptr);
        }
    }
    if (out.data[OR_SOURCE]) {
        Idx const c = Idx(m.source.size());
        for (Idx k = 0; k < c; ++k) {
            Idx const g = tc.src_group[k];
            double const bi = kBaseP3 / m.node[m.src_node[k]].u_rated / kSqrt3;
            app_out(OR_SOURCE, k, c, m.source[k].id, bi, 1.0, st.src_status[k] != 0, g,
                    g >= 0 ? &res[g].src_s[tc.src_pos[k]] : nullptr, g >= 0 ? &res[g].src_i[tc.src_pos[k]] : nullptr);
        }
    }
    int const lg_comps[4] = {OR_SYM_GEN, OR_ASYM_GEN, OR_SYM_LOAD, OR_ASYM_LOAD};
    for (int kk = 0; kk < 4; ++kk) {
        int const comp = lg_comps[kk];
        if (!out.data[comp]) continue;
        Idx const off = lg_offset(m, comp), c = m.n_lg_comp[kk];
        for (Idx k = 0; k < c; ++k) {
            Idx const l = off + k;
            Idx const g = tc.lg_group[l];
            double const bi = kBaseP3 / m.node[m.lg_node_idx[l]].u_rated / kSqrt3;
            app_out(comp, k, c, m.lg_id[l], bi, m.lg_gen[l] ? 1.0 : -1.0, st.lg_status[l] != 0, g,
                    g >= 0 ? &res[g].ld_s[tc.lg_pos[l]] : nullptr, g >= 0 ? &res[g].ld_i[tc.lg_pos[l]] : nullptr);
        }
    }
}

template <int P> Idx solve_scenario(Model const& m, State const& st, TopoCache const& tc, double err_tol,
                                    Idx max_iter, OrMutableDataset const* out, Idx scn, int method) {
    Idx const nmm = Idx(tc.topo.bus_nodes.size());
    std::vector<MathOutput<P>> res;
    Idx max_it = 0;
    for (Idx g = 0; g < nmm; ++g) {
        MathModel<P> const mm = make_math_model<P>(m, st, tc, g);
        MathInput<P> in;
        for (Idx l : tc.mm_load[g]) in.s_inj.push_back(load_s<P>(m, st, l));
        for (Idx s : tc.mm_source[g]) in.u_source.push_back(st.src_uref[s] * std::exp(cd(0.0, st.src_uang[s])));
        NRSolver<P> solver(mm);
        res.push_back(solver.run(in, err_tol, max_iter, method));
        max_it = std::max(max_it, res.back().iterations);
    }
    if (out) write_outputs<P>(m, st, tc, res, *out, scn);
    return max_it;
}

void write_msg(char* messages, Idx stride, Idx scn, std::string const& s) {
    if (!messages || stride <= 0) return;
    char* dst = messages + scn * stride;
    size_t const n = std::min<size_t>(s.size(), size_t(stride - 1));
    std::memcpy(dst, s.data(), n);
    dst[n] = 0;
}

// ------------------------------------------------------------------------------------------------
// automatic tap changing (optimizer/tap_position_optimizer.hpp, component/transformer_tap_regulator.hpp,
// main_model_impl.hpp:318-346), restated per scenario in the reference's own sequence: rank the
// regulated transformers, pilot run, iterate with fallback, exploit the neighbourhood
// ------------------------------------------------------------------------------------------------
struct MaxIterationReached : IterationDiverge {
    explicit MaxIterationReached(std::string const& what)
        : IterationDiverge("Maximum number of iterations reached! " + what + "\n") {}
};

struct RegulatorState {
    std::vector<int8_t> status;
    std::vector<double> u_set, u_band, ldc_r, ldc_x;
};

constexpr Idx kInfDist = std::numeric_limits<Idx>::max();

// tap_position_optimizer.hpp:209-432. Returns the regulator indices rank by rank.
std::vector<std::vector<Idx>> rank_regulators(Model const& m, State const& st,
                                              std::vector<OrTapRegulatorInput> const& regs,
                                              RegulatorState const& rs) {
    struct Edge {
        Idx from, to, weight, reg;
    };
    std::vector<Edge> edges;
    // add_edge<Transformer> (:209-234), add_edge<Line> (:236-252)
    for (Idx t = 0; t < Idx(m.trafo.size()); ++t) {
        if (!st.tr_fs[t] || !st.tr_ts[t]) continue;
        Idx reg = -1;
        for (Idx r = 0; r < Idx(regs.size()); ++r)
            if (rs.status[r] && regs[r].regulated_object == m.trafo[t].id) reg = r;
        if (reg >= 0) {
            Idx const ctrl = regs[reg].control_side == 0 ? m.tr_f[t] : m.tr_t[t];
            Idx const other = regs[reg].control_side == 0 ? m.tr_t[t] : m.tr_f[t];
            edges.push_back({other, ctrl, 1, reg});
        } else {
            edges.push_back({m.tr_f[t], m.tr_t[t], 0, -1});
            edges.push_back({m.tr_t[t], m.tr_f[t], 0, -1});
        }
    }
    // lines and links; generic branches are no edges (tap_position_optimizer.hpp:289)
    for (Idx l = 0; l < Idx(m.line.size() + m.link.size()); ++l) {
        if (!st.line_fs[l] || !st.line_ts[l]) continue;
        edges.push_back({m.line_f[l], m.line_t[l], 0, -1});
        edges.push_back({m.line_t[l], m.line_f[l], 0, -1});
    }
    // compressed sparse row graph: out-edges of a vertex in the order they were added
    std::stable_sort(edges.begin(), edges.end(), [](Edge const& a, Edge const& b) { return a.from < b.from; });
    Idx const n = Idx(m.node.size());
    std::vector<Idx> first(size_t(n) + 1, 0);
    for (Edge const& e : edges) ++first[size_t(e.from) + 1];
    for (Idx v = 0; v < n; ++v) first[size_t(v) + 1] += first[size_t(v)];
    std::vector<char> is_source(size_t(n), 0);
    for (Idx s = 0; s < Idx(m.source.size()); ++s) is_source[size_t(m.src_node[s])] = st.src_status[s] != 0;
    // Dijkstra from every source vertex in turn over one distance vector (process_edges_dijkstra, :308-334)
    std::vector<Idx> dist(size_t(n), kInfDist);
    for (Idx v = 0; v < n; ++v) {
        if (!is_source[size_t(v)]) continue;
        using Item = std::pair<Idx, Idx>;
        std::vector<Item> heap;
        auto cmp = [](Item const& a, Item const& b) { return a > b; };
        dist[size_t(v)] = 0;
        heap.push_back({0, v});
        while (!heap.empty()) {
            std::pop_heap(heap.begin(), heap.end(), cmp);
            auto const [d, u] = heap.back();
            heap.pop_back();
            if (d != dist[size_t(u)]) continue;
            for (Idx k = first[size_t(u)]; k < first[size_t(u) + 1]; ++k) {
                Edge const& e = edges[size_t(k)];
                if (dist[size_t(e.from)] + e.weight < dist[size_t(e.to)]) {
                    dist[size_t(e.to)] = dist[size_t(e.from)] + e.weight;
                    heap.push_back({dist[size_t(e.to)], e.to});
                    std::push_heap(heap.begin(), heap.end(), cmp);
                }
            }
        }
    }
    // get_edge_weights (:338-407)
    std::vector<std::pair<Idx, Idx>> weighted;  // (weight, regulator)
    std::vector<int32_t> invalid;
    for (Edge const& e : edges) {
        if (e.reg < 0) continue;
        Idx const a = dist[size_t(e.from)], b = dist[size_t(e.to)];
        if (a == kInfDist && b == kInfDist) continue;
        if (a == kInfDist || b == kInfDist)
            invalid.push_back(regs[size_t(e.reg)].regulated_object);
        else if (a != b - 1)
            weighted.push_back({kInfDist - 1, e.reg});
        else
            weighted.push_back({b, e.reg});
    }
    if (!invalid.empty()) {
        std::sort(invalid.begin(), invalid.end());
        invalid.erase(std::unique(invalid.begin(), invalid.end()), invalid.end());
        std::string msg = "Automatic tap changer has invalid configuration. The following transformer(s) are being "
                          "controlled from non-source side towards source side:\n  Transformer IDs: ";
        for (size_t i = 0; i < invalid.size(); ++i) {
            if (i) msg += ", ";
            msg += std::to_string(invalid[i]);
        }
        throw OtherError(msg);
    }
    // rank_transformers (:409-428)
    std::stable_sort(weighted.begin(), weighted.end(),
                     [](std::pair<Idx, Idx> const& a, std::pair<Idx, Idx> const& b) { return a.first < b.first; });
    std::vector<std::vector<Idx>> groups;
    Idx previous = std::numeric_limits<Idx>::lowest();
    for (auto const& [w, r] : weighted) {
        if (w > previous) {
            groups.emplace_back();
            previous = w;
        }
        if (std::find(groups.back().begin(), groups.back().end(), r) == groups.back().end()) groups.back().push_back(r);
    }
    return groups;
}

template <int P> struct TapOptimizer {
    // BinarySearch of the reference (:809-944)
    struct Bisection {
        int8_t lower = 0, upper = 0, current = 0;
        bool last_down = false, last_check = false, tap_reverse = false, inevitable_run = false, ctrl_at_tap = false;

        void reset(int8_t tap, int8_t tmin, int8_t tmax, bool cat) {
            last_down = last_check = inevitable_run = false;
            current = tap;
            lower = std::min(tmin, tmax);
            upper = std::max(tmin, tmax);
            tap_reverse = tmax < tmin;
            ctrl_at_tap = cat;
        }
        bool end_of_search() const { return lower >= upper; }
        int8_t search(bool prefer_higher_in) const {
            bool const prefer_higher = ctrl_at_tap != prefer_higher_in;
            int const primary = prefer_higher ? upper : lower, secondary = prefer_higher ? lower : upper;
            // std::midpoint for integers: rounded towards the first argument
            return int8_t(primary <= secondary ? primary + (secondary - primary) / 2
                                               : primary - (primary - secondary) / 2);
        }
        void adjust(bool strategy_max) {
            if (last_down)
                upper = current;
            else
                lower = current;
            if (lower < upper) current = search(strategy_max != tap_reverse);
        }
        void propose_new_pos(bool strategy_max, bool above_range) {
            bool const is_down = (above_range == tap_reverse) != ctrl_at_tap;
            if (last_check) {
                current = is_down ? lower : upper;
                inevitable_run = true;
            } else {
                last_down = is_down;
                adjust(strategy_max);
            }
        }
        void recalibrate(bool strategy_max) {
            if (tap_reverse == (ctrl_at_tap != strategy_max)) {
                lower = current;
                last_down = false;
            } else {
                upper = current;
                last_down = true;
            }
        }
        int8_t repropose_tap(bool strategy_max, bool previous_down, bool& tap_changed) {
            int8_t const tap = search((strategy_max != tap_reverse) != ctrl_at_tap);
            int const diff = int(tap) - int(current);
            if (diff == 0) {
                tap_changed = !inevitable_run;
                inevitable_run = true;
                return tap;
            }
            if ((diff == 1 && previous_down) || (diff == -1 && !previous_down)) last_check = true;
            tap_changed = true;
            current = tap;
            return tap;
        }
    };
    struct Regulated {
        Idx reg, trafo, ctrl_node;
        bool ctrl_from, ctrl_at_tap;
        int8_t tmin, tmax;
        double u_set, u_band;
        cd z;
    };

    Model const& m;
    State& st;
    TopoCache const& tc;
    double err_tol;
    Idx max_iter;
    int method;  // the requested method (0 Newton-Raphson, 2 iterative current, ...)
    int strategy;  // 1 any, 2 min, 3 max, 4 fast any
    std::vector<std::vector<Regulated>> order;
    std::vector<std::vector<Bisection>> bs;
    std::vector<uint64_t> max_range;

    std::vector<MathOutput<P>> calculate(int meth) const {
        std::vector<MathOutput<P>> res;
        for (Idx g = 0; g < Idx(tc.topo.bus_nodes.size()); ++g) {
            MathModel<P> const mm = make_math_model<P>(m, st, tc, g);
            MathInput<P> in;
            for (Idx l : tc.mm_load[g]) in.s_inj.push_back(load_s<P>(m, st, l));
            for (Idx s : tc.mm_source[g]) in.u_source.push_back(st.src_uref[s] * std::exp(cd(0.0, st.src_uang[s])));
            NRSolver<P> solver(mm);
            res.push_back(solver.run(in, err_tol, max_iter, meth));
        }
        return res;
    }
    int8_t& tap(Regulated const& r) const { return st.tr_tap[size_t(r.trafo)]; }
    void set_tap(Regulated const& r, int8_t t) const {
        if (t != kNaS) st.tr_tap[size_t(r.trafo)] = tap_limit(m.tstatic[size_t(r.trafo)], t);
    }
    static int8_t tap_up(int8_t t, Regulated const& r) {
        if (t == r.tmax) return r.tmax;
        return int8_t(r.tmin < r.tmax ? t + 1 : t - 1);
    }
    static int8_t tap_down(int8_t t, Regulated const& r) {
        if (t == r.tmin) return r.tmin;
        return int8_t(r.tmin < r.tmax ? t - 1 : t + 1);
    }
    static int8_t voltage_up(int8_t t, Regulated const& r) { return r.ctrl_at_tap ? tap_up(t, r) : tap_down(t, r); }
    static int8_t voltage_down(int8_t t, Regulated const& r) { return r.ctrl_at_tap ? tap_down(t, r) : tap_up(t, r); }

    // node state against the band (NodeState <=> TransformerTapRegulatorCalcParam, :702-729):
    // -1 below, 0 inside, 1 above, 2 unordered
    int compare(Regulated const& r, std::vector<MathOutput<P>> const& res) const {
        Idx const g = tc.topo.node_group[size_t(r.ctrl_node)];
        Val<P> const u = res[size_t(g)].u[size_t(tc.topo.node_pos[size_t(r.ctrl_node)])];
        Idx const b = m.n_lslot() + r.trafo;
        Val<P> const i = r.ctrl_from ? res[size_t(tc.br_group[size_t(b)])].br_if[size_t(tc.br_pos[size_t(b)])]
                                     : res[size_t(tc.br_group[size_t(b)])].br_it[size_t(tc.br_pos[size_t(b)])];
        double v = 0.0;
        for (int ph = 0; ph < P; ++ph) v += std::abs(u.v[ph] + r.z * i.v[ph]);
        v /= P;
        double const lower = r.u_set - 0.5 * r.u_band, upper = r.u_set + 0.5 * r.u_band;
        auto three_way = [](double a, double b) { return a < b ? -1 : a > b ? 1 : a == b ? 0 : 2; };
        int const lc = three_way(v, lower), uc = three_way(v, upper);
        return lc == uc ? lc : 0;
    }
    bool connected(Regulated const& r) const { return tc.topo.node_group[size_t(r.ctrl_node)] >= 0; }

    bool adjust_scan(Regulated const& r, std::vector<MathOutput<P>> const& res,
                     std::vector<std::pair<Regulated const*, int8_t>>& upd) const {
        if (!connected(r)) return false;
        int const c = compare(r, res);
        int8_t const t = c == 1 ? voltage_down(tap(r), r) : c == -1 ? voltage_up(tap(r), r) : tap(r);
        if (t == tap(r)) return false;
        upd.push_back({&r, t});
        return true;
    }
    bool adjust_bs(Regulated const& r, Bisection& b, std::vector<MathOutput<P>> const& res,
                   std::vector<std::pair<Regulated const*, int8_t>>& upd) const {
        bool const strategy_max = strategy == 3;
        if (!connected(r)) return false;
        if (b.end_of_search() || b.inevitable_run) return false;
        int const c = compare(r, res);
        if (c != 0) b.propose_new_pos(strategy_max, c == 1);
        if (b.current != tap(r)) {
            upd.push_back({&r, b.current});
            return true;
        }
        if (strategy == 4 && c == 0) return false;
        bool const previous_down = b.last_down;
        b.recalibrate(strategy_max);
        bool changed = false;
        int8_t const t = b.repropose_tap(strategy_max, previous_down, changed);
        if (t == tap(r) && c != 0 && !b.end_of_search()) {
            b.reset(t, r.tmin, r.tmax, b.ctrl_at_tap);
            throw MaxIterationReached("TapPositionOptimizer::binary_search: no valid tap position found between tap " +
                                      std::to_string(int(r.tmin)) + " and tap " + std::to_string(int(r.tmax)));
        }
        upd.push_back({&r, t});
        return changed;
    }

    // iterate (:1064-1110)
    std::vector<MathOutput<P>> iterate(int meth, bool binary) {
        auto result = calculate(meth);
        std::vector<uint64_t> per_rank(order.size(), 0);
        bool tap_changed = true;
        while (tap_changed) {
            tap_changed = false;
            std::vector<std::pair<Regulated const*, int8_t>> upd;
            Idx rank = 0;
            for (Idx i = 0; i < Idx(order.size()); ++i) {
                for (Idx j = 0; j < Idx(order[size_t(i)].size()); ++j) {
                    Regulated const& r = order[size_t(i)][size_t(j)];
                    bool const c = binary ? adjust_bs(r, bs[size_t(i)][size_t(j)], result, upd)
                                          : adjust_scan(r, result, upd);
                    tap_changed = c || tap_changed;
                }
                if (tap_changed) {
                    for (size_t k = size_t(rank) + 1; k < per_rank.size(); ++k) per_rank[k] = 0;
                    ++per_rank[size_t(rank)];
                    break;
                }
                ++rank;
            }
            if (tap_changed) {
                if (per_rank[size_t(rank)] > 2 * max_range[size_t(rank)])
                    throw MaxIterationReached("TapPositionOptimizer::iterate " + std::to_string(per_rank[size_t(rank)]) +
                                              " iterations reached: " + std::to_string(max_range[size_t(rank)]) +
                                              "x2 iterations in rank " + std::to_string(rank));
                for (auto const& [r, t] : upd) set_tap(*r, t);
                result = calculate(meth);
            }
        }
        return result;
    }
    // iterate_with_fallback (:1046-1062): a diverging or singular run retries after a linear-method pass
    std::vector<MathOutput<P>> iterate_with_fallback(bool binary) {
        try {
            return iterate(method, binary);
        } catch (IterationDiverge const&) {
        } catch (SparseMatrixError const&) {
        }
        (void)iterate(1, binary);
        return iterate(method, binary);
    }
    // optimize (:965-1028, pilot_run :1257-1298, exploit_neighborhood :1300-1328)
    std::vector<MathOutput<P>> optimize() {
        bool const binary = strategy != 1;
        max_range.clear();
        bs.clear();
        for (auto const& rank : order) {
            uint64_t widest = 0;
            std::vector<Bisection> row;
            for (auto const& r : rank) {
                widest = std::max<uint64_t>(widest, uint64_t(std::abs(int(r.tmax) - int(r.tmin))));
                Bisection b;
                b.reset(tap(r), r.tmin, r.tmax, r.ctrl_at_tap);
                row.push_back(b);
            }
            max_range.push_back(widest);
            bs.push_back(row);
        }
        if (strategy == 2 || strategy == 3) {
            for (auto const& rank : order)
                for (auto const& r : rank) {
                    bool const max_voltage = strategy == 3;
                    set_tap(r, (max_voltage == r.ctrl_at_tap) ? r.tmax : r.tmin);
                }
        }
        if (binary)
            for (size_t i = 0; i < order.size(); ++i)
                for (size_t j = 0; j < order[i].size(); ++j) {
                    bs[i][j].current = tap(order[i][j]);
                    bs[i][j].last_check = false;
                    bs[i][j].inevitable_run = false;
                }
        auto result = iterate_with_fallback(binary);
        if (strategy == 1 || strategy == 4) return result;
        for (auto const& rank : order)
            for (auto const& r : rank) set_tap(r, strategy == 3 ? voltage_up(tap(r), r) : voltage_down(tap(r), r));
        return iterate_with_fallback(false);
    }
};

// ------------------------------------------------------------------------------------------------
// state estimation: symmetric Newton-Raphson with the gain system solved densely
// (math_solver/newton_raphson_se_solver.hpp, measured_values.hpp, observability.hpp)
// ------------------------------------------------------------------------------------------------
struct NotObservableError : std::runtime_error {
    explicit NotObservableError(std::string const& msg)
        : std::runtime_error("Not enough measurements available for state estimation.\n" + msg + "\n") {}
};

constexpr double kInfVar = std::numeric_limits<double>::infinity();

// sensors of the model: the measured object of each, and the values of one scenario
struct SeSensorSet {
    std::vector<int32_t> vs_id, ps_id;
    std::vector<Idx> vs_node, ps_obj;
    std::vector<int> ps_term;
};
struct SeSensorValues {
    std::vector<double> vs_sigma, vs_u, vs_ang;
    std::vector<double> ps_sigma, ps_p, ps_q, ps_psig, ps_qsig;
};

// a real random variable and its inverse-variance combination (statistics.hpp:316-346)
struct RealVar {
    double value = 0.0, var = kInfVar;
};
RealVar combine_real(std::vector<RealVar> const& xs) {
    double inv = 0.0, weighted = 0.0;
    for (auto const& x : xs) {
        inv += 1.0 / x.var;
        weighted += x.value / x.var;
    }
    if (!std::isnormal(inv)) return {weighted, kInfVar};
    return {weighted / inv, 1.0 / inv};
}
// power measurement: real and imaginary parts with their own variances (statistics.hpp:348-366)
struct PowerVar {
    RealVar p, q;
    bool valid() const { return !std::isinf(p.var) && !std::isinf(q.var); }
};
PowerVar combine_power(std::vector<PowerVar> const& xs) {
    std::vector<RealVar> ps, qs;
    for (auto const& x : xs) {
        ps.push_back(x.p);
        qs.push_back(x.q);
    }
    PowerVar r{combine_real(ps), combine_real(qs)};
    if (!(std::isnormal(r.p.var) && std::isnormal(r.q.var))) r.p.var = r.q.var = kInfVar;
    return r;
}
// voltage measurement: a phasor, or a magnitude (angle NaN), with one variance (statistics.hpp:377-396)
struct VoltVar {
    cd value{0.0, 0.0};
    bool has_angle = false;
    double var = kInfVar;
};
VoltVar combine_voltage(std::vector<VoltVar> const& xs) {
    bool all_angle = true;
    for (auto const& x : xs) all_angle = all_angle && x.has_angle;
    VoltVar r;
    if (!all_angle) {
        std::vector<RealVar> mags;
        for (auto const& x : xs) mags.push_back({x.has_angle ? std::abs(x.value) : x.value.real(), x.var});
        RealVar const c = combine_real(mags);
        r.value = c.value;
        r.var = c.var;
        r.has_angle = false;
        return r;
    }
    double inv = 0.0;
    cd weighted{0.0, 0.0};
    for (auto const& x : xs) {
        inv += 1.0 / x.var;
        weighted += x.value / x.var;
    }
    r.has_angle = true;
    if (!std::isnormal(inv)) {
        r.value = weighted;
        return r;
    }
    r.value = weighted / inv;
    r.var = 1.0 / inv;
    return r;
}

// calculation parameters of one sensor in per unit (voltage_sensor.hpp sym_calc_param, power_sensor.hpp)
VoltVar voltage_calc(Model const& m, SeSensorSet const& S, SeSensorValues const& v, Idx i) {
    double const ur = m.node[S.vs_node[i]].u_rated;
    double const sigma = v.vs_sigma[i] / ur;
    VoltVar r;
    r.var = sigma * sigma;
    double const mag = v.vs_u[i] / ur;
    if (!std::isnan(v.vs_ang[i])) {
        r.value = std::polar(mag, v.vs_ang[i]);
        r.has_angle = true;
    } else {
        r.value = cd(mag, 0.0);
        r.has_angle = false;
    }
    return r;
}
double sensor_direction(int term) { return (term == 4 || term == 3) ? -1.0 : 1.0; }  // load, shunt
PowerVar power_calc(SeSensorValues const& v, SeSensorSet const& S, Idx i) {
    double const dir = sensor_direction(S.ps_term[i]);
    double const sp = v.ps_psig[i] / kBaseP3, sq = v.ps_qsig[i] / kBaseP3, ss = v.ps_sigma[i] / kBaseP3;
    double vp, vq;
    if (std::isnormal(sp) && std::isnormal(sq)) {
        vp = sp * sp;
        vq = sq * sq;
    } else {
        vp = vq = std::isnan(sp) ? ss * ss / 2.0 : kInfVar;
    }
    return PowerVar{{dir * v.ps_p[i] / kBaseP3, vp}, {dir * v.ps_q[i] / kBaseP3, vq}};
}

// the measurements of one math model, aggregated
struct SeMeasured {
    std::vector<VoltVar> volt;           // per bus (var inf: none)
    std::vector<PowerVar> inj;           // per bus, valid() when the bus has an injection measurement
    std::vector<bool> has_inj;
    std::vector<PowerVar> br_from, br_to;  // per math branch
    std::vector<bool> has_from, has_to;
    std::vector<PowerVar> shunt;         // per math shunt
    std::vector<bool> has_shunt;
    // appliances for the results: 0 disconnected, 1 unmeasured, 2 measured; unnormalised values
    std::vector<int> lg_state, src_state;
    std::vector<PowerVar> lg_val, src_val;
    std::vector<PowerVar> bus_appl;
    std::vector<int> bus_unmeasured;
    int n_angle = 0, first_volt = -1;
    double mean_shift = 0.0;
};

SeMeasured measure(Model const& m, State const& st, TopoCache const& tc, MathModel<1> const& mm, Idx g,
                   SeSensorSet const& S, SeSensorValues const& val) {
    Idx const n = mm.n_bus;
    SeMeasured r;
    Idx const nbr = Idx(mm.branch_bus.size()), nsh = Idx(mm.shunt_bus.size());
    Idx const nlg = Idx(tc.mm_load[g].size()), nsrc = Idx(tc.mm_source[g].size());
    // sensors per object of this math model, in sensor order
    std::vector<std::vector<Idx>> vs_at(n), inj_at(n), bf_at(nbr), bt_at(nbr), sh_at(nsh), lg_at(nlg), src_at(nsrc);
    for (Idx i = 0; i < Idx(S.vs_node.size()); ++i) {
        Idx const node = S.vs_node[i];
        if (tc.topo.node_group[node] == g) vs_at[tc.topo.node_pos[node]].push_back(i);
    }
    for (Idx i = 0; i < Idx(S.ps_term.size()); ++i) {
        Idx const o = S.ps_obj[i];
        switch (S.ps_term[i]) {
        case 0:
        case 1:
            if (tc.br_group[o] == g) (S.ps_term[i] == 0 ? bf_at : bt_at)[tc.br_pos[o]].push_back(i);
            break;
        case 2:
            if (tc.src_group[o] == g) src_at[tc.src_pos[o]].push_back(i);
            break;
        case 3:
            if (tc.sh_group[o] == g) sh_at[tc.sh_pos[o]].push_back(i);
            break;
        case 4:
        case 5:
            if (tc.lg_group[o] == g) lg_at[tc.lg_pos[o]].push_back(i);
            break;
        case 9:
            if (tc.topo.node_group[o] == g) inj_at[tc.topo.node_pos[o]].push_back(i);
            break;
        default:
            break;
        }
    }
    // voltages (measured_values.hpp:240-299)
    r.volt.assign(n, VoltVar{});
    double angle_sum = 0.0;
    for (Idx b = 0; b < n; ++b) {
        if (vs_at[b].empty()) continue;
        std::vector<VoltVar> xs;
        for (Idx i : vs_at[b]) xs.push_back(voltage_calc(m, S, val, i));
        VoltVar const c = combine_voltage(xs);
        if (std::isinf(c.var)) continue;
        r.volt[b] = c;
        if (r.first_volt < 0) r.first_volt = int(b);
        if (c.has_angle) {
            ++r.n_angle;
            angle_sum += std::arg(c.value * std::exp(cd(0.0, -mm.phase_shift[b])));
        }
    }
    r.mean_shift = r.n_angle > 0 ? angle_sum / double(r.n_angle) : 0.0;
    // one object: disconnected, unmeasured or measured (measured_values.hpp:510-540)
    auto object = [&](bool connected, std::vector<Idx> const& sensors, PowerVar& out) {
        if (!connected) return 0;
        if (sensors.empty()) return 1;
        std::vector<PowerVar> xs;
        for (Idx i : sensors) xs.push_back(power_calc(val, S, i));
        PowerVar const c = combine_power(xs);
        if (!c.valid()) return 1;
        out = c;
        return 2;
    };
    r.shunt.assign(nsh, PowerVar{});
    r.has_shunt.assign(nsh, false);
    r.lg_state.assign(nlg, 0);
    r.src_state.assign(nsrc, 0);
    r.lg_val.assign(nlg, PowerVar{});
    r.src_val.assign(nsrc, PowerVar{});
    r.inj.assign(n, PowerVar{});
    r.has_inj.assign(n, false);
    r.bus_appl.assign(n, PowerVar{{0.0, 0.0}, {0.0, 0.0}});
    r.bus_unmeasured.assign(n, 0);
    for (Idx k = 0; k < nsh; ++k)
        r.has_shunt[k] = object(st.sh_status[tc.mm_shunt[g][k]] != 0, sh_at[k], r.shunt[k]) == 2;
    for (Idx k = 0; k < nlg; ++k) r.lg_state[k] = object(st.lg_status[tc.mm_load[g][k]] != 0, lg_at[k], r.lg_val[k]);
    for (Idx k = 0; k < nsrc; ++k)
        r.src_state[k] = object(st.src_status[tc.mm_source[g][k]] != 0, src_at[k], r.src_val[k]);
    // injections (measured_values.hpp:318-378)
    auto add = [&](Idx bus, int state, PowerVar const& x) {
        if (state == 1) ++r.bus_unmeasured[bus];
        if (state != 2) return;
        PowerVar& a = r.bus_appl[bus];
        a.p.value += x.p.value;
        a.q.value += x.q.value;
        a.p.var += x.p.var;
        a.q.var += x.q.var;
    };
    for (Idx k = 0; k < nlg; ++k) add(mm.load_bus[k], r.lg_state[k], r.lg_val[k]);
    for (Idx k = 0; k < nsrc; ++k) add(mm.source_bus[k], r.src_state[k], r.src_val[k]);
    for (Idx b = 0; b < n; ++b) {
        std::vector<PowerVar> xs;
        for (Idx i : inj_at[b]) xs.push_back(power_calc(val, S, i));
        PowerVar const direct = combine_power(xs);
        PowerVar const& appl = r.bus_appl[b];
        if (r.bus_unmeasured[b] > 0) {
            if (direct.valid()) {
                r.inj[b] = direct;
                r.has_inj[b] = true;
            }
        } else if (!direct.valid() || appl.p.var == 0.0 || appl.q.var == 0.0) {
            r.inj[b] = appl;
            r.has_inj[b] = true;
        } else {
            r.inj[b] = combine_power({direct, appl});
            r.has_inj[b] = true;
        }
    }
    // branch sides (measured_values.hpp:380-445)
    r.br_from.assign(nbr, PowerVar{});
    r.br_to.assign(nbr, PowerVar{});
    r.has_from.assign(nbr, false);
    r.has_to.assign(nbr, false);
    for (Idx k = 0; k < nbr; ++k) {
        r.has_from[k] = object(mm.branch_bus[k][0] != -1, bf_at[k], r.br_from[k]) == 2;
        r.has_to[k] = object(mm.branch_bus[k][1] != -1, bt_at[k], r.br_to[k]) == 2;
    }
    // normalize_variance (measured_values.hpp:545-600)
    double min_var = kInfVar;
    auto consider = [&](double v) {
        if (v != 0.0) min_var = std::min(min_var, v);
    };
    for (auto const& v : r.volt)
        if (!std::isinf(v.var)) consider(v.var);
    for (Idx k = 0; k < nsh; ++k)
        if (r.has_shunt[k]) consider(r.shunt[k].p.var + r.shunt[k].q.var);
    for (Idx b = 0; b < n; ++b)
        if (r.has_inj[b]) consider(r.inj[b].p.var + r.inj[b].q.var);
    for (Idx k = 0; k < nbr; ++k) {
        if (r.has_from[k]) consider(r.br_from[k].p.var + r.br_from[k].q.var);
        if (r.has_to[k]) consider(r.br_to[k].p.var + r.br_to[k].q.var);
    }
    double const scale = 1.0 / min_var;
    auto norm = [&](PowerVar& x) {
        x.p.var *= scale;
        x.q.var *= scale;
    };
    for (auto& v : r.volt)
        if (!std::isinf(v.var)) v.var *= scale;
    for (Idx k = 0; k < nsh; ++k)
        if (r.has_shunt[k]) norm(r.shunt[k]);
    for (Idx b = 0; b < n; ++b)
        if (r.has_inj[b]) norm(r.inj[b]);
    for (Idx k = 0; k < nbr; ++k) {
        if (r.has_from[k]) norm(r.br_from[k]);
        if (r.has_to[k]) norm(r.br_to[k]);
    }
    return r;
}

// observability (observability.hpp:60-140, 644-698): throws NotObservableError; the meshed sufficient condition
// is taken as met. Returns whether the gain system may be ill-conditioned (a bus other than the last with no
// injection, flow or voltage phasor sensor), which turns on the pivot perturbation
bool check_observable(MathModel<1> const& mm, YStruct const& ys, SeMeasured const& me, bool radial) {
    Idx const n = mm.n_bus;
    if (me.first_volt < 0) throw NotObservableError("No voltage sensor found!\n");
    Idx const nnz = ys.row_indptr[n];
    std::vector<int8_t> flow(nnz, 0), vph(n, 0);
    Idx total_inj = 0;
    bool ill = false;
    for (Idx b = 0; b < n; ++b) {
        Idx const d = ys.bus_entry[b];
        if (me.has_inj[b]) {
            flow[d] = 1;
            ++total_inj;
        }
        for (Idx e = d + 1; e < ys.row_indptr[b + 1]; ++e) {
            for (Idx q = ys.y_bus_entry_indptr[e]; q < ys.y_bus_entry_indptr[e + 1]; ++q) {
                auto const& el = ys.y_bus_element[q];
                if (el.type == 4) continue;  // shunt
                Idx const br = el.idx;
                bool const connected = mm.branch_bus[br][0] != -1 && mm.branch_bus[br][1] != -1;
                if ((me.has_from[br] || me.has_to[br]) && connected) {
                    flow[e] = 1;
                    break;
                }
            }
        }
        if (!std::isinf(me.volt[b].var) && me.volt[b].has_angle) vph[b] = 1;
        if (!flow[d] && vph[b] == 0 && b != n - 1) {
            bool any = false;
            for (Idx e = d + 1; e < ys.row_indptr[b + 1] && !any; ++e) any = flow[e] != 0;
            ill = ill || !any;
        }
    }
    Idx n_flow = 0, n_vph = 0;
    for (auto f : flow) n_flow += f;
    for (auto v : vph) n_vph += v;
    if (n_vph == 0 && n_flow < n - 1)
        throw NotObservableError(
            "The total number of independent power sensors is not enough to make the grid observable.");
    if (n_vph > 0 && n_flow + n_vph < n)
        throw NotObservableError("The total number of independent power sensors and voltage phasor sensors is not "
                                 "enough to make the grid observable.");
    if ((n_vph > 1 && !radial) || total_inj > n - 2 || !radial) return ill;
    for (Idx b = 0; b + 1 < n; ++b) {
        Idx const cur = ys.bus_entry[b], up = cur + 1;
        Idx const up_bus = ys.col_indices[up];
        Idx const up_entry = ys.bus_entry[up_bus];
        if (flow[up] == 0) {
            if (flow[cur] == 1)
                std::swap(flow[up], flow[cur]);
            else if (flow[up_entry] == 1)
                std::swap(flow[up], flow[up_entry]);
            else if (vph[b] == 1)
                std::swap(flow[up], vph[b]);
            else if (vph[b + 1] == 1)
                std::swap(flow[up], vph[b + 1]);
        }
        flow[cur] = 0;
    }
    flow[ys.bus_entry[n - 1]] = 0;
    Idx n_ind = 0, n_rem = 0;
    for (auto f : flow) n_ind += f;
    for (auto v : vph) n_rem += v;
    if (n_ind < n - 1 || (n_vph > 0 && n_rem < 1))
        throw NotObservableError("The number of power, current, and voltage phasor sensors appears sufficient, but "
                                 "they are not independent enough. The system is still not observable.\n");
    return ill;
}

// dense LU with partial pivoting: solves a x = b in place (b becomes x); throws SparseMatrixError when a pivot
// is exactly zero
void dense_solve(std::vector<double>& a, std::vector<double>& b, Idx n) {
    std::vector<Idx> piv(n);
    for (Idx k = 0; k < n; ++k) {
        Idx best = k;
        for (Idx r = k + 1; r < n; ++r)
            if (std::abs(a[r * n + k]) > std::abs(a[best * n + k])) best = r;
        if (a[best * n + k] == 0.0) throw SparseMatrixError();
        if (best != k) {
            for (Idx c = 0; c < n; ++c) std::swap(a[k * n + c], a[best * n + c]);
            std::swap(b[k], b[best]);
        }
        double const d = a[k * n + k];
        for (Idx r = k + 1; r < n; ++r) {
            double const f = a[r * n + k] / d;
            if (f == 0.0) continue;
            for (Idx c = k; c < n; ++c) a[r * n + c] -= f * a[k * n + c];
            b[r] -= f * b[k];
        }
    }
    for (Idx k = n - 1; k >= 0; --k) {
        double s = b[k];
        for (Idx c = k + 1; c < n; ++c) s -= a[k * n + c] * b[c];
        b[k] = s / a[k * n + k];
    }
}

// the gain system of one Newton-Raphson state-estimation iteration, 4 unknowns per bus (theta, V, phi_p, phi_q):
// a dense matrix solved with partial pivoting, or 4x4 blocks on the LU pattern of the Y-bus solved with the
// block sparse LU, its pivot perturbation and iterative refinement (sparse_lu_solver.hpp:297-318, 514-650)
// sparse gain factorisations that perturbed a pivot, and the refinement solves they took (for the tests)
std::atomic<long long> g_se_perturbed{0}, g_se_refine_solves{0};

struct DenseGain {
    Idx n, N;
    std::vector<double> A;
    explicit DenseGain(Idx n_bus) : n(n_bus), N(4 * n_bus), A(size_t(N * N), 0.0) {}
    double& ref(Idx i, int r, Idx j, int c) { return A[size_t((4 * i + r) * N + 4 * j + c)]; }
    template <class F> void for_each_entry(F&& f) const {
        for (Idx i = 0; i < n; ++i)
            for (Idx j = 0; j < n; ++j) f(i, j);
    }
    void solve(std::vector<double>& rhs, bool) { dense_solve(A, rhs, N); }
};
struct SparseGain {
    YStruct const& ys;
    Idx n;
    std::vector<Blk<4>> lu;
    SparseGain(YStruct const& y, Idx n_bus) : ys(y), n(n_bus), lu(size_t(y.row_indptr_lu[n_bus]), Blk<4>{}) {}
    Idx entry(Idx i, Idx j) const {
        auto const b = ys.col_indices_lu.begin() + ys.row_indptr_lu[i];
        auto const e = ys.col_indices_lu.begin() + ys.row_indptr_lu[i + 1];
        auto const it = std::lower_bound(b, e, j);
        if (it == e || *it != j) throw OtherError("state estimation: gain entry outside the LU pattern");
        return Idx(it - ys.col_indices_lu.begin());
    }
    double& ref(Idx i, int r, Idx j, int c) { return at<4>(lu[size_t(entry(i, j))], r, c); }
    template <class F> void for_each_entry(F&& f) const {
        for (Idx i = 0; i < n; ++i)
            for (Idx k = ys.row_indptr_lu[i]; k < ys.row_indptr_lu[i + 1]; ++k) f(i, ys.col_indices_lu[k]);
    }
    // rhs becomes the solution
    void solve(std::vector<double>& rhs, bool use_pert) {
        Idx const nnz = Idx(lu.size());
        std::vector<Vec<4>> b(static_cast<size_t>(n)), x(static_cast<size_t>(n));
        for (Idx i = 0; i < n; ++i)
            for (int r = 0; r < 4; ++r) b[size_t(i)][size_t(r)] = rhs[size_t(4 * i + r)];
        SparseLU<4> solver(ys.row_indptr_lu, ys.col_indices_lu, ys.diag_lu);
        std::vector<Blk<4>> original;
        if (use_pert) {
            // the block-wise off-diagonal infinity norm sets the threshold (:624-650)
            original = lu;
            double norm = 0.0;
            for (Idx i = 0; i < n; ++i) {
                double row = 0.0;
                for (Idx k = ys.row_indptr_lu[i]; k < ys.row_indptr_lu[i + 1]; ++k) {
                    if (ys.col_indices_lu[k] == i) continue;
                    double bn = 0.0;
                    for (int r = 0; r < 4; ++r) {
                        double sr = 0.0;
                        for (int c = 0; c < 4; ++c) sr += std::abs(at<4>(lu[size_t(k)], r, c));
                        bn = std::max(bn, sr);
                    }
                    row += bn;
                }
                norm = std::max(norm, row);
            }
            solver.use_pert = true;
            solver.pert_thr = 1e-13 * norm;
        }
        solver.factorize(lu);
        if (solver.has_pert) ++g_se_perturbed;
        if (!solver.has_pert) {
            solver.solve(lu, b, x);
        } else {
            // iterative refinement against the unperturbed matrix: x from 0, at most six solves, until the
            // componentwise backward error (denominators capped below at 1e-4 of the largest) is at most 1e-13
            for (auto& v : x) v.fill(0.0);
            std::vector<Vec<4>> res = b, dx(static_cast<size_t>(n));
            double berr = std::numeric_limits<double>::max();
            Idx it = 0;
            while (berr > 1e-13) {
                if (it++ == 6) throw SparseMatrixError{};
                solver.solve(lu, res, dx);
                ++g_se_refine_solves;
                std::vector<Vec<4>> den(static_cast<size_t>(n));
                double max_den = 0.0;
                for (Idx i = 0; i < n; ++i) {
                    for (int r = 0; r < 4; ++r) den[size_t(i)][size_t(r)] = std::abs(b[size_t(i)][size_t(r)]);
                    for (Idx k = ys.row_indptr_lu[i]; k < ys.row_indptr_lu[i + 1]; ++k) {
                        Idx const j = ys.col_indices_lu[k];
                        for (int r = 0; r < 4; ++r)
                            for (int c = 0; c < 4; ++c)
                                den[size_t(i)][size_t(r)] +=
                                    std::abs(at<4>(original[size_t(k)], r, c)) * std::abs(x[size_t(j)][size_t(c)]);
                    }
                    for (int r = 0; r < 4; ++r) max_den = std::max(max_den, den[size_t(i)][size_t(r)]);
                }
                double const cap = 1e-4 * max_den;
                berr = 0.0;
                for (Idx i = 0; i < n; ++i)
                    for (int r = 0; r < 4; ++r) {
                        double const d = std::max(den[size_t(i)][size_t(r)], cap);
                        berr = std::max(berr, std::abs(res[size_t(i)][size_t(r)]) / d);
                        x[size_t(i)][size_t(r)] += dx[size_t(i)][size_t(r)];
                    }
                for (Idx i = 0; i < n; ++i) {
                    res[size_t(i)] = b[size_t(i)];
                    for (Idx k = ys.row_indptr_lu[i]; k < ys.row_indptr_lu[i + 1]; ++k) {
                        Idx const j = ys.col_indices_lu[k];
                        for (int r = 0; r < 4; ++r)
                            for (int c = 0; c < 4; ++c)
                                res[size_t(i)][size_t(r)] -= at<4>(original[size_t(k)], r, c) * x[size_t(j)][size_t(c)];
                    }
                }
            }
        }
        (void)nnz;
        for (Idx i = 0; i < n; ++i)
            for (int r = 0; r < 4; ++r) rhs[size_t(4 * i + r)] = x[size_t(i)][size_t(r)];
    }
};

// the dense gain system instead of the sparse one (a second restatement, to check the first)
bool g_se_dense = false;

// Newton-Raphson state estimation of one math model; returns the iteration count, u in per unit
// (newton_raphson_se_solver.hpp:170-214)
template <class Gain>
Idx se_solve_on(MathModel<1> const& mm, YStruct const& ys, SeMeasured const& me, Idx slack, bool use_pert,
                double err_tol, Idx max_iter, std::vector<cd>& u) {
    (void)ys;
    Idx const n = mm.n_bus;
    // dense Y-bus: branches then shunts (y_bus.hpp)
    std::vector<cd> Y(n * n, cd{0.0, 0.0});
    std::vector<std::array<cd, 4>> by(mm.branch_bus.size());
    for (size_t b = 0; b < mm.branch_bus.size(); ++b) {
        for (int w = 0; w < 4; ++w) by[b][w] = mm.branch_param[b].v[w].m[0];
        Idx const f = mm.branch_bus[b][0], t = mm.branch_bus[b][1];
        if (f >= 0) Y[f * n + f] += by[b][0];
        if (f >= 0 && t >= 0) {
            Y[f * n + t] += by[b][1];
            Y[t * n + f] += by[b][2];
        }
        if (t >= 0) Y[t * n + t] += by[b][3];
    }
    for (size_t k = 0; k < mm.shunt_bus.size(); ++k) Y[mm.shunt_bus[k] * n + mm.shunt_bus[k]] += mm.shunt_param[k].m[0];
    Idx const vref = !std::isinf(me.volt[slack].var) ? slack : me.first_volt;
    bool const has_angle = me.n_angle > 0;
    // state: theta, V, phi_p, phi_q per bus (initialize_unknown :219-236)
    std::vector<double> th(n), v(n, 1.0), php(n, 0.0), phq(n, 0.0);
    u.assign(n, cd{});
    for (Idx b = 0; b < n; ++b) {
        th[b] = me.mean_shift + mm.phase_shift[b];
        if (!std::isinf(me.volt[b].var)) {
            if (me.volt[b].has_angle) th[b] = std::arg(me.volt[b].value);
            v[b] = me.volt[b].has_angle ? std::abs(me.volt[b].value) : me.volt[b].value.real();
        }
        u[b] = std::polar(v[b], th[b]);
    }
    Idx const N = 4 * n;
    // measurement Jacobian of one bus: dP/dtheta, dP/dV, dQ/dtheta, dQ/dV
    struct J {
        double pt, pv, qt, qv;
    };
    auto jac = [](cd hm, cd nl) { return J{hm.imag(), nl.real(), -hm.real(), nl.imag()}; };
    auto jac_diag = [](cd fv, cd f) { return J{-f.imag(), fv.real(), f.real(), fv.imag()}; };
    auto plus = [](J a, J b) { return J{a.pt + b.pt, a.pv + b.pv, a.qt + b.qt, a.qv + b.qv}; };
    double max_dev = std::numeric_limits<double>::max();
    Idx num_iter = 0;
    while (max_dev > err_tol || num_iter == 0) {
        if (num_iter++ == max_iter)
            throw IterationDiverge("Iteration failed to converge after " + std::to_string(max_iter) +
                                   " iterations! Max deviation: " + fmt_double(max_dev) +
                                   ", error tolerance: " + fmt_double(err_tol) + ".\n");
        Gain G = [&] {
            if constexpr (std::is_same_v<Gain, DenseGain>)
                return DenseGain(n);
            else
                return SparseGain(ys, n);
        }();
        std::vector<double> rhs(N, 0.0);
        auto a = [&](Idx i, int r, Idx j, int c) -> double& { return G.ref(i, r, j, c); };
        // weighted measurement term: G(x, y) += H_x^T W H_y, eta(x) += H_x^T W (z - f)
        auto wls = [&](std::vector<std::pair<Idx, J>> const& hs, cd f, PowerVar const& z) {
            double const wp = 1.0 / z.p.var, wq = 1.0 / z.q.var;
            double const dp = z.p.value - f.real(), dq = z.q.value - f.imag();
            for (auto const& [x, hx] : hs) {
                for (auto const& [y, hy] : hs) {
                    a(x, 0, y, 0) += wp * hx.pt * hy.pt + wq * hx.qt * hy.qt;
                    a(x, 0, y, 1) += wp * hx.pt * hy.pv + wq * hx.qt * hy.qv;
                    a(x, 1, y, 0) += wp * hx.pv * hy.pt + wq * hx.qv * hy.qt;
                    a(x, 1, y, 1) += wp * hx.pv * hy.pv + wq * hx.qv * hy.qv;
                }
                rhs[4 * x] += wp * hx.pt * dp + wq * hx.qt * dq;
                rhs[4 * x + 1] += wp * hx.pv * dp + wq * hx.qv * dq;
            }
        };
        // voltages (:647-679)
        for (Idx b = 0; b < n; ++b) {
            if (std::isinf(me.volt[b].var)) continue;
            double const w = 1.0 / me.volt[b].var;
            double const mag = me.volt[b].has_angle ? std::abs(me.volt[b].value) : me.volt[b].value.real();
            a(b, 1, b, 1) += w;
            rhs[4 * b + 1] += w * (mag - v[b]);
            if (me.volt[b].has_angle) {
                a(b, 0, b, 0) += w;
                rhs[4 * b] += w * (std::arg(me.volt[b].value) - th[b]);
            } else if (b == vref && !has_angle) {
                a(b, 0, b, 0) += w;
                rhs[4 * b] += w * (0.0 - th[b]);
            }
        }
        // branch power at one side: the measured side's bus chi, the other psi (:473-523)
        for (size_t b = 0; b < mm.branch_bus.size(); ++b) {
            Idx const f = mm.branch_bus[b][0], t = mm.branch_bus[b][1];
            if (f < 0 || t < 0 || f == t) continue;
            for (int side = 0; side < 2; ++side) {
                if (!(side == 0 ? me.has_from[b] : me.has_to[b])) continue;
                Idx const chi = side == 0 ? f : t, psi = side == 0 ? t : f;
                cd const yxx = side == 0 ? by[b][0] : by[b][3], yxm = side == 0 ? by[b][1] : by[b][2];
                cd const hm1 = std::conj(yxx) * u[chi] * std::conj(u[chi]);
                cd const hm2 = std::conj(yxm) * u[chi] * std::conj(u[psi]);
                cd const fx = hm1 + hm2;
                J const h_chi = plus(jac(hm1, hm1 / v[chi]), jac_diag(fx / v[chi], fx));
                J const h_psi = jac(hm2, hm2 / v[psi]);
                wls({{chi, h_chi}, {psi, h_psi}}, fx, side == 0 ? me.br_from[b] : me.br_to[b]);
            }
        }
        // shunt power (:450-471)
        for (size_t k = 0; k < mm.shunt_bus.size(); ++k) {
            if (!me.has_shunt[k]) continue;
            Idx const b = mm.shunt_bus[k];
            cd const hm = std::conj(-mm.shunt_param[k].m[0]) * u[b] * std::conj(u[b]);
            J const h = plus(jac(hm, hm / v[b]), jac_diag(hm / v[b], hm));
            wls({{b, h}}, hm, me.shunt[k]);
        }
        // Y-bus rows without structural zeros
        std::vector<std::vector<std::pair<Idx, cd>>> yrow(static_cast<size_t>(n));
        for (Idx i = 0; i < n; ++i)
            for (Idx j = 0; j < n; ++j)
                if (Y[size_t(i * n + j)] != cd{0.0, 0.0}) yrow[size_t(i)].emplace_back(j, Y[size_t(i * n + j)]);
        // injection constraints with Lagrange multipliers (:397-448, 583-617); rows 2, 3 of bus i: Q, R
        for (Idx i = 0; i < n; ++i) {
            if (!me.has_inj[i]) {
                a(i, 2, i, 2) = -1.0;
                a(i, 3, i, 3) = -1.0;
                continue;
            }
            cd f_inj{0.0, 0.0};
            J diag{0, 0, 0, 0};
            for (auto const& [j, y] : yrow[size_t(i)]) {
                cd const hm = std::conj(y) * u[i] * std::conj(u[j]);
                J const q = jac(hm, hm / v[j]);
                a(i, 2, j, 0) += q.pt;
                a(i, 2, j, 1) += q.pv;
                a(i, 3, j, 0) += q.qt;
                a(i, 3, j, 1) += q.qv;
                J const qd = jac_diag(hm / v[i], hm);
                diag = plus(diag, qd);
                f_inj += hm;
            }
            a(i, 2, i, 0) += diag.pt;
            a(i, 2, i, 1) += diag.pv;
            a(i, 3, i, 0) += diag.qt;
            a(i, 3, i, 1) += diag.qv;
            rhs[4 * i + 2] += me.inj[i].p.value - f_inj.real();
            rhs[4 * i + 3] += me.inj[i].q.value - f_inj.imag();
            a(i, 2, i, 2) = -me.inj[i].p.var;
            a(i, 3, i, 3) = -me.inj[i].q.var;
        }
        // Q^T and the Lagrange terms
        G.for_each_entry([&](Idx i, Idx j) {
            a(i, 0, j, 2) = a(j, 2, i, 0);
            a(i, 0, j, 3) = a(j, 3, i, 0);
            a(i, 1, j, 2) = a(j, 2, i, 1);
            a(i, 1, j, 3) = a(j, 3, i, 1);
            rhs[4 * i] += a(i, 0, j, 2) * php[j] + a(i, 0, j, 3) * phq[j];
            rhs[4 * i + 1] += a(i, 1, j, 2) * php[j] + a(i, 1, j, 3) * phq[j];
        });
        G.solve(rhs, use_pert);
        // iterate_unknown (:777-821)
        double const offset = has_angle ? 0.0 : th[slack] + rhs[4 * slack];
        max_dev = 0.0;
        for (Idx b = 0; b < n; ++b) {
            th[b] += rhs[4 * b] - offset;
            v[b] += rhs[4 * b + 1];
            if (me.has_inj[b]) {
                if (me.inj[b].p.var == 0.0) php[b] += rhs[4 * b + 2];
                if (me.inj[b].q.var == 0.0) phq[b] += rhs[4 * b + 3];
            }
            cd const nu = std::polar(v[b], th[b]);
            max_dev = std::max(max_dev, std::abs(nu - u[b]));
            u[b] = nu;
        }
    }
    return num_iter;
}

Idx se_solve(MathModel<1> const& mm, YStruct const& ys, SeMeasured const& me, Idx slack, bool use_pert,
             double err_tol, Idx max_iter, std::vector<cd>& u) {
    return g_se_dense ? se_solve_on<DenseGain>(mm, ys, me, slack, use_pert, err_tol, max_iter, u)
                      : se_solve_on<SparseGain>(mm, ys, me, slack, use_pert, err_tol, max_iter, u);
}

// one scenario: every math model estimated, outputs written (main_core/output.hpp:189-280)
Idx se_scenario(Model const& m, State const& st, TopoCache const& tc, SeSensorSet const& S,
                SeSensorValues const& val, double err_tol, Idx max_iter, OrMutableDataset const& out,
                OrVoltageSensorOutputSym* vs_out, OrPowerSensorOutputSym* ps_out, Idx scn) {
    Idx const nmm = Idx(tc.topo.bus_nodes.size());
    std::vector<MathOutput<1>> res(nmm);
    std::vector<SeMeasured> meas(nmm);
    std::vector<MathModel<1>> models(nmm);
    std::vector<std::vector<cd>> bus_s(nmm);
    Idx iters = 0;
    for (Idx g = 0; g < nmm; ++g) {
        MathModel<1>& mm = models[g];
        mm = make_math_model<1>(m, st, tc, g);
        Idx const n = mm.n_bus;
        SeMeasured& me = meas[g];
        me = measure(m, st, tc, mm, g, S, val);
        std::vector<std::vector<Idx>> sh_per_bus(n);
        for (size_t k = 0; k < mm.shunt_bus.size(); ++k) sh_per_bus[mm.shunt_bus[k]].push_back(Idx(k));
        YStruct const ys = build_ystruct(n, mm.branch_bus, sh_per_bus, mm.fill_in);
        bool const use_pert = check_observable(mm, ys, me, tc.topo.radial[g]);
        Idx const slack = tc.topo.slack[g];
        std::vector<cd> u;
        iters = std::max(iters, se_solve(mm, ys, me, slack, use_pert, err_tol, max_iter, u));
        // results (common_solver_functions.hpp:449-456, y_bus.hpp:482-546, measured_values.hpp:124-150)
        MathOutput<1>& r = res[g];
        r.u.resize(n);
        for (Idx b = 0; b < n; ++b) r.u[b].v[0] = u[b];
        std::vector<cd> s(n, cd{});
        {
            std::vector<cd> Y(n * n, cd{});
            for (size_t b = 0; b < mm.branch_bus.size(); ++b) {
                Idx const f = mm.branch_bus[b][0], t = mm.branch_bus[b][1];
                auto const& bp = mm.branch_param[b];
                if (f >= 0) Y[f * n + f] += bp.v[0].m[0];
                if (f >= 0 && t >= 0) {
                    Y[f * n + t] += bp.v[1].m[0];
                    Y[t * n + f] += bp.v[2].m[0];
                }
                if (t >= 0) Y[t * n + t] += bp.v[3].m[0];
            }
            for (size_t k = 0; k < mm.shunt_bus.size(); ++k)
                Y[mm.shunt_bus[k] * n + mm.shunt_bus[k]] += mm.shunt_param[k].m[0];
            for (Idx i = 0; i < n; ++i) {
                cd cur{0.0, 0.0};
                for (Idx j = 0; j < n; ++j) cur += Y[i * n + j] * u[j];
                s[i] = std::conj(cur) * u[i];
            }
        }
        bus_s[g] = s;
        r.bus_injection.resize(n);
        for (Idx b = 0; b < n; ++b) r.bus_injection[b].v[0] = s[b];
        Idx const nbr = Idx(mm.branch_bus.size());
        r.br_sf.resize(nbr);
        r.br_st.resize(nbr);
        r.br_if.resize(nbr);
        r.br_it.resize(nbr);
        for (Idx b = 0; b < nbr; ++b) {
            Idx const f = mm.branch_bus[b][0], t = mm.branch_bus[b][1];
            cd const uf = f >= 0 ? u[f] : cd{}, ut = t >= 0 ? u[t] : cd{};
            auto const& bp = mm.branch_param[b];
            cd const i_f = bp.v[0].m[0] * uf + bp.v[1].m[0] * ut, i_t = bp.v[2].m[0] * uf + bp.v[3].m[0] * ut;
            r.br_if[b].v[0] = i_f;
            r.br_it[b].v[0] = i_t;
            r.br_sf[b].v[0] = uf * std::conj(i_f);
            r.br_st[b].v[0] = ut * std::conj(i_t);
        }
        Idx const nsh = Idx(mm.shunt_bus.size());
        r.sh_s.resize(nsh);
        r.sh_i.resize(nsh);
        for (Idx k = 0; k < nsh; ++k) {
            cd const ub = u[mm.shunt_bus[k]];
            cd const i = -mm.shunt_param[k].m[0] * ub;
            r.sh_i[k].v[0] = i;
            r.sh_s[k].v[0] = ub * std::conj(i);
        }
        Idx const nlg = Idx(mm.load_bus.size()), nsrc = Idx(mm.source_bus.size());
        r.ld_s.assign(nlg, Val<1>{});
        r.ld_i.assign(nlg, Val<1>{});
        r.src_s.assign(nsrc, Val<1>{});
        r.src_i.assign(nsrc, Val<1>{});
        for (Idx b = 0; b < n; ++b) {
            PowerVar const& ap = me.bus_appl[b];
            if (me.bus_unmeasured[b] > 0) {
                cd const each = (s[b] - cd(ap.p.value, ap.q.value)) / double(me.bus_unmeasured[b]);
                for (Idx k = 0; k < nlg; ++k)
                    if (mm.load_bus[k] == b)
                        r.ld_s[k].v[0] = me.lg_state[k] == 2 ? cd(me.lg_val[k].p.value, me.lg_val[k].q.value)
                                         : me.lg_state[k] == 1 ? each
                                                               : cd{};
                for (Idx k = 0; k < nsrc; ++k)
                    if (mm.source_bus[k] == b)
                        r.src_s[k].v[0] = me.src_state[k] == 2 ? cd(me.src_val[k].p.value, me.src_val[k].q.value)
                                          : me.src_state[k] == 1 ? each
                                                                 : cd{};
            } else {
                cd const delta = cd(ap.p.value, ap.q.value) - s[b];
                double const mu_p = delta.real() / ap.p.var, mu_q = delta.imag() / ap.q.var;
                for (Idx k = 0; k < nlg; ++k)
                    if (mm.load_bus[k] == b && me.lg_state[k] == 2)
                        r.ld_s[k].v[0] = cd(me.lg_val[k].p.value - me.lg_val[k].p.var * mu_p,
                                            me.lg_val[k].q.value - me.lg_val[k].q.var * mu_q);
                for (Idx k = 0; k < nsrc; ++k)
                    if (mm.source_bus[k] == b && me.src_state[k] == 2)
                        r.src_s[k].v[0] = cd(me.src_val[k].p.value - me.src_val[k].p.var * mu_p,
                                             me.src_val[k].q.value - me.src_val[k].q.var * mu_q);
            }
        }
        for (Idx k = 0; k < nlg; ++k) r.ld_i[k].v[0] = std::conj(r.ld_s[k].v[0] / u[mm.load_bus[k]]);
        for (Idx k = 0; k < nsrc; ++k) r.src_i[k].v[0] = std::conj(r.src_s[k].v[0] / u[mm.source_bus[k]]);
    }
    write_outputs<1>(m, st, tc, res, out, scn);
    // node p, q: the bus injection (node.hpp get_output with the solver's bus_injection)
    if (out.data[OR_NODE]) {
        Idx const nn = Idx(m.node.size());
        for (Idx i = 0; i < nn; ++i) {
            Idx const g = tc.topo.node_group[i];
            if (g < 0) continue;
            auto& o = static_cast<OrNodeOutputSym*>(out.data[OR_NODE])[scn * nn + i];
            cd const sv = bus_s[g][tc.topo.node_pos[i]];
            o.p = kBaseP3 * sv.real();
            o.q = kBaseP3 * sv.imag();
        }
    }
    // sensors (voltage_sensor.hpp get_sym_output, power_sensor.hpp get_generic_output)
    Idx const nvs = Idx(S.vs_id.size()), nps = Idx(S.ps_id.size());
    for (Idx i = 0; vs_out && i < nvs; ++i) {
        auto& o = vs_out[scn * nvs + i];
        o = OrVoltageSensorOutputSym{};
        o.id = S.vs_id[i];
        Idx const node = S.vs_node[i], g = tc.topo.node_group[node];
        if (g < 0) continue;
        o.energized = 1;
        cd const uc = res[g].u[tc.topo.node_pos[node]].v[0];
        VoltVar const mv = voltage_calc(m, S, val, i);
        double const ur = m.node[node].u_rated;
        if (mv.has_angle) {
            o.u_residual = (std::abs(mv.value) - std::abs(uc)) * ur;
            o.u_angle_residual = std::arg(std::exp(cd(0.0, std::arg(mv.value) - std::arg(uc))));
        } else {
            o.u_residual = (mv.value.real() - std::abs(uc)) * ur;
            o.u_angle_residual = kNan;
        }
    }
    for (Idx i = 0; ps_out && i < nps; ++i) {
        auto& o = ps_out[scn * nps + i];
        o = OrPowerSensorOutputSym{};
        o.id = S.ps_id[i];
        Idx const obj = S.ps_obj[i];
        Idx g = -1, pos = -1;
        switch (S.ps_term[i]) {
        case 0:
        case 1: g = tc.br_group[obj]; pos = tc.br_pos[obj]; break;
        case 2: g = tc.src_group[obj]; pos = tc.src_pos[obj]; break;
        case 3: g = tc.sh_group[obj]; pos = tc.sh_pos[obj]; break;
        case 4:
        case 5: g = tc.lg_group[obj]; pos = tc.lg_pos[obj]; break;
        case 9: g = tc.topo.node_group[obj]; pos = g >= 0 ? tc.topo.node_pos[obj] : -1; break;
        default: break;
        }
        if (g < 0) continue;
        o.energized = 1;
        cd sc{};
        switch (S.ps_term[i]) {
        case 0: sc = res[g].br_sf[pos].v[0]; break;
        case 1: sc = res[g].br_st[pos].v[0]; break;
        case 2: sc = res[g].src_s[pos].v[0]; break;
        case 3: sc = res[g].sh_s[pos].v[0]; break;
        case 4:
        case 5: sc = res[g].ld_s[pos].v[0]; break;
        case 9: sc = bus_s[g][pos]; break;
        default: break;
        }
        PowerVar const pm = power_calc(val, S, i);
        double const dir = sensor_direction(S.ps_term[i]);
        cd const resid = (cd(pm.p.value, pm.q.value) - sc) * dir * kBaseP3;
        o.p_residual = resid.real();
        o.q_residual = resid.imag();
    }
    return iters;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// C API
// ------------------------------------------------------------------------------------------------
extern "C" int or_power_flow(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                             int64_t n_scenarios, int symmetric, double err_tol, int64_t max_iter, int64_t threading,
                             OrMutableDataset const* output, int64_t* iterations, int32_t* status, char* messages,
                             int64_t msg_stride) {
    return or_power_flow_method(input, system_frequency, update, n_scenarios, symmetric, err_tol, max_iter,
                                threading, 0, output, iterations, status, messages, msg_stride);
}

extern "C" int or_power_flow_method(OrConstDataset const* input, double system_frequency,
                                    OrConstDataset const* update, int64_t n_scenarios, int symmetric, double err_tol,
                                    int64_t max_iter, int64_t threading, int method, OrMutableDataset const* output,
                                    int64_t* iterations, int32_t* status, char* messages, int64_t msg_stride) {
    Model model;
    try {
        model = build_model(*input, system_frequency);
    } catch (std::exception const& e) {
        write_msg(messages, msg_stride, 0, e.what());
        return -1;
    }
    if (!update) n_scenarios = 1;
    // base topology, built once (the cache run)
    TopoCache base_tc = build_topo_cache(model, model.base);
    // job_dispatch.hpp:166-172
    Idx n_threads;
    if (threading < 0 || threading == 1 || !update)
        n_threads = 1;
    else if (threading == 0) {
        Idx const hw = Idx(std::thread::hardware_concurrency());
        n_threads = hw < 2 ? 1 : std::min(hw, n_scenarios);
    } else
        n_threads = std::min<Idx>(threading, n_scenarios);
    n_threads = std::max<Idx>(n_threads, 1);
    auto worker = [&](Idx t) {
        bool cache_valid = true;
        TopoCache own;
        for (Idx s = t; s < n_scenarios; s += n_threads) {
            State st = model.base;
            bool topo_changed = false, param_changed = false;
            try {
                if (update) apply_update(model, st, *update, s, topo_changed, param_changed);
                TopoCache const* tc = &base_tc;
                if (topo_changed || !cache_valid) {
                    // topology rebuilt on a status change, and again on the restore that follows it
                    own = build_topo_cache(model, st);
                    tc = &own;
                    cache_valid = !topo_changed;
                }
                Idx const it = symmetric ? solve_scenario<1>(model, st, *tc, err_tol, max_iter, output, s, method)
                                         : solve_scenario<3>(model, st, *tc, err_tol, max_iter, output, s, method);
                if (iterations) iterations[s] = it;
                if (status) status[s] = 0;
                write_msg(messages, msg_stride, s, "");
            } catch (SparseMatrixError const& e) {
                if (iterations) iterations[s] = 0;
                if (status) status[s] = 1;
                write_msg(messages, msg_stride, s, e.what());
                cache_valid = !topo_changed;
            } catch (IterationDiverge const& e) {
                if (iterations) iterations[s] = 0;
                if (status) status[s] = 2;
                write_msg(messages, msg_stride, s, e.what());
                cache_valid = !topo_changed;
            } catch (std::exception const& e) {
                if (iterations) iterations[s] = 0;
                if (status) status[s] = 3;
                write_msg(messages, msg_stride, s, e.what());
                cache_valid = !topo_changed;
            }
        }
    };
    if (n_threads == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (Idx t = 0; t < n_threads; ++t) th.emplace_back(worker, t);
        for (auto& x : th) x.join();
    }
    return 0;
}

extern "C" int or_power_flow_tap(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                                 int64_t n_scenarios, int symmetric, double err_tol, int64_t max_iter,
                                 int64_t threading, int method, int strategy, OrTapRegulatorInput const* regulators,
                                 int64_t n_regulator, OrTapRegulatorUpdate const* reg_update,
                                 int64_t reg_update_count, OrMutableDataset const* output, int8_t* tap_out,
                                 int64_t* iterations, int32_t* status, char* messages, int64_t msg_stride) {
    Model model;
    std::vector<OrTapRegulatorInput> regs(regulators, regulators + n_regulator);
    std::vector<Idx> reg_trafo(size_t(n_regulator), -1);
    std::vector<double> reg_u_rated(size_t(n_regulator), 0.0);
    try {
        model = build_model(*input, system_frequency);
        // main_core/input.hpp:168-241
        std::set<Idx> regulated;
        for (Idx r = 0; r < n_regulator; ++r) {
            auto it = model.ids.find(regs[size_t(r)].regulated_object);
            if (it == model.ids.end())
                throw OtherError("The id cannot be found: " + std::to_string(regs[size_t(r)].regulated_object) + "\n");
            if (it->second.first != OR_TRANSFORMER)
                throw OtherError("transformer_tap_regulator regulator is not supported for object with ID " +
                                 std::to_string(regs[size_t(r)].regulated_object));
            if (!regulated.insert(it->second.second).second)
                throw OtherError("There are objects regulated by more than one regulator. Maximum one regulator is "
                                 "allowed.");
            reg_trafo[size_t(r)] = it->second.second;
            Idx const t = it->second.second;
            reg_u_rated[size_t(r)] =
                model.node[size_t(regs[size_t(r)].control_side == 0 ? model.tr_f[size_t(t)] : model.tr_t[size_t(t)])]
                    .u_rated;
        }
    } catch (std::exception const& e) {
        write_msg(messages, msg_stride, 0, e.what());
        return -1;
    }
    if (!update) n_scenarios = 1;
    TopoCache base_tc = build_topo_cache(model, model.base);
    Idx n_threads = 1;
    if (update && threading == 0) {
        Idx const hw = Idx(std::thread::hardware_concurrency());
        n_threads = hw < 2 ? 1 : std::min(hw, n_scenarios);
    } else if (update && threading > 1) {
        n_threads = std::min<Idx>(threading, n_scenarios);
    }
    double const base_power = symmetric ? kBaseP3 : kBaseP3 / 3.0;
    auto run = [&](auto phases, Idx s) {
        constexpr int P = decltype(phases)::value;
        State st = model.base;
        bool topo_changed = false, param_changed = false;
        if (update) apply_update(model, st, *update, s, topo_changed, param_changed);
        RegulatorState rs;
        for (auto const& r : regs) {
            rs.status.push_back(r.status != 0);
            rs.u_set.push_back(r.u_set);
            rs.u_band.push_back(r.u_band);
            rs.ldc_r.push_back(r.line_drop_compensation_r);
            rs.ldc_x.push_back(r.line_drop_compensation_x);
        }
        for (Idx k = 0; reg_update && k < reg_update_count; ++k) {
            OrTapRegulatorUpdate const& u = reg_update[s * reg_update_count + k];
            Idx r = k;
            if (u.id != kNaID) {
                r = -1;
                for (Idx q = 0; q < n_regulator; ++q)
                    if (regs[size_t(q)].id == u.id) r = q;
                if (r < 0) throw OtherError("The id cannot be found: " + std::to_string(u.id) + "\n");
            }
            if (u.status != kNaS) rs.status[size_t(r)] = u.status != 0;
            if (!is_nan(u.u_set)) rs.u_set[size_t(r)] = u.u_set;
            if (!is_nan(u.u_band)) rs.u_band[size_t(r)] = u.u_band;
            if (!is_nan(u.line_drop_compensation_r)) rs.ldc_r[size_t(r)] = u.line_drop_compensation_r;
            if (!is_nan(u.line_drop_compensation_x)) rs.ldc_x[size_t(r)] = u.line_drop_compensation_x;
        }
        TopoCache own;
        TopoCache const* tc = &base_tc;
        if (topo_changed) {
            own = build_topo_cache(model, st);
            tc = &own;
        }
        TapOptimizer<P> opt{model, st, *tc, err_tol, max_iter, method, strategy, {}, {}, {}};
        for (auto const& rank : rank_regulators(model, st, regs, rs)) {
            opt.order.emplace_back();
            for (Idx r : rank) {
                typename TapOptimizer<P>::Regulated g{};
                g.reg = r;
                g.trafo = reg_trafo[size_t(r)];
                g.ctrl_from = regs[size_t(r)].control_side == 0;
                g.ctrl_node = g.ctrl_from ? model.tr_f[size_t(g.trafo)] : model.tr_t[size_t(g.trafo)];
                g.ctrl_at_tap = regs[size_t(r)].control_side == model.trafo[size_t(g.trafo)].tap_side;
                g.tmin = model.trafo[size_t(g.trafo)].tap_min;
                g.tmax = model.trafo[size_t(g.trafo)].tap_max;
                // transformer_tap_regulator.hpp:76-88
                double const ur = reg_u_rated[size_t(r)];
                g.u_set = rs.u_set[size_t(r)] / ur;
                g.u_band = rs.u_band[size_t(r)] / ur;
                g.z = cd(is_nan(rs.ldc_r[size_t(r)]) ? 0.0 : rs.ldc_r[size_t(r)],
                         is_nan(rs.ldc_x[size_t(r)]) ? 0.0 : rs.ldc_x[size_t(r)]) /
                      (ur * ur / base_power);
                opt.order.back().push_back(g);
            }
        }
        auto const res = opt.optimize();
        if (output) write_outputs<P>(model, st, *tc, res, *output, s);
        Idx it = 0;
        for (auto const& r : res) it = std::max(it, r.iterations);
        if (tap_out) {
            for (Idx r = 0; r < n_regulator; ++r) tap_out[s * n_regulator + r] = kNaS;
            for (auto const& rank : opt.order)
                for (auto const& g : rank) tap_out[s * n_regulator + g.reg] = st.tr_tap[size_t(g.trafo)];
        }
        return it;
    };
    auto worker = [&](Idx t) {
        for (Idx s = t; s < n_scenarios; s += n_threads) {
            int32_t code = 0;
            std::string msg;
            Idx it = 0;
            try {
                it = symmetric ? run(std::integral_constant<int, 1>{}, s) : run(std::integral_constant<int, 3>{}, s);
            } catch (MaxIterationReached const& e) {
                code = 4;
                msg = e.what();
            } catch (SparseMatrixError const& e) {
                code = 1;
                msg = e.what();
            } catch (IterationDiverge const& e) {
                code = 2;
                msg = e.what();
            } catch (std::exception const& e) {
                msg = e.what();
                code = msg.rfind("Automatic tap changer", 0) == 0 ? 4 : 3;
            }
            if (iterations) iterations[s] = code ? 0 : it;
            if (status) status[s] = code;
            write_msg(messages, msg_stride, s, msg);
        }
    };
    if (n_threads == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (Idx t = 0; t < n_threads; ++t) th.emplace_back(worker, t);
        for (auto& x : th) x.join();
    }
    return 0;
}

extern "C" int64_t or_minimum_degree_ordering(int64_t n_keys, int64_t const* keys, int64_t const* indptr,
                                              int64_t const* neighbours, int64_t* alpha_out, int64_t* n_alpha_out,
                                              int64_t* fills_out, int64_t max_fills) {
    AdjMap d;
    for (int64_t i = 0; i < n_keys; ++i) d[keys[i]].assign(neighbours + indptr[i], neighbours + indptr[i + 1]);
    auto [alpha, fills] = minimum_degree_ordering(std::move(d));
    for (size_t i = 0; i < alpha.size(); ++i) alpha_out[i] = alpha[i];
    *n_alpha_out = int64_t(alpha.size());
    if (int64_t(fills.size()) > max_fills) return -1;
    for (size_t i = 0; i < fills.size(); ++i) {
        fills_out[2 * i] = fills[i].first;
        fills_out[2 * i + 1] = fills[i].second;
    }
    return int64_t(fills.size());
}

extern "C" int64_t or_build_topology(int64_t n_node, int64_t n_branch, int64_t const* branch_from,
                                     int64_t const* branch_to, int8_t const* branch_from_status,
                                     int8_t const* branch_to_status, double const* branch_phase_shift,
                                     int64_t n_source, int64_t const* source_node, int8_t const* source_status,
                                     int64_t* node_group, int64_t* node_pos, int64_t* n_math_models,
                                     int64_t* fill_in_out, int64_t max_fill) {
    std::vector<CompBranch> br;
    for (int64_t b = 0; b < n_branch; ++b)
        br.push_back({branch_from[b], branch_to[b], branch_from_status[b] != 0, branch_to_status[b] != 0,
                      branch_phase_shift[b]});
    std::vector<Idx> sn(source_node, source_node + n_source);
    std::vector<bool> son;
    for (int64_t s = 0; s < n_source; ++s) son.push_back(source_status[s] != 0);
    auto const res = build_topology(n_node, br, sn, son);
    for (int64_t i = 0; i < n_node; ++i) {
        node_group[i] = res.node_group[i];
        node_pos[i] = res.node_pos[i];
    }
    *n_math_models = int64_t(res.bus_nodes.size());
    int64_t k = 0;
    for (size_t g = 0; g < res.fill_in.size(); ++g)
        for (auto const& f : res.fill_in[g]) {
            if (k >= max_fill) return -1;
            fill_in_out[3 * k] = int64_t(g);
            fill_in_out[3 * k + 1] = f[0];
            fill_in_out[3 * k + 2] = f[1];
            ++k;
        }
    return k;
}

extern "C" int64_t or_ybus_structure(int64_t n_bus, int64_t n_branch, int64_t const* branch_bus, int64_t n_shunt,
                                     int64_t const* shunt_bus, int64_t n_fill, int64_t const* fill_in,
                                     int64_t nnz_lu_max, int64_t* row_indptr_lu, int64_t* col_indices_lu,
                                     int64_t* diag_lu, int64_t* map_lu_y_bus, int64_t* lu_transpose_entry,
                                     int64_t* row_indptr, int64_t* col_indices, int64_t* bus_entry,
                                     int64_t* y_bus_entry_indptr, int64_t* nnz_out) {
    std::vector<std::array<Idx, 2>> bb(n_branch), fi(n_fill);
    for (int64_t b = 0; b < n_branch; ++b) bb[b] = {branch_bus[2 * b], branch_bus[2 * b + 1]};
    for (int64_t f = 0; f < n_fill; ++f) fi[f] = {fill_in[2 * f], fill_in[2 * f + 1]};
    std::vector<std::vector<Idx>> spb(n_bus);
    for (int64_t s = 0; s < n_shunt; ++s) spb[shunt_bus[s]].push_back(s);
    auto const ys = build_ystruct(n_bus, bb, spb, fi);
    int64_t const nnz_lu = ys.row_indptr_lu.back();
    if (nnz_lu > nnz_lu_max) return -1;
    std::copy(ys.row_indptr_lu.begin(), ys.row_indptr_lu.end(), row_indptr_lu);
    std::copy(ys.col_indices_lu.begin(), ys.col_indices_lu.end(), col_indices_lu);
    std::copy(ys.diag_lu.begin(), ys.diag_lu.end(), diag_lu);
    std::copy(ys.map_lu_y_bus.begin(), ys.map_lu_y_bus.end(), map_lu_y_bus);
    std::copy(ys.lu_transpose_entry.begin(), ys.lu_transpose_entry.end(), lu_transpose_entry);
    std::copy(ys.row_indptr.begin(), ys.row_indptr.end(), row_indptr);
    std::copy(ys.col_indices.begin(), ys.col_indices.end(), col_indices);
    std::copy(ys.bus_entry.begin(), ys.bus_entry.end(), bus_entry);
    std::copy(ys.y_bus_entry_indptr.begin(), ys.y_bus_entry_indptr.end(), y_bus_entry_indptr);
    *nnz_out = ys.row_indptr.back();
    return nnz_lu;
}

namespace {
template <int NB> int block_lu_solve_impl(int64_t n, int64_t const* ip, int64_t const* ci, int64_t const* dg,
                                          double* data, double const* rhs, double* x) {
    std::vector<Idx> ipv(ip, ip + n + 1), civ(ci, ci + ip[n]), dgv(dg, dg + n);
    std::vector<Blk<NB>> lu(ip[n]);
    for (int64_t k = 0; k < ip[n]; ++k)
        for (int i = 0; i < NB * NB; ++i) lu[k][i] = data[k * NB * NB + i];
    std::vector<Vec<NB>> b(n), xx(n);
    for (int64_t r = 0; r < n; ++r)
        for (int i = 0; i < NB; ++i) b[r][i] = rhs[r * NB + i];
    SparseLU<NB> solver(ipv, civ, dgv);
    try {
        solver.factorize(lu);
    } catch (SparseMatrixError const&) {
        return 1;
    } catch (std::exception const&) {
        return 2;
    }
    solver.solve(lu, b, xx);
    for (int64_t r = 0; r < n; ++r)
        for (int i = 0; i < NB; ++i) x[r * NB + i] = xx[r][i];
    for (int64_t k = 0; k < ip[n]; ++k)
        for (int i = 0; i < NB * NB; ++i) data[k * NB * NB + i] = lu[k][i];
    return 0;
}

template <int P>
int64_t nr_math_model_impl(int64_t n_bus, int64_t n_branch, int64_t const* branch_bus, double const* branch_param,
                           int64_t n_shunt, int64_t const* shunt_bus, double const* shunt_param, int64_t n_source,
                           int64_t const* source_bus, double const* source_y, double const* u_source,
                           int64_t n_load, int64_t const* load_bus, int8_t const* load_type, double const* s_inj,
                           int64_t n_fill, int64_t const* fill_in, double err_tol, int64_t max_iter, double* u,
                           double* branch_flow, double* shunt_flow, double* load_flow, double* source_flow,
                           double* bus_injection) {
    MathModel<P> mm;
    mm.n_bus = n_bus;
    auto ten = [](double const* p) {
        Ten<P> t;
        for (int i = 0; i < P * P; ++i) t.m[i] = cd(p[2 * i], p[2 * i + 1]);
        return t;
    };
    for (int64_t b = 0; b < n_branch; ++b) {
        mm.branch_bus.push_back({branch_bus[2 * b], branch_bus[2 * b + 1]});
        BranchParam<P> bp;
        for (int k = 0; k < 4; ++k) bp.v[k] = ten(branch_param + (b * 4 + k) * 2 * P * P);
        mm.branch_param.push_back(bp);
    }
    for (int64_t s = 0; s < n_shunt; ++s) {
        mm.shunt_bus.push_back(shunt_bus[s]);
        mm.shunt_param.push_back(ten(shunt_param + s * 2 * P * P));
    }
    MathInput<P> in;
    for (int64_t s = 0; s < n_source; ++s) {
        mm.source_bus.push_back(source_bus[s]);
        mm.source_y1.push_back(cd(source_y[4 * s], source_y[4 * s + 1]));
        mm.source_y0.push_back(cd(source_y[4 * s + 2], source_y[4 * s + 3]));
        in.u_source.push_back(cd(u_source[2 * s], u_source[2 * s + 1]));
    }
    for (int64_t l = 0; l < n_load; ++l) {
        mm.load_bus.push_back(load_bus[l]);
        mm.load_type.push_back(load_type[l]);
        Val<P> v;
        for (int ph = 0; ph < P; ++ph) v.v[ph] = cd(s_inj[(l * P + ph) * 2], s_inj[(l * P + ph) * 2 + 1]);
        in.s_inj.push_back(v);
    }
    for (int64_t f = 0; f < n_fill; ++f) mm.fill_in.push_back({fill_in[2 * f], fill_in[2 * f + 1]});
    NRSolver<P> solver(mm);
    MathOutput<P> out;
    try {
        out = solver.run(in, err_tol, max_iter);
    } catch (SparseMatrixError const&) {
        return -1;
    } catch (IterationDiverge const&) {
        return -2;
    }
    auto put = [](double* dst, std::vector<Val<P>> const& v, int64_t stride_vals, int64_t offset) {
        if (!dst) return;
        for (size_t i = 0; i < v.size(); ++i)
            for (int ph = 0; ph < P; ++ph) {
                dst[((i * stride_vals + offset) * P + ph) * 2] = v[i].v[ph].real();
                dst[((i * stride_vals + offset) * P + ph) * 2 + 1] = v[i].v[ph].imag();
            }
    };
    put(u, out.u, 1, 0);
    put(bus_injection, out.bus_injection, 1, 0);
    put(branch_flow, out.br_sf, 4, 0);
    put(branch_flow, out.br_st, 4, 1);
    put(branch_flow, out.br_if, 4, 2);
    put(branch_flow, out.br_it, 4, 3);
    put(shunt_flow, out.sh_s, 2, 0);
    put(shunt_flow, out.sh_i, 2, 1);
    put(load_flow, out.ld_s, 2, 0);
    put(load_flow, out.ld_i, 2, 1);
    put(source_flow, out.src_s, 2, 0);
    put(source_flow, out.src_i, 2, 1);
    return out.iterations;
}
}  // namespace

extern "C" int or_block_lu_solve(int64_t nb, int64_t n, int64_t const* row_indptr, int64_t const* col_indices,
                                 int64_t const* diag_lu, double* data, double const* rhs, double* x) {
    switch (nb) {
    case 1: return block_lu_solve_impl<1>(n, row_indptr, col_indices, diag_lu, data, rhs, x);
    case 2: return block_lu_solve_impl<2>(n, row_indptr, col_indices, diag_lu, data, rhs, x);
    case 6: return block_lu_solve_impl<6>(n, row_indptr, col_indices, diag_lu, data, rhs, x);
    default: return -1;
    }
}

extern "C" int64_t or_nr_math_model(int p, int64_t n_bus, int64_t n_branch, int64_t const* branch_bus,
                                    double const* branch_param, int64_t n_shunt, int64_t const* shunt_bus,
                                    double const* shunt_param, int64_t n_source, int64_t const* source_bus,
                                    double const* source_y, double const* u_source, int64_t n_load,
                                    int64_t const* load_bus, int8_t const* load_type, double const* s_inj,
                                    int64_t n_fill, int64_t const* fill_in, double err_tol, int64_t max_iter,
                                    double* u, double* branch_flow, double* shunt_flow, double* load_flow,
                                    double* source_flow, double* bus_injection) {
    if (p == 1)
        return nr_math_model_impl<1>(n_bus, n_branch, branch_bus, branch_param, n_shunt, shunt_bus, shunt_param,
                                     n_source, source_bus, source_y, u_source, n_load, load_bus, load_type, s_inj,
                                     n_fill, fill_in, err_tol, max_iter, u, branch_flow, shunt_flow, load_flow,
                                     source_flow, bus_injection);
    return nr_math_model_impl<3>(n_bus, n_branch, branch_bus, branch_param, n_shunt, shunt_bus, shunt_param, n_source,
                                 source_bus, source_y, u_source, n_load, load_bus, load_type, s_inj, n_fill, fill_in,
                                 err_tol, max_iter, u, branch_flow, shunt_flow, load_flow, source_flow, bus_injection);
}

extern "C" void or_set_se_dense(int dense) { g_se_dense = dense != 0; }
extern "C" void or_se_counters(long long* perturbed, long long* refine_solves, int reset) {
    if (perturbed) *perturbed = g_se_perturbed.load();
    if (refine_solves) *refine_solves = g_se_refine_solves.load();
    if (reset) {
        g_se_perturbed = 0;
        g_se_refine_solves = 0;
    }
}

extern "C" int or_state_estimation(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                                   int64_t n_scenarios, double err_tol, int64_t max_iter, int64_t threading,
                                   OrSymVoltageSensorInput const* vs, int64_t n_vs,
                                   OrSymVoltageSensorUpdate const* vs_upd, int64_t vs_upd_count,
                                   OrSymPowerSensorInput const* ps, int64_t n_ps, OrSymPowerSensorUpdate const* ps_upd,
                                   int64_t ps_upd_count, OrMutableDataset const* output,
                                   OrVoltageSensorOutputSym* vs_out, OrPowerSensorOutputSym* ps_out,
                                   int64_t* iterations, int32_t* status, char* messages, int64_t msg_stride) {
    Model model;
    SeSensorSet S;
    SeSensorValues base;
    try {
        model = build_model(*input, system_frequency);
        // sensors: the measured object must exist and be of the terminal type's kind (main_core/input.hpp:84-129)
        for (Idx i = 0; i < n_vs; ++i) {
            S.vs_id.push_back(vs[i].id);
            S.vs_node.push_back(node_idx(model, vs[i].measured_object));
            base.vs_sigma.push_back(vs[i].u_sigma);
            base.vs_u.push_back(vs[i].u_measured);
            base.vs_ang.push_back(vs[i].u_angle_measured);
        }
        for (Idx i = 0; i < n_ps; ++i) {
            auto it = model.ids.find(ps[i].measured_object);
            if (it == model.ids.end())
                throw OtherError("The id cannot be found: " + std::to_string(ps[i].measured_object) + "\n");
            int const comp = it->second.first;
            Idx obj = it->second.second;
            int const term = ps[i].measured_terminal_type;
            bool ok = false;
            if (term == 0 || term == 1) {
                ok = comp == OR_LINE || comp == OR_GENERIC_BRANCH || comp == OR_TRANSFORMER;
                if (comp == OR_GENERIC_BRANCH) obj += Idx(model.line.size() + model.link.size());
                if (comp == OR_TRANSFORMER) obj += model.n_lslot();
            } else if (term == 2) {
                ok = comp == OR_SOURCE;
            } else if (term == 3) {
                ok = comp == OR_SHUNT;
            } else if (term == 4) {
                ok = comp == OR_SYM_LOAD || comp == OR_ASYM_LOAD;
                obj += lg_offset(model, comp);
            } else if (term == 5) {
                ok = comp == OR_SYM_GEN || comp == OR_ASYM_GEN;
                obj += lg_offset(model, comp);
            } else if (term == 9) {
                ok = comp == OR_NODE;
            }
            if (!ok) throw OtherError("Wrong type for object with id " + std::to_string(ps[i].measured_object) + "\n");
            S.ps_id.push_back(ps[i].id);
            S.ps_term.push_back(term);
            S.ps_obj.push_back(obj);
            base.ps_sigma.push_back(ps[i].power_sigma);
            base.ps_p.push_back(ps[i].p_measured);
            base.ps_q.push_back(ps[i].q_measured);
            base.ps_psig.push_back(ps[i].p_sigma);
            base.ps_qsig.push_back(ps[i].q_sigma);
        }
    } catch (std::exception const& e) {
        write_msg(messages, msg_stride, 0, e.what());
        return -1;
    }
    if (!update) n_scenarios = 1;
    std::map<int32_t, Idx> vs_index, ps_index;
    for (Idx i = 0; i < n_vs; ++i) vs_index[vs[i].id] = i;
    for (Idx i = 0; i < n_ps; ++i) ps_index[ps[i].id] = i;
    // job_dispatch.hpp:166-172
    Idx n_threads;
    if (threading < 0 || threading == 1 || !update)
        n_threads = 1;
    else if (threading == 0) {
        Idx const hw = Idx(std::thread::hardware_concurrency());
        n_threads = hw < 2 ? 1 : std::min(hw, n_scenarios);
    } else
        n_threads = std::min<Idx>(threading, n_scenarios);
    n_threads = std::max<Idx>(n_threads, 1);
    auto worker = [&](Idx t) {
    for (Idx s = t; s < n_scenarios; s += n_threads) {
        State st = model.base;
        SeSensorValues val = base;
        try {
            bool topo_changed = false, param_changed = false;
            if (update) apply_update(model, st, *update, s, topo_changed, param_changed);
            // sensor updates: NaN keeps the value (voltage_sensor.hpp:88-100, power_sensor.hpp:94-105)
            auto find = [](std::map<int32_t, Idx> const& idx, int32_t id, Idx row) -> Idx {
                if (id == kNaID) return row;
                auto it = idx.find(id);
                if (it == idx.end()) throw OtherError("The id cannot be found: " + std::to_string(id) + "\n");
                return it->second;
            };
            for (Idx r = 0; vs_upd && r < vs_upd_count; ++r) {
                auto const& u = vs_upd[s * vs_upd_count + r];
                Idx const i = find(vs_index, u.id, r);
                if (!std::isnan(u.u_measured)) val.vs_u[i] = u.u_measured;
                if (!std::isnan(u.u_angle_measured)) val.vs_ang[i] = u.u_angle_measured;
                if (!std::isnan(u.u_sigma)) val.vs_sigma[i] = u.u_sigma;
            }
            for (Idx r = 0; ps_upd && r < ps_upd_count; ++r) {
                auto const& u = ps_upd[s * ps_upd_count + r];
                Idx const i = find(ps_index, u.id, r);
                if (!std::isnan(u.p_measured)) val.ps_p[i] = u.p_measured;
                if (!std::isnan(u.q_measured)) val.ps_q[i] = u.q_measured;
                if (!std::isnan(u.power_sigma)) val.ps_sigma[i] = u.power_sigma;
                if (!std::isnan(u.p_sigma)) val.ps_psig[i] = u.p_sigma;
                if (!std::isnan(u.q_sigma)) val.ps_qsig[i] = u.q_sigma;
            }
            TopoCache const tc = build_topo_cache(model, st);
            Idx const it = se_scenario(model, st, tc, S, val, err_tol, max_iter, *output, vs_out, ps_out, s);
            if (iterations) iterations[s] = it;
            if (status) status[s] = 0;
            write_msg(messages, msg_stride, s, "");
        } catch (SparseMatrixError const& e) {
            if (iterations) iterations[s] = 0;
            if (status) status[s] = 1;
            write_msg(messages, msg_stride, s, e.what());
        } catch (IterationDiverge const& e) {
            if (iterations) iterations[s] = 0;
            if (status) status[s] = 2;
            write_msg(messages, msg_stride, s, e.what());
        } catch (std::exception const& e) {
            if (iterations) iterations[s] = 0;
            if (status) status[s] = 3;
            write_msg(messages, msg_stride, s, e.what());
        }
    }
    };
    if (n_threads == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (Idx t = 0; t < n_threads; ++t) th.emplace_back(worker, t);
        for (auto& x : th) x.join();
    }
    return 0;
}
