// model.cpp -- component model and per-unit parameters (see model.hpp for the reference files).
#include "model.hpp"

#include <cmath>
#include <limits>
#include <string>

namespace pgm_amd {

namespace {
constexpr double kNan = std::numeric_limits<double>::quiet_NaN();
bool isnan_(double x) { return std::isnan(x); }

Tensor3 scalar_tensor(cd y, bool symmetric) {
    Tensor3 t;
    if (symmetric) {
        t.m[0] = y;
    } else {
        for (int i = 0; i < 3; ++i) t.m[i * 4] = y;
    }
    return t;
}
Tensor3 sm_tensor(cd s, cd mm) {
    Tensor3 t;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t.m[r * 3 + c] = r == c ? s : mm;
    return t;
}

// one-sequence pi model (branch.hpp:196-226)
std::array<cd, 4> pi_model(cd y_series, cd y_shunt, cd tap_ratio, bool fs, bool ts) {
    std::array<cd, 4> p{};
    double const tap = std::abs(tap_ratio);
    if (fs && ts) {
        p[3] = y_series + 0.5 * y_shunt;
        p[0] = (1.0 / tap / tap) * p[3];
        p[1] = (-1.0 / std::conj(tap_ratio)) * y_series;
        p[2] = (-1.0 / tap_ratio) * y_series;
    } else if (fs || ts) {
        cd const sh = std::abs(y_shunt) < kNumTol ? cd{} : 0.5 * y_shunt + 1.0 / (1.0 / y_series + 2.0 / y_shunt);
        if (fs) p[0] = (1.0 / tap / tap) * sh;
        if (ts) p[3] = sh;
    }
    return p;
}

double tap_adjust(double tap, double tmin, double tmax, double tnom, double xk, double xk_min, double xk_max) {
    if (tap <= std::max(tnom, tmax) && tap >= std::min(tnom, tmax)) {
        if (tmax == tnom) return xk;
        return xk + (tap - tnom) * ((xk_max - xk) / (tmax - tnom));
    }
    if (tmin == tnom) return xk;
    return xk + (tap - tnom) * ((xk_min - xk) / (tmin - tnom));
}

struct TrafoY {
    cd ys, ysh, y0sh;
    double k;
};
TrafoY trafo_y(TrafoStatic const& t, IntS tap) {
    double const base_y_to = t.base_i_to * t.base_i_to / kBaseP1;
    double u1 = t.u1, u2 = t.u2;
    double const dv = t.tap_direction * (tap - t.tap_nom) * t.tap_size;
    if (t.tap_side == 0)
        u1 += dv;
    else
        u2 += dv;
    double const k = (u1 / u2) / t.nominal_ratio;
    double const uk = tap_adjust(tap, t.tap_min, t.tap_max, t.tap_nom, t.uk, t.uk_min, t.uk_max);
    double const pk = tap_adjust(tap, t.tap_min, t.tap_max, t.tap_nom, t.pk, t.pk_min, t.pk_max);
    double const zabs = std::abs(uk) * u2 * u2 / t.sn;
    double const zr = pk * u2 * u2 / t.sn / t.sn;
    double const zi2 = zabs * zabs - zr * zr;
    cd const z(zr, (uk >= 0 ? 1.0 : -1.0) * (zi2 > 0.0 ? std::sqrt(zi2) : 0.0));
    auto shunt = [&](double i0, double p0) {
        double const yabs = i0 * t.sn / u2 / u2;
        double const yr = p0 / u2 / u2;
        double const yi2 = yabs * yabs - yr * yr;
        return cd(yr, yi2 > 0.0 ? -std::sqrt(yi2) : 0.0) / base_y_to;
    };
    return {(1.0 / z) / base_y_to, shunt(t.i0, t.p0), shunt(t.i0_zs, t.p0_zs), k};
}

Tensor3 matmul3(Tensor3 const& a, Tensor3 const& b) {
    Tensor3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd s = 0;
            for (int k = 0; k < 3; ++k) s += a.m[i * 3 + k] * b.m[k * 3 + j];
            r.m[i * 3 + j] = s;
        }
    return r;
}
}  // namespace

Idx Model::count(int comp) const {
    switch (comp) {
    case C_NODE: return n_node();
    case C_LINE: return Idx(line.size());
    case C_LINK: return Idx(link.size());
    case C_GENERIC_BRANCH: return Idx(gbranch.size());
    case C_TRANSFORMER: return Idx(trafo.size());
    case C_SHUNT: return Idx(shunt.size());
    case C_SOURCE: return Idx(source.size());
    case C_SYM_GEN: return lg_count[0];
    case C_ASYM_GEN: return lg_count[1];
    case C_SYM_LOAD: return lg_count[2];
    case C_ASYM_LOAD: return lg_count[3];
    case C_VOLTAGE_SENSOR: return Idx(vs_id.size());
    case C_POWER_SENSOR: return Idx(ps_id.size());
    default: return 0;
    }
}
ID Model::comp_id(int comp, Idx i) const {
    switch (comp) {
    case C_NODE: return node[i].id;
    case C_LINE: return line[i].id;
    case C_LINK: return link[i].id;
    case C_GENERIC_BRANCH: return gbranch[i].id;
    case C_TRANSFORMER: return trafo[i].id;
    case C_SHUNT: return shunt[i].id;
    case C_SOURCE: return source[i].id;
    case C_SYM_GEN: return lg_id[lg_offset[0] + i];
    case C_ASYM_GEN: return lg_id[lg_offset[1] + i];
    case C_SYM_LOAD: return lg_id[lg_offset[2] + i];
    case C_ASYM_LOAD: return lg_id[lg_offset[3] + i];
    case C_VOLTAGE_SENSOR: return vs_id[i];
    case C_POWER_SENSOR: return ps_id[i];
    default: return na_IntID;
    }
}

Idx find_component(Model const& m, int comp, ID id) {
    auto it = m.ids.find(id);
    if (it == m.ids.end()) throw PowerGridError("The id cannot be found: " + std::to_string(id) + "\n");
    if (it->second.first != comp) throw PowerGridError("Wrong type for object with id " + std::to_string(id) + "\n");
    return it->second.second;
}

IntS tap_limit(TrafoStatic const& t, IntS tap) {
    tap = std::min<IntS>(tap, std::max(t.tap_max, t.tap_min));
    tap = std::max<IntS>(tap, std::min(t.tap_max, t.tap_min));
    return tap;
}

double branch_phase_shift(Model const& m, Idx b) {
    return b < m.n_line() ? m.lstat[b].shift : m.tstat[b - m.n_line()].clock * kDeg30;
}

Model build_model(PGM_ConstDataset const& in, double f) {
    Model m;
    m.freq = f;
    std::vector<char> rec(512);
    auto each = [&](int comp, auto&& fn) {
        auto const* buf = in.find(comp_names[comp]);
        if (!buf) return;
        for (Idx i = 0; i < buf->total_elements; ++i) {
            buf->read_record(i, rec.data());
            fn(rec.data());
        }
    };
    for (auto const& b : in.buffers) {
        bool known = b.meta->name == kTapRegulatorName || b.meta->name == kFaultName ||
                     b.meta->name == kVoltageSensorName || b.meta->name == kPowerSensorName;
        for (int c = 0; c < N_COMP; ++c) known |= b.meta->name == comp_names[c];
        if (!known) throw PowerGridError("component " + b.meta->name + " is not supported by this build");
    }
    auto reg = [&](ID id, int comp, Idx i) {
        if (!m.ids.emplace(id, std::make_pair(comp, i)).second)
            throw PowerGridError("Conflicting id detected: " + std::to_string(id) + "\n");
    };
    std::unordered_map<ID, Idx> node_idx;
    auto nidx = [&](ID id) {
        auto it = node_idx.find(id);
        if (it == node_idx.end()) throw PowerGridError("The id cannot be found: " + std::to_string(id) + "\n");
        return it->second;
    };
    each(C_NODE, [&](char const* r) {
        NodeInput n;
        std::memcpy(&n, r, sizeof n);
        reg(n.id, C_NODE, m.n_node());
        node_idx[n.id] = m.n_node();
        m.node.push_back(n);
    });
    each(C_LINE, [&](char const* r) {
        LineInput l;
        std::memcpy(&l, r, sizeof l);
        reg(l.id, C_LINE, m.n_line());
        Idx const fi = nidx(l.from_node), ti = nidx(l.to_node);
        double const u1 = m.node[fi].u_rated, u2 = m.node[ti].u_rated;
        if (std::abs(u1 - u2) > kNumTol)
            throw PowerGridError("Conflicting voltage for line " + std::to_string(l.id) + "\n voltage at from node " +
                                 std::to_string(l.from_node) + " is " + std::to_string(u1) + "\n voltage at to node " +
                                 std::to_string(l.to_node) + " is " + std::to_string(u2) + "\n");
        LineStatic s;
        s.base_i = kBaseP3 / u1 / kSqrt3;
        double const base_y = s.base_i / (u1 / kSqrt3);
        s.y1s = 1.0 / cd(l.r1, l.x1) / base_y;
        s.y1sh = 2.0 * kPi * f * l.c1 / base_y * cd(l.tan1, 1.0);
        s.y0s = 1.0 / cd(l.r0, l.x0) / base_y;
        s.y0sh = 2.0 * kPi * f * l.c0 / base_y * cd(l.tan0, 1.0);
        s.i_n = l.i_n;
        s.base_i_to = s.base_i;
        m.line.push_back(l);
        m.line_f.push_back(fi);
        m.line_t.push_back(ti);
        m.lstat.push_back(s);
        m.base.line_fs.push_back(l.from_status != 0);
        m.base.line_ts.push_back(l.to_status != 0);
    });
    // link.hpp:19-36: y_series = y_link = (1 + 1j) * 1e6 S at 10 kV in per unit, no shunt, tap 1, phase
    // shift 0, loading 0; base currents from the rated voltages of its two nodes (no voltage check)
    each(C_LINK, [&](char const* r) {
        LinkInput l;
        std::memcpy(&l, r, sizeof l);
        reg(l.id, C_LINK, Idx(m.link.size()));
        Idx const fi = nidx(l.from_node), ti = nidx(l.to_node);
        double const g_link = 1e6 / (kBaseP3 / 10e3 / 10e3);
        LineStatic s;
        s.y1s = s.y0s = cd(g_link, g_link);
        s.y1sh = s.y0sh = cd{};
        s.base_i = kBaseP3 / m.node[fi].u_rated / kSqrt3;
        s.base_i_to = kBaseP3 / m.node[ti].u_rated / kSqrt3;
        s.i_n = std::numeric_limits<double>::infinity();
        m.link.push_back(l);
        m.line_f.push_back(fi);
        m.line_t.push_back(ti);
        m.lstat.push_back(s);
        m.base.line_fs.push_back(l.from_status != 0);
        m.base.line_ts.push_back(l.to_status != 0);
    });
    // generic_branch.hpp:24-60: y1 in per unit of the to side, ratio k e^(j theta) (k NaN: 1, theta NaN: 0,
    // theta taken modulo 2 pi), loading max s / sn (0 without sn); no zero-sequence model
    each(C_GENERIC_BRANCH, [&](char const* r) {
        GenericBranchInput g;
        std::memcpy(&g, r, sizeof g);
        reg(g.id, C_GENERIC_BRANCH, Idx(m.gbranch.size()));
        Idx const fi = nidx(g.from_node), ti = nidx(g.to_node);
        double const u1 = m.node[fi].u_rated, u2 = m.node[ti].u_rated;
        LineStatic s;
        s.base_i = kBaseP3 / u1 / kSqrt3;
        s.base_i_to = kBaseP3 / u2 / kSqrt3;
        double const base_y = s.base_i_to / (u2 / kSqrt3);
        s.y1s = 1.0 / cd(g.r1, g.x1) / base_y;
        s.y1sh = cd(g.g1, g.b1) / base_y;
        s.y0s = s.y1s;
        s.y0sh = s.y1sh;
        double const k = std::isnan(g.k) ? 1.0 : g.k;
        s.shift = std::isnan(g.theta) ? 0.0 : std::fmod(g.theta, 2.0 * kPi);
        s.tap = k * std::exp(cd(0.0, s.shift));
        s.i_n = std::numeric_limits<double>::infinity();
        s.sn = std::isnan(g.sn) ? std::numeric_limits<double>::infinity() : g.sn;
        m.gbranch.push_back(g);
        m.line_f.push_back(fi);
        m.line_t.push_back(ti);
        m.lstat.push_back(s);
        m.base.line_fs.push_back(g.from_status != 0);
        m.base.line_ts.push_back(g.to_status != 0);
    });
    each(C_TRANSFORMER, [&](char const* r) {
        TransformerInput t;
        std::memcpy(&t, r, sizeof t);
        reg(t.id, C_TRANSFORMER, Idx(m.trafo.size()));
        Idx const fi = nidx(t.from_node), ti = nidx(t.to_node);
        double const u1r = m.node[fi].u_rated, u2r = m.node[ti].u_rated;
        TrafoStatic s{};
        s.u1 = t.u1;
        s.u2 = t.u2;
        s.sn = t.sn;
        s.tap_size = t.tap_size;
        s.uk = t.uk;
        s.pk = t.pk;
        s.i0 = t.i0;
        s.p0 = t.p0;
        s.i0_zs = isnan_(t.i0_zero_sequence) ? s.i0 : t.i0_zero_sequence;
        s.p0_zs = isnan_(t.p0_zero_sequence) ? s.p0 + s.pk * (s.i0_zs * s.i0_zs - s.i0 * s.i0) : t.p0_zero_sequence;
        s.winding_from = t.winding_from;
        s.winding_to = t.winding_to;
        s.tap_side = t.tap_side;
        s.tap_min = t.tap_min;
        s.tap_max = t.tap_max;
        s.tap_nom = t.tap_nom == na_IntS ? IntS{0} : t.tap_nom;
        s.tap_direction = t.tap_max > t.tap_min ? 1 : -1;
        s.uk_min = isnan_(t.uk_min) ? s.uk : t.uk_min;
        s.uk_max = isnan_(t.uk_max) ? s.uk : t.uk_max;
        s.pk_min = isnan_(t.pk_min) ? s.pk : t.pk_min;
        s.pk_max = isnan_(t.pk_max) ? s.pk : t.pk_max;
        s.base_i_from = kBaseP3 / u1r / kSqrt3;
        s.base_i_to = kBaseP3 / u2r / kSqrt3;
        s.nominal_ratio = u1r / u2r;
        auto zpu = [](double rr, double xx, double u) {
            rr = isnan_(rr) ? 0 : rr;
            xx = isnan_(xx) ? 0 : xx;
            double const bz = u * u / kBaseP3;
            return cd(rr / bz, xx / bz);
        };
        s.z_ground_from = zpu(t.r_grounding_from, t.x_grounding_from, u1r);
        s.z_ground_to = zpu(t.r_grounding_to, t.x_grounding_to, u2r);
        bool const even = (t.clock % 2) == 0;
        bool const from_wye = t.winding_from == 0 || t.winding_from == 1;
        bool const to_wye = t.winding_to == 0 || t.winding_to == 1;
        if (even != (from_wye == to_wye))
            throw PowerGridError("Invalid clock for transformer " + std::to_string(t.id) + ", clock " +
                                 std::to_string(int(t.clock)) + "\n");
        s.clock = IntS(((t.clock % 12) + 12) % 12);
        IntS tap = t.tap_pos == na_IntS ? (t.tap_nom == na_IntS ? IntS{0} : t.tap_nom) : t.tap_pos;
        m.trafo.push_back(t);
        m.tr_f.push_back(fi);
        m.tr_t.push_back(ti);
        m.tstat.push_back(s);
        m.base.tr_tap.push_back(tap_limit(s, tap));
        m.base.tr_fs.push_back(t.from_status != 0);
        m.base.tr_ts.push_back(t.to_status != 0);
    });
    each(C_SHUNT, [&](char const* r) {
        ShuntInput s;
        std::memcpy(&s, r, sizeof s);
        reg(s.id, C_SHUNT, Idx(m.shunt.size()));
        m.shunt.push_back(s);
        m.sh_node.push_back(nidx(s.node));
        m.base.sh_status.push_back(s.status != 0);
        m.base.sh_g1.push_back(s.g1);
        m.base.sh_b1.push_back(s.b1);
        m.base.sh_g0.push_back(s.g0);
        m.base.sh_b0.push_back(s.b0);
    });
    each(C_SOURCE, [&](char const* r) {
        SourceInput s;
        std::memcpy(&s, r, sizeof s);
        reg(s.id, C_SOURCE, Idx(m.source.size()));
        m.source.push_back(s);
        m.src_node.push_back(nidx(s.node));
        m.base.src_status.push_back(s.status != 0);
        m.base.src_uref.push_back(s.u_ref);
        m.base.src_uang.push_back(isnan_(s.u_ref_angle) ? 0.0 : s.u_ref_angle);
        m.base.src_sk.push_back(isnan_(s.sk) ? 1e10 : s.sk);
        m.base.src_rx.push_back(isnan_(s.rx_ratio) ? 0.1 : s.rx_ratio);
        m.base.src_z01.push_back(isnan_(s.z01_ratio) ? 1.0 : s.z01_ratio);
    });
    int const lg_comps[4] = {C_SYM_GEN, C_ASYM_GEN, C_SYM_LOAD, C_ASYM_LOAD};
    for (int k = 0; k < 4; ++k) {
        int const comp = lg_comps[k];
        bool const asym = comp == C_ASYM_GEN || comp == C_ASYM_LOAD;
        m.lg_offset[k] = m.n_lg();
        Idx cnt = 0;
        each(comp, [&](char const* r) {
            std::array<double, 3> p{kNan, kNan, kNan}, q{kNan, kNan, kNan};
            ID id, node;
            IntS status, type;
            if (asym) {
                AsymLoadGenInput a;
                std::memcpy(&a, r, sizeof a);
                id = a.id;
                node = a.node;
                status = a.status;
                type = a.type;
                for (int ph = 0; ph < 3; ++ph) {
                    p[ph] = a.p_specified[ph];
                    q[ph] = a.q_specified[ph];
                }
            } else {
                SymLoadGenInput a;
                std::memcpy(&a, r, sizeof a);
                id = a.id;
                node = a.node;
                status = a.status;
                type = a.type;
                p[0] = a.p_specified;
                q[0] = a.q_specified;
            }
            reg(id, comp, cnt);
            m.lg_id.push_back(id);
            m.lg_type.push_back(type);
            m.lg_asym.push_back(asym);
            m.lg_gen.push_back(comp == C_SYM_GEN || comp == C_ASYM_GEN);
            m.lg_node.push_back(nidx(node));
            m.base.lg_status.push_back(status != 0);
            m.base.lg_p.push_back(p);
            m.base.lg_q.push_back(q);
            ++cnt;
        });
        m.lg_count[k] = cnt;
    }
    // faults (main_core/input.hpp:164-167: the faulted object must exist and be a node)
    if (auto const* buf = in.find(kFaultName)) {
        for (Idx i = 0; i < buf->total_elements; ++i) {
            FaultInput fl;
            buf->read_record(i, rec.data());
            std::memcpy(&fl, rec.data(), sizeof fl);
            reg(fl.id, C_FAULT, Idx(m.fault.size()));
            (void)find_component(m, C_NODE, fl.fault_object);
            m.fault.push_back(fl);
        }
    }
    // transformer tap regulators (main_core/input.hpp:168-213, 235-241): the regulated object is a
    // transformer, the control side one of its two sides, each transformer regulated at most once
    if (auto const* buf = in.find(kTapRegulatorName)) {
        std::vector<char> regulated(m.trafo.size(), 0);
        for (Idx i = 0; i < buf->total_elements; ++i) {
            buf->read_record(i, rec.data());
            TapRegulatorInput r;
            std::memcpy(&r, rec.data(), sizeof r);
            auto it = m.ids.find(r.regulated_object);
            if (it == m.ids.end())
                throw PowerGridError("The id cannot be found: " + std::to_string(r.regulated_object) + "\n");
            if (it->second.first != C_TRANSFORMER)
                throw PowerGridError(std::string(kTapRegulatorName) +
                                     " regulator is not supported for object with ID " +
                                     std::to_string(r.regulated_object));
            if (r.control_side != 0 && r.control_side != 1)
                throw PowerGridError("transformer_tap_regulator item retrieval is not implemented for " +
                                     std::to_string(int(r.control_side)) + "!\n");
            Idx const t = it->second.second;
            reg(r.id, C_TAP_REGULATOR, Idx(m.regulator.size()));
            if (regulated[t])
                throw PowerGridError(
                    "There are objects regulated by more than one regulator. Maximum one regulator is allowed.");
            regulated[t] = 1;
            m.regulator.push_back(r);
            m.reg_trafo.push_back(t);
            m.reg_u_rated.push_back(m.node[r.control_side == 0 ? m.tr_f[t] : m.tr_t[t]].u_rated);
            m.base.reg_status.push_back(r.status != 0);
            m.base.reg_u_set.push_back(r.u_set);
            m.base.reg_u_band.push_back(r.u_band);
            m.base.reg_ldc_r.push_back(r.line_drop_compensation_r);
            m.base.reg_ldc_x.push_back(r.line_drop_compensation_x);
        }
    }
    // voltage sensors: the measured object is a node (main_core/input.hpp:84-86)
    if (auto const* buf = in.find(kVoltageSensorName)) {
        for (Idx i = 0; i < buf->total_elements; ++i) {
            buf->read_record(i, rec.data());
            SymVoltageSensorInput v;
            std::memcpy(&v, rec.data(), sizeof v);
            Idx const node = find_component(m, C_NODE, v.measured_object);
            reg(v.id, C_VOLTAGE_SENSOR, Idx(m.vs_id.size()));
            m.vs_id.push_back(v.id);
            m.vs_node.push_back(node);
            m.base.vs_sigma.push_back(v.u_sigma);
            m.base.vs_u.push_back(v.u_measured);
            m.base.vs_angle.push_back(v.u_angle_measured);
        }
    }
    // power sensors: the measured object must be of the kind the terminal type names, never a link
    // (main_core/input.hpp:87-129)
    if (auto const* buf = in.find(kPowerSensorName)) {
        for (Idx i = 0; i < buf->total_elements; ++i) {
            buf->read_record(i, rec.data());
            SymPowerSensorInput ps;
            std::memcpy(&ps, rec.data(), sizeof ps);
            auto it = m.ids.find(ps.measured_object);
            if (it == m.ids.end())
                throw PowerGridError("The id cannot be found: " + std::to_string(ps.measured_object) + "\n");
            int const comp = it->second.first;
            Idx obj = it->second.second;
            if (comp == C_LINK) throw PowerGridError("PowerSensor measurement is not supported for object of type Link");
            auto wrong = [&] { return PowerGridError("Wrong type for object with id " + std::to_string(ps.measured_object) + "\n"); };
            switch (ps.measured_terminal_type) {
            case T_BRANCH_FROM:
            case T_BRANCH_TO:
                if (comp == C_TRANSFORMER) {
                    obj += m.n_line();
                } else if (comp == C_LINE || comp == C_GENERIC_BRANCH) {
                    int slot = comp;
                    line_slot(m, slot, obj);
                } else {
                    throw wrong();
                }
                break;
            case T_SHUNT:
                if (comp != C_SHUNT) throw wrong();
                break;
            case T_SOURCE:
                if (comp != C_SOURCE) throw wrong();
                break;
            case T_LOAD:
                if (comp != C_SYM_LOAD && comp != C_ASYM_LOAD) throw wrong();
                obj += m.lg_offset[comp == C_SYM_LOAD ? 2 : 3];
                break;
            case T_GENERATOR:
                if (comp != C_SYM_GEN && comp != C_ASYM_GEN) throw wrong();
                obj += m.lg_offset[comp == C_SYM_GEN ? 0 : 1];
                break;
            case T_NODE:
                if (comp != C_NODE) throw wrong();
                break;
            default:
                throw PowerGridError("measured terminal type " + std::to_string(int(ps.measured_terminal_type)) +
                                     " of sym_power_sensor is not supported by this build\n");
            }
            reg(ps.id, C_POWER_SENSOR, Idx(m.ps_id.size()));
            m.ps_id.push_back(ps.id);
            m.ps_term.push_back(ps.measured_terminal_type);
            m.ps_obj.push_back(obj);
            m.base.ps_sigma.push_back(ps.power_sigma);
            m.base.ps_p.push_back(ps.p_measured);
            m.base.ps_q.push_back(ps.q_measured);
            m.base.ps_psig.push_back(ps.p_sigma);
            m.base.ps_qsig.push_back(ps.q_sigma);
        }
    }
    return m;
}

void apply_voltage_sensor_update(ComponentState& st, Idx i, SymVoltageSensorUpdate const& u) {
    if (!std::isnan(u.u_measured)) st.vs_u[i] = u.u_measured;
    if (!std::isnan(u.u_angle_measured)) st.vs_angle[i] = u.u_angle_measured;
    if (!std::isnan(u.u_sigma)) st.vs_sigma[i] = u.u_sigma;
}
void apply_power_sensor_update(ComponentState& st, Idx i, SymPowerSensorUpdate const& u) {
    if (!std::isnan(u.p_measured)) st.ps_p[i] = u.p_measured;
    if (!std::isnan(u.q_measured)) st.ps_q[i] = u.q_measured;
    if (!std::isnan(u.power_sigma)) st.ps_sigma[i] = u.power_sigma;
    if (!std::isnan(u.p_sigma)) st.ps_psig[i] = u.p_sigma;
    if (!std::isnan(u.q_sigma)) st.ps_qsig[i] = u.q_sigma;
}

void apply_regulator_update(ComponentState& st, Idx i, TapRegulatorUpdate const& u) {
    if (u.status != na_IntS) st.reg_status[i] = u.status != 0;
    if (!std::isnan(u.u_set)) st.reg_u_set[i] = u.u_set;
    if (!std::isnan(u.u_band)) st.reg_u_band[i] = u.u_band;
    if (!std::isnan(u.line_drop_compensation_r)) st.reg_ldc_r[i] = u.line_drop_compensation_r;
    if (!std::isnan(u.line_drop_compensation_x)) st.reg_ldc_x[i] = u.line_drop_compensation_x;
}

BranchY branch_param(Model const& m, ComponentState const& st, Idx b, bool symmetric) {
    BranchY r;
    Idx const nl = m.n_line();
    if (b < nl) {
        auto const& s = m.lstat[b];
        bool const fs = st.line_fs[b], ts = st.line_ts[b];
        auto const p1 = pi_model(s.y1s, s.y1sh, s.tap, fs, ts);
        if (symmetric) {
            for (int i = 0; i < 4; ++i) r.v[i].m[0] = p1[i];
        } else {
            if (b >= generic_slot0(m))
                throw PowerGridError("The functionality is not implemented. asymmetric generic_branch\n");
            auto const p0 = pi_model(s.y0s, s.y0sh, 1.0, fs, ts);
            for (int i = 0; i < 4; ++i) r.v[i] = sm_tensor((2.0 * p1[i] + p0[i]) / 3.0, (p0[i] - p1[i]) / 3.0);
        }
        return r;
    }
    Idx const ti = b - nl;
    auto const& t = m.tstat[ti];
    bool const fs = st.tr_fs[ti], ts = st.tr_ts[ti];
    auto const ty = trafo_y(t, st.tr_tap[ti]);
    auto const p1 = pi_model(ty.ys, ty.ysh, ty.k * std::exp(cd(0, t.clock * kDeg30)), fs, ts);
    if (symmetric) {
        for (int i = 0; i < 4; ++i) r.v[i].m[0] = p1[i];
        return r;
    }
    // transformer.hpp asym_calc_param: sequence admittances mapped back to phase quantities
    auto const p2 = pi_model(ty.ys, ty.ysh, ty.k * std::exp(cd(0, -t.clock * kDeg30)), fs, ts);
    std::array<cd, 4> p0{};
    constexpr IntS wye_n = 1, delta = 2, zigzag_n = 4;
    double const k = ty.k;
    if (t.winding_from == wye_n && t.winding_to == wye_n) {
        double const ps0 = (t.clock == 2 || t.clock == 6 || t.clock == 10) ? 6.0 * kDeg30 : 0.0;
        cd const z0s = 1.0 / ty.ys + 3.0 * (t.z_ground_to + t.z_ground_from / k / k);
        p0 = pi_model(1.0 / z0s, ty.y0sh, k * std::exp(cd(0, ps0)), fs, ts);
    } else if (t.winding_from == wye_n && fs) {
        cd y0 = ty.y0sh;
        if (t.winding_to == delta) y0 += ty.ys;
        if (y0 != cd(0.0, 0.0)) p0[0] = (1.0 / (1.0 / y0 + 3.0 * t.z_ground_from / k / k)) / k / k;
    } else if (t.winding_to == wye_n && ts) {
        cd y0 = ty.y0sh;
        if (t.winding_from == delta) y0 += ty.ys;
        if (y0 != cd(0.0, 0.0)) p0[3] = 1.0 / (1.0 / y0 + 3.0 * t.z_ground_to);
    }
    if (t.winding_from == zigzag_n && fs) p0[0] = (1.0 / ((1.0 / ty.ys) * 0.1 + 3.0 * t.z_ground_from / k / k)) / k / k;
    if (t.winding_to == zigzag_n && ts) p0[3] = 1.0 / ((1.0 / ty.ys) * 0.1 + 3.0 * t.z_ground_to);
    double const low_b = -1e-8 * t.sn / kBaseP3 / t.uk;
    auto avail = [](IntS a, IntS b) { return a == wye_n ? (b == wye_n || b == delta) : a == zigzag_n; };
    if (!avail(t.winding_from, t.winding_to) && fs) p0[0] += cd(0.0, low_b);
    if (!avail(t.winding_to, t.winding_from) && ts) p0[3] += cd(0.0, low_b);
    Tensor3 sm, smi;
    cd const a[9] = {1.0, 1.0, 1.0, 1.0, kA2, kA, 1.0, kA, kA2};
    cd const ai[9] = {1.0, 1.0, 1.0, 1.0, kA, kA2, 1.0, kA2, kA};
    for (int i = 0; i < 9; ++i) {
        sm.m[i] = a[i];
        smi.m[i] = ai[i] / 3.0;
    }
    for (int i = 0; i < 4; ++i) {
        Tensor3 d;
        d.m[0] = p0[i];
        d.m[4] = p1[i];
        d.m[8] = p2[i];
        r.v[i] = matmul3(matmul3(sm, d), smi);
    }
    return r;
}

Tensor3 shunt_param(Model const& m, ComponentState const& st, Idx s, bool symmetric) {
    if (!st.sh_status[s]) return Tensor3{};
    double const u = m.node[m.sh_node[s]].u_rated;
    double const base_y = (kBaseP3 / u / kSqrt3) / (u / kSqrt3);
    cd const y1 = cd(st.sh_g1[s], st.sh_b1[s]) / base_y;
    if (symmetric) return scalar_tensor(y1, true);
    cd const y0 = cd(st.sh_g0[s], st.sh_b0[s]) / base_y;
    return sm_tensor((2.0 * y1 + y0) / 3.0, (y0 - y1) / 3.0);
}

void source_param(ComponentState const& st, Idx s, cd& y1, cd& y0) {
    double const z_abs = kBaseP3 / st.src_sk[s];
    double const rx = st.src_rx[s];
    double const x1 = z_abs / std::sqrt(rx * rx + 1.0);
    y1 = 1.0 / cd(x1 * rx, x1);
    y0 = y1 / st.src_z01[s];
}

std::array<cd, 3> load_power(Model const& m, ComponentState const& st, Idx l, bool symmetric) {
    std::array<cd, 3> r{};
    if (!st.lg_status[l]) return r;
    double const dir = m.lg_gen[l] ? 1.0 : -1.0;
    if (!m.lg_asym[l]) {
        cd const s = dir * cd(st.lg_p[l][0], st.lg_q[l][0]) / kBaseP3;
        if (std::isnan(s.real()) || std::isnan(s.imag())) return {cd(kNan, kNan), cd(kNan, kNan), cd(kNan, kNan)};
        return {s, s, s};
    }
    for (int ph = 0; ph < 3; ++ph) r[ph] = dir * cd(st.lg_p[l][ph], st.lg_q[l][ph]) / kBaseP1;
    if (symmetric) r[0] = (r[0] + r[1] + r[2]) / 3.0;
    return r;
}

}  // namespace pgm_amd
