// model.hpp -- host-side component model of the power-flow path.
//
// Holds the caller's input records, the id index, the per-unit parameters and the mutable component
// state (statuses, taps, powers) that updates act on. Per-unit formulas follow the reference
// components: line.hpp:26-36, branch.hpp:196-245, transformer.hpp:32-320, transformer_utils.hpp,
// source.hpp:40-48, load_gen.hpp:86-139, shunt.hpp:18-100, common.hpp:81-107.
#pragma once

#include "dataset.hpp"
#include "records.hpp"

#include <array>
#include <complex>
#include <unordered_map>
#include <vector>

namespace pgm_amd {

using cd = std::complex<double>;

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt3 = 1.7320508075688772935;
constexpr double kDeg30 = kPi / 6.0;
constexpr double kNumTol = 1e-8;
constexpr double kBaseP3 = 1e6;
constexpr double kBaseP1 = 1e6 / 3.0;
inline cd const kA{-0.5, kSqrt3 / 2.0};
inline cd const kA2{-0.5, -kSqrt3 / 2.0};

// 3x3 complex tensor, row-major; the symmetric case uses only element 0
struct Tensor3 {
    std::array<cd, 9> m{};
};
// branch admittance parameters ff, ft, tf, tt
struct BranchY {
    std::array<Tensor3, 4> v{};
};

// lines, behind them the links (y1 = y0 = y_link, no shunt; base currents of both ends, i_n infinite so
// that the loading is 0), then the generic branches (ratio k e^(j theta), loading max s / sn; symmetric only)
struct LineStatic {
    cd y1s, y1sh, y0s, y0sh;
    double base_i, i_n, base_i_to;
    cd tap{1.0, 0.0};
    double shift = 0.0;
    double sn = 0.0;  // generic branch: rated power (infinite when not given)
};
struct TrafoStatic {
    double u1, u2, sn, tap_size, uk, pk, i0, p0, i0_zs, p0_zs, uk_min, uk_max, pk_min, pk_max;
    IntS winding_from, winding_to, clock, tap_side, tap_min, tap_max, tap_nom, tap_direction;
    double base_i_from, base_i_to, nominal_ratio;
    cd z_ground_from, z_ground_to;
};

// mutable state touched by updates
struct ComponentState {
    std::vector<IntS> line_fs, line_ts;
    std::vector<IntS> tr_fs, tr_ts, tr_tap;
    std::vector<IntS> sh_status;
    std::vector<double> sh_g1, sh_b1, sh_g0, sh_b0;
    std::vector<IntS> src_status;
    std::vector<double> src_uref, src_uang, src_sk, src_rx, src_z01;
    // load_gens in container order: sym_gen, asym_gen, sym_load, asym_load
    std::vector<IntS> lg_status;
    std::vector<std::array<double, 3>> lg_p, lg_q;
    // transformer tap regulators
    std::vector<IntS> reg_status;
    std::vector<double> reg_u_set, reg_u_band, reg_ldc_r, reg_ldc_x;
    // sym_voltage_sensor: sigma and magnitude in V, angle in rad, as given (NaN: not given)
    std::vector<double> vs_sigma, vs_u, vs_angle;
    // sym_power_sensor: sigmas and powers in W / var, as given
    std::vector<double> ps_sigma, ps_p, ps_q, ps_psig, ps_qsig;
};

struct Model {
    double freq = 50.0;
    std::vector<NodeInput> node;
    std::vector<LineInput> line;
    std::vector<LinkInput> link;
    std::vector<GenericBranchInput> gbranch;
    std::vector<FaultInput> fault;  // kept for the output; a power flow does not use it
    std::vector<TransformerInput> trafo;
    std::vector<ShuntInput> shunt;
    std::vector<SourceInput> source;
    // load_gens flattened in container order
    std::vector<ID> lg_id;
    std::vector<IntS> lg_type;
    std::vector<bool> lg_asym, lg_gen;
    std::vector<Idx> lg_node;
    std::array<Idx, 4> lg_count{};  // sym_gen, asym_gen, sym_load, asym_load
    std::array<Idx, 4> lg_offset{};
    std::unordered_map<ID, std::pair<int, Idx>> ids;  // id -> (component slot, index within the component)
    std::vector<Idx> line_f, line_t, tr_f, tr_t, sh_node, src_node;
    std::vector<LineStatic> lstat;
    std::vector<TrafoStatic> tstat;
    // transformer tap regulators: the regulated transformer and the rated voltage of its control-side node
    std::vector<TapRegulatorInput> regulator;
    std::vector<Idx> reg_trafo;
    std::vector<double> reg_u_rated;
    // sensors (state estimation): the measured node of every voltage sensor; the terminal type and the
    // measured object of every power sensor (node, branch slot, shunt, source or load_gen index)
    std::vector<ID> vs_id, ps_id;
    std::vector<Idx> vs_node, ps_obj;
    std::vector<int> ps_term;
    ComponentState base;

    Idx n_node() const { return Idx(node.size()); }
    // branch slots: lines, then links and generic branches (line slots too), then transformers
    Idx n_line() const { return Idx(line_f.size()); }
    Idx n_branch() const { return Idx(line_f.size() + trafo.size()); }
    Idx n_lg() const { return Idx(lg_id.size()); }
    Idx count(int comp) const;
    ID comp_id(int comp, Idx i) const;
};

Model build_model(PGM_ConstDataset const& input, double system_frequency);

// per-unit parameters
BranchY branch_param(Model const& m, ComponentState const& st, Idx branch, bool symmetric);
Tensor3 shunt_param(Model const& m, ComponentState const& st, Idx shunt, bool symmetric);
void source_param(ComponentState const& st, Idx s, cd& y1, cd& y0);
// injected power of a load_gen in per unit (per phase for asymmetric), zero when switched off
std::array<cd, 3> load_power(Model const& m, ComponentState const& st, Idx l, bool symmetric);
IntS tap_limit(TrafoStatic const& t, IntS tap);
double branch_phase_shift(Model const& m, Idx branch);

// lookup of an update row: component index, or throws IDNotFound / IDWrongType
Idx find_component(Model const& m, int comp, ID id);
// the state slot an update row acts on: a link's rows act on its line slot behind the lines
inline void line_slot(Model const& m, int& comp, Idx& i) {
    if (comp == C_LINK) {
        comp = C_LINE;
        i += Idx(m.line.size());
    } else if (comp == C_GENERIC_BRANCH) {
        comp = C_LINE;
        i += Idx(m.line.size() + m.link.size());
    }
}
// first line slot of the generic branches (they are no edges of the tap changer's ranking graph)
inline Idx generic_slot0(Model const& m) { return Idx(m.line.size() + m.link.size()); }
// apply one transformer_tap_regulator update record (transformer_tap_regulator.hpp:38-46: NaN / na keep
// the current value)
void apply_regulator_update(ComponentState& st, Idx i, TapRegulatorUpdate const& u);
// apply one sensor update record (voltage_sensor.hpp:88-100, power_sensor.hpp:94-105: NaN keeps the value)
void apply_voltage_sensor_update(ComponentState& st, Idx i, SymVoltageSensorUpdate const& u);
void apply_power_sensor_update(ComponentState& st, Idx i, SymPowerSensorUpdate const& u);

}  // namespace pgm_amd
