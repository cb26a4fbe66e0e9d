// records.hpp -- row-based dataset records of the power-flow path.
//
// The layouts are the reference's dataset layouts (code_generation/data/attribute_classes/
// {input,update,output}.json; structs in power_grid_model_c/power_grid_model/include/power_grid_model/
// auxiliary/{input,update,output}.hpp). They have to be byte-identical because callers hand the same
// numpy / C buffers to either library; the C compiler's natural alignment produces the same offsets.
#pragma once

#include <cstdint>
#include <limits>

namespace pgm_amd {

using ID = int32_t;
using IntS = int8_t;
using Idx = int64_t;

constexpr IntS na_IntS = std::numeric_limits<IntS>::min();
constexpr ID na_IntID = std::numeric_limits<ID>::min();

// ---------- input ----------
struct NodeInput {
    ID id;
    double u_rated;
};
struct LineInput {
    ID id, from_node, to_node;
    IntS from_status, to_status;
    double r1, x1, c1, tan1, r0, x0, c0, tan0, i_n;
};
// link: a branch of very high admittance between two nodes (component/link.hpp)
struct LinkInput {
    ID id, from_node, to_node;
    IntS from_status, to_status;
};
// generic branch: a pi model with off-nominal ratio k and angle shift theta (component/generic_branch.hpp)
struct GenericBranchInput {
    ID id, from_node, to_node;
    IntS from_status, to_status;
    double r1, x1, g1, b1, k, theta, sn;
};
struct TransformerInput {
    ID id, from_node, to_node;
    IntS from_status, to_status;
    double u1, u2, sn, uk, pk, i0, p0, i0_zero_sequence, p0_zero_sequence;
    IntS winding_from, winding_to, clock, tap_side, tap_pos, tap_min, tap_max, tap_nom;
    double tap_size, uk_min, uk_max, pk_min, pk_max;
    double r_grounding_from, x_grounding_from, r_grounding_to, x_grounding_to;
};
struct ShuntInput {
    ID id, node;
    IntS status;
    double g1, b1, g0, b0;
};
struct SourceInput {
    ID id, node;
    IntS status;
    double u_ref, u_ref_angle, sk, rx_ratio, z01_ratio;
};
struct SymLoadGenInput {
    ID id, node;
    IntS status, type;
    double p_specified, q_specified;
};
struct AsymLoadGenInput {
    ID id, node;
    IntS status, type;
    double p_specified[3], q_specified[3];
};
struct TapRegulatorInput {
    ID id, regulated_object;
    IntS status, control_side;
    double u_set, u_band, line_drop_compensation_r, line_drop_compensation_x;
};

// fault: input of short-circuit calculations; a power flow only checks it and reports it not energised
struct FaultInput {
    ID id;
    IntS status, fault_type, fault_phase;
    ID fault_object;
    double r_f, x_f;
};

// symmetric voltage and power sensors, inputs of state estimation (auxiliary/input.hpp VoltageSensorInput,
// PowerSensorInput); measured_terminal_type follows MeasuredTerminalType (common/enum.hpp:43-54)
struct SymVoltageSensorInput {
    ID id, measured_object;
    double u_sigma, u_measured, u_angle_measured;
};
struct SymPowerSensorInput {
    ID id, measured_object;
    IntS measured_terminal_type;
    double power_sigma, p_measured, q_measured, p_sigma, q_sigma;
};

// ---------- update ----------
struct BranchUpdate {
    ID id;
    IntS from_status, to_status;
};
struct TransformerUpdate {
    ID id;
    IntS from_status, to_status, tap_pos;
};
struct ShuntUpdate {
    ID id;
    IntS status;
    double g1, b1, g0, b0;
};
struct SourceUpdate {
    ID id;
    IntS status;
    double u_ref, u_ref_angle, sk, rx_ratio, z01_ratio;
};
struct SymLoadGenUpdate {
    ID id;
    IntS status;
    double p_specified, q_specified;
};
struct AsymLoadGenUpdate {
    ID id;
    IntS status;
    double p_specified[3], q_specified[3];
};
struct TapRegulatorUpdate {
    ID id;
    IntS status;
    double u_set, u_band, line_drop_compensation_r, line_drop_compensation_x;
};

struct SymVoltageSensorUpdate {
    ID id;
    double u_sigma, u_measured, u_angle_measured;
};
struct SymPowerSensorUpdate {
    ID id;
    double power_sigma, p_measured, q_measured, p_sigma, q_sigma;
};

// ---------- output ----------
template <int P> struct NodeOutput {
    ID id;
    IntS energized;
    double u_pu[P], u[P], u_angle[P], p[P], q[P];
};
template <int P> struct BranchOutput {
    ID id;
    IntS energized;
    double loading;
    double p_from[P], q_from[P], i_from[P], s_from[P], p_to[P], q_to[P], i_to[P], s_to[P];
};
template <int P> struct ApplianceOutput {
    ID id;
    IntS energized;
    double p[P], q[P], i[P], s[P], pf[P];
};
// fault output of a power flow (fault.hpp:70-78): id and energized 0
struct FaultOutput {
    ID id;
    IntS energized;
};
// the same record for symmetric and asymmetric output
struct TapRegulatorOutput {
    ID id;
    IntS energized, tap_pos;
};

// state-estimation outputs of the sensors (auxiliary/output.hpp VoltageSensorOutput, PowerSensorOutput)
template <int P> struct VoltageSensorOutput {
    ID id;
    IntS energized;
    double u_residual[P], u_angle_residual[P];
};
template <int P> struct PowerSensorOutput {
    ID id;
    IntS energized;
    double p_residual[P], q_residual[P];
};

static_assert(sizeof(SymVoltageSensorInput) == 32 && sizeof(SymVoltageSensorUpdate) == 32);
static_assert(sizeof(SymPowerSensorInput) == 56 && sizeof(SymPowerSensorUpdate) == 48);
static_assert(sizeof(VoltageSensorOutput<1>) == 24 && sizeof(PowerSensorOutput<1>) == 24);
static_assert(sizeof(NodeInput) == 16);
static_assert(sizeof(LineInput) == 88);
static_assert(sizeof(LinkInput) == 16);
static_assert(sizeof(FaultInput) == 32 && sizeof(FaultOutput) == 8);
static_assert(sizeof(GenericBranchInput) == 72);
static_assert(sizeof(TransformerInput) == 168);
static_assert(sizeof(SourceInput) == 56);
static_assert(sizeof(SymLoadGenInput) == 32);
static_assert(sizeof(AsymLoadGenInput) == 64);
static_assert(sizeof(SymLoadGenUpdate) == 24);
static_assert(sizeof(AsymLoadGenUpdate) == 56);
static_assert(sizeof(NodeOutput<1>) == 48 && sizeof(NodeOutput<3>) == 128);
static_assert(sizeof(BranchOutput<1>) == 80 && sizeof(BranchOutput<3>) == 208);
static_assert(sizeof(ApplianceOutput<1>) == 48 && sizeof(ApplianceOutput<3>) == 128);
static_assert(sizeof(TapRegulatorInput) == 48 && sizeof(TapRegulatorUpdate) == 40 && sizeof(TapRegulatorOutput) == 8);

// component slots, in the reference container order (all_components.hpp:37-40), except that the link and
// the generic branch come last: their rows live behind the lines in the branch arrays, so no slot before
// them moves
enum Comp : int {
    C_NODE = 0,
    C_LINE,
    C_TRANSFORMER,
    C_SHUNT,
    C_SOURCE,
    C_SYM_GEN,
    C_ASYM_GEN,
    C_SYM_LOAD,
    C_ASYM_LOAD,
    C_LINK,
    C_GENERIC_BRANCH,
    N_COMP
};
inline char const* const comp_names[N_COMP] = {"node",     "line",      "transformer", "shunt",
                                               "source",   "sym_gen",   "asym_gen",    "sym_load",
                                               "asym_load", "link",     "generic_branch"};
// transformer_tap_regulator: not part of a math model (no slot above); it steers the tap changer
inline char const* const kTapRegulatorName = "transformer_tap_regulator";
constexpr int C_TAP_REGULATOR = N_COMP;  // slot of its ids in the id index
inline char const* const kFaultName = "fault";
constexpr int C_FAULT = N_COMP + 1;  // slot of fault ids in the id index
// sensors: inputs of state estimation only (no slot in the power-flow math model)
inline char const* const kVoltageSensorName = "sym_voltage_sensor";
constexpr int C_VOLTAGE_SENSOR = N_COMP + 2;
inline char const* const kPowerSensorName = "sym_power_sensor";
constexpr int C_POWER_SENSOR = N_COMP + 3;
// MeasuredTerminalType
enum Terminal : int {
    T_BRANCH_FROM = 0,
    T_BRANCH_TO = 1,
    T_SOURCE = 2,
    T_SHUNT = 3,
    T_LOAD = 4,
    T_GENERATOR = 5,
    T_BRANCH3_1 = 6,
    T_BRANCH3_2 = 7,
    T_BRANCH3_3 = 8,
    T_NODE = 9
};

}  // namespace pgm_amd
