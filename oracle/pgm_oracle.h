/*
 * pgm_oracle.h -- CPU FP64 restatement of power-grid-model's Newton-Raphson power-flow path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is called by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py. Nothing in the product
 * (power-grid-model_amd/) links, loads or calls it.
 *
 * Every function restates the reference algorithm; each source file cites the reference
 * file:line it follows. Record layouts below are the reference's row-based dataset layouts
 * (code_generation/data/attribute_classes/{input,update,output}.json), restated as plain C
 * structs so the same numpy buffers can be handed to the oracle and to the product.
 */
#ifndef PGM_ORACLE_H
#define PGM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------- input records ---------- */
typedef struct { int32_t id; double u_rated; } OrNodeInput;
typedef struct {
    int32_t id, from_node, to_node; int8_t from_status, to_status;
    double r1, x1, c1, tan1, r0, x0, c0, tan0, i_n;
} OrLineInput;
typedef struct { int32_t id, from_node, to_node; int8_t from_status, to_status; } OrLinkInput;
typedef struct {
    int32_t id, from_node, to_node; int8_t from_status, to_status;
    double r1, x1, g1, b1, k, theta, sn;
} OrGenericBranchInput;
typedef struct {
    int32_t id, from_node, to_node; int8_t from_status, to_status;
    double u1, u2, sn, uk, pk, i0, p0, i0_zero_sequence, p0_zero_sequence;
    int8_t winding_from, winding_to, clock, tap_side, tap_pos, tap_min, tap_max, tap_nom;
    double tap_size, uk_min, uk_max, pk_min, pk_max;
    double r_grounding_from, x_grounding_from, r_grounding_to, x_grounding_to;
} OrTransformerInput;
typedef struct { int32_t id, node; int8_t status; double g1, b1, g0, b0; } OrShuntInput;
typedef struct { int32_t id, node; int8_t status; double u_ref, u_ref_angle, sk, rx_ratio, z01_ratio; } OrSourceInput;
typedef struct { int32_t id, node; int8_t status, type; double p_specified, q_specified; } OrSymLoadGenInput;
typedef struct { int32_t id, node; int8_t status, type; double p_specified[3], q_specified[3]; } OrAsymLoadGenInput;
typedef struct {
    int32_t id, regulated_object; int8_t status, control_side;
    double u_set, u_band, line_drop_compensation_r, line_drop_compensation_x;
} OrTapRegulatorInput;

typedef struct { int32_t id, measured_object; double u_sigma, u_measured, u_angle_measured; } OrSymVoltageSensorInput;
typedef struct {
    int32_t id, measured_object; int8_t measured_terminal_type;
    double power_sigma, p_measured, q_measured, p_sigma, q_sigma;
} OrSymPowerSensorInput;

/* ---------- update records ---------- */
typedef struct { int32_t id; int8_t from_status, to_status; } OrBranchUpdate;
typedef struct { int32_t id; int8_t from_status, to_status, tap_pos; } OrTransformerUpdate;
typedef struct { int32_t id; int8_t status; double g1, b1, g0, b0; } OrShuntUpdate;
typedef struct { int32_t id; int8_t status; double u_ref, u_ref_angle, sk, rx_ratio, z01_ratio; } OrSourceUpdate;
typedef struct { int32_t id; int8_t status; double p_specified, q_specified; } OrSymLoadGenUpdate;
typedef struct { int32_t id; int8_t status; double p_specified[3], q_specified[3]; } OrAsymLoadGenUpdate;
typedef struct {
    int32_t id; int8_t status; double u_set, u_band, line_drop_compensation_r, line_drop_compensation_x;
} OrTapRegulatorUpdate;

typedef struct { int32_t id; double u_sigma, u_measured, u_angle_measured; } OrSymVoltageSensorUpdate;
typedef struct { int32_t id; double power_sigma, p_measured, q_measured, p_sigma, q_sigma; } OrSymPowerSensorUpdate;

/* ---------- output records ---------- */
typedef struct { int32_t id; int8_t energized; double u_pu, u, u_angle, p, q; } OrNodeOutputSym;
typedef struct { int32_t id; int8_t energized; double u_pu[3], u[3], u_angle[3], p[3], q[3]; } OrNodeOutputAsym;
typedef struct {
    int32_t id; int8_t energized; double loading;
    double p_from, q_from, i_from, s_from, p_to, q_to, i_to, s_to;
} OrBranchOutputSym;
typedef struct {
    int32_t id; int8_t energized; double loading;
    double p_from[3], q_from[3], i_from[3], s_from[3], p_to[3], q_to[3], i_to[3], s_to[3];
} OrBranchOutputAsym;
typedef struct { int32_t id; int8_t energized; double p, q, i, s, pf; } OrApplianceOutputSym;
typedef struct { int32_t id; int8_t energized; double p[3], q[3], i[3], s[3], pf[3]; } OrApplianceOutputAsym;

typedef struct { int32_t id; int8_t energized; double u_residual, u_angle_residual; } OrVoltageSensorOutputSym;
typedef struct { int32_t id; int8_t energized; double p_residual, q_residual; } OrPowerSensorOutputSym;

/* component slots, in the reference's container order (all_components.hpp:37-40), the link and the generic
 * branch last */
enum OrComp {
    OR_NODE = 0, OR_LINE, OR_TRANSFORMER, OR_SHUNT, OR_SOURCE, OR_SYM_GEN, OR_ASYM_GEN, OR_SYM_LOAD, OR_ASYM_LOAD,
    OR_LINK, OR_GENERIC_BRANCH, OR_N_COMP
};

/* one dataset: per component a row buffer and a count (per scenario for update/output batches) */
typedef struct {
    void const* data[OR_N_COMP];
    int64_t count[OR_N_COMP];
} OrConstDataset;
typedef struct {
    void* data[OR_N_COMP];
    int64_t count[OR_N_COMP];
} OrMutableDataset;

/* Batch power flow, Newton-Raphson, FP64.
 * update: NULL for a single calculation; otherwise uniform batch, update->count = rows per scenario.
 * output: row buffers of n_scenarios * n_component rows (sym or asym output records).
 * iterations[n_scenarios]: max NR iteration count over the math models of each scenario (0 on error).
 * status[n_scenarios]: 0 ok, 1 SparseMatrixError, 2 IterationDiverge, 3 other error.
 * messages: n_scenarios * msg_stride bytes, error text per scenario (may be NULL).
 * Returns 0, or a negative number for an error that is not per scenario (message in messages[0..]). */
int or_power_flow(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                  int64_t n_scenarios, int symmetric, double err_tol, int64_t max_iter, int64_t threading,
                  OrMutableDataset const* output, int64_t* iterations, int32_t* status, char* messages,
                  int64_t msg_stride);

/* or_power_flow with a calculation method: 0 Newton-Raphson, 1 linear (linear_pf_solver.hpp),
 * 2 iterative current, 3 linear current (iterative_current_pf_solver.hpp). */
int or_power_flow_method(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                         int64_t n_scenarios, int symmetric, double err_tol, int64_t max_iter, int64_t threading,
                         int method, OrMutableDataset const* output, int64_t* iterations, int32_t* status,
                         char* messages, int64_t msg_stride);

/* or_power_flow_method with automatic tap changing (optimizer/tap_position_optimizer.hpp): strategy 1 any
 * valid tap, 2 min voltage tap, 3 max voltage tap, 4 fast any tap. regulators: n_regulator
 * transformer_tap_regulator input records; reg_update: NULL, or n_scenarios * reg_update_count update
 * records. The output is that of the last power flow at the final taps; tap_out[n_scenarios * n_regulator]
 * receives the final tap of every regulated transformer (-128 for regulators that did not regulate).
 * status 4: the optimizer failed (MaxIterationReached, invalid regulator configuration). */
int or_power_flow_tap(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                      int64_t n_scenarios, int symmetric, double err_tol, int64_t max_iter, int64_t threading,
                      int method, int strategy, OrTapRegulatorInput const* regulators, int64_t n_regulator,
                      OrTapRegulatorUpdate const* reg_update, int64_t reg_update_count,
                      OrMutableDataset const* output, int8_t* tap_out, int64_t* iterations, int32_t* status,
                      char* messages, int64_t msg_stride);
/* Batch Newton-Raphson state estimation, symmetric, FP64 (math_solver/newton_raphson_se_solver.hpp with
 * measured_values.hpp and observability.hpp; the gain system is solved densely). The sensors come beside the
 * datasets: vs / ps input records, vs_upd / ps_upd NULL or n_scenarios * *_upd_count update records; outputs as
 * or_power_flow (symmetric records) plus vs_out / ps_out, n_scenarios * n_vs / n_ps records.
 * status: 0 ok, 1 SparseMatrixError, 2 IterationDiverge, 3 other error (NotObservableError included). */
int or_state_estimation(OrConstDataset const* input, double system_frequency, OrConstDataset const* update,
                        int64_t n_scenarios, double err_tol, int64_t max_iter, int64_t threading,
                        OrSymVoltageSensorInput const* vs, int64_t n_vs, OrSymVoltageSensorUpdate const* vs_upd,
                        int64_t vs_upd_count, OrSymPowerSensorInput const* ps, int64_t n_ps,
                        OrSymPowerSensorUpdate const* ps_upd, int64_t ps_upd_count, OrMutableDataset const* output,
                        OrVoltageSensorOutputSym* vs_out, OrPowerSensorOutputSym* ps_out, int64_t* iterations,
                        int32_t* status, char* messages, int64_t msg_stride);

/* The state estimation's gain system: 1 solves it densely (partial pivoting), 0 (the default) with the block
 * sparse LU, pivot perturbation and iterative refinement as the reference (sparse_lu_solver.hpp). */
void or_set_se_dense(int dense);
/* Sparse gain factorisations that perturbed a pivot and the refinement solves they took since the last reset. */
void or_se_counters(long long* perturbed, long long* refine_solves, int reset);

/* Minimum-degree ordering (sparse_ordering.hpp:183-219). The adjacency map is given as sorted keys with
 * CSR neighbour lists. alpha_out has room for n_vertices (all vertices), fills_out for 2*max_fills.
 * Returns the number of fills, or -1 if max_fills is too small. n_alpha_out receives the order length. */
int64_t or_minimum_degree_ordering(int64_t n_keys, int64_t const* keys, int64_t const* indptr,
                                   int64_t const* neighbours, int64_t* alpha_out, int64_t* n_alpha_out,
                                   int64_t* fills_out, int64_t max_fills);

/* Component-level topology (topology.hpp:196-368): branches (from,to,from_status,to_status,phase_shift),
 * sources (node,status). Outputs node group/pos, number of math models, and per model the fill-ins.
 * fill_in_out receives (group, from, to) triples. Returns the number of fill-ins or -1 on overflow. */
int64_t or_build_topology(int64_t n_node, int64_t n_branch, int64_t const* branch_from, int64_t const* branch_to,
                          int8_t const* branch_from_status, int8_t const* branch_to_status,
                          double const* branch_phase_shift, int64_t n_source, int64_t const* source_node,
                          int8_t const* source_status, int64_t* node_group, int64_t* node_pos,
                          int64_t* n_math_models, int64_t* fill_in_out, int64_t max_fill);

/* Y-bus structure (y_bus.hpp:122-293) of one math model. Outputs must hold n_bus+1 / nnz_lu_max entries.
 * Returns nnz_lu, or -1 on overflow. */
int64_t or_ybus_structure(int64_t n_bus, int64_t n_branch, int64_t const* branch_bus, int64_t n_shunt,
                          int64_t const* shunt_bus, int64_t n_fill, int64_t const* fill_in, int64_t nnz_lu_max,
                          int64_t* row_indptr_lu, int64_t* col_indices_lu, int64_t* diag_lu, int64_t* map_lu_y_bus,
                          int64_t* lu_transpose_entry, int64_t* row_indptr, int64_t* col_indices,
                          int64_t* bus_entry, int64_t* y_bus_entry_indptr, int64_t* nnz_out);

/* Block sparse LU factorise + solve (sparse_lu_solver.hpp:346-495, 769-827) with block size nb
 * (1, 2 or 6), data blocks column-major. Returns 0 on success, 1 on SparseMatrixError. */
int or_block_lu_solve(int64_t nb, int64_t n, int64_t const* row_indptr, int64_t const* col_indices,
                      int64_t const* diag_lu, double* data, double const* rhs, double* x);

/* Newton-Raphson on a hand-built math model (newton_raphson_pf_solver.hpp, iterative_pf_solver.hpp).
 * Complex numbers as (re, im) pairs. phases p = 1 (sym) or 3 (asym); tensors are p*p complex, row-major.
 * branch_param: n_branch * 4 tensors (ff, ft, tf, tt); shunt_param: n_shunt tensors;
 * source_y: n_source * 2 complex (y1, y0); u_source: n_source complex; s_inj: n_load * p complex.
 * Outputs: u (n_bus*p), branch flows s_f,s_t,i_f,i_t (n_branch*p each), shunt s,i, load s,i, source s,i,
 * bus_injection (n_bus*p). Returns iterations >= 1, or -1 SparseMatrixError, -2 IterationDiverge. */
int64_t or_nr_math_model(int p, int64_t n_bus, int64_t n_branch, int64_t const* branch_bus,
                         double const* branch_param, int64_t n_shunt, int64_t const* shunt_bus,
                         double const* shunt_param, int64_t n_source, int64_t const* source_bus,
                         double const* source_y, double const* u_source, int64_t n_load, int64_t const* load_bus,
                         int8_t const* load_type, double const* s_inj, int64_t n_fill, int64_t const* fill_in,
                         double err_tol, int64_t max_iter, double* u, double* branch_flow, double* shunt_flow,
                         double* load_flow, double* source_flow, double* bus_injection);

#ifdef __cplusplus
}
#endif

#endif
