// kernels.hpp -- device data of one group (one topology, a set of scenarios) and the launchers of
// the Newton-Raphson kernels in kernels.hip.
//
// Layout in HBM: every per-scenario array is scenario-interleaved, index (item * S_pad + s), so the 64
// lanes of a wave (64 consecutive scenarios) read and write contiguous 512-byte (FP64) lines. Arrays
// shared by all scenarios of the group (pattern, admittances, schedule) are read with uniform loads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgm_amd {

constexpr int kWaveLanes = 64;  // lanes of a gfx950 wavefront: one scenario (or task) each
// INIT: linear system (NR initial guess, linear method); NR: Newton-Raphson step; RAW: block LU test
// hook; ICF: iterative current, factorise Y + y_src and solve the current residual; RESID: iterative
// current with the stored factors (u += x after both)
// MODE_NRA: the Newton-Raphson Jacobian assembled and factorised at the converged state of converged
// lanes, every block kept (the adjoint solves with it)
// MODE_REFINE: a pass of iterative refinement with the stored factors (forward: substitution of the residual
// held in the forward vectors; backward: back substitution, the correction added to the accumulated step)
enum SweepMode : int { MODE_INIT = 0, MODE_NR = 1, MODE_RAW = 2, MODE_ICF = 3, MODE_RESID = 4, MODE_NRA = 5,
                       MODE_REFINE = 6, MODE_RAWP = 7 };
// RAWP: a raw system whose pivot perturbation, once it happened in a pair, holds for the pair's later pivots
// (the reference's per-solver flag; the state estimation's gain systems)
// REFINE: the pair's Newton-Raphson step came from a factorisation with perturbed pivots and waits for its
// iterative refinement (it takes no part in the sweeps meanwhile)
enum DoneState : int { ACTIVE = 0, CONVERGED = 1, FAILED = 2, REFINE = 3 };
// pivot perturbation of FP32 factors: in the full-pivot factorisation of a diagonal block, a chosen pivot below
// this fraction of the row's scale (the largest entry of its assembled diagonal Jacobian block, before
// elimination; the rows are close to diagonally dominant) takes that magnitude with its own sign, and the step
// is refined in FP64. 2^-14, about 512 FP32 epsilons, as the reference's threshold (1e-13 of its matrix norm)
// is about 450 FP64 epsilons: a pivot that small relative to its row is a cancellation of much larger eliminated
// terms, which the stored FP32 blocks carry only to about that relative level
constexpr double kPivotPerturbation = 6.103515625e-05;  // 2^-14
// the refinement stops once the backward error of the step is at most this (the reference's convergence
// criterion of its refinement) and fails after this many passes (one initial solve and five refinements)
constexpr double kRefineTol = 1e-13;
// refinement of every step on request (no perturbed pivot): down to this backward error, near the FP64
// rounding level, so that the steps are those of an FP64 factorisation; the step is applied after the last
// pass whatever its backward error
constexpr double kRefineTolSteps = 1e-15;
constexpr int kRefineMaxPasses = 6;
enum ErrCode : int { ERR_NONE = 0, ERR_SPARSE = 1, ERR_DIVERGE = 2 };

struct DevGroup {
    int P = 1;         // phases
    int n = 0;         // buses
    int E = 0;         // LU entries
    int n_mm = 0;      // math models
    int S = 0;         // scenarios in the group
    int S_pad = 0;     // S rounded up to 64
    int fp32 = 0;      // factor storage precision
    // structure (shared); row p's entries: row_ptr[p], + stride, ... < row_end[p], stride =
    // row_stride[p] (nullptr: 1)
    int const* row_ptr = nullptr;
    int const* row_end = nullptr;
    int const* row_stride = nullptr;
    int const* col = nullptr;
    int const* diag = nullptr;
    int const* transpose = nullptr;
    int8_t const* fill = nullptr;
    double const* ydata = nullptr;  // E * P*P * 2
    // union groups (scenarios that differ only in branch statuses share one superset structure): Y
    // values per lane [E][P*P*2][S_pad] instead of ydata, and buses without a source in a lane
    // [n][S_pad] (solved as identity rows, reported as not energised)
    double const* ylane = nullptr;
    uint8_t const* dead = nullptr;
    int const* schur_ptr = nullptr;
    int const* schur = nullptr;     // 3 ints per update
    int const* task_ptr = nullptr;
    int const* task_pivot = nullptr;
    int const* task_mm = nullptr;
    int const* level_task = nullptr;  // tasks listed level by level
    int8_t const* keep = nullptr;     // per pivot: one U entry, no off-diagonal Schur target
    // iterative current: phase shift and math model of every bus, first bus of every math model,
    // bus of every source
    double const* phase = nullptr;
    int const* bus_mm = nullptr;
    int const* mm_src_ptr = nullptr;  // sources of every math model: mm_src[mm_src_ptr[k] ...]
    int const* mm_src = nullptr;
    int const* src_bus = nullptr;
    int const* load_ptr = nullptr;
    int8_t const* load_type = nullptr;
    int const* src_ptr = nullptr;
    // per scenario
    double* u = nullptr;        // [n][P][2][S_pad]
    double* xv = nullptr;       // [n][P][2][S_pad] (theta, V)
    double* y = nullptr;        // [n][2P][S_pad]
    void* lu = nullptr;         // [E][4P^2][S_pad] double or float
    uint32_t* perm = nullptr;   // [n][S_pad]
    double const* s_inj = nullptr;  // [n_load][P][2][S_pad]
    double const* u_src = nullptr;  // [n_src][2][S_pad]
    double const* y_src = nullptr;  // [n_src][4][S_pad] (y1 re, im, y0 re, im)
    int* done = nullptr;            // [n_mm][S_pad]
    int* err = nullptr;             // [n_mm][S_pad]
    int* iters = nullptr;           // [n_mm][S_pad]
    unsigned long long* devmax = nullptr;  // [n_mm][S_pad], bits of a non-negative double
    double* lastdev = nullptr;      // [n_mm][S_pad]
    // pivot perturbation and iterative refinement (FP32 factors): a flag per pair that the factorisation of
    // this iteration perturbed a pivot; per pair the largest denominator and the backward error of the last
    // residual (bits of non-negative doubles); the accumulated step and the residual's denominators (|rhs| +
    // |J| |x|), [n][2P][S_pad], made when a refinement first happens
    int* pert = nullptr;
    unsigned long long* rmax = nullptr;
    unsigned long long* berr = nullptr;
    double* xacc = nullptr;
    double* rden = nullptr;
    int refine_all = 0;  // 1: every FP32 step is held and refined, perturbed or not
    // raw block LU (test hook and the state-estimation gain matrix): input blocks and rhs; FP64 raw systems
    // perturb pivots below pthr[pair] when pthr is given, and flag the pair in pert
    double const* raw_a = nullptr;  // [E][4P^2][S_pad]
    double const* raw_b = nullptr;  // [n][2P][S_pad]
    double const* pthr = nullptr;   // [n_mm][S_pad]
    // union groups whose lanes close branches that the shared structure has open (symmetric, radial
    // structure): every such branch adds two off-pattern admittances (a, b) and (b, a) in its lane. The
    // sweeps factorise the matrix without them; a low-rank correction between the forward and the
    // backward sweep (Woodbury) makes the solve exact. Per lane: the extra entries (row and column bus,
    // admittance, and the Jacobian or admittance block the forward sweep computes for it) and the
    // distinct buses they touch; per bus: whether any lane has an extra entry in that row, and the
    // chain of pivots from the bus to the root of its elimination tree (node, U entry of the previous
    // chain node in the node's column)
    int8_t const* xbus = nullptr;   // [n]
    int const* xcnt = nullptr;      // [S_pad]
    int const* xrow = nullptr;      // [S_pad][kExtraMax] bus
    int const* xcol = nullptr;      // [S_pad][kExtraMax] bus
    double const* xy = nullptr;     // [S_pad][kExtraMax][2]
    double* xblk = nullptr;         // [S_pad][kExtraMax][4]
    int const* xm = nullptr;        // [S_pad] distinct buses
    int const* xb = nullptr;        // [S_pad][kExtraMax]
    int const* ch_off = nullptr;    // [n] first chain step of the bus (-1: no chain)
    int const* ch_len = nullptr;    // [n]
    int const* ch_node = nullptr;   // chain steps
    int const* ch_ue = nullptr;     // U entry (previous node, node); -1 at the first step
    int ch_max = 0;                 // longest chain
    double* xst = nullptr;          // [lanes with extras][kExtraMax][ch_max][8]: S and T of every step
};
// extra (off-pattern) entries per lane of a union group: two per branch closed against the structure
constexpr int kExtraMax = 8;
// one block per lane with extra entries: (group, lane, slot of the lane's scratch)
struct ExtraWork {
    int group, lane, slot;
};
// low-rank correction of the union lanes with extra entries, between the forward and backward sweeps
// (MODE_INIT: admittance blocks, MODE_NR: Jacobian blocks)
void launch_lowrank(DevGroup const* groups, ExtraWork const* work, int n_work, bool fp32, hipStream_t st);

// ids of an update buffer in HBM (4-byte ids every `stride` bytes): *flags |= 1 when any id is NA, 2 when any
// is a number, 4 when some scenario's id differs from scenario 0's at the same position (eps rows per
// scenario from row row0; eps <= 0: no comparison). *flags must start at 0.
void launch_id_scan(char const* base, long long stride, long long total, long long row0, long long eps, int* flags,
                    hipStream_t st);
// union groups: per-member rows src[member][cols] to the lane-interleaved dst[cols][S_pad] (lane_member:
// the member of every lane, -1 for padding lanes, which get 0)
void launch_expand_member_rows(uint8_t const* src, int const* lane_member, int cols, int S_pad, uint8_t* dst,
                               hipStream_t st);
// union groups: Y values of every lane from the shared values plus per-lane patches of the entries
// that switched branches touch (patch q: lane, entry and P*P*2 values)
void launch_expand_ylane(double const* ydata, int E, int PP, int S_pad, double* ylane, int n_patch,
                         int const* patch_lane, int const* patch_entry, double const* patch_val, hipStream_t st);

// one wave of work: (group, task, block of 64 scenarios); packed groups: (group, first index into the
// level's task list, number of tasks)
struct WorkItem {
    int group, task, block;
};
// per (group, math model, scenario) iteration state, contiguous over the whole batch
struct PairState {
    long long n = 0;
    int* done = nullptr;
    int* err = nullptr;
    int* iters = nullptr;
    unsigned long long* devmax = nullptr;  // bits of a non-negative double
    double* lastdev = nullptr;
    int* pert = nullptr;                   // as in DevGroup
    unsigned long long* rmax = nullptr;
    unsigned long long* berr = nullptr;
};

// one sweep over the work items of one level (groups: device table of DevGroup)
// kernel kinds: plain groups (one task and 64 scenarios per wave), union groups (per-lane Y values and
// dead-bus rows) and packed groups (at most 32 scenarios: several tasks of one level per wave)
enum : int { KIND_PLAIN = 0, KIND_UNION = 1, KIND_PACKED = 2 };
void launch_forward(DevGroup const* groups, WorkItem const* items, int n_items, int P, bool fp32, int mode,
                    hipStream_t st, int kind = KIND_PLAIN);
// 4x4-block raw systems (state-estimation gain matrix, FP64, plain groups): forward (factorise and forward
// substitution, or with refine the forward substitution of the residual in y) or backward sweep (solution
// in y, or with refine added to xacc)
void launch_gain_sweep(DevGroup const* groups, WorkItem const* items, int n_items, bool refine, bool fwd,
                       hipStream_t st);
void launch_backward(DevGroup const* groups, WorkItem const* items, int n_items, int P, bool fp32, int mode,
                     hipStream_t st, int kind = KIND_PLAIN);
// iterative refinement of the pairs in REFINE (FP32 factors with perturbed pivots), as the reference refines a
// solve with perturbed pivots (sparse_lu_solver.hpp:514-649): with x the accumulated step, the residual
// r = rhs - J x of the Newton-Raphson system at the present voltages in FP64 (J applied without being stored)
// goes into the forward vectors and its componentwise backward error |r| / max(|rhs| + |J| |x|, 1e-4 max) into
// each pair's berr and the largest over all pairs into *gmax (bits); MODE_REFINE sweeps then solve the
// correction with the stored factors. bus_offsets as for launch_bus_results.
void launch_refine_start(PairState const& ps, DevGroup const* groups, long long const* bus_offsets, int n_groups,
                         long long total, int P, hipStream_t st);
void launch_refine_residual(PairState const& ps, DevGroup const* groups, long long const* bus_offsets, int n_groups,
                            long long total, int P, unsigned long long* gmax, hipStream_t st);
// the refined step applied to the state of the REFINE pairs (theta += dtheta, V += V dV), the deviation
// folded; then per pair: a pair with a perturbed pivot whose backward error is still above kRefineTol fails
// (SparseMatrixError), otherwise the convergence decision of converge_kernel, active pairs counted in *active
void launch_refine_finish(PairState const& ps, DevGroup const* groups, long long const* bus_offsets, int n_groups,
                          long long total, int P, int max_iter, double err_tol, int* active, hipStream_t st);
// fold the deviation of iteration `it`, mark converged / diverged, count the still active pairs
// it_dev: read the iteration number from this device counter instead of `it` (captured iterations)
void launch_converge(PairState const& ps, int it, int max_iter, double err_tol, int* active, hipStream_t st,
                     int const* it_dev = nullptr);
// one thread: ++*it_dev, *active = 0 (the start of a captured iteration)
void launch_begin_iteration(int* it_dev, int* active, hipStream_t st);
// one thread: *p = v (a fill that graph capture records as a kernel node)
void launch_set_int(int* p, int v, hipStream_t st);
void launch_init_state(PairState const& ps, hipStream_t st);
// one plain copy kernel over n floats (HBM counter calibration); 0 or -1
int copy_probe(long long n);
// the device address and size of this GPU's copy of the allocation-free entry point's fixed scratch region
// (located once per device); 0, or -1 when it cannot be located
int scratch_region(void** ptr, size_t* bytes);
// iterative current flat start: every bus of an active pair at the averaged source voltage of its math
// model, rotated by the bus's phase shift (iterative_current_pf_solver.hpp make_flat_start)
void launch_ic_start(int P, DevGroup const* groups, long long const* bus_offsets, int n_groups, long long total,
                     hipStream_t st);
// pairs flagged linear that are still active after the linear solve: converged after one solve
void launch_retire_linear(PairState const& ps, uint8_t const* pair_linear, hipStream_t st);

// outputs of a group, written as the caller's records at out[scenario_index[s] * n_comp + i]
struct DevOutput {
    int P = 1;
    int S = 0, S_pad = 0;
    int const* scenario_index = nullptr;  // global scenario of each lane
    int const* ok = nullptr;              // per lane: 1 when every math model converged
    int linear = 0;                       // linear method: load results as const_y (linear_pf_solver.hpp)
    // buses / Y rows (Y entries = non-fill LU entries)
    int n = 0;
    int const* row_ptr = nullptr;
    int const* row_end = nullptr;
    int const* row_stride = nullptr;  // as in DevGroup
    int const* bus_mm = nullptr;     // -1 for an unused bus label
    int const* col = nullptr;
    int8_t const* fill = nullptr;
    double const* ydata = nullptr;
    double const* ylane = nullptr;   // union groups, as in DevGroup
    uint8_t const* dead = nullptr;   // union groups, as in DevGroup
    double const* u = nullptr;
    // loads (math order)
    int n_load = 0;
    int const* load_bus = nullptr;
    int8_t const* load_type = nullptr;
    double const* s_inj = nullptr;
    // sources (math order)
    int n_src = 0;
    int const* src_bus = nullptr;
    int const* src_ptr = nullptr;  // per bus
    int const* load_ptr = nullptr; // per bus
    int const* shunt_ptr = nullptr; // per bus
    double const* u_src = nullptr;
    double const* y_src = nullptr;
    double* src_si = nullptr;   // [n_src][2][P][2][S_pad]
    double* bus_inj = nullptr;  // [n][P][2][S_pad] appliance injection sum per bus
    // shunts
    int n_shunt = 0;
    int const* shunt_bus = nullptr;
    double const* shunt_y = nullptr;  // P*P*2 per shunt
    double* shunt_si = nullptr;       // [n_shunt][2][P][2][S_pad]
};

// Newton-Raphson state estimation (newton_raphson_se_solver.hpp) of one group of scenarios sharing a
// topology: unknowns per bus theta, V and the Lagrange multipliers phi_p, phi_q; the gain matrix has one
// 4x4 block (column-major) per LU entry of the Y-bus pattern. Per-lane arrays are [item][...][S_pad].
struct SeDev {
    int n = 0, n_mm = 0;
    long long S_pad = 0;
    int const* row_ptr = nullptr;
    int const* col = nullptr;
    int const* diag = nullptr;
    int const* transpose = nullptr;
    int8_t const* fill = nullptr;
    int const* bus_mm = nullptr;
    double const* phase = nullptr;   // phase shift of every bus
    // Y-bus elements of every LU entry: el[el_ptr[k] ...], kind in the low 2 bits (0 shunt, 1 branch whose
    // from side is the entry's row, 2 branch whose to side is the entry's row), index above
    int const* el_ptr = nullptr;
    int const* el = nullptr;
    double const* yent = nullptr;    // admittance of every LU entry (re, im), zero for fill-ins
    double const* ybr = nullptr;     // per math branch yff, yft, ytf, ytt (re, im)
    double const* ysh = nullptr;     // per shunt (re, im)
    // aggregated measurements, variances normalised per math model (measured_values.hpp)
    int8_t const* v_flag = nullptr;  // [n][S_pad] 0 none, 1 magnitude only, 2 phasor
    double const* v_val = nullptr;   // [n][3][S_pad] re, im, variance
    int8_t const* inj_flag = nullptr;  // [n][S_pad]
    double const* inj_val = nullptr;   // [n][4][S_pad] p, q, variance p, variance q
    int8_t const* bf_flag = nullptr;   // [n_branch][S_pad] branch from-side power
    double const* bf_val = nullptr;    // [n_branch][4][S_pad]
    int8_t const* bt_flag = nullptr;   // to side
    double const* bt_val = nullptr;
    int8_t const* sh_flag = nullptr;   // [n_shunt][S_pad]
    double const* sh_val = nullptr;    // [n_shunt][4][S_pad]
    // per pair (math model, lane)
    int const* slack = nullptr;        // [n_mm] slack bus
    int const* vref = nullptr;         // [n_mm][S_pad] bus of the virtual angle measurement
    int8_t const* has_angle = nullptr; // [n_mm][S_pad]
    int8_t const* use_pert = nullptr;  // [n_mm][S_pad] observable and possibly ill-conditioned
    int* done = nullptr;
    unsigned long long* devmax = nullptr;
    unsigned long long* norm = nullptr;  // bits: block-wise off-diagonal infinity norm of the gain matrix
    double* pthr = nullptr;              // perturbation threshold of the factorisation
    int* pert = nullptr;                 // the factorisation perturbed a pivot: the solve is refined
    unsigned long long* rmax = nullptr;  // refinement: largest denominator, backward error (bits)
    unsigned long long* berr = nullptr;
    double* offset = nullptr;            // angle offset of the update
    // state and the linear system
    double* x = nullptr;      // [n][4][S_pad] theta, V, phi_p, phi_q
    double* u = nullptr;      // [n][2][S_pad]
    double* raw_a = nullptr;  // [E][16][S_pad]
    double* raw_b = nullptr;  // [n][4][S_pad]
    double* y = nullptr;      // [n][4][S_pad] solution of the factor sweep, residual during a refinement
    double* xacc = nullptr;   // [n][4][S_pad] refined solution
    double* rden = nullptr;   // [n][4][S_pad] refinement denominators
};
// per (bus, lane): initial state (newton_raphson_se_solver.hpp:219-236) from the per-pair mean angle shift
void launch_se_init(SeDev const& g, double const* mean_shift, hipStream_t st);
// pair flags of an iteration: norm, pert cleared
void launch_se_pair_reset(SeDev const& g, hipStream_t st);
// gain matrix and right-hand side of every row (prepare_matrix_and_rhs), then Q^T blocks, the Lagrange
// terms and the per-pair matrix norm (fill_qt, process_lagrange_multiplier), then the thresholds
void launch_se_gain(SeDev const& g, hipStream_t st);
// refinement of pairs in REFINE: xacc = y (begin); residual r = b - A xacc into y with denominators (the
// pairs' largest denominator and backward error restart at zero); backward error per pair
void launch_se_refine_begin(SeDev const& g, hipStream_t st);
void launch_se_residual(SeDev const& g, hipStream_t st);
void launch_se_berr(SeDev const& g, hipStream_t st);
// after the solve: angle offset per pair, then the update of every bus and the deviation per pair
// (iterate_unknown)
void launch_se_update(SeDev const& g, hipStream_t st);

// record writers: one thread per (lane, component); map gives the math index of every component (-1
// when it is not energised in this group)
struct NodeRec {
    int n_comp;
    int const* id;
    double const* u_rated;
    int const* bus;
};
struct BranchRec {
    int n_comp;
    int const* id;
    int const* math;           // math branch or -1
    int const* from_bus;       // per math branch
    int const* to_bus;
    double const* branch_y;    // per math branch 4 * P*P * 2
    // union groups: per lane and math branch the side statuses (bit 0 from, bit 1 to), and per math
    // branch the admittances of the four status combinations [4][4 * P*P * 2]; nullptr otherwise
    uint8_t const* lane_status;
    double const* branch_y_var;
    double const* base_i_from;
    double const* base_i_to;
    double const* rating;      // line: i_n, transformer: sn
    int is_line;
};
struct ApplRec {
    int n_comp;
    int const* id;
    int const* math;      // math index into si array or -1
    int const* bus;       // bus of the appliance or -1 (union groups check it against the dead mask)
    double const* base_i;
    double direction;
    int8_t const* status; // per component per lane [n_comp][S_pad], or nullptr (use status_const)
    int8_t const* status_const;
    double const* si;     // [n_math][2][P][2][S_pad]; nullptr for load_gens, whose power and current the
                          // record writer forms from the injection and the bus voltage
};
// components per record tile (64 lanes x TN records staged in at most 24 KiB of LDS)
constexpr int record_tile(size_t record_bytes) {
    return int(24576 / (64 * record_bytes)) > 0 ? int(24576 / (64 * record_bytes)) : 1;
}
// flat launches over all groups: offsets[g] = first flat index of group g (n_groups + 1 entries); the
// record writers count tiles (ceil(S_pad / W) x ceil(n_comp / (64 record_tile / W)) per group, with
// W = min(S_pad, 64) lanes per tile)
void launch_bus_results(int P, DevOutput const* outs, long long const* offsets, int n_groups, long long total,
                        hipStream_t st);
// adjoint of a symmetric Newton-Raphson batch (plain groups): objective per scenario
//   L = sum over nodes (u_pu - 1)^2 + sum over lines (p_from + p_to) [MW],
// gradients by every load_gen's p and q [W, var] and every line's r1 and x1 [ohm]
struct DevAdjoint {
    int n_lg = 0, n_line = 0;
    int const* bus_line_ptr = nullptr;  // [n + 1] lines incident to each bus
    int const* bus_line = nullptr;      // math branch of each incidence
    int const* node_w = nullptr;        // [n] nodes of each bus
    int const* br_from = nullptr;       // [n_br] bus of each side (-1: open)
    int const* br_to = nullptr;
    double const* br_y = nullptr;       // [n_br][4][2] ff, ft, tf, tt
    double const* br_dr = nullptr;      // [n_br][4][2] derivatives by r1 (lines; zero otherwise)
    double const* br_dx = nullptr;      // [n_br][4][2] derivatives by x1
    int const* lg_math = nullptr;       // [n_lg] math load index, -1 when not in the group
    int const* load_bus = nullptr;      // [n_load] bus of each math load
    int const* line_math = nullptr;     // [n_line] math branch of each line, -1 when not in the group
    int8_t const* lg_gen = nullptr;     // [n_lg] 1: generator (direction +1)
    int8_t const* load_status = nullptr;  // [n_lg][S_pad] status of every load_gen in every lane
    double* grad_p = nullptr;           // [n_scenarios][n_lg]
    double* grad_q = nullptr;
    double* grad_r = nullptr;           // [n_scenarios][n_line]
    double* grad_x = nullptr;
};
// dL/dx of the converged state into the groups' forward vectors (theta and ln V columns)
void launch_adjoint_rhs(DevGroup const* groups, DevAdjoint const* adj, long long const* bus_offsets, int n_groups,
                        long long total, hipStream_t st);
// J^T lambda = dL/dx with the factors of MODE_NRA: the U^T sweep (forward order), the L^T sweep
// (backward order); lambda ends in the forward vectors
void launch_adjoint_sweep(DevGroup const* groups, WorkItem const* items, int n_items, bool fp32, bool ut_sweep,
                          hipStream_t st);
// gradients of every load_gen and line of every lane: offsets over (n_lg + n_line) * S_pad per group
void launch_adjoint_grad(DevGroup const* groups, DevOutput const* outs, DevAdjoint const* adj,
                         long long const* offsets, int n_groups, long long total, hipStream_t st);

// load_gen updates applied on the device: one update buffer whose scenarios all list the same
// components in the same order. Fields are addressed as base + row * stride (row = scenario * n_elem +
// element); a null field is absent (the base value stays). asym: the record holds three phases.
struct LoadUpdateSource {
    int n_elem = 0;
    int const* comp = nullptr;  // [n_elem] load_gen index (sym_gen, asym_gen, sym_load, asym_load order)
    int asym = 0;
    char const* status = nullptr;
    long long status_stride = 0;
    char const* p = nullptr;
    long long p_stride = 0;
    char const* q = nullptr;
    long long q_stride = 0;
};
// base state of every load_gen and its place in a group
struct LoadTable {
    int n_lg = 0;
    int P = 1;
    int symmetric = 1;
    int8_t const* status = nullptr;  // [n_lg]
    double const* p = nullptr;       // [n_lg][3]
    double const* q = nullptr;       // [n_lg][3]
    int8_t const* gen = nullptr;     // [n_lg] 1: generator (direction +1)
    int8_t const* asym = nullptr;    // [n_lg] 1: asymmetric component
    int const* math = nullptr;       // [n_lg] math load index in the group, -1 when not in it
};
// every lane of a group gets the base injections (lanes >= S: zero) and statuses
void launch_load_broadcast(double* s_inj, int8_t* lstat, double const* base_s, int8_t const* base_status, int n_load,
                           int n_lg, int P, int S, int S_pad, hipStream_t st);
// apply one update source to the lanes of a group (lane -> scenario through scenario_index)
void launch_load_scatter(double* s_inj, int8_t* lstat, int S, int S_pad, int const* scenario_index,
                         LoadUpdateSource const& src, LoadTable const& tab, hipStream_t st);

// attribute columns a record writer produces instead of whole records (columnar output): byte offset
// and size of the attribute in the record, and the column [n_scenarios][n_comp] (n = 0: whole records)
constexpr int kMaxColumns = 16;
struct ColumnSet {
    int n = 0;
    int bus_sums = 1;  // 0: the bus sums (node p, q) were not formed in this run; leave them out
    int off[kMaxColumns] = {};
    int size[kMaxColumns] = {};
    char* dst[kMaxColumns] = {};
};
void launch_node_records(int P, DevOutput const* outs, NodeRec const* recs, long long const* offsets, int n_groups,
                         long long total, void* out, ColumnSet const& cols, hipStream_t st);
void launch_branch_records(int P, DevOutput const* outs, BranchRec const* recs, long long const* offsets,
                           int n_groups, long long total, void* out, ColumnSet const& cols, hipStream_t st);
void launch_appliance_records(int P, DevOutput const* outs, ApplRec const* recs, long long const* offsets,
                              int n_groups, long long total, void* out, ColumnSet const& cols, hipStream_t st);

}  // namespace pgm_amd
