// engine.hpp -- device-resident batch of power-flow scenarios.
#pragma once

#include "dataset.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "topology.hpp"

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace pgm_amd {
class ChunkPool;
class UploadCache;
class PlanCache;
// schedule parameters of a batch (0: the default, or the PGM_AMD_* environment override):
//   plain_waves   waves a plain group's schedule aims at per level (tasks x 64-scenario blocks)
//   union_tasks   largest number of tasks per level of a union group's schedule
//   packed_target target number of parallel tasks of a packed group's schedule
struct Tuning {
    int plain_waves = 0;
    int union_tasks = 0;
    int packed_target = 0;
};
}  // namespace pgm_amd

// the model, and the device memory that calculations on it keep between calls: released chunks of
// earlier batches (reused by the next calculation instead of new hipMallocs) and update buffers the caller
// made resident in HBM. A copy of the model starts with neither.
struct PGM_PowerGridModel {
    pgm_amd::Model model;
    std::shared_ptr<pgm_amd::ChunkPool> pool;
    std::unique_ptr<pgm_amd::UploadCache> resident;
    // schedule parameters picked by the autotuner, per batch shape (tune_key)
    std::map<std::string, pgm_amd::Tuning> tuned;
    // host plans (topology, ordering, Y-bus, schedule) of the structural states met so far; emptied when
    // the model is updated
    std::shared_ptr<pgm_amd::PlanCache> plans;
    PGM_PowerGridModel();
    PGM_PowerGridModel(PGM_PowerGridModel const& other);
    ~PGM_PowerGridModel();
};

namespace pgm_amd {

// device chunks kept for reuse: callers that create one batch after another (the tap changer's rounds)
// hand the chunks of a finished batch to the next instead of freeing and allocating tens of GB again
class ChunkPool {
  public:
    ChunkPool() = default;
    ChunkPool(ChunkPool const&) = delete;
    ChunkPool& operator=(ChunkPool const&) = delete;
    ~ChunkPool();
    // a free chunk of at least `bytes` (the smallest such), or nullptr
    std::pair<void*, size_t> get(size_t bytes);
    void put(void* p, size_t bytes) { free_.emplace_back(p, bytes); }

  private:
    std::vector<std::pair<void*, size_t>> free_;
};

// device copies of caller buffers that several batches upload unchanged (the tap changer's rounds all read
// the caller's load updates): keyed by host address and size, freed with the cache
class UploadCache {
  public:
    UploadCache() = default;
    UploadCache(UploadCache const&) = delete;
    UploadCache& operator=(UploadCache const&) = delete;
    ~UploadCache();
    void const* get(void const* host, size_t bytes);
    // the device copy of exactly this host range, or nullptr (nothing is uploaded)
    void const* find(void const* host, size_t bytes) const;
    size_t size() const { return dev_.size(); }

  private:
    std::map<std::pair<void const*, size_t>, void*> dev_;
};

// topology and plan of one structural state (and schedule target), shared between batches of one model:
// the reference keeps its topology and solvers between calculations while the topology is unchanged
// (main_model_impl.hpp is_topology_up_to_date_); here they are kept per structural state
struct PlanEntry {
    Topology topo;
    Plan plan;
};
class PlanCache {
  public:
    std::shared_ptr<PlanEntry const> find(std::string const& key);
    void put(std::string const& key, std::shared_ptr<PlanEntry const> e);
    void clear();

  private:
    std::mutex mu_;
    std::map<std::string, std::shared_ptr<PlanEntry const>> map_;
};

// bump allocator over large device chunks: one batch makes thousands of small arrays (one set per
// topology group) and frees them together (or returns them to its pool)
class Arena {
  public:
    Arena() = default;
    Arena(Arena const&) = delete;
    Arena& operator=(Arena const&) = delete;
    ~Arena();
    void* take(size_t bytes);
    size_t reserved() const { return reserved_; }
    void set_pool(std::shared_ptr<ChunkPool> pool) { pool_ = std::move(pool); }

  private:
    std::vector<std::pair<void*, size_t>> chunks_;
    char* cur_ = nullptr;
    size_t left_ = 0;
    size_t reserved_ = 0;
    std::shared_ptr<ChunkPool> pool_;  // shared: a batch may outlive the model that lent it
};

// apply one update record of component `comp` to a state; `structural` is set when it changes a branch
// or source status, a tap or a shunt -- the state that defines the topology group of a scenario
bool apply_update_record(Model const& m, ComponentState& st, int comp, Idx index, void const* record,
                         bool& structural);

struct Group;

struct Batch {
    Model model;
    Tuning tuning;  // the schedule parameters the batch was built with (0: default)
    bool symmetric = true;
    int P = 1;
    double err_tol = 1e-8;
    Idx max_iter = 20;
    bool fp32 = false;
    bool refine_steps = false;  // every FP32 Newton-Raphson step refined in FP64
    bool linear = false;   // calculation_method = linear
    bool current = false;  // iterative current or linear current
    bool is_batch = false;
    Idx n_scenarios = 1;
    hipStream_t stream = nullptr;
    hipStream_t stream_u = nullptr;  // union groups
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_lead = nullptr;
    Arena arena;
    std::vector<std::unique_ptr<Group>> groups;
    // device tables over the groups
    DevGroup* d_groups = nullptr;
    DevOutput* d_outputs = nullptr;
    NodeRec* d_node_recs = nullptr;
    BranchRec* d_branch_recs[4] = {nullptr, nullptr, nullptr, nullptr};  // lines, transformers, links, generic
    ApplRec* d_appl_recs[N_COMP] = {};
    // work items per level (forward in increasing, backward in decreasing level order)
    std::vector<WorkItem*> d_items;
    std::vector<int> n_items;
    std::vector<double> level_bytes_full;  // algorithmic forward bytes per level, every lane iterating
    std::vector<WorkItem*> d_items_u;      // union groups, per level
    std::vector<int> n_items_u;
    ExtraWork* d_xwork = nullptr;          // union lanes with extra entries (low-rank correction)
    int n_xwork = 0;
    std::vector<WorkItem*> d_items_p;      // packed groups, per level
    std::vector<int> n_items_p;
    PairState pairs;
    long long real_pairs = 0;
    long long total_lanes = 0;
    uint8_t* d_pair_linear = nullptr;  // per pair: solved by the linear method (nullptr: none)
    int* d_active = nullptr;
    int* h_active = nullptr;  // pinned
    int* d_ok = nullptr;      // per lane of every group
    long long* d_bus_off = nullptr;
    long long bus_total = 0;
    long long* d_rec_off[N_COMP] = {};
    long long rec_total[N_COMP] = {};
    // output records on the device, per component slot: n_scenarios * count records
    void* out[N_COMP] = {};
    size_t rec_size[N_COMP] = {};
    // output selection (set_output): the components a run produces; a component given as attribute
    // columns writes only those columns, each [n_scenarios][count] on the device
    bool want[N_COMP] = {true, true, true, true, true, true, true, true, true, true, true};
    struct DevColumn {
        PGM_MetaAttribute const* attr;
        void* dev;
    };
    std::vector<DevColumn> cols[N_COMP];
    std::vector<void*> col_mem;
    // results of the last run
    std::vector<int> status;  // 0 ok, 1 sparse, 2 diverge, 3 update error
    std::vector<Idx> iterations;
    std::vector<std::string> messages;
    double kernel_ms = 0.0;
    double fwd_launch_ms = 0.0;
    Idx fwd_launches = 0;
    Idx fwd_timed = 0;
    double fwd_bytes = 0.0;     // algorithmic bytes of the timed forward-sweep launches
    double fwd_total_ms = 0.0;  // their summed durations
    std::vector<hipEvent_t> event_pool;
    bool ran = false;
    // Newton-Raphson iterations as hipGraph launches: one captured iteration (every level of every kind
    // of group, the low-rank correction, the convergence check), instantiated twice so that the host can
    // read one iteration's active count and event times while the next one runs. d_iter is the device
    // iteration counter the captured check reads; h_act[2] the pinned active counts of the two parities.
    bool graph_tried = false;
    bool graph_ok = false;
    hipGraphExec_t nr_exec[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> nr_ev[2];  // per parity: start and end event of every plain forward level
    hipEvent_t ev_iter[2] = {nullptr, nullptr};
    int* d_iter = nullptr;
    int* h_act = nullptr;  // pinned, 2 entries
    int graph_launches = 0;  // iterations launched by the last run (the last may have found every lane done)
    // pivot perturbation (FP32 factors): steps held by the sweeps are refined in FP64 and applied; returns the
    // pairs that go on iterating. refinements / refine_passes: counts of the last run
    int refine_held_steps();
    bool refine_ready = false;
    unsigned long long* d_gmax = nullptr;
    int refinements = 0;
    int refine_passes = 0;
    void destroy_graphs();

    Batch() = default;
    ~Batch();
    void release();
    void run();
    // adjoint of the last run (symmetric Newton-Raphson, plain groups): per scenario the gradients of
    // L = sum (u_pu - 1)^2 + sum line losses [MW] by every load_gen's p and q and every line's r1 and x1,
    // into host arrays [n_scenarios][n_lg] and [n_scenarios][n_line]
    void adjoint(double* grad_p, double* grad_q, double* grad_r, double* grad_x);
    DevAdjoint* d_adj = nullptr;
    DevGroup* d_groups_adj = nullptr;  // the group table with FP64 factor storage (mixed-precision batches)
    double* d_grad = nullptr;  // grad_p, grad_q, grad_r, grad_x back to back
    long long* d_adj_off = nullptr;
    long long adj_total = 0;
    void set_output(PGM_MutableDataset const& output);
    void fetch(PGM_MutableDataset const& output, PGM_Handle* handle);
};

// pool: where the batch takes its device chunks from and returns them to (nullptr: allocate and free);
// uploads: device copies of update buffers shared with other batches (nullptr: the batch uploads its own)
// what a batch may borrow from its caller: pool, where the batch takes its device chunks from and returns
// them to (nullptr: allocate and free); uploads, device copies of update buffers shared with other
// batches (nullptr: the batch uploads its own); resident, update buffers the caller made resident in HBM
// (looked up only, never filled); tuning, schedule parameters; plans, host plans of structural states
struct BatchContext {
    std::shared_ptr<ChunkPool> pool;
    UploadCache* uploads = nullptr;
    UploadCache const* resident = nullptr;
    Tuning const* tuning = nullptr;
    PlanCache* plans = nullptr;
};
std::unique_ptr<Batch> create_batch(Model const& model, PGM_Options const& opt, PGM_ConstDataset const* update,
                                    BatchContext const& ctx = {});
// the context of a calculation on a model: its pool, resident buffers, plans and tuned parameters
BatchContext model_context(PGM_PowerGridModel const& model, PGM_Options const& opt, PGM_ConstDataset const* update);
// the shape a tuning applies to: scenarios, symmetry, factor precision, method
std::string tune_key(PGM_Options const& opt, PGM_ConstDataset const* update);
// one measured candidate of the autotuner
struct TuneTrial {
    Tuning t;
    double ms;  // device time of one run (HIP events over the whole run)
};
// try schedule parameters on this batch (each candidate: build the batch, run it twice, keep the second
// run's device time), coordinate by coordinate from the defaults; the best is stored in model.tuned
Tuning autotune(PGM_PowerGridModel& model, PGM_Options const& opt, PGM_ConstDataset const* update,
                std::vector<TuneTrial>& trials);
// device copies of every load_gen update buffer of `update` (rows, or the status / p / q columns) in
// `cache`; returns the bytes uploaded
size_t make_resident(Model const& model, PGM_ConstDataset const& update, UploadCache& cache);

// test hook: batched block LU on the device
int device_block_lu(Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices, Idx const* diag_lu, Idx batch,
                    double const* data, double const* rhs, double* x, int32_t* status, bool fp32);
// the same in the fixed device scratch, with no device allocation and no stream creation (default stream);
// throws ScratchTooSmall before any kernel runs when the batch does not fit
int device_block_lu_noalloc(Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices, Idx const* diag_lu,
                            Idx batch, double const* data, double const* rhs, double* x, int32_t* status, bool fp32);

std::string format_double(double x);
// throws PowerGridError when no HIP device is visible
void ensure_device();
void batch_plan_sizes(Batch const& b, double& buses, double& entries);

}  // namespace pgm_amd

struct PGM_AmdBatch {
    std::unique_ptr<pgm_amd::Batch> b;
};
