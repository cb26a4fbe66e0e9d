// engine.hpp -- device-resident batch of power-flow scenarios.
#pragma once

#include "dataset.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "topology.hpp"

#include <memory>
#include <string>
#include <vector>

struct PGM_PowerGridModel {
    pgm_amd::Model model;
};

namespace pgm_amd {

class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { alloc(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(DeviceBuffer const&) = delete;
    DeviceBuffer& operator=(DeviceBuffer const&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    void alloc(size_t bytes);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr_); }
    void* get() const { return ptr_; }
    size_t bytes() const { return bytes_; }

  private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// apply one update record of component `comp` to a state; returns true if it changed the topology
// (branch or source status) or the admittances (tap, shunt) -- these define the group of a scenario
bool apply_update_record(Model const& m, ComponentState& st, int comp, Idx index, void const* record,
                         bool& structural);

struct Group;

struct Batch {
    Model model;
    bool symmetric = true;
    int P = 1;
    double err_tol = 1e-8;
    Idx max_iter = 20;
    bool fp32 = false;
    bool is_batch = false;
    Idx n_scenarios = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_fwd0 = nullptr, ev_fwd1 = nullptr;
    std::vector<std::unique_ptr<Group>> groups;
    // output records on the device, per component slot: n_scenarios * count records
    DeviceBuffer out[N_COMP];
    size_t rec_size[N_COMP] = {};
    // results of the last run
    std::vector<int> status;          // 0 ok, 1 sparse, 2 diverge
    std::vector<Idx> iterations;
    std::vector<std::string> messages;
    double kernel_ms = 0.0;
    double fwd_launch_ms = 0.0;
    Idx fwd_launches = 0;
    Idx fwd_timed = 0;
    std::vector<hipEvent_t> event_pool;
    static constexpr size_t kMaxTimedLaunches = 4096;
    bool ran = false;

    Batch() = default;
    ~Batch();
    void run();
    void fetch(PGM_MutableDataset const& output, PGM_Handle* handle);
};

std::unique_ptr<Batch> create_batch(Model const& model, PGM_Options const& opt, PGM_ConstDataset const* update);

// test hook: batched block LU on the device
int device_block_lu(Idx nb, Idx n, Idx const* row_indptr, Idx const* col_indices, Idx const* diag_lu, Idx batch,
                    double const* data, double const* rhs, double* x, int32_t* status, bool fp32);

std::string format_double(double x);
void batch_plan_sizes(Batch const& b, double& buses, double& entries);

}  // namespace pgm_amd

struct PGM_AmdBatch {
    std::unique_ptr<pgm_amd::Batch> b;
};
