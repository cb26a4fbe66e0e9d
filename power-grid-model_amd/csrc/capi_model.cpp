// capi_model.cpp -- model and calculation entry points of the C ABI (model.h of the reference,
// power_grid_model_c/power_grid_model_c/src/model.cpp:48-359), plus the PGM_amd_* extensions.
#include "engine.hpp"
#include "se.hpp"
#include "tap_changer.hpp"

#include "../../include/power_grid_model_amd.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

using namespace pgm_amd;

namespace {
// a permanent update of the model's base state (main_core/update.hpp permanent_update_t): every component the
// power flow and the state estimation read, the sensors included
void update_model_permanent(PGM_PowerGridModel& model, PGM_ConstDataset const& update) {
    if (update.meta->name != "update") throw PowerGridError("update dataset must be of type update\n");
    Model& m = model.model;
    ComponentState st = m.base;
    std::vector<char> rec;
    for (auto const& buf : update.buffers) {
        int comp = -1;
        for (int c = 0; c < N_COMP; ++c)
            if (buf.meta->name == comp_names[c]) comp = c;
        rec.resize(buf.meta->size);
        if (buf.meta->name == kTapRegulatorName) {
            for (Idx i = buf.begin(0); i < buf.end(0); ++i) {
                buf.read_record(i, rec.data());
                TapRegulatorUpdate u;
                std::memcpy(&u, rec.data(), sizeof u);
                Idx const idx = u.id == na_IntID ? i - buf.begin(0) : find_component(m, C_TAP_REGULATOR, u.id);
                apply_regulator_update(st, idx, u);
            }
            continue;
        }
        if (buf.meta->name == kVoltageSensorName || buf.meta->name == kPowerSensorName) {
            bool const vs = buf.meta->name == kVoltageSensorName;
            for (Idx i = buf.begin(0); i < buf.end(0); ++i) {
                buf.read_record(i, rec.data());
                ID id;
                std::memcpy(&id, rec.data(), 4);
                Idx const idx =
                    id == na_IntID ? i - buf.begin(0) : find_component(m, vs ? C_VOLTAGE_SENSOR : C_POWER_SENSOR, id);
                if (vs) {
                    SymVoltageSensorUpdate u;
                    std::memcpy(&u, rec.data(), sizeof u);
                    apply_voltage_sensor_update(st, idx, u);
                } else {
                    SymPowerSensorUpdate u;
                    std::memcpy(&u, rec.data(), sizeof u);
                    apply_power_sensor_update(st, idx, u);
                }
            }
            continue;
        }
        for (Idx i = buf.begin(0); i < buf.end(0); ++i) {
            buf.read_record(i, rec.data());
            ID id;
            std::memcpy(&id, rec.data(), 4);
            Idx idx = id == na_IntID ? i - buf.begin(0) : find_component(m, comp, id);
            int slot = comp;
            line_slot(m, slot, idx);
            bool structural = false;
            apply_update_record(m, st, slot, idx, rec.data(), structural);
        }
    }
    m.base = st;
    // plans and tuned schedules were made for the previous base state
    model.plans->clear();
    model.tuned.clear();
}

// scenario i of a batch dataset as a single dataset (auxiliary/dataset.hpp get_individual_scenario)
PGM_ConstDataset scenario_of(PGM_ConstDataset const& d, Idx i) {
    PGM_ConstDataset r;
    r.meta = d.meta;
    for (auto const& b : d.buffers) {
        ComponentBuffer c;
        c.meta = b.meta;
        Idx const b0 = b.begin(i), n = b.end(i) - b0;
        c.elements_per_scenario = n;
        c.total_elements = n;
        if (b.data) c.data = static_cast<char*>(b.data) + size_t(b0) * b.meta->size;
        for (auto const& [a, ptr] : b.columns) c.columns.emplace_back(a, static_cast<char*>(ptr) + size_t(b0) * a->size);
        r.buffers.push_back(std::move(c));
    }
    return r;
}

// scenarios [s0, s0 + n) of a batch output dataset (auxiliary/dataset.hpp get_slice_scenario)
PGM_MutableDataset output_rows(PGM_MutableDataset const& d, Idx s0, Idx n) {
    PGM_MutableDataset r;
    r.meta = d.meta;
    r.is_batch = true;
    r.batch_size = n;
    for (auto const& b : d.buffers) {
        if (b.elements_per_scenario < 0) throw PowerGridError("output buffers of a batch must be uniform\n");
        ComponentBuffer c;
        c.meta = b.meta;
        c.elements_per_scenario = b.elements_per_scenario;
        c.total_elements = n * b.elements_per_scenario;
        Idx const b0 = s0 * b.elements_per_scenario;
        if (b.data) c.data = static_cast<char*>(b.data) + size_t(b0) * b.meta->size;
        for (auto const& [a, ptr] : b.columns) c.columns.emplace_back(a, static_cast<char*>(ptr) + size_t(b0) * a->size);
        r.buffers.push_back(std::move(c));
    }
    return r;
}

void record_error(PGM_Handle* h, char const* msg) {
    if (!h) return;
    h->err_code = PGM_regular_error;
    h->err_msg = msg;
}
template <class F> void guarded(PGM_Handle* h, F&& f) {
    if (h) h->clear();
    try {
        f();
    } catch (std::exception const& e) {
        record_error(h, e.what());
    } catch (...) {
        record_error(h, "Unknown error!\n");
    }
}

// one calculation, or one batch dimension (model.cpp:336-359 of the reference C API)
void calculate_one(PGM_Handle* handle, PGM_PowerGridModel& model, PGM_Options const& opt,
                   PGM_MutableDataset const& output, PGM_ConstDataset const* batch) {
    if (opt.calculation_type == PGM_state_estimation) {
        run_state_estimation(model.model, opt, output, batch, handle, model.pool);
        return;
    }
    write_null_fault_output(model.model, output);
    // automatic tap changing (main_model_impl.hpp:318-346, optimizer/tap_position_optimizer.hpp)
    if (opt.tap_changing_strategy != 0) {
        calculate_with_tap_changer(model.model, opt, output, batch, handle);
        return;
    }
    write_null_regulator_output(model.model, output);
    auto b = create_batch(model.model, opt, batch, model_context(model, opt, batch));
    // only the requested components (and, for columnar buffers, attributes) are produced
    b->set_output(output);
    b->run();
    b->fetch(output, handle);
}

// batches over the cartesian product of several update datasets (model.cpp:285-334 of the reference C API):
// every scenario of the first dimension updates a copy of the model permanently, and the remaining dimensions
// run on that copy as a batch into the output rows of its stride. Failures of a stride count for its own
// scenarios (MDBatchExceptionHandler, :230-264).
void calculate_dimensions(PGM_Handle* handle, PGM_PowerGridModel& model, PGM_Options const& opt,
                          PGM_MutableDataset const& output, PGM_ConstDataset const* batch) {
    if (!batch || !batch->next) {
        calculate_one(handle, model, opt, output, batch);
        return;
    }
    Idx stride = 1;
    for (auto const* d = batch->next; d; d = d->next) stride *= d->batch_size;
    std::vector<Idx> failed;
    std::vector<std::string> msgs;
    std::string last;
    for (Idx i = 0; i < batch->batch_size; ++i) {
        PGM_Handle local;
        PGM_MutableDataset const rows = output_rows(output, i * stride, stride);
        try {
            PGM_PowerGridModel copy{model};
            copy.pool = model.pool;  // the strides run one after another: device chunks pass between them
            update_model_permanent(copy, scenario_of(*batch, i));
            calculate_dimensions(&local, copy, opt, rows, batch->next);
        } catch (std::exception const& e) {
            local.err_code = PGM_regular_error;
            local.err_msg = e.what();
        }
        if (local.err_code == PGM_batch_error) {
            for (size_t k = 0; k < local.failed_scenarios.size(); ++k) {
                failed.push_back(local.failed_scenarios[k] + i * stride);
                msgs.push_back(local.batch_errs[k]);
            }
            last = local.err_msg;
        } else if (local.err_code != PGM_no_error) {
            for (Idx k = 0; k < stride; ++k) {
                failed.push_back(i * stride + k);
                msgs.push_back(local.err_msg);
            }
            last = local.err_msg;
        }
    }
    if (!failed.empty() && handle) {
        handle->err_code = PGM_batch_error;
        handle->err_msg = last;
        handle->failed_scenarios = std::move(failed);
        handle->batch_errs = std::move(msgs);
    }
}
}  // namespace

extern "C" {

PGM_PowerGridModel* PGM_create_model(PGM_Handle* handle, double system_frequency,
                                     PGM_ConstDataset const* input_dataset) noexcept {
    PGM_PowerGridModel* out = nullptr;
    guarded(handle, [&] {
        if (input_dataset->meta->name != "input") throw PowerGridError("input dataset must be of type input\n");
        if (input_dataset->is_batch) throw PowerGridError("input dataset cannot be a batch\n");
        auto* m = new PGM_PowerGridModel{};
        try {
            m->model = build_model(*input_dataset, system_frequency);
        } catch (...) {
            delete m;
            throw;
        }
        out = m;
    });
    return out;
}

void PGM_update_model(PGM_Handle* handle, PGM_PowerGridModel* model, PGM_ConstDataset const* update) noexcept {
    guarded(handle, [&] { update_model_permanent(*model, *update); });
}

PGM_PowerGridModel* PGM_copy_model(PGM_Handle* handle, PGM_PowerGridModel const* model) noexcept {
    PGM_PowerGridModel* out = nullptr;
    guarded(handle, [&] { out = new PGM_PowerGridModel{*model}; });
    return out;
}

void PGM_get_indexer(PGM_Handle* handle, PGM_PowerGridModel const* model, char const* component, PGM_Idx size,
                     PGM_ID const* ids, PGM_Idx* indexer) noexcept {
    guarded(handle, [&] {
        int comp = -1;
        for (int c = 0; c < N_COMP; ++c)
            if (std::strcmp(component, comp_names[c]) == 0) comp = c;
        if (std::strcmp(component, kTapRegulatorName) == 0) comp = C_TAP_REGULATOR;
        if (std::strcmp(component, kFaultName) == 0) comp = C_FAULT;
        if (comp < 0) throw PowerGridError(std::string("unknown component ") + component);
        for (PGM_Idx i = 0; i < size; ++i) indexer[i] = find_component(model->model, comp, ids[i]);
    });
}

void PGM_calculate(PGM_Handle* handle, PGM_PowerGridModel* model, PGM_Options const* opt,
                   PGM_MutableDataset const* output_dataset, PGM_ConstDataset const* batch_dataset) noexcept {
    guarded(handle, [&] {
        if (batch_dataset && (!batch_dataset->is_batch || !output_dataset->is_batch))
            throw PowerGridError(
                "If batch_dataset is provided. Both batch_dataset and output_dataset should be a batch!\n");
        if (batch_dataset) {
            Idx total = 1;
            for (auto const* d = batch_dataset; d; d = d->next) total *= d->batch_size;
            if (total == 0) return;
            if (batch_dataset->next && output_dataset->batch_size != total)
                throw PowerGridError("The output dataset must hold every scenario of the cartesian product of the "
                                     "update datasets\n");
        }
        calculate_dimensions(handle, *model, *opt, *output_dataset, batch_dataset);
    });
}

void PGM_destroy_model(PGM_PowerGridModel* model) noexcept { delete model; }

void PGM_amd_tap_regulator_ranks(PGM_Handle* handle, PGM_PowerGridModel const* model, PGM_Idx* ranks) noexcept {
    guarded(handle, [&] {
        auto const r = tap_regulator_ranks(model->model, model->model.base);
        std::copy(r.begin(), r.end(), ranks);
    });
}

PGM_AmdBatch* PGM_amd_batch_create(PGM_Handle* handle, PGM_PowerGridModel const* model, PGM_Options const* opt,
                                   PGM_ConstDataset const* batch_dataset) noexcept {
    PGM_AmdBatch* out = nullptr;
    guarded(handle, [&] {
        auto b = create_batch(model->model, *opt, batch_dataset, model_context(*model, *opt, batch_dataset));
        out = new PGM_AmdBatch{std::move(b)};
    });
    return out;
}
PGM_Idx PGM_amd_autotune(PGM_Handle* handle, PGM_PowerGridModel* model, PGM_Options const* opt,
                         PGM_ConstDataset const* batch_dataset, PGM_Idx* params, double* trials_out,
                         PGM_Idx max_trials) noexcept {
    PGM_Idx n = 0;
    guarded(handle, [&] {
        std::vector<TuneTrial> trials;
        Tuning const best = autotune(*model, *opt, batch_dataset, trials);
        params[0] = best.plain_waves;
        params[1] = best.union_tasks;
        params[2] = best.packed_target;
        n = PGM_Idx(trials.size());
        for (PGM_Idx i = 0; i < std::min(n, max_trials); ++i) {
            trials_out[4 * i] = trials[size_t(i)].t.plain_waves;
            trials_out[4 * i + 1] = trials[size_t(i)].t.union_tasks;
            trials_out[4 * i + 2] = trials[size_t(i)].t.packed_target;
            trials_out[4 * i + 3] = trials[size_t(i)].ms;
        }
    });
    return n;
}
PGM_Idx PGM_amd_model_make_resident(PGM_Handle* handle, PGM_PowerGridModel* model,
                                    PGM_ConstDataset const* update_dataset) noexcept {
    PGM_Idx bytes = 0;
    guarded(handle, [&] {
        if (update_dataset->meta->name != "update") throw PowerGridError("update dataset must be of type update\n");
        bytes = PGM_Idx(make_resident(model->model, *update_dataset, *model->resident));
    });
    return bytes;
}
void PGM_amd_model_release_device_memory(PGM_PowerGridModel* model) noexcept {
    model->resident = std::make_unique<pgm_amd::UploadCache>();
    model->pool = std::make_shared<pgm_amd::ChunkPool>();
}
void PGM_amd_batch_run(PGM_Handle* handle, PGM_AmdBatch* batch) noexcept {
    guarded(handle, [&] { batch->b->run(); });
}
void PGM_amd_batch_adjoint(PGM_Handle* handle, PGM_AmdBatch* batch, double* grad_p, double* grad_q, double* grad_r1,
                           double* grad_x1) noexcept {
    guarded(handle, [&] { batch->b->adjoint(grad_p, grad_q, grad_r1, grad_x1); });
}
void PGM_amd_batch_set_output(PGM_Handle* handle, PGM_AmdBatch* batch, PGM_MutableDataset const* output) noexcept {
    guarded(handle, [&] { batch->b->set_output(*output); });
}
void PGM_amd_batch_fetch(PGM_Handle* handle, PGM_AmdBatch* batch, PGM_MutableDataset const* output) noexcept {
    guarded(handle, [&] { batch->b->fetch(*output, handle); });
}
void PGM_amd_batch_destroy(PGM_AmdBatch* batch) noexcept { delete batch; }
PGM_Idx PGM_amd_batch_n_scenarios(PGM_AmdBatch const* batch) noexcept { return batch->b->n_scenarios; }
void PGM_amd_batch_iterations(PGM_AmdBatch const* batch, PGM_Idx* iterations) noexcept {
    std::copy(batch->b->iterations.begin(), batch->b->iterations.end(), iterations);
}
void PGM_amd_batch_stats(PGM_AmdBatch const* batch, double* out) noexcept {
    auto const& b = *batch->b;
    double it_sum = 0;
    for (auto x : b.iterations) it_sum += double(x);
    double buses = 0, entries = 0;
    batch_plan_sizes(b, buses, entries);
    out[0] = b.kernel_ms;
    out[1] = it_sum;
    out[2] = buses;
    out[3] = entries;
    out[4] = double(b.groups.size());
    out[5] = b.fwd_launch_ms;
    out[6] = double(b.fwd_launches);
    out[7] = buses;
    out[8] = b.fwd_bytes;
    out[9] = b.fwd_total_ms;
    out[10] = double(b.fwd_timed);
    out[11] = double(b.graph_launches);
    out[12] = double(b.refinements);
    out[13] = double(b.refine_passes);
}

PGM_Idx PGM_amd_inspect_topology(PGM_Handle* handle, PGM_PowerGridModel const* model, PGM_Idx* node_group,
                                 PGM_Idx* node_pos, PGM_Idx* fills, PGM_Idx max_fill) noexcept {
    PGM_Idx result = -1;
    guarded(handle, [&] {
        Model const& m = model->model;
        Topology const t = build_topology(m, m.base);
        for (Idx i = 0; i < m.n_node(); ++i) {
            node_group[i] = t.node_group[i];
            node_pos[i] = t.node_pos[i];
        }
        PGM_Idx k = 0;
        for (size_t g = 0; g < t.models.size(); ++g)
            for (auto const& f : t.models[g].fill_in) {
                if (k >= max_fill) {
                    result = -1;
                    return;
                }
                fills[3 * k] = PGM_Idx(g);
                fills[3 * k + 1] = f[0];
                fills[3 * k + 2] = f[1];
                ++k;
            }
        result = k;
    });
    return result;
}

void PGM_amd_plan_stats(PGM_Handle* handle, PGM_PowerGridModel const* model, PGM_Idx symmetric,
                        PGM_Idx target_parallel_tasks, PGM_Idx packed_lanes, PGM_Idx* out) noexcept {
    guarded(handle, [&] {
        Model const& m = model->model;
        Topology const t = build_topology(m, m.base);
        std::vector<Idx> all(t.models.size());
        for (size_t k = 0; k < all.size(); ++k) all[k] = Idx(k);
        Plan const p = build_plan(m, m.base, t, all, symmetric != 0, int(target_parallel_tasks));
        // position of every pivot: (level, task, index within the task)
        std::vector<int> level(p.n, -1), task(p.n, -1), pos(p.n, -1);
        Idx crit = 0;
        for (size_t l = 0; l + 1 < p.level_ptr.size(); ++l) {
            Idx longest = 0;
            for (int q = p.level_ptr[l]; q < p.level_ptr[l + 1]; ++q) {
                int const tk = p.level_task[q];
                longest = std::max<Idx>(longest, p.task_ptr[tk + 1] - p.task_ptr[tk]);
                for (int i = p.task_ptr[tk]; i < p.task_ptr[tk + 1]; ++i) {
                    level[p.task_pivot[i]] = int(l);
                    task[p.task_pivot[i]] = tk;
                    pos[p.task_pivot[i]] = i;
                }
            }
            crit += longest;
        }
        auto before = [&](int a, int b) {  // pivot a is complete before pivot b starts
            if (level[a] < 0 || level[b] < 0) return false;
            if (level[a] < level[b]) return true;
            return task[a] == task[b] && pos[a] < pos[b];
        };
        bool valid = true;
        auto row_of = [&](int e) {
            return int(std::upper_bound(p.row_ptr.begin(), p.row_ptr.end(), e) - p.row_ptr.begin()) - 1;
        };
        for (Idx r = 0; r < p.n; ++r) {
            for (int k = p.row_ptr[r]; k < p.diag[r]; ++k) valid = valid && before(p.col[k], int(r));
            for (int q = p.schur_ptr[r]; q < p.schur_ptr[r + 1]; ++q) {
                int const le = p.schur[q][1], ue = p.schur[q][2];
                int const k = row_of(ue);  // the U entry (k, j) lives in row k, its pivot
                valid = valid && before(k, int(r)) && p.col[le] == k;
            }
        }
        if (packed_lanes > 0) {
            // the layout of a packed group of packed_lanes scenarios must stay consistent
            Plan q = p;
            try {
                interleave_plan(q, int(kWaveLanes / std::clamp<Idx>(packed_lanes, 1, kWaveLanes)));
                check_plan(q);
            } catch (std::logic_error const&) {
                valid = false;
            }
        }
        out[0] = p.n;
        out[1] = p.E;
        out[2] = Idx(p.level_ptr.size()) - 1;
        out[3] = Idx(p.task_mm.size());
        out[4] = Idx(p.schur.size());
        out[5] = crit;
        out[6] = valid ? 1 : 0;
    });
}

PGM_Idx PGM_amd_minimum_degree_ordering(PGM_Idx n_keys, PGM_Idx const* keys, PGM_Idx const* indptr,
                                        PGM_Idx const* neighbours, PGM_Idx* alpha_out, PGM_Idx* n_alpha_out,
                                        PGM_Idx* fills_out, PGM_Idx max_fills) noexcept {
    try {
        std::vector<Idx> k(keys, keys + n_keys);
        std::vector<std::vector<Idx>> adj(n_keys);
        for (PGM_Idx i = 0; i < n_keys; ++i) adj[i].assign(neighbours + indptr[i], neighbours + indptr[i + 1]);
        auto [alpha, fl] = minimum_degree(k, adj);
        std::copy(alpha.begin(), alpha.end(), alpha_out);
        *n_alpha_out = PGM_Idx(alpha.size());
        if (PGM_Idx(fl.size()) > max_fills) return -1;
        for (size_t i = 0; i < fl.size(); ++i) {
            fills_out[2 * i] = fl[i].first;
            fills_out[2 * i + 1] = fl[i].second;
        }
        return PGM_Idx(fl.size());
    } catch (...) {
        return -1;
    }
}

int PGM_amd_block_lu_solve(PGM_Idx nb, PGM_Idx n, PGM_Idx const* row_indptr, PGM_Idx const* col_indices,
                           PGM_Idx const* diag_lu, PGM_Idx batch, double const* data, double const* rhs, double* x,
                           int32_t* status, PGM_Idx precision) noexcept {
    try {
        return device_block_lu(nb, n, row_indptr, col_indices, diag_lu, batch, data, rhs, x, status, precision == 1);
    } catch (...) {
        return -1;
    }
}
int PGM_amd_block_lu_solve_noalloc(PGM_Idx nb, PGM_Idx n, PGM_Idx const* row_indptr, PGM_Idx const* col_indices,
                                   PGM_Idx const* diag_lu, PGM_Idx batch, double const* data, double const* rhs,
                                   double* x, int32_t* status, PGM_Idx precision) noexcept {
    try {
        return device_block_lu_noalloc(nb, n, row_indptr, col_indices, diag_lu, batch, data, rhs, x, status,
                                       precision == 1);
    } catch (pgm_amd::ScratchTooSmall const& e) {
        std::fprintf(stderr, "PGM_amd_block_lu_solve_noalloc: %s", e.what());
        return -2;
    } catch (std::exception const& e) {
        std::fprintf(stderr, "PGM_amd_block_lu_solve_noalloc: %s", e.what());
        return -1;
    } catch (...) {
        return -1;
    }
}
}  // extern "C"

int PGM_amd_copy_probe(PGM_Idx n_floats) noexcept {
    try {
        pgm_amd::ensure_device();
        return pgm_amd::copy_probe(n_floats);
    } catch (...) {
        return -1;
    }
}
