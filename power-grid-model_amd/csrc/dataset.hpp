// dataset.hpp -- metadata tables, datasets and the handle of the C ABI.
//
// Behaviour follows the reference C API: meta_data.h / meta_data.cpp (dataset, component and attribute
// lookup), dataset.h / dataset.cpp and auxiliary/dataset.hpp:398-520 (row or columnar buffers, uniform
// or sparse batches via indptr), handle.hpp:20-89 (error code, message, failed scenarios).
#pragma once

#include "records.hpp"

#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct PGM_MetaAttribute {
    std::string name;
    int ctype;  // PGM_CType
    size_t offset;
    size_t size;
};
struct PGM_MetaComponent {
    std::string name;
    size_t size;
    size_t alignment;
    std::vector<PGM_MetaAttribute> attributes;
    PGM_MetaAttribute const* find(std::string const& n) const {
        for (auto const& a : attributes)
            if (a.name == n) return &a;
        return nullptr;
    }
};
struct PGM_MetaDataset {
    std::string name;
    std::vector<PGM_MetaComponent> components;
    PGM_MetaComponent const* find(std::string const& n) const {
        for (auto const& c : components)
            if (c.name == n) return &c;
        return nullptr;
    }
};

namespace pgm_amd {

struct PowerGridError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// a request larger than a fixed device region
struct ScratchTooSmall : PowerGridError {
    using PowerGridError::PowerGridError;
};
struct DatasetError : PowerGridError {
    explicit DatasetError(std::string const& m) : PowerGridError("Dataset error: " + m) {}
};

std::vector<PGM_MetaDataset> const& meta_datasets();
PGM_MetaDataset const& meta_dataset(std::string const& name);

// write the null value (NaN / na) of an attribute into a record
void set_null(PGM_MetaAttribute const& a, void* record);
void set_null_records(PGM_MetaComponent const& c, void* ptr, Idx n);

struct ComponentBuffer {
    PGM_MetaComponent const* meta = nullptr;
    Idx elements_per_scenario = -1;
    Idx total_elements = 0;
    Idx const* indptr = nullptr;
    void* data = nullptr;  // row-based buffer, or nullptr when columnar
    std::vector<std::pair<PGM_MetaAttribute const*, void*>> columns;

    Idx begin(Idx scenario) const {
        return elements_per_scenario >= 0 ? scenario * elements_per_scenario : indptr[scenario];
    }
    Idx end(Idx scenario) const {
        return elements_per_scenario >= 0 ? (scenario + 1) * elements_per_scenario : indptr[scenario + 1];
    }
    bool columnar() const { return data == nullptr; }
    // copy record i (global row index) into dst (size meta->size), filling missing columns with null
    void read_record(Idx i, void* dst) const;
    // write a full record into row i (row-based or columnar)
    void write_record(Idx i, void const* src) const;
};

struct DatasetInfoData {
    std::string name;
    bool is_batch = false;
    Idx batch_size = 1;
    std::vector<std::string> comp_names;
    std::vector<Idx> eps, total;
    std::vector<std::vector<std::string>> attr_indications;
};

}  // namespace pgm_amd

struct PGM_DatasetInfo {
    pgm_amd::DatasetInfoData d;
};

struct PGM_ConstDataset {
    PGM_MetaDataset const* meta = nullptr;
    bool is_batch = false;
    pgm_amd::Idx batch_size = 1;
    std::vector<pgm_amd::ComponentBuffer> buffers;
    PGM_ConstDataset const* next = nullptr;  // next cartesian product dimension
    mutable PGM_DatasetInfo info;
    pgm_amd::ComponentBuffer const* find(std::string const& n) const {
        for (auto const& b : buffers)
            if (b.meta->name == n) return &b;
        return nullptr;
    }
    void refresh_info() const;
};

struct PGM_MutableDataset {
    PGM_MetaDataset const* meta = nullptr;
    bool is_batch = false;
    pgm_amd::Idx batch_size = 1;
    std::vector<pgm_amd::ComponentBuffer> buffers;
    mutable PGM_DatasetInfo info;
    pgm_amd::ComponentBuffer const* find(std::string const& n) const {
        for (auto const& b : buffers)
            if (b.meta->name == n) return &b;
        return nullptr;
    }
    void refresh_info() const;
};

struct PGM_Handle {
    pgm_amd::Idx err_code = 0;
    std::string err_msg;
    std::vector<pgm_amd::Idx> failed_scenarios;
    std::vector<std::string> batch_errs;
    mutable std::vector<char const*> batch_errs_c;
    void clear() {
        err_code = 0;
        err_msg.clear();
        failed_scenarios.clear();
        batch_errs.clear();
        batch_errs_c.clear();
    }
};

struct PGM_Options {
    pgm_amd::Idx calculation_type = 0;
    pgm_amd::Idx calculation_method = -128;
    pgm_amd::Idx symmetric = 1;
    double err_tol = 1e-8;
    pgm_amd::Idx max_iter = 20;
    pgm_amd::Idx threading = -1;
    pgm_amd::Idx short_circuit_voltage_scaling = 0;
    pgm_amd::Idx tap_changing_strategy = 0;
    pgm_amd::Idx experimental_features = 0;
    pgm_amd::Idx factor_precision = 0;
    // 1: every topology group of a batch is a plain group (no union or packed groups), the layout the
    // adjoint runs on
    pgm_amd::Idx adjoint_layout = 0;
    // 1: every Newton-Raphson step of an FP32 factorisation is refined in FP64 against the FP64 Jacobian (not
    // only the steps whose factorisation perturbed a pivot): the iterates are then the FP64 path's
    pgm_amd::Idx refine_steps = 0;
};
