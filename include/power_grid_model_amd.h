/*
 * power_grid_model_amd.h -- C ABI of the MI355X batched Newton-Raphson power-flow library.
 *
 * The PGM_* entry points are the drop-in boundary: same names, arguments, ownership and error
 * behaviour as the reference's C API (power_grid_model_c/power_grid_model_c/include/power_grid_model_c/
 * *.h), for the power-flow path. Each declaration cites the reference declaration it replaces.
 * Only plain C types cross this boundary.
 *
 * Scope: calculation_type = power flow, calculation_method = default or Newton-Raphson, symmetric or
 * asymmetric, components node, line, transformer, shunt, source, sym_gen, asym_gen, sym_load,
 * asym_load. Other requests fail with PGM_regular_error and a message naming what is unsupported.
 *
 * PGM_amd_* entry points have no counterpart in the reference: they expose the device-resident batch
 * (for callers that keep inputs in HBM across calls), precision selection and per-scenario iteration
 * counts.
 */
#ifndef POWER_GRID_MODEL_AMD_H
#define POWER_GRID_MODEL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define PGM_AMD_NOEXCEPT noexcept
extern "C" {
#else
#define PGM_AMD_NOEXCEPT
#endif

#define PGM_AMD_API __attribute__((visibility("default")))

/* basics.h:47-75 */
typedef int64_t PGM_Idx;
typedef int32_t PGM_ID;
typedef struct PGM_PowerGridModel PGM_PowerGridModel;
typedef struct PGM_Handle PGM_Handle;
typedef struct PGM_Options PGM_Options;
typedef struct PGM_MetaAttribute PGM_MetaAttribute;
typedef struct PGM_MetaComponent PGM_MetaComponent;
typedef struct PGM_MetaDataset PGM_MetaDataset;
typedef struct PGM_ConstDataset PGM_ConstDataset;
typedef struct PGM_MutableDataset PGM_MutableDataset;
typedef struct PGM_WritableDataset PGM_WritableDataset;
typedef struct PGM_DatasetInfo PGM_DatasetInfo;

/* basics.h:80-140 */
enum PGM_CalculationType { PGM_power_flow = 0, PGM_state_estimation = 1, PGM_short_circuit = 2 };
enum PGM_CalculationMethod {
    PGM_default_method = -128,
    PGM_linear = 0,
    PGM_newton_raphson = 1,
    PGM_iterative_linear = 2,
    PGM_iterative_current = 3,
    PGM_linear_current = 4,
    PGM_iec60909 = 5
};
enum PGM_SymmetryType { PGM_asymmetric = 0, PGM_symmetric = 1 };
enum PGM_ErrorCode { PGM_no_error = 0, PGM_regular_error = 1, PGM_batch_error = 2, PGM_serialization_error = 3 };
enum PGM_CType { PGM_int32 = 0, PGM_int8 = 1, PGM_double = 2, PGM_double3 = 3 };

/* handle.h:27-100 */
PGM_AMD_API PGM_Handle* PGM_create_handle(void) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_handle(PGM_Handle* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_error_code(PGM_Handle const* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_error_message(PGM_Handle const* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_n_failed_scenarios(PGM_Handle const* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx const* PGM_failed_scenarios(PGM_Handle const* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const** PGM_batch_errors(PGM_Handle const* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_clear_error(PGM_Handle* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_version(void) PGM_AMD_NOEXCEPT;

/* options.h:25-120 */
PGM_AMD_API PGM_Options* PGM_create_options(PGM_Handle* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_options(PGM_Options* opt) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_calculation_type(PGM_Handle* handle, PGM_Options* opt, PGM_Idx type) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_calculation_method(PGM_Handle* handle, PGM_Options* opt,
                                            PGM_Idx method) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_symmetric(PGM_Handle* handle, PGM_Options* opt, PGM_Idx sym) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_err_tol(PGM_Handle* handle, PGM_Options* opt, double err_tol) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_max_iter(PGM_Handle* handle, PGM_Options* opt, PGM_Idx max_iter) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_threading(PGM_Handle* handle, PGM_Options* opt, PGM_Idx threading) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_short_circuit_voltage_scaling(PGM_Handle* handle, PGM_Options* opt,
                                                       PGM_Idx short_circuit_voltage_scaling) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_tap_changing_strategy(PGM_Handle* handle, PGM_Options* opt,
                                               PGM_Idx tap_changing_strategy) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_set_experimental_features(PGM_Handle* handle, PGM_Options* opt,
                                               PGM_Idx experimental_features) PGM_AMD_NOEXCEPT;

/* meta_data.h:32-190 */
PGM_AMD_API PGM_Idx PGM_meta_n_datasets(PGM_Handle* handle) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MetaDataset const* PGM_meta_get_dataset_by_idx(PGM_Handle* handle, PGM_Idx idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MetaDataset const* PGM_meta_get_dataset_by_name(PGM_Handle* handle,
                                                                 char const* dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_meta_dataset_name(PGM_Handle* handle, PGM_MetaDataset const* dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_meta_n_components(PGM_Handle* handle, PGM_MetaDataset const* dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MetaComponent const* PGM_meta_get_component_by_idx(PGM_Handle* handle,
                                                                   PGM_MetaDataset const* dataset,
                                                                   PGM_Idx idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MetaComponent const* PGM_meta_get_component_by_name(PGM_Handle* handle, char const* dataset,
                                                                    char const* component) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_meta_component_name(PGM_Handle* handle,
                                                PGM_MetaComponent const* component) PGM_AMD_NOEXCEPT;
PGM_AMD_API size_t PGM_meta_component_size(PGM_Handle* handle, PGM_MetaComponent const* component) PGM_AMD_NOEXCEPT;
PGM_AMD_API size_t PGM_meta_component_alignment(PGM_Handle* handle,
                                                PGM_MetaComponent const* component) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_meta_n_attributes(PGM_Handle* handle, PGM_MetaComponent const* component) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MetaAttribute const* PGM_meta_get_attribute_by_idx(PGM_Handle* handle,
                                                                   PGM_MetaComponent const* component,
                                                                   PGM_Idx idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MetaAttribute const* PGM_meta_get_attribute_by_name(PGM_Handle* handle, char const* dataset,
                                                                    char const* component,
                                                                    char const* attribute) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_meta_attribute_name(PGM_Handle* handle,
                                                PGM_MetaAttribute const* attribute) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_meta_attribute_ctype(PGM_Handle* handle, PGM_MetaAttribute const* attribute) PGM_AMD_NOEXCEPT;
PGM_AMD_API size_t PGM_meta_attribute_offset(PGM_Handle* handle, PGM_MetaAttribute const* attribute) PGM_AMD_NOEXCEPT;
PGM_AMD_API int PGM_is_little_endian(PGM_Handle* handle) PGM_AMD_NOEXCEPT;

/* buffer.h:24-118 */
PGM_AMD_API void* PGM_create_buffer(PGM_Handle* handle, PGM_MetaComponent const* component,
                                    PGM_Idx size) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_buffer(void* ptr) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_buffer_set_nan(PGM_Handle* handle, PGM_MetaComponent const* component, void* ptr,
                                    PGM_Idx buffer_offset, PGM_Idx size) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_buffer_set_value(PGM_Handle* handle, PGM_MetaAttribute const* attribute, void* buffer_ptr,
                                      void const* src_ptr, PGM_Idx buffer_offset, PGM_Idx size,
                                      PGM_Idx src_stride) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_buffer_get_value(PGM_Handle* handle, PGM_MetaAttribute const* attribute,
                                      void const* buffer_ptr, void* dest_ptr, PGM_Idx buffer_offset, PGM_Idx size,
                                      PGM_Idx dest_stride) PGM_AMD_NOEXCEPT;

/* dataset.h:27-333 */
PGM_AMD_API char const* PGM_dataset_info_name(PGM_Handle* handle, PGM_DatasetInfo const* info) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_is_batch(PGM_Handle* handle, PGM_DatasetInfo const* info) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_batch_size(PGM_Handle* handle, PGM_DatasetInfo const* info) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_n_components(PGM_Handle* handle,
                                                  PGM_DatasetInfo const* info) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_dataset_info_component_name(PGM_Handle* handle, PGM_DatasetInfo const* info,
                                                        PGM_Idx component_idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_elements_per_scenario(PGM_Handle* handle, PGM_DatasetInfo const* info,
                                                           PGM_Idx component_idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_total_elements(PGM_Handle* handle, PGM_DatasetInfo const* info,
                                                    PGM_Idx component_idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_has_attribute_indications(PGM_Handle* handle, PGM_DatasetInfo const* info,
                                                               PGM_Idx component_idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_dataset_info_n_attribute_indications(PGM_Handle* handle, PGM_DatasetInfo const* info,
                                                             PGM_Idx component_idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_dataset_info_attribute_name(PGM_Handle* handle, PGM_DatasetInfo const* info,
                                                        PGM_Idx component_idx,
                                                        PGM_Idx attribute_idx) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_ConstDataset* PGM_create_dataset_const(PGM_Handle* handle, char const* dataset, PGM_Idx is_batch,
                                                       PGM_Idx batch_size) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_ConstDataset* PGM_create_dataset_const_from_mutable(
    PGM_Handle* handle, PGM_MutableDataset const* mutable_dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_dataset_const(PGM_ConstDataset* dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_dataset_const_add_buffer(PGM_Handle* handle, PGM_ConstDataset* dataset, char const* component,
                                              PGM_Idx elements_per_scenario, PGM_Idx total_elements,
                                              PGM_Idx const* indptr, void const* data) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_dataset_const_add_attribute_buffer(PGM_Handle* handle, PGM_ConstDataset* dataset,
                                                        char const* component, char const* attribute,
                                                        void const* data) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_DatasetInfo const* PGM_dataset_const_get_info(PGM_Handle* handle,
                                                              PGM_ConstDataset const* dataset) PGM_AMD_NOEXCEPT;
/* dataset.h PGM_dataset_const_set_next_cartesian_product_dimension: a batch over the cartesian product of the
 * chained datasets; PGM_calculate runs it as model.cpp:285-334 (first dimension outermost, each of its scenarios
 * a permanent update of a model copy, the rest one batch on that copy) */
PGM_AMD_API void PGM_dataset_const_set_next_cartesian_product_dimension(
    PGM_Handle* handle, PGM_ConstDataset* dataset, PGM_ConstDataset const* next_dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_MutableDataset* PGM_create_dataset_mutable(PGM_Handle* handle, char const* dataset,
                                                           PGM_Idx is_batch, PGM_Idx batch_size) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_dataset_mutable(PGM_MutableDataset* dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_dataset_mutable_add_buffer(PGM_Handle* handle, PGM_MutableDataset* dataset,
                                                char const* component, PGM_Idx elements_per_scenario,
                                                PGM_Idx total_elements, PGM_Idx const* indptr,
                                                void* data) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_dataset_mutable_add_attribute_buffer(PGM_Handle* handle, PGM_MutableDataset* dataset,
                                                          char const* component, char const* attribute,
                                                          void* data) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_DatasetInfo const* PGM_dataset_mutable_get_info(PGM_Handle* handle,
                                                                PGM_MutableDataset const* dataset) PGM_AMD_NOEXCEPT;

/* dataset.h:241-269: writable datasets only come out of a deserializer, which this build does not
 * provide; these fail with PGM_regular_error (the reference's Python package binds them at import) */
PGM_AMD_API PGM_DatasetInfo const* PGM_dataset_writable_get_info(PGM_Handle* handle,
                                                                 PGM_WritableDataset const* dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_dataset_writable_set_buffer(PGM_Handle* handle, PGM_WritableDataset* dataset,
                                                 char const* component, PGM_Idx* indptr, void* data) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_dataset_writable_set_attribute_buffer(PGM_Handle* handle, PGM_WritableDataset* dataset,
                                                           char const* component, char const* attribute,
                                                           void* data) PGM_AMD_NOEXCEPT;

/* serialization.h:29-114: JSON / msgpack (de)serialization is out of this build's scope; every entry
 * point fails with PGM_regular_error naming it, so callers that bind the whole API still load */
typedef struct PGM_Serializer PGM_Serializer;
typedef struct PGM_Deserializer PGM_Deserializer;
PGM_AMD_API PGM_Deserializer* PGM_create_deserializer_from_binary_buffer(PGM_Handle* handle, char const* data,
                                                                         PGM_Idx size,
                                                                         PGM_Idx serialization_format) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Deserializer* PGM_create_deserializer_from_null_terminated_string(
    PGM_Handle* handle, char const* data_string, PGM_Idx serialization_format) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_WritableDataset* PGM_deserializer_get_dataset(PGM_Handle* handle,
                                                              PGM_Deserializer* deserializer) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_deserializer_parse_to_buffer(PGM_Handle* handle, PGM_Deserializer* deserializer) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_deserializer(PGM_Deserializer* deserializer) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Serializer* PGM_create_serializer(PGM_Handle* handle, PGM_ConstDataset const* dataset,
                                                  PGM_Idx serialization_format) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_serializer_get_to_binary_buffer(PGM_Handle* handle, PGM_Serializer* serializer,
                                                     PGM_Idx use_compact_list, char const** data,
                                                     PGM_Idx* size) PGM_AMD_NOEXCEPT;
PGM_AMD_API char const* PGM_serializer_get_to_zero_terminated_string(PGM_Handle* handle, PGM_Serializer* serializer,
                                                                     PGM_Idx use_compact_list,
                                                                     PGM_Idx indent) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_serializer(PGM_Serializer* serializer) PGM_AMD_NOEXCEPT;

/* model.h:33-130 */
PGM_AMD_API PGM_PowerGridModel* PGM_create_model(PGM_Handle* handle, double system_frequency,
                                                 PGM_ConstDataset const* input_dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_update_model(PGM_Handle* handle, PGM_PowerGridModel* model,
                                  PGM_ConstDataset const* update_dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_PowerGridModel* PGM_copy_model(PGM_Handle* handle, PGM_PowerGridModel const* model) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_get_indexer(PGM_Handle* handle, PGM_PowerGridModel const* model, char const* component,
                                 PGM_Idx size, PGM_ID const* ids, PGM_Idx* indexer) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_calculate(PGM_Handle* handle, PGM_PowerGridModel* model, PGM_Options const* opt,
                               PGM_MutableDataset const* output_dataset,
                               PGM_ConstDataset const* batch_dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_destroy_model(PGM_PowerGridModel* model) PGM_AMD_NOEXCEPT;

/* ---------- extensions (no counterpart in the reference) ---------- */

/* Factorisation precision of the Newton-Raphson linear solves: 0 = FP64 (default, plain path),
 * 1 = FP32 factor and triangular solves with FP64 mismatch and state update (mixed path). */
PGM_AMD_API void PGM_amd_set_factor_precision(PGM_Handle* handle, PGM_Options* opt,
                                              PGM_Idx precision) PGM_AMD_NOEXCEPT;

/* 1: batches built with these options keep every topology group as a plain group (no union groups that
 * share one structure between switching states, no packed groups of a few scenarios), the layout
 * PGM_amd_batch_adjoint runs on; 0 (default): the fastest layout. Results are the same to FP64 rounding. */
PGM_AMD_API void PGM_amd_set_adjoint_layout(PGM_Handle* handle, PGM_Options* opt, PGM_Idx value) PGM_AMD_NOEXCEPT;

/* 1: every Newton-Raphson step of an FP32 factorisation (factor_precision 1) is refined in FP64 against the
 * FP64 Jacobian, as the reference refines solves after pivot perturbation (sparse_lu_solver.hpp:514-649), so
 * the iterates follow the FP64 path; 0 (default): only steps whose FP32 factorisation perturbed a pivot are
 * refined. No counterpart in the reference, which has no FP32 factorisation. */
PGM_AMD_API void PGM_amd_set_step_refinement(PGM_Handle* handle, PGM_Options* opt, PGM_Idx value) PGM_AMD_NOEXCEPT;

/* Device-resident batch: build the batch (topology grouping, plans, device upload of every
 * scenario's update data) once, then run it any number of times with the inputs already in HBM.
 * PGM_calculate is build + run + fetch + destroy. Returns NULL on error (message in handle). */
typedef struct PGM_AmdBatch PGM_AmdBatch;
PGM_AMD_API PGM_AmdBatch* PGM_amd_batch_create(PGM_Handle* handle, PGM_PowerGridModel const* model,
                                               PGM_Options const* opt,
                                               PGM_ConstDataset const* batch_dataset) PGM_AMD_NOEXCEPT;
/* Make the load_gen update buffers of an update dataset resident in HBM for this model: every
 * sym_gen / asym_gen / sym_load / asym_load buffer (the row buffer, or its status, p_specified and
 * q_specified columns) is copied to the device once. Later PGM_calculate and PGM_amd_batch_create calls
 * on this model whose update datasets hold the same host buffers (same address and size) read the device
 * copies instead of uploading them again; the host buffers must not change meanwhile. Returns the bytes
 * uploaded. For callers that solve one scenario set again and again (studies, serving). */
PGM_AMD_API PGM_Idx PGM_amd_model_make_resident(PGM_Handle* handle, PGM_PowerGridModel* model,
                                                PGM_ConstDataset const* update_dataset) PGM_AMD_NOEXCEPT;
/* Free the device memory the model keeps between calculations: the resident update buffers and the
 * device chunks of finished batches (PGM_calculate and destroyed PGM_AmdBatch objects hand their chunks
 * back to the model, and the next calculation reuses them instead of allocating). */
PGM_AMD_API void PGM_amd_model_release_device_memory(PGM_PowerGridModel* model) PGM_AMD_NOEXCEPT;
/* Autotuner of the schedule (no counterpart in the reference): builds and runs this batch with candidate
 * schedule parameters, coordinate by coordinate from the defaults, and keeps the fastest for the model.
 * Later PGM_calculate and PGM_amd_batch_create calls on the model with the same shape (number of
 * scenarios, symmetry, factor precision, method) use it. Parameters: [0] waves a plain group's schedule
 * aims at per level, [1] largest number of tasks per level of a union group, [2] target number of parallel
 * tasks of a packed group. trials_out receives up to max_trials rows of (plain waves, union tasks, packed
 * target, device ms of one run). Returns the number of trials, or 0 on error (message in handle). */
PGM_AMD_API PGM_Idx PGM_amd_autotune(PGM_Handle* handle, PGM_PowerGridModel* model, PGM_Options const* opt,
                                     PGM_ConstDataset const* batch_dataset, PGM_Idx* params, double* trials_out,
                                     PGM_Idx max_trials) PGM_AMD_NOEXCEPT;
/* Run every scenario to convergence on the device; returns when the device work is complete. */
PGM_AMD_API void PGM_amd_batch_run(PGM_Handle* handle, PGM_AmdBatch* batch) PGM_AMD_NOEXCEPT;
/* Select what the following runs produce on the device, from the layout of an output dataset: only
 * its components are computed, and a component given as attribute buffers (columnar) writes only
 * those attributes, as device columns. Without this call a batch produces every component as whole
 * records. The reference filters output the same way through the output dataset it is given
 * (PGM_calculate, model.h:116-118; columnar buffers: PGM_dataset_mutable_add_attribute_buffer,
 * dataset.h:312-323); here the filter also removes device work. Fetch then accepts datasets with the
 * same components and a subset of the selected columns. */
PGM_AMD_API void PGM_amd_batch_set_output(PGM_Handle* handle, PGM_AmdBatch* batch,
                                          PGM_MutableDataset const* output) PGM_AMD_NOEXCEPT;
/* Adjoint gradients of the last run (no counterpart in the reference, which has no derivatives): for
 * every scenario, L = sum over energised nodes (u_pu - 1)^2 + sum over lines (p_from + p_to) in MW, and
 * its gradients by every load_gen's p_specified and q_specified (per phase for asymmetric ones) and by
 * every line's r1 and x1. The Newton-Raphson Jacobian is factorised again at the converged state and
 * J^T lambda = dL/dx is solved with those factors. Arrays are [n_scenarios][n_load_gen] (load_gens in
 * the order sym_gen, asym_gen, sym_load, asym_load, input order within each) and
 * [n_scenarios][n_line]. Symmetric Newton-Raphson batches whose topology groups each hold more than 32
 * scenarios; failed scenarios get zeros. */
PGM_AMD_API void PGM_amd_batch_adjoint(PGM_Handle* handle, PGM_AmdBatch* batch, double* grad_p, double* grad_q,
                                       double* grad_r1, double* grad_x1) PGM_AMD_NOEXCEPT;
/* Copy results into the caller's output dataset and report per-scenario errors like PGM_calculate. */
PGM_AMD_API void PGM_amd_batch_fetch(PGM_Handle* handle, PGM_AmdBatch* batch,
                                     PGM_MutableDataset const* output_dataset) PGM_AMD_NOEXCEPT;
PGM_AMD_API void PGM_amd_batch_destroy(PGM_AmdBatch* batch) PGM_AMD_NOEXCEPT;
PGM_AMD_API PGM_Idx PGM_amd_batch_n_scenarios(PGM_AmdBatch const* batch) PGM_AMD_NOEXCEPT;
/* Per-scenario Newton-Raphson iteration count of the last run (max over the scenario's math models;
 * 0 for failed scenarios). iterations must hold n_scenarios entries. */
PGM_AMD_API void PGM_amd_batch_iterations(PGM_AmdBatch const* batch, PGM_Idx* iterations) PGM_AMD_NOEXCEPT;
/* Statistics of the last run: out[0] = device time of the NR kernels in ms (HIP events),
 * out[1] = sum over scenarios of iterations, out[2] = number of bus blocks factorised per scenario
 * (sum over groups), out[3] = LU entries per scenario, out[4] = number of topology groups,
 * out[5] = average duration in ms of one forward-sweep launch, out[6] = number of forward-sweep
 * launches, out[7] = buses per scenario, out[8] = algorithmic bytes of the timed forward-sweep
 * launches (factor written and read once, 2 E b, plus 32 P bytes per bus, for the scenarios still
 * iterating), out[9] = their summed duration in ms (HIP events), out[10] = number of timed launches,
 * out[11] = Newton-Raphson iterations launched as graphs (0: one launch per level), out[12] = times the
 * FP32 factorisation perturbed pivots and the held steps were refined in FP64, out[13] = residuals
 * formed by those refinements. out must hold 14 entries. */
PGM_AMD_API void PGM_amd_batch_stats(PGM_AmdBatch const* batch, double* out) PGM_AMD_NOEXCEPT;

/* Host-only inspection of the product's topology for a model (no GPU needed): per node the math
 * model index (-1 if not energised) and the bus position; fills receives (model, bus_a, bus_b) triples.
 * Returns the number of fill-ins, or -1 if max_fill is too small. */
PGM_AMD_API PGM_Idx PGM_amd_inspect_topology(PGM_Handle* handle, PGM_PowerGridModel const* model,
                                             PGM_Idx* node_group, PGM_Idx* node_pos, PGM_Idx* fills,
                                             PGM_Idx max_fill) PGM_AMD_NOEXCEPT;

/* Host-only statistics and validity check of the device factorisation schedule of a model's base
 * topology (no GPU needed). out[0] buses, out[1] LU entries, out[2] levels, out[3] tasks, out[4] Schur
 * updates, out[5] critical path in pivots (sum over levels of the longest task), out[6] 1 when every
 * dependency of every pivot is scheduled before it (and, for packed_lanes > 0, the interleaved layout
 * of a packed group of that many scenarios is consistent), else 0. */
PGM_AMD_API void PGM_amd_plan_stats(PGM_Handle* handle, PGM_PowerGridModel const* model, PGM_Idx symmetric,
                                    PGM_Idx target_parallel_tasks, PGM_Idx packed_lanes,
                                    PGM_Idx* out) PGM_AMD_NOEXCEPT;

/* Host-only minimum-degree ordering on an adjacency map given as sorted keys + CSR neighbour lists
 * (same contract as sparse_ordering.hpp:183 minimum_degree_ordering). Returns the number of fills or -1. */
PGM_AMD_API PGM_Idx PGM_amd_minimum_degree_ordering(PGM_Idx n_keys, PGM_Idx const* keys, PGM_Idx const* indptr,
                                                    PGM_Idx const* neighbours, PGM_Idx* alpha_out,
                                                    PGM_Idx* n_alpha_out, PGM_Idx* fills_out,
                                                    PGM_Idx max_fills) PGM_AMD_NOEXCEPT;

/* GPU block-sparse LU factorise + solve of `batch` independent systems sharing one pattern (test hook
 * for the device LU). nb in {2, 6}; data: batch x nnz blocks (column-major nb*nb), rhs/x: batch x n x nb.
 * precision 0 = FP64, 1 = FP32. status[batch]: 0 ok, 1 singular. Returns 0, or -1 on a launch error. */
PGM_AMD_API int PGM_amd_block_lu_solve(PGM_Idx nb, PGM_Idx n, PGM_Idx const* row_indptr,
                                       PGM_Idx const* col_indices, PGM_Idx const* diag_lu, PGM_Idx batch,
                                       double const* data, double const* rhs, double* x, int32_t* status,
                                       PGM_Idx precision) PGM_AMD_NOEXCEPT;

/* The same block LU factorise + solve with no device allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Everything on the device (structure, input blocks, factors, vectors, per-system state) lives in a
 * fixed file-scope device region of 64 MiB per GPU, held until the process ends, and the work runs on the
 * default stream. Returns -2 (and prints why on stderr) when the batch does not fit the region, before any
 * kernel runs; otherwise as PGM_amd_block_lu_solve. Calls on one GPU share its region: calls that overlap in
 * time on one GPU race and must be serialised by the caller. */
PGM_AMD_API int PGM_amd_block_lu_solve_noalloc(PGM_Idx nb, PGM_Idx n, PGM_Idx const* row_indptr,
                                               PGM_Idx const* col_indices, PGM_Idx const* diag_lu, PGM_Idx batch,
                                               double const* data, double const* rhs, double* x, int32_t* status,
                                               PGM_Idx precision) PGM_AMD_NOEXCEPT;

/* Test hook with no counterpart in the reference: one plain copy kernel over n_floats floats (reads and
 * writes 4 * n_floats bytes of HBM), the reference kernel that calibrates rocprofv3 FETCH_SIZE /
 * WRITE_SIZE counters. Allocates and frees its own buffers. Returns 0, or -1 on a HIP error. */
PGM_AMD_API int PGM_amd_copy_probe(PGM_Idx n_floats) PGM_AMD_NOEXCEPT;

/* Test hook of the automatic tap changer (PGM_calculate with a tap changing strategy,
 * optimizer/tap_position_optimizer.hpp:209-432 rank_transformers): the rank of every transformer_tap_regulator
 * of the model in its input state, in input order; -1 for regulators that are off or cannot be reached
 * from a source. Host work only (no device needed). ranks must hold one entry per regulator. */
PGM_AMD_API void PGM_amd_tap_regulator_ranks(PGM_Handle* handle, PGM_PowerGridModel const* model,
                                             PGM_Idx* ranks) PGM_AMD_NOEXCEPT;

#ifdef __cplusplus
}
#endif

#endif
