// se.hpp -- batched Newton-Raphson state estimation.
#pragma once

#include <memory>

#include "dataset.hpp"
#include "model.hpp"

namespace pgm_amd {

class ChunkPool;

// state estimation of the model's state (update == nullptr) or of every scenario of a batch update: the
// results go into output; failed scenarios are reported through the handle as PGM_calculate reports them
// (main_model_impl.hpp calculate_state_estimation; job_dispatch.hpp:208-224)
// pool: device chunks of the model, reused across calls (may be null)
void run_state_estimation(Model const& model, PGM_Options const& options, PGM_MutableDataset const& output,
                          PGM_ConstDataset const* update, PGM_Handle* handle,
                          std::shared_ptr<ChunkPool> pool = nullptr);

}  // namespace pgm_amd
