// tap_changer.hpp -- automatic tap changing over batched power flows.
//
// Behaviour follows the reference optimizer (power_grid_model_c/power_grid_model/include/power_grid_model/
// optimizer/tap_position_optimizer.hpp, transformer_tap_regulator.hpp, main_model_impl.hpp:318-346): per
// scenario, regulated transformers are ranked by their distance from the sources, and each rank is
// adjusted until the (line-drop compensated) voltage of every control node lies in its band. The
// power flows of all scenarios that wait for one run as one device batch per round.
#pragma once

#include "engine.hpp"

#include <vector>

namespace pgm_amd {

// PGM_calculate with a tap changing strategy (1 any valid tap, 2 min voltage tap, 3 max voltage tap,
// 4 fast any tap)
void calculate_with_tap_changer(Model const& model, PGM_Options const& opt, PGM_MutableDataset const& output,
                                PGM_ConstDataset const* update, PGM_Handle* handle);

// transformer_tap_regulator output without tap changing: every regulator reports no tap position
// (main_core/output.hpp:381-396 with an empty optimizer output)
void write_null_regulator_output(Model const& model, PGM_MutableDataset const& output);
// fault output of a power flow: every fault reported not energised (fault.hpp:70-78)
void write_null_fault_output(Model const& model, PGM_MutableDataset const& output);

// rank of each regulator in the base state: 0 for transformers next to a source, growing with the number
// of regulated transformers in between; a regulator whose control side is also reached through a closer
// regulated transformer gets the last rank; -1 for regulators that are off or out of reach of any source.
// Throws PowerGridError for transformers controlled from the non-source side towards the source side.
std::vector<Idx> tap_regulator_ranks(Model const& model, ComponentState const& st);

}  // namespace pgm_amd
