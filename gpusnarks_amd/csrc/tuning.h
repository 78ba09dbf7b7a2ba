// Pass schedules: the default heuristic per (field, log2 n), overridable per
// size by the autotuner (gsn_tuning_set / a tuning file loaded at start-up).
// A transform that runs alone and one fused into a batch (many transforms
// sharing each launch) may prefer different tile shapes; batch = true looks
// up the batch table first.
#pragma once

#include <vector>

#include "gsn_launch.h"

namespace gsn {

std::vector<PassCfg> default_passes(int field, unsigned log_n);
std::vector<PassCfg> choose_passes(int field, unsigned log_n, bool batch = false);

}  // namespace gsn
