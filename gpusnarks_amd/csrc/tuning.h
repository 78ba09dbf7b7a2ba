// Pass schedules: the default heuristic per (field, log2 n), overridable per
// size by the autotuner (gsn_tuning_set / a tuning file loaded at start-up).
#pragma once

#include <vector>

#include "gsn_launch.h"

namespace gsn {

std::vector<PassCfg> default_passes(int field, unsigned log_n);
std::vector<PassCfg> choose_passes(int field, unsigned log_n);

}  // namespace gsn
