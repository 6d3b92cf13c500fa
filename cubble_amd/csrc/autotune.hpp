// autotune.hpp -- launch-shape autotuner over an engine's foams.
#pragma once

#include <cstdint>

#include "cubble_amd.h"
#include "engine.hpp"

namespace cubble_amd {

// Times candidate launch shapes on the engine's own foams (state restored
// afterwards), applies the fastest, and stores it in the optional cache.
int autotune(Engine &e, uint32_t flags, const cubble_autotune_opts *o, cubble_autotune_report *rep);

}  // namespace cubble_amd
