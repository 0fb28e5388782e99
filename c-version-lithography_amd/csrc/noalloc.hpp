// Allocation-free entry points (noalloc.hip): their scratch is a static device
// array of kNoAllocScratchBytes per GPU.
#pragma once
#include <cstddef>

#include "plan.hpp"

namespace litho {

// 64 MiB: reinit_sdf of up to 4 Mi pixels per call (16 bytes per pixel), e.g.
// one 2048 x 2048 tile or ten 640 x 640 Micros.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;

// Looks up this device's address of the static scratch array (once per
// device; the device must be current).  Called by litho_ctx_create.
void noalloc_init(int device);

// reinit_sdf (level_set_utils.cpp:290-319) on the static scratch: no device
// allocation.  Calls that overlap on one GPU share the array and race.
void ls_reinit_sdf_noalloc(Context& ctx, int batch, int rows, int cols, const float* phi, double dx, double dy,
                           float* sdf);

}  // namespace litho
