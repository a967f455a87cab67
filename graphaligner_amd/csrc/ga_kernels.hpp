// Host-callable launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "ga_device_types.h"

namespace gad {

// One thread per request; arenaOffsets[i] locates request i's workspace.
hipError_t launchExtend(hipStream_t stream, uint32_t nreq, const DevGraph& g, const ExtParams& p, const ExtRequest* reqs, const char* seqbuf, uint64_t* tracebuf, ExtResult* results, unsigned char* arenas, const uint64_t* arenaOffsets, const ArenaCaps& caps, SliceSummary* sums, const uint64_t* sumOffsets);

// Workspace bytes of one request of length len (capacities for slices and band
// items scale with len; the others come from caps).
uint64_t requestArenaBytes(uint32_t len, const ArenaCaps& caps);

} // namespace gad
