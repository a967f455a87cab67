// Host-callable launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "ga_device_types.h"

namespace gad {

hipError_t launchExtend(hipStream_t stream, uint32_t numWaves, const DevGraph& g, const ExtParams& p, const ExtRequest* reqs, uint32_t nreq, const char* seqbuf, uint64_t* tracebuf, ExtResult* results, unsigned char* arenas, uint64_t arenaStride, const ArenaCaps& caps, uint32_t* counter, SliceSummary* sums, const uint64_t* sumOffsets);

uint64_t waveArenaBytes(const ArenaCaps& caps);

} // namespace gad
