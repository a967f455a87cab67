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

hipError_t launchSeedLookup(hipStream_t stream, const char* reads, const uint64_t* readOff, const uint32_t* readLen, uint32_t nreads, const SeedTableSlot* table, uint64_t mask, uint32_t k, uint32_t w, uint64_t maxCount, SeedMatch* out, uint32_t* outCount);

// Copies matches[readOff[i], + count[i]) of every read i to out[outOff[i], ...).
hipError_t launchGatherMatches(hipStream_t stream, const SeedMatch* matches, const uint64_t* readOff, const uint32_t* count, const uint64_t* outOff, uint32_t nreads, SeedMatch* out);

// Copies src[srcOff[i], + count[i]) to dst[dstOff[i], ...) for every range i.
hipError_t launchGatherU64(hipStream_t stream, const uint64_t* src, const uint64_t* srcOff, const uint32_t* count, const uint64_t* dstOff, uint32_t nranges, uint64_t* dst);

hipError_t launchRevcomp(hipStream_t stream, const char* reads, char* rev, const uint64_t* readOff, const uint32_t* readLen, uint32_t nreads, uint32_t maxLen);

} // namespace gad
