// gfx950 kernels: seed extension with one extension per thread. Requests are
// sorted longest first so the 64 threads of a wave carry similar lengths.
#define GA_THREAD_PER_EXTENSION 1
#include <hip/hip_runtime.h>
#include "ga_extend_core.h"
#include "ga_kernels.hpp"

namespace gad {

__global__ void __launch_bounds__(64) extendThreadKernel(DevGraph g, ExtParams p, const ExtRequest* __restrict__ reqs, uint32_t nreq, const char* __restrict__ seqbuf, uint64_t* __restrict__ tracebuf, ExtResult* __restrict__ results, unsigned char* __restrict__ arenas, const uint64_t* __restrict__ arenaOffsets, ArenaCaps caps, SliceSummary* __restrict__ sums, const uint64_t* __restrict__ sumOffsets)
{
	uint32_t idx = blockIdx.x * 64 + threadIdx.x;
	if (idx >= nreq) return;
	ExtRequest r = reqs[idx];
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(r.len);
	c.maxItems = requestItems(r.len, caps);
	ExtResult out;
	runRequest(g, p, r, seqbuf, tracebuf, arenas + arenaOffsets[idx], c, sums ? sums + sumOffsets[idx] : nullptr, out);
	results[idx] = out;
}

hipError_t launchExtend(hipStream_t stream, uint32_t nreq, const DevGraph& g, const ExtParams& p, const ExtRequest* reqs, const char* seqbuf, uint64_t* tracebuf, ExtResult* results, unsigned char* arenas, const uint64_t* arenaOffsets, const ArenaCaps& caps, SliceSummary* sums, const uint64_t* sumOffsets)
{
	uint32_t blocks = (nreq + 63) / 64;
	hipLaunchKernelGGL(extendThreadKernel, dim3(blocks), dim3(64), 0, stream, g, p, reqs, nreq, seqbuf, tracebuf, results, arenas, arenaOffsets, caps, sums, sumOffsets);
	return hipGetLastError();
}

uint64_t requestArenaBytes(uint32_t len, const ArenaCaps& caps)
{
	ArenaCaps c = caps;
	c.maxSlices = requestSlices(len);
	c.maxItems = requestItems(len, caps);
	return arenaBytes(c);
}

} // namespace gad
