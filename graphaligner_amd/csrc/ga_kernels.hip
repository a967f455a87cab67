// gfx950 kernels: the persistent seed-extension kernel (one wavefront per
// extension, extensions pulled from a device work counter).
#include <hip/hip_runtime.h>
#include "ga_extend_core.h"
#include "ga_kernels.hpp"

namespace gad {

__global__ void __launch_bounds__(64) extendKernel(DevGraph g, ExtParams p, const ExtRequest* __restrict__ reqs, uint32_t nreq, const char* __restrict__ seqbuf, uint64_t* __restrict__ tracebuf, ExtResult* __restrict__ results, unsigned char* __restrict__ arenas, uint64_t arenaStride, ArenaCaps caps, uint32_t* __restrict__ counter, SliceSummary* __restrict__ sums, const uint64_t* __restrict__ sumOffsets)
{
	unsigned char* arena = arenas + (uint64_t)blockIdx.x * arenaStride;
	while (true)
	{
		uint32_t idx = 0;
		if (threadIdx.x == 0) idx = atomicAdd(counter, 1u);
		idx = __shfl(idx, 0);
		if (idx >= nreq) break;
		ExtRequest r = reqs[idx];
		ExtResult out;
		runRequest(g, p, r, seqbuf, tracebuf, arena, caps, sums ? sums + sumOffsets[idx] : nullptr, out);
		if (threadIdx.x == 0) results[idx] = out;
	}
}

hipError_t launchExtend(hipStream_t stream, uint32_t numWaves, const DevGraph& g, const ExtParams& p, const ExtRequest* reqs, uint32_t nreq, const char* seqbuf, uint64_t* tracebuf, ExtResult* results, unsigned char* arenas, uint64_t arenaStride, const ArenaCaps& caps, uint32_t* counter, SliceSummary* sums, const uint64_t* sumOffsets)
{
	hipLaunchKernelGGL(extendKernel, dim3(numWaves), dim3(64), 0, stream, g, p, reqs, nreq, seqbuf, tracebuf, results, arenas, arenaStride, caps, counter, sums, sumOffsets);
	return hipGetLastError();
}

uint64_t waveArenaBytes(const ArenaCaps& caps) { return arenaBytes(caps); }

} // namespace gad
