// Plain device copy of a known number of bytes, used to calibrate the HBM
// traffic counters (FETCH_SIZE, WRITE_SIZE) that profile the extension kernel:
// the counters' reading for this copy over its true byte count is the scale
// applied to the extension kernel's readings.
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ void calibCopyKernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n)
{
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" int ga_calib_copy(uint64_t bytes, int repeats, double* msPerCopy)
{
	uint64_t n = bytes / sizeof(uint4);
	if (n == 0 || repeats < 1) return 1;
	uint4 *a = nullptr, *b = nullptr;
	if (hipMalloc(&a, n * sizeof(uint4)) != hipSuccess) return 2;
	if (hipMalloc(&b, n * sizeof(uint4)) != hipSuccess)
	{
		(void)hipFree(a);
		return 2;
	}
	int rc = 0;
	hipEvent_t e0, e1;
	if (hipMemset(a, 1, n * sizeof(uint4)) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = 3;
	if (rc == 0)
	{
		hipEventRecord(e0, nullptr);
		for (int r = 0; r < repeats; r++) hipLaunchKernelGGL(calibCopyKernel, dim3(4096), dim3(256), 0, nullptr, a, b, n);
		hipEventRecord(e1, nullptr);
		if (hipEventSynchronize(e1) != hipSuccess) rc = 4;
		float ms = 0;
		hipEventElapsedTime(&ms, e0, e1);
		*msPerCopy = ms / repeats;
		hipEventDestroy(e0);
		hipEventDestroy(e1);
	}
	(void)hipFree(a);
	(void)hipFree(b);
	return rc;
}
