/* A plain device copy of known size.  Its profiler byte counters
 * (FETCH_SIZE, WRITE_SIZE per dispatch) give the factor that turns the same
 * counters of the renderer's kernels into HBM bytes, and its event-timed
 * rate is the achievable copy bandwidth of the device. */
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

__global__ void __launch_bounds__(256) srp_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
	size_t stride = (size_t) gridDim.x * blockDim.x;
	for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		dst[i] = src[i];
}

/* `reps` copies of `bytes` (rounded down to 16) between two fresh device
 * buffers on the null stream.  msPerCopy receives the mean event-timed
 * duration of one copy.  Returns 0, or a negative code on a HIP error. */
extern "C" int srpBenchCopyProbe(size_t bytes, int reps, double* msPerCopy)
{
	size_t n = bytes / sizeof(uint4);
	if (n == 0 || reps <= 0) return -1;
	uint4 *a = NULL, *b = NULL;
	if (hipMalloc((void**) &a, n * sizeof(uint4)) != hipSuccess) return -2;
	if (hipMalloc((void**) &b, n * sizeof(uint4)) != hipSuccess)
	{
		(void) hipFree(a);
		return -2;
	}
	int rc = 0;
	hipEvent_t e0, e1;
	if (hipMemset(a, 0x5A, n * sizeof(uint4)) != hipSuccess || hipEventCreate(&e0) != hipSuccess)
		rc = -3;
	else if (hipEventCreate(&e1) != hipSuccess)
	{
		(void) hipEventDestroy(e0);
		rc = -3;
	}
	else
	{
		const unsigned blocks = 8192;  /* 32 workgroups per CU, grid-stride beyond */
		srp_copy_probe<<<blocks, 256>>>(a, b, n);  /* warm-up */
		(void) hipEventRecord(e0, 0);
		for (int r = 0; r < reps; r++)
			srp_copy_probe<<<blocks, 256>>>(a, b, n);
		(void) hipEventRecord(e1, 0);
		if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess)
			rc = -4;
		float ms = 0.0f;
		(void) hipEventElapsedTime(&ms, e0, e1);
		if (msPerCopy) *msPerCopy = (double) ms / reps;
		(void) hipEventDestroy(e0);
		(void) hipEventDestroy(e1);
	}
	(void) hipFree(a);
	(void) hipFree(b);
	return rc;
}
