/* Vertex shaders of shader_list.h (part 0) and the terrain's neural
 * fragment shader; the other fragment shaders are in device_shaders_1..4.hip. */
#define SC_DEVICE_PART 0
#include "device_shaders_part.h"

/* The terrain's neural material (GPU only): the network of
 * srp/device/neural.hpp evaluated on the fragment's surface features; the
 * uniform's `network` points to its SrpNnDevice view. */
#include <srp/device/neural.hpp>

static __device__ __forceinline__ void terrainNeuralColor(const float rgb[3], SRPFragmentShaderOut* out)
{
	for (int c = 0; c < 3; c++)
		out->color[c] = rgb[c];
	out->color[3] = 1.0f;
}

static __device__ void fsTerrainNeuralLane(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out)
{
	const UniTerrain* u = (const UniTerrain*) in->uniform;
	float f[8], rgb[3];
	terrainFeatures((const VarTerrain*) in->varyings, &u->bounds, f);
	srp::nn::inferLane((const SrpNnDevice*) u->network, f, rgb);
	terrainNeuralColor(rgb, out);
}

static __device__ void fsTerrainNeuralWave(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out, bool active)
{
	const UniTerrain* u = (const UniTerrain*) in->uniform;
	float f[8], rgb[3];
	terrainFeatures((const VarTerrain*) in->varyings, &u->bounds, f);
	if (!active)
		for (int k = 0; k < 8; k++) f[k] = 0.0f;
	srp::nn::inferWave((const SrpNnDevice*) u->network, f, rgb);
	terrainNeuralColor(rgb, out);
}

SRP_HIP_FRAGMENT_SHADER_WAVE(sc_fsTerrainNeural, fsTerrainNeuralLane, fsTerrainNeuralWave, sizeof(UniTerrain))

/* The procedural material alone, dense over decoded training features (the
 * surface point and normal of each captured fragment): what the plain
 * shader costs per fragment without the rasteriser around it. */
static __global__ __launch_bounds__(256) void k_terrainDense(const float* feat, size_t n, TerrainBounds b, float* rgb)
{
	const size_t i = (size_t) blockIdx.x * 256 + threadIdx.x;
	if (i >= n)
		return;
	const float* f = feat + i * 8;
	const vec3 p = VEC3(f[0] / b.inv[0] + b.lo[0], f[1] / b.inv[1] + b.lo[1], f[2] / b.inv[2] + b.lo[2]);
	const vec3 nrm = VEC3(f[3], f[4], f[5]);
	float c[4];
	terrainShade(p, nrm, c);
	rgb[i * 3 + 0] = c[0];
	rgb[i * 3 + 1] = c[1];
	rgb[i * 3 + 2] = c[2];
}

/* Device ms per dense pass of the procedural material over `n` samples of
 * host features[n][8] (mean of `reps` after a warm-up), on the library
 * stream; -1 on a device error. */
extern "C" double srpBenchTerrainDenseMs(const float* features, size_t n, const float bounds[6], int reps)
{
	hipStream_t s = (hipStream_t) srpHipGetStream();
	float *dF = nullptr, *dRgb = nullptr;
	if (hipMalloc((void**) &dF, n * 8 * sizeof(float)) != hipSuccess || hipMalloc((void**) &dRgb, n * 3 * sizeof(float)) != hipSuccess)
		return -1.0;
	TerrainBounds b;
	for (int k = 0; k < 3; k++) { b.lo[k] = bounds[k]; b.inv[k] = bounds[3 + k]; }
	double ms = -1.0;
	hipEvent_t e0, e1;
	if (hipMemcpyAsync(dF, features, n * 8 * sizeof(float), hipMemcpyHostToDevice, s) == hipSuccess &&
	    hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess)
	{
		const unsigned blocks = (unsigned) ((n + 255) / 256);
		hipLaunchKernelGGL(k_terrainDense, dim3(blocks), dim3(256), 0, s, (const float*) dF, n, b, dRgb);
		hipEventRecord(e0, s);
		for (int r = 0; r < reps; r++)
			hipLaunchKernelGGL(k_terrainDense, dim3(blocks), dim3(256), 0, s, (const float*) dF, n, b, dRgb);
		hipEventRecord(e1, s);
		float t = 0.0f;
		if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess)
			ms = (double) t / reps;
		hipEventDestroy(e0);
		hipEventDestroy(e1);
	}
	hipStreamSynchronize(s);
	hipFree(dF);
	hipFree(dRgb);
	return ms;
}
