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
