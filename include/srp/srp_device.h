/* HIP shader modules: include this in the .hip file that defines the
 * program's shaders, then register each shader with the macros below.
 *
 *   SRP_HIP_VERTEX_SHADER(symbol, fn, uniformSize)
 *   SRP_HIP_FRAGMENT_SHADER(symbol, fn, uniformSize)
 *
 * `fn` is a __device__ (or __host__ __device__) function with the
 * reference's shader signature (include/srp/shaders.h):
 *   void fn(SRPVertexShaderIn*, SRPVertexShaderOut*)
 *   void fn(SRPFragmentShaderIn*, SRPFragmentShaderOut*)
 * `symbol` becomes an extern "C" host function of that same signature.
 * Host C code uses it exactly where a reference program puts its shader
 * function pointer (SRPVertexShader.shader / SRPFragmentShader.shader);
 * the library maps it to the kernels instantiated here with `fn` inlined.
 * `uniformSize` is sizeof the uniform struct (its bytes are copied to the
 * device at every draw; an SRPTexture* inside it is already a device
 * handle).  FP contraction is switched off for the rest of the file, as
 * bit-exact results require.
 *
 *   SRP_HIP_FRAGMENT_SHADER_WAVE(symbol, fn, waveFn, uniformSize)
 *
 * registers a fragment shader that can also shade a whole wave of 64
 * fragments together (matrix-core work such as the neural shader of
 * srp/device/neural.hpp): waveFn(in, out, active) is called by all 64
 * lanes of the batched triangle rasteriser at once, `active` false on
 * lanes without a fragment (their inputs are zeroed and their output is
 * discarded); fn serves the kernels whose lanes are not convergent (lines,
 * points, the unbatched rasteriser).  In single triangle draws without
 * depth output such a shader is deferred: the rasteriser resolves
 * visibility first, and waveFn then shades only the fragment that wins
 * each pixel, with the same inputs it would have had (one evaluation per
 * visible pixel instead of one per fragment that passes the depth test;
 * the frame is the same, srpHipSetDeferredShading). */
#pragma once

#include "srp/srp.h"
#include "srp/vec.h"
#include "srp/mat.h"
#include "srp/srp_hip.h"
#include "srp/device/pipeline.hpp"

#pragma clang fp contract(off)

extern "C" void srpHipShaderCalledOnHost(const char* name);

#define SRP_HIP_VERTEX_SHADER(symbol, fn, uniformSize)                                            \
	struct SrpVsFunctor_##symbol                                                                  \
	{                                                                                             \
		__device__ __forceinline__ void operator()(SRPVertexShaderIn* i, SRPVertexShaderOut* o) const \
		{                                                                                         \
			fn(i, o);                                                                             \
		}                                                                                         \
	};                                                                                            \
	extern "C" void symbol(SRPVertexShaderIn*, SRPVertexShaderOut*) { srpHipShaderCalledOnHost(#symbol); } \
	static void srpVsLaunch_##symbol(const SrpVsLaunch* l, void* s)                               \
	{                                                                                             \
		srp::dev::launchVS<SrpVsFunctor_##symbol>(l, s);                                          \
	}                                                                                             \
	static const SrpVsEntry srpVsEntry_##symbol = {SRP_ABI_VERSION, (size_t) (uniformSize), srpVsLaunch_##symbol}; \
	static const int srpVsReg_##symbol = (srpHipRegisterVertexShader((const void*) &symbol, &srpVsEntry_##symbol), 0);

#define SRP_HIP_FS_REGISTER_(symbol, uniformSize, resolveFn)                                      \
	extern "C" void symbol(SRPFragmentShaderIn*, SRPFragmentShaderOut*) { srpHipShaderCalledOnHost(#symbol); } \
	static void srpFsRaster_##symbol(const SrpRasterLaunch* l, void* s)                           \
	{                                                                                             \
		srp::dev::launchRaster<SrpFsFunctor_##symbol>(l, s);                                      \
	}                                                                                             \
	static void srpFsFrags_##symbol(const SrpFragLaunch* l, void* s)                              \
	{                                                                                             \
		srp::dev::launchFrags<SrpFsFunctor_##symbol>(l, s);                                       \
	}                                                                                             \
	static const SrpFsEntry srpFsEntry_##symbol = {SRP_ABI_VERSION, (size_t) (uniformSize),       \
	                                               srpFsRaster_##symbol, srpFsFrags_##symbol, resolveFn}; \
	static const int srpFsReg_##symbol = (srpHipRegisterFragmentShader((const void*) &symbol, &srpFsEntry_##symbol), 0);

#define SRP_HIP_FRAGMENT_SHADER_WAVE(symbol, fn, waveFn, uniformSize)                             \
	struct SrpFsFunctor_##symbol                                                                  \
	{                                                                                             \
		static constexpr bool kWave = true;                                                       \
		__device__ __forceinline__ void operator()(SRPFragmentShaderIn* i, SRPFragmentShaderOut* o) const \
		{                                                                                         \
			fn(i, o);                                                                             \
		}                                                                                         \
		__device__ __forceinline__ void wave(SRPFragmentShaderIn* i, SRPFragmentShaderOut* o, bool active) const \
		{                                                                                         \
			waveFn(i, o, active);                                                                 \
		}                                                                                         \
	};                                                                                            \
	static void srpFsResolve_##symbol(const SrpRasterLaunch* l, void* s)                          \
	{                                                                                             \
		srp::dev::launchResolve<SrpFsFunctor_##symbol>(l, s);                                     \
	}                                                                                             \
	SRP_HIP_FS_REGISTER_(symbol, uniformSize, srpFsResolve_##symbol)

#define SRP_HIP_FRAGMENT_SHADER(symbol, fn, uniformSize)                                          \
	struct SrpFsFunctor_##symbol                                                                  \
	{                                                                                             \
		__device__ __forceinline__ void operator()(SRPFragmentShaderIn* i, SRPFragmentShaderOut* o) const \
		{                                                                                         \
			fn(i, o);                                                                             \
		}                                                                                         \
	};                                                                                            \
	SRP_HIP_FS_REGISTER_(symbol, uniformSize, nullptr)
