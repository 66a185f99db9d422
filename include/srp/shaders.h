/* Shader I/O contract.  Interface of reference include/srp/shaders.h:19-97:
 * identical member order, types and meaning, so shaders written for the
 * reference compile unchanged (as HIP device functions here, see srp_hip.h). */
#pragma once

#include <stdbool.h>
#include "srp/vertex.h"

typedef struct SRPUniform SRPUniform;

typedef struct SRPVertexShaderIn
{
	SRPUniform* uniform;  /**< the program's uniform block */
	SRPVertex* vertex;    /**< this vertex inside the vertex buffer */
	size_t vertexID;      /**< vertex index (not the stream position) */
} SRPVertexShaderIn;

typedef struct SRPVertexShaderOut
{
	union {
		float clipPosition[4];  /**< written by the shader */
		float ndcPosition[4];   /**< same storage after the perspective divide */
	};
	SRPVarying* varyings;       /**< where the shader writes its varyings */
} SRPVertexShaderOut;

typedef struct SRPVertexShader
{
	void (*shader)(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
	size_t nVaryings;              /**< entries in varyingsInfo */
	SRPVaryingInfo* varyingsInfo;  /**< layout of the varying bytes */
	size_t varyingsSize;           /**< bytes of varyings per vertex */
} SRPVertexShader;

typedef struct SRPFragmentShaderIn
{
	SRPUniform* uniform;
	SRPInterpolated* varyings;  /**< interpolated varyings of this fragment */
	float fragCoord[4];         /**< x+0.5, y+0.5, depth, 1/(interpolated 1/w) */
	bool frontFacing;
	size_t primitiveID;         /**< index of the assembled primitive in its draw */
} SRPFragmentShaderIn;

typedef struct SRPFragmentShaderOut
{
	float color[4];   /**< RGBA in [0,1], clamped and packed to RGBA8 */
	float fragDepth;  /**< honoured when not NaN and mayOverwriteDepth is set */
} SRPFragmentShaderOut;

typedef struct SRPFragmentShader
{
	void (*shader)(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
	bool mayOverwriteDepth;  /**< false: the depth test runs before the shader */
} SRPFragmentShader;

typedef struct SRPShaderProgram
{
	SRPUniform* uniform;
	SRPVertexShader* vs;
	SRPFragmentShader* fs;
} SRPShaderProgram;
