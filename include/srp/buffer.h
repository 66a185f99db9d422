/* Vertex/index buffers and the draw calls.  Interface of reference
 * include/srp/buffer.h:18-92 (entry points replaced: srpDrawVertexBuffer
 * core/buffer.c:53-59, srpDrawIndexBuffer core/buffer.c:130-136). */
#pragma once

#include "srp/framebuffer.h"
#include "srp/shaders.h"
#include "srp/detail/qualifiers.h"

typedef enum SRPPrimitive
{
	SRP_PRIM_POINTS,
	SRP_PRIM_LINES,
	SRP_PRIM_LINE_STRIP,
	SRP_PRIM_LINE_LOOP,
	SRP_PRIM_TRIANGLES,
	SRP_PRIM_TRIANGLE_STRIP,
	SRP_PRIM_TRIANGLE_FAN
} SRPPrimitive;

typedef struct SRPVertexBuffer SRPVertexBuffer;
typedef struct SRPIndexBuffer SRPIndexBuffer;

SRP_EXTERN_C_BEGIN
SRPVertexBuffer* srpNewVertexBuffer(void);
void srpFreeVertexBuffer(SRPVertexBuffer* this_);
/** Deep-copies `nBytesData` bytes; vertices are `nBytesPerVertex` apart. */
void srpVertexBufferCopyData(
	SRPVertexBuffer* this_, size_t nBytesPerVertex, size_t nBytesData, const void* data);
void srpDrawVertexBuffer(
	const SRPVertexBuffer* this_, const SRPFramebuffer* fb, const SRPShaderProgram* sp,
	SRPPrimitive primitive, size_t startIndex, size_t count);

SRPIndexBuffer* srpNewIndexBuffer(void);
/** `indicesType` is one of SRP_UINT8/16/32/64. */
void srpIndexBufferCopyData(
	SRPIndexBuffer* this_, SRPType indicesType, size_t nBytesData, const void* data);
void srpFreeIndexBuffer(SRPIndexBuffer* this_);
void srpDrawIndexBuffer(
	const SRPIndexBuffer* this_, const SRPVertexBuffer* vb, const SRPFramebuffer* fb,
	const SRPShaderProgram* sp, SRPPrimitive primitive, size_t startIndex, size_t count);
SRP_EXTERN_C_END
