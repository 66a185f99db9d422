/* Global render state.  Interface of reference include/srp/context.h:20-194.
 *
 * As with the reference, the user program defines `SRPContext srpContext;`
 * and calls srpNewContext(&srpContext) before anything else.  The struct
 * layout is kept member for member; the `arena` slot holds this library's
 * device runtime handle. */
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include "srp/message_callback.h"
#include "srp/arena.h"
#include "srp/detail/qualifiers.h"

typedef enum SRPProvokingVertexMode
{
	SRP_PROVOKING_VERTEX_FIRST,
	SRP_PROVOKING_VERTEX_LAST    /**< default */
} SRPProvokingVertexMode;

typedef enum SRPWinding
{
	SRP_WINDING_CCW,  /**< default */
	SRP_WINDING_CW
} SRPWinding;

typedef enum SRPFace
{
	SRP_FACE_NONE,
	SRP_FACE_FRONT,           /**< lines and points are always front-facing */
	SRP_FACE_BACK,
	SRP_FACE_FRONT_AND_BACK
} SRPFace;

typedef enum SRPPolygonMode
{
	SRP_POLYGON_MODE_FILL,   /**< default */
	SRP_POLYGON_MODE_LINE,
	SRP_POLYGON_MODE_POINT
} SRPPolygonMode;

typedef enum SRPCompareOp
{
	SRP_COMPARE_NEVER,
	SRP_COMPARE_ALWAYS,
	SRP_COMPARE_LESS,
	SRP_COMPARE_LEQUAL,
	SRP_COMPARE_GREATER,
	SRP_COMPARE_GEQUAL,
	SRP_COMPARE_EQUAL,
	SRP_COMPARE_NOTEQUAL
} SRPCompareOp;

typedef enum {
	SRP_STENCIL_KEEP,
	SRP_STENCIL_ZERO,
	SRP_STENCIL_REPLACE,
	SRP_STENCIL_INCR,       /**< saturates at 255 */
	SRP_STENCIL_INCR_WRAP,
	SRP_STENCIL_DECR,       /**< saturates at 0 */
	SRP_STENCIL_DECR_WRAP,
	SRP_STENCIL_INVERT
} SRPStencilOp;

typedef struct SRPRasterState
{
	SRPWinding frontFace;
	SRPFace cullFace;
	SRPPolygonMode polygonMode;
	float pointSize;              /**< side of a rasterised point, in pixels */
} SRPRasterState;

typedef struct SRPScissorState {
	bool enabled;
	size_t x;       /**< left edge of the scissor box */
	size_t y;       /**< top edge of the scissor box */
	size_t width;
	size_t height;
} SRPScissorState;

typedef struct SRPStencilFaceState {
	SRPCompareOp func;     /**< (ref & mask) func (stored & mask) */
	uint8_t ref;
	uint8_t mask;
	uint8_t writeMask;     /**< bits the stencil ops may change */
	SRPStencilOp sfailOp;  /**< stencil test failed */
	SRPStencilOp dfailOp;  /**< stencil passed, depth failed */
	SRPStencilOp passOp;   /**< both passed */
} SRPStencilFaceState;

typedef struct SRPStencilState {
	bool enabled;
	SRPStencilFaceState front;
	SRPStencilFaceState back;
} SRPStencilState;

typedef struct SRPDepthState
{
	bool testEnable;
	bool writeEnable;        /**< only effective while the test is enabled */
	SRPCompareOp compareOp;
} SRPDepthState;

typedef struct SRPContext
{
	SRPMessageCallback messageCallback;
	SRPProvokingVertexMode provokingVertexMode;
	SRPRasterState raster;
	SRPScissorState scissor;
	SRPStencilState stencil;
	SRPDepthState depth;
	SRPArena* arena;
} SRPContext;

SRP_EXTERN_C_BEGIN
void srpNewContext(SRPContext* pContext);
void srpSetMessageCallback(SRPMessageCallback callback);
void srpProvokingVertexMode(SRPProvokingVertexMode mode);
void srpRasterCullFace(SRPFace face);
void srpRasterFrontFace(SRPWinding face);
void srpRasterPolygonMode(SRPPolygonMode mode);
void srpRasterPointSize(float size);
void srpScissorTest(bool enable);
void srpScissorOptions(size_t x, size_t y, size_t width, size_t height);
/** Enables the stencil test.  Like the reference (core/context.c:96-99) the
 *  argument is ignored: the call always enables the test. */
void srpStencilTest(bool enable);
void srpStencilFunc(SRPCompareOp func, uint8_t ref, uint8_t mask);
void srpStencilFuncSeparate(SRPFace face, SRPCompareOp func, uint8_t ref, uint8_t mask);
void srpStencilOp(SRPStencilOp sfail, SRPStencilOp dfail, SRPStencilOp pass);
void srpStencilOpSeparate(SRPFace face, SRPStencilOp sfail, SRPStencilOp dfail, SRPStencilOp pass);
void srpStencilWriteMask(uint8_t mask);
void srpStencilWriteMaskSeparate(SRPFace face, uint8_t mask);
void srpDepthTest(bool enable);
void srpDepthWrite(bool enable);
void srpDepthCompareOp(SRPCompareOp op);

/** Defined by the user program (reference include/srp/context.h:194). */
extern SRPContext srpContext;
SRP_EXTERN_C_END
