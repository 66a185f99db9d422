/* CPU oracle for the srp draw pipeline — TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded restatement of the reference renderer's draw path
 * (kitrofimov/srp) that implements the same public C API (include/srp) with
 * host function-pointer shaders.  Only the tests, the smoke check and the
 * benchmark's CPU-baseline leg link it, always as the checker; the product
 * path (srp_amd) never calls into it.
 *
 * Every function cites the reference file:line whose arithmetic it restates.
 * Floating-point expressions keep the reference's operand order, float/double
 * promotions and roundings; the file must be compiled with
 * -ffp-contract=off and without -ffast-math (ISO C semantics, as the
 * reference's -std=c2x build has).
 *
 * Deliberate definitions where the reference has undefined behaviour:
 *  - Sutherland-Hodgman output up to 16 vertices / 14 triangles per input
 *    triangle (the reference's 6/4-entry buffers overflow for 7+ vertices,
 *    pipeline/clipping.c:83-84, primitive_assembly.c:59-60);
 *  - a line fragment whose linear pixel index y*W+x falls outside the
 *    framebuffer is dropped after the covered count (line.c:58-59 can round
 *    to x==W or y==H; the reference writes out of bounds there);
 *  - new framebuffers have a zeroed stencil buffer (the reference mallocs it;
 *    its tests rely on fresh zero pages);
 *  - colour channels that are NaN pack to 0 (x86 truncation of NaN gives 0).
 *
 * Parity pinning of each part is listed in the tests (tests/test_oracle.py). */
#include <math.h>
#include <stdarg.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SRP_INCLUDE_VEC
#define SRP_INCLUDE_MAT
#include "srp/srp.h"
#include "image_decode.h"
#include "png_encode.h"

/* vec/mat helpers: same definitions the library exports. */
#define SRP_MATH_FN
#include "srp/detail/vec_impl.h"
#include "srp/detail/mat_impl.h"
#undef SRP_MATH_FN

#define ORC_EPS 1e-9  /* reference math/utils.h:41 */
#define ORC_ROUGHLY_ZERO(x) (fabs(x) <= ORC_EPS)

/* ------------------------------------------------------------------ stats */

typedef struct OracleStats
{
	uint64_t covered;     /* emitFragment calls (fragment.c:63) */
	uint64_t shaded;      /* fragment shader invocations (fragment.c:98) */
	uint64_t primitives;  /* assembled primitives (post clip/cull) */
	uint64_t draws;
} OracleStats;

static OracleStats g_stats;
/* pixels of the triangles' bounding boxes walked by the raster loop
 * (triangle.c:73-111 visits each once) */
static uint64_t g_bboxPixels;

void srpOracleResetStats(void) { memset(&g_stats, 0, sizeof(g_stats)); g_bboxPixels = 0; }
uint64_t srpOracleBboxPixels(void) { return g_bboxPixels; }
void srpOracleGetStats(uint64_t out[4])
{
	out[0] = g_stats.covered; out[1] = g_stats.shaded;
	out[2] = g_stats.primitives; out[3] = g_stats.draws;
}

/* ------------------------------------------------------ message callback */

/* utils/message_callback.c:18-35 */
static void orcMessage(SRPMessageType type, SRPMessageSeverity sev, const char* fn, const char* fmt, ...)
{
	if (srpContext.messageCallback.func == NULL)
		return;
	char text[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(text, sizeof(text), fmt, ap);
	va_end(ap);
	srpContext.messageCallback.func(type, sev, fn, text, srpContext.messageCallback.userParameter);
}

/* ---------------------------------------------------------------- context */

/* core/context.c:15-51 (defaults) */
void srpNewContext(SRPContext* c)
{
	c->messageCallback.func = NULL;
	c->messageCallback.userParameter = NULL;
	c->provokingVertexMode = SRP_PROVOKING_VERTEX_LAST;
	c->raster.frontFace = SRP_WINDING_CCW;
	c->raster.cullFace = SRP_FACE_NONE;
	c->raster.polygonMode = SRP_POLYGON_MODE_FILL;
	c->raster.pointSize = 1.0f;
	memset(&c->scissor, 0, sizeof(c->scissor));
	c->depth.testEnable = false;
	c->depth.writeEnable = true;
	c->depth.compareOp = SRP_COMPARE_GREATER;
	SRPStencilFaceState face = {
		.func = SRP_COMPARE_ALWAYS, .ref = 0, .mask = 0xFF, .writeMask = 0xFF,
		.sfailOp = SRP_STENCIL_KEEP, .dfailOp = SRP_STENCIL_KEEP, .passOp = SRP_STENCIL_KEEP
	};
	c->stencil.enabled = false;
	c->stencil.front = face;
	c->stencil.back = face;
	c->arena = NULL;
}

/* core/context.c:53-180 */
void srpSetMessageCallback(SRPMessageCallback cb) { srpContext.messageCallback = cb; }
void srpProvokingVertexMode(SRPProvokingVertexMode m) { srpContext.provokingVertexMode = m; }
void srpRasterCullFace(SRPFace f) { srpContext.raster.cullFace = f; }
void srpRasterFrontFace(SRPWinding w) { srpContext.raster.frontFace = w; }
void srpRasterPolygonMode(SRPPolygonMode m) { srpContext.raster.polygonMode = m; }
void srpRasterPointSize(float s) { srpContext.raster.pointSize = s; }
void srpScissorTest(bool e) { srpContext.scissor.enabled = e; }
void srpScissorOptions(size_t x, size_t y, size_t w, size_t h)
{
	srpContext.scissor.x = x; srpContext.scissor.y = y;
	srpContext.scissor.width = w; srpContext.scissor.height = h;
}
/* context.c:96-99: the argument is ignored, the test is always enabled. */
void srpStencilTest(bool e) { (void) e; srpContext.stencil.enabled = true; }

static SRPStencilFaceState* orcFace(SRPFace f)
{
	return (f == SRP_FACE_FRONT) ? &srpContext.stencil.front : &srpContext.stencil.back;
}

void srpStencilFunc(SRPCompareOp func, uint8_t ref, uint8_t mask)
{
	SRPStencilFaceState* s[2] = {&srpContext.stencil.front, &srpContext.stencil.back};
	for (int i = 0; i < 2; i++) { s[i]->func = func; s[i]->ref = ref; s[i]->mask = mask; }
}
/* context.c:111-123: NONE ignored; FRONT_AND_BACK sets both, then the back face again. */
void srpStencilFuncSeparate(SRPFace face, SRPCompareOp func, uint8_t ref, uint8_t mask)
{
	if (face == SRP_FACE_NONE) return;
	if (face == SRP_FACE_FRONT_AND_BACK) srpStencilFunc(func, ref, mask);
	SRPStencilFaceState* s = orcFace(face);
	s->func = func; s->ref = ref; s->mask = mask;
}
void srpStencilOp(SRPStencilOp sfail, SRPStencilOp dfail, SRPStencilOp pass)
{
	SRPStencilFaceState* s[2] = {&srpContext.stencil.front, &srpContext.stencil.back};
	for (int i = 0; i < 2; i++) { s[i]->sfailOp = sfail; s[i]->dfailOp = dfail; s[i]->passOp = pass; }
}
void srpStencilOpSeparate(SRPFace face, SRPStencilOp sfail, SRPStencilOp dfail, SRPStencilOp pass)
{
	if (face == SRP_FACE_NONE) return;
	if (face == SRP_FACE_FRONT_AND_BACK) srpStencilOp(sfail, dfail, pass);
	SRPStencilFaceState* s = orcFace(face);
	s->sfailOp = sfail; s->dfailOp = dfail; s->passOp = pass;
}
void srpStencilWriteMask(uint8_t m)
{
	srpContext.stencil.front.writeMask = m;
	srpContext.stencil.back.writeMask = m;
}
void srpStencilWriteMaskSeparate(SRPFace face, uint8_t m)
{
	if (face == SRP_FACE_NONE) return;
	if (face == SRP_FACE_FRONT_AND_BACK) srpStencilWriteMask(m);
	orcFace(face)->writeMask = m;
}
void srpDepthTest(bool e) { srpContext.depth.testEnable = e; }
void srpDepthWrite(bool e) { srpContext.depth.writeEnable = e; }
void srpDepthCompareOp(SRPCompareOp op) { srpContext.depth.compareOp = op; }

/* ---------------------------------------------------------------- buffers */

struct SRPVertexBuffer { size_t stride, count, capacity; uint8_t* data; };
struct SRPIndexBuffer { SRPType type; size_t elem, count, capacity; uint8_t* data; };

/* utils/type.c:15-40 (signed integer types are not index types there either) */
static size_t orcSizeofIndex(SRPType t)
{
	switch (t)
	{
		case SRP_UINT8: return 1;
		case SRP_UINT16: return 2;
		case SRP_UINT32: return 4;
		case SRP_UINT64: return 8;
		case SRP_FLOAT: return sizeof(float);
		case SRP_DOUBLE: return sizeof(double);
		default:
			orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpSizeofType", "Unknown type (%i)", t);
			return 0;
	}
}

SRPVertexBuffer* srpNewVertexBuffer(void) { return (SRPVertexBuffer*) calloc(1, sizeof(SRPVertexBuffer)); }
void srpFreeVertexBuffer(SRPVertexBuffer* b) { if (b) { free(b->data); free(b); } }
/* core/buffer.c:30-45 */
void srpVertexBufferCopyData(SRPVertexBuffer* b, size_t stride, size_t nBytes, const void* data)
{
	if (nBytes > b->capacity) { free(b->data); b->data = (uint8_t*) malloc(nBytes); b->capacity = nBytes; }
	b->stride = stride;
	b->count = nBytes / stride;
	memcpy(b->data, data, nBytes);
}
SRPIndexBuffer* srpNewIndexBuffer(void)
{
	SRPIndexBuffer* b = (SRPIndexBuffer*) calloc(1, sizeof(SRPIndexBuffer));
	b->type = SRP_UINT8;
	b->elem = 1;
	return b;
}
void srpFreeIndexBuffer(SRPIndexBuffer* b) { if (b) { free(b->data); free(b); } }
/* core/buffer.c:77-94 */
void srpIndexBufferCopyData(SRPIndexBuffer* b, SRPType type, size_t nBytes, const void* data)
{
	if (nBytes > b->capacity) { free(b->data); b->data = (uint8_t*) malloc(nBytes); b->capacity = nBytes; }
	b->type = type;
	b->elem = orcSizeofIndex(type);
	b->count = nBytes / b->elem;
	memcpy(b->data, data, nBytes);
}

/* core/buffer.c:102-128 */
static uint64_t orcIndexAt(const SRPIndexBuffer* b, size_t i)
{
	const uint8_t* p = b->data + i * b->elem;
	switch (b->type)
	{
		case SRP_UINT8: return *p;
		case SRP_UINT16: { uint16_t v; memcpy(&v, p, 2); return v; }
		case SRP_UINT32: { uint32_t v; memcpy(&v, p, 4); return v; }
		case SRP_UINT64: { uint64_t v; memcpy(&v, p, 8); return v; }
		default:
			orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "indexIndexBuffer", "Unexpected type (%i)", b->type);
			return 0;
	}
}

/* ------------------------------------------------------------ framebuffer */

SRPFramebuffer* srpNewFramebuffer(size_t w, size_t h)
{
	SRPFramebuffer* fb = (SRPFramebuffer*) malloc(sizeof(SRPFramebuffer));
	fb->width = w; fb->height = h; fb->size = w * h;
	fb->color = (uint32_t*) malloc(sizeof(uint32_t) * fb->size);
	fb->depth = (float*) malloc(sizeof(float) * fb->size);
	fb->stencil = (uint8_t*) calloc(fb->size, 1);
	return fb;
}
void srpFreeFramebuffer(SRPFramebuffer* fb) { if (fb) { free(fb->color); free(fb->depth); free(fb->stencil); free(fb); } }

/* The test harness's frame writer (reference tests/utils/save.c:4-33). */
int srpFramebufferSaveImage(const SRPFramebuffer* fb, const char* path)
{
	if (!fb || !path || fb->width > (1u << 24) || fb->height > (1u << 24))
		return -1;
	return srpPngSaveColorBuffer(path, (int) fb->width, (int) fb->height, fb->color);
}
/* core/framebuffer.c:57-62: colour 0, depth -1, stencil untouched. */
void srpFramebufferClear(const SRPFramebuffer* fb)
{
	memset(fb->color, 0, fb->size * sizeof(uint32_t));
	for (size_t i = 0; i < fb->size; i++) fb->depth[i] = -1.0f;
}

/* core/framebuffer.c:48-55: x = (W/2)(ndc+1), y = (H/2)(1-ndc), in double. */
static void orcNdcToScreen(const SRPFramebuffer* fb, const float* ndc, float* ss)
{
	ss[0] = (float) (((double) (float) fb->width / 2.) * ((double) ndc[0] + 1.));
	ss[1] = (float) (((double) (float) fb->height / 2.) * (1. - (double) ndc[1]));
	ss[2] = ndc[2];
}

/* --------------------------------------------------------------- textures */

struct SRPTexture { uint8_t* rgb; int w, h, wM1, hM1; SRPTextureWrappingMode wrapX, wrapY; };

/* core/texture.c:27-48 (three channels requested) */
SRPTexture* srpNewTexture(const char* path, SRPTextureWrappingMode wx, SRPTextureWrappingMode wy)
{
	SRPTexture* t = (SRPTexture*) malloc(sizeof(SRPTexture));
	const char* err = NULL;
	t->rgb = srpImageLoadRGB(path, &t->w, &t->h, &err);
	if (t->rgb == NULL)
	{
		orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpNewTexture", "Failed to load image `%s`: %s", path, err);
		free(t);
		return NULL;
	}
	t->wM1 = t->w - 1; t->hM1 = t->h - 1;
	t->wrapX = wx; t->wrapY = wy;
	return t;
}

/* Test hook: a texture from caller RGB8 bytes (no file). */
SRPTexture* srpOracleTextureFromRGB(const uint8_t* rgb, int w, int h, SRPTextureWrappingMode wx, SRPTextureWrappingMode wy)
{
	SRPTexture* t = (SRPTexture*) malloc(sizeof(SRPTexture));
	t->rgb = (uint8_t*) malloc((size_t) w * h * 3);
	memcpy(t->rgb, rgb, (size_t) w * h * 3);
	t->w = w; t->h = h; t->wM1 = w - 1; t->hM1 = h - 1;
	t->wrapX = wx; t->wrapY = wy;
	return t;
}

void srpFreeTexture(SRPTexture* t) { if (t) { free(t->rgb); free(t); } }

/* core/texture.c:56-81 */
void srpTextureGetFilteredColor(const SRPTexture* t, float u, float v, float out[4])
{
	if (u < 0 || u > 1)
		u = (t->wrapX == TW_REPEAT) ? (float) (u - floor(u)) : (float) fmax(0.0, fmin(1.0, u));
	if (v < 0 || v > 1)
		v = (t->wrapY == TW_REPEAT) ? (float) (v - floor(v)) : (float) fmax(0.0, fmin(1.0, v));
	const float x = t->wM1 * u;
	const float y = t->hM1 * (1 - v);
	const size_t xi = (size_t) (x + 0.5);
	const size_t yi = (size_t) (y + 0.5);
	const uint8_t* texel = t->rgb + (xi + yi * (size_t) t->w) * 3;
	const float inv255 = (float) (1. / 255.);
	out[0] = texel[0] * inv255;
	out[1] = texel[1] * inv255;
	out[2] = texel[2] * inv255;
	out[3] = 1.0f;
}

/* core/texture.c:83-137 */
int srpTextureGet(SRPTexture* t, SRPTextureParameter p)
{
	if (p == SRP_TEXTURE_WRAPPING_MODE_X) return t->wrapX;
	if (p == SRP_TEXTURE_WRAPPING_MODE_Y) return t->wrapY;
	orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpTextureGet", "Unknown texture parameter (%i)", p);
	return -1;
}
void srpTextureSet(SRPTexture* t, SRPTextureParameter p, int v)
{
	if (p != SRP_TEXTURE_WRAPPING_MODE_X && p != SRP_TEXTURE_WRAPPING_MODE_Y)
	{
		orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpTextureSet", "Unknown texture parameter (%i)", p);
		return;
	}
	if (p == SRP_TEXTURE_WRAPPING_MODE_X) t->wrapX = (SRPTextureWrappingMode) v;
	else t->wrapY = (SRPTextureWrappingMode) v;
	if (v != TW_REPEAT && v != TW_CLAMP_TO_EDGE)
		orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpTextureSet",
		           "Unknown texture wrapping mode (%i). Falling back to TW_REPEAT", v);
}

/* ------------------------------------------------------ per-draw scratch */

/* Bump allocator for clip-generated varyings; reset after each draw
 * (the reference's arena, memory/arena.c:40-116, 8-byte aligned). */
typedef struct Scratch { uint8_t** blocks; size_t nBlocks; } Scratch;
static Scratch g_scratch;

static void* orcScratch(size_t n)
{
	/* Pad so that descriptor-driven reads past a short varying block stay in bounds. */
	uint8_t* p = (uint8_t*) calloc(1, n + 256);
	g_scratch.blocks = (uint8_t**) realloc(g_scratch.blocks, (g_scratch.nBlocks + 1) * sizeof(uint8_t*));
	g_scratch.blocks[g_scratch.nBlocks++] = p;
	return p;
}
static void orcScratchReset(void)
{
	for (size_t i = 0; i < g_scratch.nBlocks; i++) free(g_scratch.blocks[i]);
	free(g_scratch.blocks);
	g_scratch.blocks = NULL;
	g_scratch.nBlocks = 0;
}

/* ---------------------------------------------------------- interpolation */

static size_t orcElemSize(SRPType t)
{
	switch (t)
	{
		case SRP_FLOAT: case SRP_INT32: case SRP_UINT32: return 4;
		case SRP_DOUBLE: case SRP_INT64: case SRP_UINT64: return 8;
		case SRP_INT8: case SRP_UINT8: return 1;
		case SRP_INT16: case SRP_UINT16: return 2;
		default: return 0;
	}
}

/* pipeline/interpolation.c:93-163.  `invW == NULL` is the clipping case:
 * every floating varying is blended affinely (including FLAT ones). */
static void orcInterpolate(const SRPVertexShaderOut* v, size_t n, const float* w, const float* invW,
                           float rec, const SRPShaderProgram* sp, void* out)
{
	uint8_t* dst = (uint8_t*) out;
	size_t off = 0;
	const size_t pv = (srpContext.provokingVertexMode == SRP_PROVOKING_VERTEX_FIRST) ? 0 : n - 1;
	for (size_t a = 0; a < sp->vs->nVaryings; a++)
	{
		const SRPVaryingInfo* info = &sp->vs->varyingsInfo[a];
		bool persp = info->interpolationMode == SRP_INTERPOLATION_MODE_PERSPECTIVE;
		bool affine = info->interpolationMode == SRP_INTERPOLATION_MODE_AFFINE;
		if (invW == NULL) { persp = false; affine = true; }
		const uint8_t* av[3];
		for (size_t i = 0; i < n; i++) av[i] = (const uint8_t*) v[i].varyings + off;

		const size_t es = orcElemSize(info->type);
		if (info->type == SRP_FLOAT)
		{
			for (size_t e = 0; e < info->nItems; e++)
			{
				float val = 0.;
				if (persp)
				{
					for (size_t i = 0; i < n; i++) { float x; memcpy(&x, av[i] + 4 * e, 4); val += x * invW[i] * w[i]; }
					val *= rec;
				}
				else if (affine)
					for (size_t i = 0; i < n; i++) { float x; memcpy(&x, av[i] + 4 * e, 4); val += x * w[i]; }
				else
					memcpy(&val, av[pv] + 4 * e, 4);
				memcpy(dst + 4 * e, &val, 4);
			}
		}
		else if (info->type == SRP_DOUBLE)
		{
			for (size_t e = 0; e < info->nItems; e++)
			{
				double val = 0.;
				if (persp)
				{
					for (size_t i = 0; i < n; i++) { double x; memcpy(&x, av[i] + 8 * e, 8); val += x * invW[i] * w[i]; }
					val *= rec;
				}
				else if (affine)
					for (size_t i = 0; i < n; i++) { double x; memcpy(&x, av[i] + 8 * e, 8); val += x * w[i]; }
				else
					memcpy(&val, av[pv] + 8 * e, 8);
				memcpy(dst + 8 * e, &val, 8);
			}
		}
		else if (es != 0)
			memcpy(dst, av[pv], es * info->nItems);
		else
			orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "interpolateAttributes", "Unexpected type (%i)", info->type);
		dst += es * info->nItems;
		off += es * info->nItems;
	}
}

/* ------------------------------------------------------------ vertex stage */

typedef struct Cache
{
	size_t base, size;
	bool* valid;
	SRPVertexShaderOut* out;
	uint8_t* varyings;
} Cache;

typedef struct DrawIn
{
	const SRPIndexBuffer* ib;
	const SRPVertexBuffer* vb;
	const SRPFramebuffer* fb;
	const SRPShaderProgram* sp;
	SRPPrimitive prim;
	size_t start, count;
} DrawIn;

static size_t orcVertexIndex(const DrawIn* d, size_t stream)
{
	return d->ib ? (size_t) orcIndexAt(d->ib, stream) : stream;
}

/* pipeline/vertex_processing.c:55-74 */
static void orcRunVS(const DrawIn* d, size_t vi, void* varyings, SRPVertexShaderOut* out)
{
	SRPVertexShaderIn in;
	in.uniform = d->sp->uniform;
	in.vertex = (SRPVertex*) (d->vb->data + vi * d->vb->stride);
	in.vertexID = vi;
	memset(out->clipPosition, 0, sizeof(out->clipPosition));
	out->varyings = (SRPVarying*) varyings;
	d->sp->vs->shader(&in, out);
}

/* vertex_processing.c:23-36, 90-112: a dense cache over [min, max] vertex index. */
static void orcCacheInit(Cache* c, const DrawIn* d)
{
	size_t lo = SIZE_MAX, hi = 0;
	if (d->ib)
		for (size_t i = 0; i < d->count; i++)
		{
			const size_t vi = orcVertexIndex(d, d->start + i);
			if (vi < lo) lo = vi;
			if (vi > hi) hi = vi;
		}
	else { lo = d->start; hi = d->start + d->count - 1; }
	c->base = lo;
	c->size = hi - lo + 1;
	c->valid = (bool*) calloc(c->size, sizeof(bool));
	c->out = (SRPVertexShaderOut*) calloc(c->size, sizeof(SRPVertexShaderOut));
	c->varyings = (uint8_t*) calloc(c->size * d->sp->vs->varyingsSize + 256, 1);
}

static void orcCacheFree(Cache* c) { free(c->valid); free(c->out); free(c->varyings); }

/* vertex_processing.c:38-53 */
static SRPVertexShaderOut orcFetch(Cache* c, const DrawIn* d, size_t vi)
{
	const size_t k = vi - c->base;
	if (!c->valid[k])
	{
		orcRunVS(d, vi, c->varyings + k * d->sp->vs->varyingsSize, &c->out[k]);
		c->valid[k] = true;
	}
	return c->out[k];
}

/* vertex_processing.c:76-88: invW = 1/w in double, rounded to float; xyz *= invW; w = 1. */
static float orcDivide(SRPVertexShaderOut* v)
{
	const float invW = (float) (1.0 / (double) v->clipPosition[3]);
	v->ndcPosition[0] *= invW;
	v->ndcPosition[1] *= invW;
	v->ndcPosition[2] *= invW;
	v->ndcPosition[3] = 1.0f;
	return invW;
}

/* ---------------------------------------------------------------- clipping */

#define ORC_MAX_POLY 16

/* clipping.c:260-277 (plane order L, R, B, T, N, F) */
static float orcPlaneDist(const SRPVertexShaderOut* v, int p)
{
	const float x = v->clipPosition[0], y = v->clipPosition[1], z = v->clipPosition[2], w = v->clipPosition[3];
	switch (p)
	{
		case 0: return x + w;
		case 1: return w - x;
		case 2: return y + w;
		case 3: return w - y;
		case 4: return z + w;
		default: return w - z;
	}
}

/* clipping.c:121-137 */
static unsigned orcOutcode(const SRPVertexShaderOut* v)
{
	unsigned code = 0;
	for (int p = 0; p < 6; p++)
		code |= (orcPlaneDist(v, p) < 0) ? (1u << p) : 0u;
	return code;
}

/* clipping.c:244-258: position a(1-t)+bt; varyings blended affinely (floats)
 * or copied from the pair's provoking vertex (integers). */
static SRPVertexShaderOut orcLerpVertex(const SRPVertexShaderOut* a, const SRPVertexShaderOut* b, float t,
                                        const SRPShaderProgram* sp)
{
	SRPVertexShaderOut r;
	for (int i = 0; i < 4; i++)
		r.clipPosition[i] = a->clipPosition[i] * (1 - t) + b->clipPosition[i] * t;
	r.varyings = (SRPVarying*) orcScratch(sp->vs->varyingsSize);
	const SRPVertexShaderOut pair[2] = {*a, *b};
	const float w[2] = {1 - t, t};
	orcInterpolate(pair, 2, w, NULL, 0.f, sp, r.varyings);
	return r;
}

/* clipping.c:199-242: one Sutherland-Hodgman pass.  An edge crossing the
 * plane with |da-db| <= 1e-9 emits nothing at all (clipping.c:226-227). */
static size_t orcClipPass(const SRPVertexShaderOut* in, size_t n, int plane, const SRPShaderProgram* sp,
                          SRPVertexShaderOut* out)
{
	size_t m = 0;
	for (size_t i = 0; i < n; i++)
	{
		const SRPVertexShaderOut* cur = &in[i];
		const SRPVertexShaderOut* nxt = &in[(i + 1) % n];
		const float da = orcPlaneDist(cur, plane);
		const float db = orcPlaneDist(nxt, plane);
		const bool curIn = da >= 0, nxtIn = db >= 0;
		if (curIn && nxtIn)
			out[m++] = *nxt;
		else if (curIn || nxtIn)
		{
			if (ORC_ROUGHLY_ZERO(da - db))
				continue;
			const float t = da / (da - db);
			out[m++] = orcLerpVertex(cur, nxt, t, sp);
			if (!curIn && nxtIn)
				out[m++] = *nxt;
		}
		if (m + 2 > ORC_MAX_POLY) { fprintf(stderr, "oracle: clip polygon overflow\n"); abort(); }
	}
	return m;
}

/* clipping.c:68-119: returns the fan triangles (as vertex triples). */
static size_t orcClipTriangle(const SRPVertexShaderOut tri[3], const SRPShaderProgram* sp,
                              SRPVertexShaderOut out[][3])
{
	const unsigned c0 = orcOutcode(&tri[0]), c1 = orcOutcode(&tri[1]), c2 = orcOutcode(&tri[2]);
	if ((c0 | c1 | c2) == 0)
	{
		for (int i = 0; i < 3; i++) out[0][i] = tri[i];
		return 1;
	}
	if ((c0 & c1 & c2) != 0)
		return 0;
	SRPVertexShaderOut bufA[ORC_MAX_POLY], bufB[ORC_MAX_POLY];
	SRPVertexShaderOut* src = bufA;
	SRPVertexShaderOut* dst = bufB;
	for (int i = 0; i < 3; i++) src[i] = tri[i];
	size_t n = 3;
	for (int p = 0; p < 6; p++)
	{
		n = (n == 0) ? 0 : orcClipPass(src, n, p, sp, dst);
		if (n == 0)
			return 0;
		SRPVertexShaderOut* t = src; src = dst; dst = t;
	}
	size_t k = 0;
	for (size_t i = 1; i + 1 < n; i++, k++)
	{
		out[k][0] = src[0];
		out[k][1] = src[i];
		out[k][2] = src[i + 1];
	}
	return k;
}

/* clipping.c:139-183: both endpoints re-derived from the original pair. */
static bool orcClipLine(SRPVertexShaderOut v[2], const SRPShaderProgram* sp)
{
	const unsigned c0 = orcOutcode(&v[0]), c1 = orcOutcode(&v[1]);
	if ((c0 | c1) == 0) return false;
	if ((c0 & c1) != 0) return true;
	float t0 = 0., t1 = 1.;
	for (int p = 0; p < 6; p++)
	{
		const float da = orcPlaneDist(&v[0], p);
		const float db = orcPlaneDist(&v[1], p);
		if (da < 0. && db < 0.) return true;
		if (da < 0. || db < 0.)
		{
			if (ORC_ROUGHLY_ZERO(da - db)) continue;
			const float t = da / (da - db);
			if (da < 0.) t0 = (t0 > t) ? t0 : t;
			else t1 = (t1 > t) ? t : t1;
			if (t0 > t1) return true;
		}
	}
	const SRPVertexShaderOut a = v[0], b = v[1];
	if (t0 > 0.) v[0] = orcLerpVertex(&a, &b, t0, sp);
	if (t1 < 1.) v[1] = orcLerpVertex(&a, &b, t1, sp);
	return false;
}

/* clipping.c:185-197 */
static bool orcClipPoint(const SRPVertexShaderOut* v)
{
	const float x = v->clipPosition[0], y = v->clipPosition[1], z = v->clipPosition[2], w = v->clipPosition[3];
	return (x < -w || x > w) || (y < -w || y > w) || (z < -w || z > w);
}

/* ----------------------------------------------------------- primitives */

typedef struct OTri
{
	SRPVertexShaderOut v[3];
	float invW[3];
	float ss[3][3];
	float edge[3][3];
	bool tl[3];
	float minX, minY, maxX, maxY;
	float lambda[3], dldx[3], dldy[3];
	bool front;
	size_t id;
} OTri;

typedef struct OLine { SRPVertexShaderOut v[2]; float invW[2]; float ss[2][3]; size_t id; } OLine;
typedef struct OPoint { SRPVertexShaderOut v; size_t id; } OPoint;

static float orcCross(const float* a, const float* b) { return a[0] * b[1] - a[1] * b[0]; }

/* raster/triangle.c:113-160 (+ shouldCullTriangle :162-182, winding :184-195,
 * barycentrics :197-226, top-left :235-238). */
static bool orcSetupTriangle(OTri* t, const SRPFramebuffer* fb)
{
	for (int i = 0; i < 3; i++) t->invW[i] = orcDivide(&t->v[i]);

	const float* p0 = t->v[0].ndcPosition;
	const float* p1 = t->v[1].ndcPosition;
	const float* p2 = t->v[2].ndcPosition;
	const float e0[2] = {p1[0] - p0[0], p1[1] - p0[1]};
	const float e1[2] = {p2[0] - p0[0], p2[1] - p0[1]};
	const float area = orcCross(e0, e1);
	const bool ccw = area > 0;
	if (srpContext.raster.cullFace == SRP_FACE_FRONT_AND_BACK)
		return false;
	const bool front = ((area > 0) && srpContext.raster.frontFace == SRP_WINDING_CCW)
	                || ((area < 0) && srpContext.raster.frontFace == SRP_WINDING_CW);
	const bool cull = (front && srpContext.raster.cullFace == SRP_FACE_FRONT)
	               || (!front && srpContext.raster.cullFace == SRP_FACE_BACK);
	t->front = front;
	if (cull)
		return false;
	if (!ccw)
	{
		const SRPVertexShaderOut tv = t->v[1]; t->v[1] = t->v[2]; t->v[2] = tv;
		const float tw = t->invW[1]; t->invW[1] = t->invW[2]; t->invW[2] = tw;
	}
	for (int i = 0; i < 3; i++) orcNdcToScreen(fb, t->v[i].ndcPosition, t->ss[i]);
	for (int i = 0; i < 3; i++)
		for (int k = 0; k < 3; k++)
			t->edge[i][k] = t->ss[(i + 1) % 3][k] - t->ss[i][k];

	const float areaX2 = (float) fabs(orcCross(t->edge[0], t->edge[2]));
	if (ORC_ROUGHLY_ZERO(areaX2))
		return false;

#define ORC_MIN(a, b) ((a) > (b) ? (b) : (a))
#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
	const float mnx = ORC_MIN(t->ss[0][0], ORC_MIN(t->ss[1][0], t->ss[2][0]));
	const float mny = ORC_MIN(t->ss[0][1], ORC_MIN(t->ss[1][1], t->ss[2][1]));
	const float mxx = ORC_MAX(t->ss[0][0], ORC_MAX(t->ss[1][0], t->ss[2][0]));
	const float mxy = ORC_MAX(t->ss[0][1], ORC_MAX(t->ss[1][1], t->ss[2][1]));
	t->minX = (float) ORC_MAX(floor(mnx), 0.0);
	t->minY = (float) ORC_MAX(floor(mny), 0.0);
	t->maxX = (float) ORC_MIN(ceil(mxx), (double) fb->width);
	t->maxY = (float) ORC_MIN(ceil(mxy), (double) fb->height);

	const float px = (float) (t->minX + 0.5);
	const float py = (float) (t->minY + 0.5);
	float rel[3][2];
	for (int i = 0; i < 3; i++) { rel[i][0] = px - t->ss[i][0]; rel[i][1] = py - t->ss[i][1]; }
	/* lambda_i is the area opposite vertex i: edge (i+1) seen from vertex i+1. */
	for (int i = 0; i < 3; i++)
	{
		const int j = (i + 1) % 3;
		t->lambda[i] = orcCross(rel[j], t->edge[j]) / areaX2;
		t->dldx[i] = t->edge[j][1] / areaX2;
		t->dldy[i] = -t->edge[j][0] / areaX2;
	}
	for (int i = 0; i < 3; i++)
		t->tl[i] = ((t->edge[i][0] > 0) && ORC_ROUGHLY_ZERO(t->edge[i][1])) || (t->edge[i][1] < 0);
	return true;
}

/* raster/line.c:79-86 (the reference's third, out-of-range iteration only
 * scribbles fields that are rewritten before use). */
static void orcSetupLine(OLine* l, const SRPFramebuffer* fb)
{
	for (int i = 0; i < 2; i++) l->invW[i] = orcDivide(&l->v[i]);
	for (int i = 0; i < 2; i++) orcNdcToScreen(fb, l->v[i].ndcPosition, l->ss[i]);
}

/* ------------------------------------------------------- fragment stage */

/* fragment.c:156-178 and :225-245 */
static bool orcCompare(SRPCompareOp op, double lhs, double rhs)
{
	switch (op)
	{
		case SRP_COMPARE_NEVER: return false;
		case SRP_COMPARE_ALWAYS: return true;
		case SRP_COMPARE_LESS: return lhs < rhs;
		case SRP_COMPARE_LEQUAL: return lhs <= rhs;
		case SRP_COMPARE_GREATER: return lhs > rhs;
		case SRP_COMPARE_GEQUAL: return lhs >= rhs;
		case SRP_COMPARE_EQUAL: return lhs == rhs;
		case SRP_COMPARE_NOTEQUAL: return lhs != rhs;
		default: return false;
	}
}

/* fragment.c:204-223 */
static uint8_t orcStencilOp(SRPStencilOp op, uint8_t s, uint8_t ref)
{
	switch (op)
	{
		case SRP_STENCIL_KEEP: return s;
		case SRP_STENCIL_ZERO: return 0;
		case SRP_STENCIL_REPLACE: return ref;
		case SRP_STENCIL_INCR: return (s < 255) ? (uint8_t) (s + 1) : 255;
		case SRP_STENCIL_INCR_WRAP: return (uint8_t) (s + 1);
		case SRP_STENCIL_DECR: return (s > 0) ? (uint8_t) (s - 1) : 0;
		case SRP_STENCIL_DECR_WRAP: return (uint8_t) (s - 1);
		case SRP_STENCIL_INVERT: return (uint8_t) ~s;
		default:
			orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "applyStencilOp",
			           "Unknown stencil operation (%i), fallback to SRP_STENCIL_KEEP", op);
			return s;
	}
}

static void orcApplyStencil(uint8_t* s, SRPStencilOp op, const SRPStencilFaceState* st)
{
	const uint8_t v = orcStencilOp(op, *s, st->ref);
	*s = (uint8_t) ((*s & ~st->writeMask) | (v & st->writeMask));
}

/* core/color.c:14-23: clamp(c*255) truncated, R in the most significant byte. */
static uint32_t orcPack(const float c[4])
{
	uint32_t r = 0;
	for (int i = 0; i < 4; i++)
	{
		const float v = c[i] * 255;
		const float cl = (v < 0) ? 0.f : ((v > 255) ? 255.f : v);
		const uint32_t b = (cl != cl) ? 0u : (uint32_t) (uint8_t) cl;
		r |= b << (24 - 8 * i);
	}
	return r;
}

/* raster/fragment.c:63-125 */
static void orcEmit(const SRPFramebuffer* fb, const SRPShaderProgram* sp, int x, int y, SRPFragmentShaderIn* in)
{
	g_stats.covered++;
	const SRPContext* c = &srpContext;
	if (c->scissor.enabled)
	{
		const size_t ux = (size_t) x, uy = (size_t) y;
		if (ux < c->scissor.x || ux >= c->scissor.x + c->scissor.width ||
		    uy < c->scissor.y || uy >= c->scissor.y + c->scissor.height)
			return;
	}
	const size_t idx = (size_t) y * fb->width + (size_t) x;
	if (x < 0 || y < 0 || idx >= fb->size)
		return;
	uint32_t* pc = &fb->color[idx];
	float* pd = &fb->depth[idx];
	uint8_t* ps = &fb->stencil[idx];

	const bool early = !sp->fs->mayOverwriteDepth;
	const SRPStencilFaceState* st = in->frontFacing ? &c->stencil.front : &c->stencil.back;
	const float stored = *pd;
	float depth = in->fragCoord[2];

	if (c->stencil.enabled && !orcCompare(st->func, (double) (st->ref & st->mask), (double) (*ps & st->mask)))
	{
		orcApplyStencil(ps, st->sfailOp, st);
		return;
	}
	if (early && c->depth.testEnable && !orcCompare(c->depth.compareOp, depth, stored))
	{
		if (c->stencil.enabled) orcApplyStencil(ps, st->dfailOp, st);
		return;
	}
	SRPFragmentShaderOut out;
	memset(&out, 0, sizeof(out));
	out.fragDepth = NAN;
	g_stats.shaded++;
	sp->fs->shader(in, &out);
	if (!early)
	{
		if (!isnan(out.fragDepth)) depth = out.fragDepth;
		if (c->depth.testEnable && !orcCompare(c->depth.compareOp, depth, stored))
		{
			if (c->stencil.enabled) orcApplyStencil(ps, st->dfailOp, st);
			return;
		}
	}
	if (c->stencil.enabled) orcApplyStencil(ps, st->passOp, st);
	*pc = orcPack(out.color);
	if (c->depth.testEnable && c->depth.writeEnable)
		*pd = depth;
}

/* triangle.c:73-111 + interpolation.c:34-47: barycentrics walk the bounding
 * box by repeated float addition (dldx per pixel, dldy per row). */
static void orcRasterTriangle(OTri* t, const SRPFramebuffer* fb, const SRPShaderProgram* sp, void* ibuf)
{
	float row[3] = {t->lambda[0], t->lambda[1], t->lambda[2]};
	float lam[3] = {row[0], row[1], row[2]};
	if (t->maxX > t->minX && t->maxY > t->minY)
		g_bboxPixels += (uint64_t) (t->maxX - t->minX) * (uint64_t) (t->maxY - t->minY);
	for (size_t y = (size_t) t->minY; y < t->maxY; y++)
	{
		for (size_t x = (size_t) t->minX; x < t->maxX; x++)
		{
			bool inside = true;
			for (int i = 0; i < 3; i++)
				if (!((lam[i] > 0.) || (ORC_ROUGHLY_ZERO(lam[i]) && t->tl[i]))) { inside = false; break; }
			if (inside)
			{
				const float* w = lam;
				const float rec = 1 / (t->invW[0] * w[0] + t->invW[1] * w[1] + t->invW[2] * w[2]);
				const float depth = t->v[0].ndcPosition[2] * t->invW[0] * w[0]
				                  + t->v[1].ndcPosition[2] * t->invW[1] * w[1]
				                  + t->v[2].ndcPosition[2] * t->invW[2] * w[2];
				orcInterpolate(t->v, 3, w, t->invW, rec, sp, ibuf);
				SRPFragmentShaderIn in;
				in.uniform = sp->uniform;
				in.varyings = (SRPInterpolated*) ibuf;
				in.fragCoord[0] = (float) (x + 0.5);
				in.fragCoord[1] = (float) (y + 0.5);
				in.fragCoord[2] = depth;
				in.fragCoord[3] = rec;
				in.frontFacing = t->front;
				in.primitiveID = t->id;
				orcEmit(fb, sp, (int) x, (int) y, &in);
			}
			for (int i = 0; i < 3; i++) lam[i] += t->dldx[i];
		}
		for (int i = 0; i < 3; i++) { row[i] += t->dldy[i]; lam[i] = row[i]; }
	}
}

/* raster/line.c:34-77, :88-96; interpolation.c:49-60 (DDA, weights {1-t, t}) */
static void orcRasterLine(OLine* l, const SRPFramebuffer* fb, const SRPShaderProgram* sp, void* ibuf)
{
	const float dx = l->ss[1][0] - l->ss[0][0];
	const float dy = l->ss[1][1] - l->ss[0][1];
	int steps = (int) ceil(fmax(fabs(dx), fabs(dy)));
	if (steps == 0) steps = 1;
	const float xInc = dx / steps, yInc = dy / steps;
	const float tInc = (float) (1. / steps);
	float x = l->ss[0][0], y = l->ss[0][1], t = 0.;
	for (int i = 0; i <= steps; i++)
	{
		const int px = (int) round(x), py = (int) round(y);
		const float w[2] = {1 - t, t};
		const float rec = 1 / (l->invW[0] * w[0] + l->invW[1] * w[1]);
		const float depth = l->v[0].ndcPosition[2] * l->invW[0] * w[0] + l->v[1].ndcPosition[2] * l->invW[1] * w[1];
		orcInterpolate(l->v, 2, w, l->invW, rec, sp, ibuf);
		SRPFragmentShaderIn in;
		in.uniform = sp->uniform;
		in.varyings = (SRPInterpolated*) ibuf;
		in.fragCoord[0] = (float) (px + 0.5);
		in.fragCoord[1] = (float) (py + 0.5);
		in.fragCoord[2] = depth;
		in.fragCoord[3] = rec;
		in.frontFacing = true;
		in.primitiveID = l->id;
		orcEmit(fb, sp, px, py, &in);
		x += xInc; y += yInc; t += tInc;
	}
}

/* raster/point.c:32-118 */
static void orcRasterPoint(OPoint* p, const SRPFramebuffer* fb, const SRPShaderProgram* sp)
{
	const float size = srpContext.raster.pointSize;
	float ss[3];
	orcNdcToScreen(fb, p->v.ndcPosition, ss);
	const float half = (float) (size * 0.5);
	const float mnx = ss[0] - half, mny = ss[1] - half;
	const float mxx = ss[0] + half, mxy = ss[1] + half;
	int x0 = (int) floor(mnx), x1 = (int) floor(mxx);
	int y0 = (int) floor(mny), y1 = (int) floor(mxy);
	if (x1 < 0 || y1 < 0 || x0 >= (int) fb->width || y0 >= (int) fb->height)
		return;
	if (x0 < 0) x0 = 0;
	if (y0 < 0) y0 = 0;
	if (x1 >= (int) fb->width) x1 = (int) fb->width - 1;
	if (y1 >= (int) fb->height) y1 = (int) fb->height - 1;
	for (int y = y0; y <= y1; y++)
		for (int x = x0; x <= x1; x++)
		{
			const float px = (float) (x + 0.5), py = (float) (y + 0.5);
			if (px < mnx || px >= mxx || py < mny || py >= mxy)
				continue;
			SRPFragmentShaderIn in;
			in.uniform = sp->uniform;
			in.varyings = (SRPInterpolated*) p->v.varyings;
			in.fragCoord[0] = px;
			in.fragCoord[1] = py;
			in.fragCoord[2] = p->v.ndcPosition[2];
			in.fragCoord[3] = p->v.ndcPosition[3];
			in.frontFacing = true;
			in.primitiveID = p->id;
			orcEmit(fb, sp, x, y, &in);
		}
}

/* ------------------------------------------------------------- assembly */

typedef struct PrimList { void* items; size_t n, cap, size; } PrimList;

static void* orcPush(PrimList* l)
{
	if (l->n == l->cap)
	{
		l->cap = l->cap ? 2 * l->cap : 1024;
		l->items = realloc(l->items, l->cap * l->size);
	}
	return (uint8_t*) l->items + (l->n++) * l->size;
}

/* topology.c:24-48 */
static void orcTriStream(SRPPrimitive prim, size_t base, size_t k, size_t out[3])
{
	if (prim == SRP_PRIM_TRIANGLES) { out[0] = base + 3 * k; out[1] = base + 3 * k + 1; out[2] = base + 3 * k + 2; }
	else if (prim == SRP_PRIM_TRIANGLE_STRIP)
	{
		const bool odd = (k % 2) == 1;
		out[0] = base + k + (odd ? 1 : 0);
		out[1] = base + k + (odd ? 0 : 1);
		out[2] = base + k + 2;
	}
	else { out[0] = base; out[1] = base + k + 1; out[2] = base + k + 2; }
}

/* topology.c:62-83 */
static void orcLineStream(SRPPrimitive prim, size_t base, size_t k, size_t count, size_t out[2])
{
	if (prim == SRP_PRIM_LINES) { out[0] = base + 2 * k; out[1] = base + 2 * k + 1; }
	else if (prim == SRP_PRIM_LINE_STRIP) { out[0] = base + k; out[1] = base + k + 1; }
	else { out[0] = base + k; out[1] = base + (k + 1) % count; }
}

/* draw.c:84-125 + primitive_assembly.c:37-135 */
static void orcDrawTriangles(const DrawIn* d)
{
	if (srpContext.raster.cullFace == SRP_FACE_FRONT_AND_BACK)
		return;
	if (d->prim == SRP_PRIM_TRIANGLES && d->count % 3 != 0)
		orcMessage(SRP_MESSAGE_WARNING, SRP_MESSAGE_SEVERITY_LOW, "warnOnExcessVertexCount",
		           "Vertex count not divisible by 3 when drawing SRP_PRIM_TRIANGLES. The last %i vertex/vertices will be ignored\n",
		           (int) (d->count % 3));
	const size_t nTri = (d->prim == SRP_PRIM_TRIANGLES) ? d->count / 3 : (d->count >= 3 ? d->count - 2 : 0);
	if (nTri == 0)
		return;
	Cache cache;
	orcCacheInit(&cache, d);
	const SRPPolygonMode mode = srpContext.raster.polygonMode;
	PrimList tris = {NULL, 0, 0, sizeof(OTri)};
	PrimList lines = {NULL, 0, 0, sizeof(OLine)};
	PrimList points = {NULL, 0, 0, sizeof(OPoint)};
	size_t id = 0;
	for (size_t k = 0; k < nTri; k++)
	{
		size_t s[3];
		orcTriStream(d->prim, d->start, k, s);
		SRPVertexShaderOut v[3];
		for (int i = 0; i < 3; i++) v[i] = orcFetch(&cache, d, orcVertexIndex(d, s[i]));
		SRPVertexShaderOut fan[ORC_MAX_POLY][3];
		const size_t nc = orcClipTriangle(v, d->sp, fan);
		for (size_t c = 0; c < nc; c++)
		{
			if (mode == SRP_POLYGON_MODE_FILL)
			{
				OTri* t = (OTri*) orcPush(&tris);
				memcpy(t->v, fan[c], sizeof(t->v));
				if (!orcSetupTriangle(t, d->fb)) { tris.n--; continue; }
				t->id = id++;
			}
			else if (mode == SRP_POLYGON_MODE_LINE)
				for (int j = 0; j < 3; j++)
				{
					OLine* l = (OLine*) orcPush(&lines);
					l->v[0] = fan[c][j];
					l->v[1] = fan[c][(j + 1) % 3];
					orcSetupLine(l, d->fb);
					l->id = id++;
				}
			else if (mode == SRP_POLYGON_MODE_POINT)
				for (int j = 0; j < 3; j++)
				{
					OPoint* p = (OPoint*) orcPush(&points);
					p->v = fan[c][j];
					orcDivide(&p->v);
					p->id = id++;
				}
			else
				orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "assembleTrianglesGeneric",
				           "Unexpected srpContext.raster.polygonMode (%i)", mode);
		}
	}
	g_stats.primitives += id;
	void* ibuf = calloc(1, d->sp->vs->varyingsSize + 256);
	for (size_t i = 0; i < tris.n; i++) orcRasterTriangle(&((OTri*) tris.items)[i], d->fb, d->sp, ibuf);
	for (size_t i = 0; i < lines.n; i++) orcRasterLine(&((OLine*) lines.items)[i], d->fb, d->sp, ibuf);
	for (size_t i = 0; i < points.n; i++) orcRasterPoint(&((OPoint*) points.items)[i], d->fb, d->sp);
	free(ibuf);
	free(tris.items); free(lines.items); free(points.items);
	orcCacheFree(&cache);
}

/* draw.c:127-146 + primitive_assembly.c:137-178 */
static void orcDrawLines(const DrawIn* d)
{
	if (d->prim == SRP_PRIM_LINES && d->count % 2 != 0)
		orcMessage(SRP_MESSAGE_WARNING, SRP_MESSAGE_SEVERITY_LOW, "warnOnExcessVertexCount",
		           "Odd vertex count when drawing SRP_PRIM_LINES. The last vertex will be ignored\n");
	size_t nLines;
	if (d->prim == SRP_PRIM_LINES) nLines = d->count / 2;
	else if (d->prim == SRP_PRIM_LINE_STRIP) nLines = (d->count != 1) ? d->count - 1 : 0;
	else nLines = (d->count != 1) ? d->count : 0;
	if (nLines == 0)
		return;
	Cache cache;
	orcCacheInit(&cache, d);
	OLine* lines = (OLine*) calloc(nLines, sizeof(OLine));
	size_t id = 0;
	for (size_t k = 0; k < nLines; k++)
	{
		OLine* l = &lines[id];
		size_t s[2];
		orcLineStream(d->prim, d->start, k, d->count, s);
		for (int i = 0; i < 2; i++) l->v[i] = orcFetch(&cache, d, orcVertexIndex(d, s[i]));
		if (orcClipLine(l->v, d->sp))
			continue;
		orcSetupLine(l, d->fb);
		l->id = id++;
	}
	g_stats.primitives += id;
	void* ibuf = calloc(1, d->sp->vs->varyingsSize + 256);
	for (size_t i = 0; i < id; i++) orcRasterLine(&lines[i], d->fb, d->sp, ibuf);
	free(ibuf);
	free(lines);
	orcCacheFree(&cache);
}

/* draw.c:148-165 + primitive_assembly.c:221-254 (no vertex cache) */
static void orcDrawPoints(const DrawIn* d)
{
	if (srpContext.raster.pointSize <= 0.)
		return;
	OPoint* pts = (OPoint*) calloc(d->count, sizeof(OPoint));
	uint8_t* vary = (uint8_t*) calloc(d->count * d->sp->vs->varyingsSize + 256, 1);
	size_t id = 0;
	for (size_t k = 0; k < d->count; k++)
	{
		OPoint* p = &pts[id];
		orcRunVS(d, orcVertexIndex(d, d->start + k), vary + k * d->sp->vs->varyingsSize, &p->v);
		if (orcClipPoint(&p->v))
			continue;
		orcDivide(&p->v);
		p->id = id++;
	}
	g_stats.primitives += id;
	for (size_t i = 0; i < id; i++) orcRasterPoint(&pts[i], d->fb, d->sp);
	free(vary);
	free(pts);
}

/* draw.c:63-82, :167-189 */
static void orcDraw(const DrawIn* d)
{
	if (d->count == 0)
		return;
	const size_t end = d->start + d->count - 1;
	const size_t size = d->ib ? d->ib->count : d->vb->count;
	if (end >= size)
	{
		orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "checkOOB",
		           d->ib ? "Attempt to OOB access index buffer (read) at indices %zu-%zu (size: %zu)\n"
		                 : "Attempt to OOB access vertex buffer (read) at indices %zu-%zu (size: %zu)\n",
		           d->start, end, size);
		return;
	}
	g_stats.draws++;
	switch (d->prim)
	{
		case SRP_PRIM_TRIANGLES: case SRP_PRIM_TRIANGLE_STRIP: case SRP_PRIM_TRIANGLE_FAN:
			orcDrawTriangles(d); break;
		case SRP_PRIM_LINES: case SRP_PRIM_LINE_STRIP: case SRP_PRIM_LINE_LOOP:
			orcDrawLines(d); break;
		case SRP_PRIM_POINTS:
			orcDrawPoints(d); break;
		default:
			orcMessage(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "Unknown primitive type: %i", d->prim);
	}
	orcScratchReset();
}

void srpDrawVertexBuffer(const SRPVertexBuffer* vb, const SRPFramebuffer* fb, const SRPShaderProgram* sp,
                         SRPPrimitive prim, size_t start, size_t count)
{
	const DrawIn d = {NULL, vb, fb, sp, prim, start, count};
	orcDraw(&d);
}

void srpDrawIndexBuffer(const SRPIndexBuffer* ib, const SRPVertexBuffer* vb, const SRPFramebuffer* fb,
                        const SRPShaderProgram* sp, SRPPrimitive prim, size_t start, size_t count)
{
	const DrawIn d = {ib, vb, fb, sp, prim, start, count};
	orcDraw(&d);
}
