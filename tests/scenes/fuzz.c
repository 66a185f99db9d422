/* Randomised parity scene: random render state, primitive type, index
 * type, clip-space geometry (including straddling, degenerate, grid-snapped
 * and w<=0 vertices), varying types and interpolation modes, several draws
 * into one small framebuffer with random initial depth and stencil.
 * Uses the public API only (runs on the oracle, the compiled reference and
 * the GPU library alike). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "shaders.h"
#include "scenes.h"

void sc_vsFuzz(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsFuzz(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_fsFuzzDepth(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);

static uint64_t fzNext(uint64_t* s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static double fzU(uint64_t* s) { return (double) (fzNext(s) >> 11) * (1.0 / 9007199254740992.0); }
static int fzI(uint64_t* s, int n) { return (int) (fzNext(s) % (uint64_t) n); }
static float fzF(uint64_t* s, float lo, float hi) { return (float) (lo + (hi - lo) * fzU(s)); }

static float fzCoord(uint64_t* s)
{
	const int k = fzI(s, 10);
	if (k < 3) return (float) (fzI(s, 33) - 16) / 8.0f;   /* grid-snapped: shared edges, axis-aligned */
	if (k < 4) return fzF(s, -4.0f, 4.0f);                /* far outside */
	return fzF(s, -1.4f, 1.4f);
}

static SRPStencilFaceState fzFace(uint64_t* s)
{
	SRPStencilFaceState f;
	f.func = (SRPCompareOp) fzI(s, 8);
	f.ref = (uint8_t) fzNext(s);
	f.mask = (fzI(s, 2) == 0) ? 0xFF : (uint8_t) fzNext(s);
	f.writeMask = (fzI(s, 2) == 0) ? 0xFF : (uint8_t) fzNext(s);
	f.sfailOp = (SRPStencilOp) fzI(s, 8);
	f.dfailOp = (SRPStencilOp) fzI(s, 8);
	f.passOp = (SRPStencilOp) fzI(s, 8);
	return f;
}

int srpSceneFuzz(const SceneArgs* a, SceneOut* o)
{
	uint64_t s = a->seed * 0x2545F4914F6CDD1Dull + 17;
	const size_t W = a->width ? (size_t) a->width : (size_t) (24 + fzI(&s, 180));
	const size_t H = a->height ? (size_t) a->height : (size_t) (24 + fzI(&s, 180));

	srpNewContext(&srpContext);
	SRPFramebuffer* fb = srpNewFramebuffer(W, H);
	srpFramebufferClear(fb);
	for (size_t i = 0; i < fb->size; i++)
	{
		fb->depth[i] = (fzI(&s, 4) == 0) ? fzF(&s, -1.0f, 1.0f) : -1.0f;
		fb->stencil[i] = (uint8_t) ((fzI(&s, 3) == 0) ? fzNext(&s) : 0);
	}

	const int nVerts = 1 + fzI(&s, 160);
	VtxFuzz* v = (VtxFuzz*) calloc((size_t) nVerts, sizeof(VtxFuzz));
	for (int i = 0; i < nVerts; i++)
	{
		float w = fzF(&s, 0.25f, 2.0f);
		if (fzI(&s, 25) == 0) w = fzF(&s, -1.0f, 0.1f);
		const float x = fzCoord(&s), y = fzCoord(&s), z = fzF(&s, -1.3f, 1.3f);
		v[i].clip = VEC4(x * w, y * w, z * w, w);
		for (int k = 0; k < 3; k++) v[i].var.f[k] = fzF(&s, -0.5f, 1.5f);
		v[i].var.d = fzU(&s) * 3.0 - 1.0;
		v[i].var.i32 = (int32_t) fzNext(&s);
		v[i].var.u16 = (uint16_t) fzNext(&s);
		v[i].var.i8 = (int8_t) fzNext(&s);
		v[i].var.u64 = fzNext(&s);
		v[i].var.g = fzF(&s, 0.0f, 1.0f);
	}
	/* Exact duplicates make zero-area and shared-edge cases. */
	for (int i = 1; i < nVerts; i++)
		if (fzI(&s, 12) == 0) v[i].clip = v[fzI(&s, i)].clip;

	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxFuzz), (size_t) nVerts * sizeof(VtxFuzz), v);

	const int nIdx = 1 + fzI(&s, 300);
	const SRPType types[4] = {SRP_UINT8, SRP_UINT16, SRP_UINT32, SRP_UINT64};
	const SRPType itype = types[fzI(&s, 4)];
	const int vmax = (itype == SRP_UINT8 && nVerts > 256) ? 256 : nVerts;
	uint8_t* ibytes = (uint8_t*) calloc((size_t) nIdx, 8);
	for (int i = 0; i < nIdx; i++)
	{
		const uint64_t vi = (uint64_t) fzI(&s, vmax);
		switch (itype)
		{
			case SRP_UINT8: ibytes[i] = (uint8_t) vi; break;
			case SRP_UINT16: ((uint16_t*) ibytes)[i] = (uint16_t) vi; break;
			case SRP_UINT32: ((uint32_t*) ibytes)[i] = (uint32_t) vi; break;
			default: ((uint64_t*) ibytes)[i] = vi; break;
		}
	}
	const size_t isz = (itype == SRP_UINT8) ? 1 : (itype == SRP_UINT16) ? 2 : (itype == SRP_UINT32) ? 4 : 8;
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpIndexBufferCopyData(ib, itype, (size_t) nIdx * isz, ibytes);

	SRPVaryingInfo info[7] = {
		{3, SRP_FLOAT, (SRPInterpolationMode) fzI(&s, 3)},
		{1, SRP_DOUBLE, (SRPInterpolationMode) fzI(&s, 3)},
		{1, SRP_INT32, (SRPInterpolationMode) fzI(&s, 3)},
		{1, SRP_UINT16, (SRPInterpolationMode) fzI(&s, 3)},
		{1, SRP_INT8, (SRPInterpolationMode) fzI(&s, 3)},
		{1, SRP_UINT64, (SRPInterpolationMode) fzI(&s, 3)},
		{1, SRP_FLOAT, (SRPInterpolationMode) fzI(&s, 3)},
	};
	SRPVertexShader vs;
	vs.shader = sc_vsFuzz; vs.nVaryings = 7; vs.varyingsInfo = info; vs.varyingsSize = sizeof(VarFuzz);
	SRPFragmentShader fs;
	const bool late = fzI(&s, 3) == 0;
	fs.shader = late ? sc_fsFuzzDepth : sc_fsFuzz;
	fs.mayOverwriteDepth = late;
	SRPShaderProgram sp;
	sp.uniform = NULL; sp.vs = &vs; sp.fs = &fs;

	const int nDraws = 1 + fzI(&s, 4);
	for (int d = 0; d < nDraws; d++)
	{
		srpRasterCullFace((SRPFace) ((fzI(&s, 10) == 0) ? SRP_FACE_FRONT_AND_BACK : fzI(&s, 3)));
		srpRasterFrontFace((SRPWinding) fzI(&s, 2));
		const int pm = fzI(&s, 10);
		srpRasterPolygonMode(pm < 6 ? SRP_POLYGON_MODE_FILL : (pm < 8 ? SRP_POLYGON_MODE_LINE : SRP_POLYGON_MODE_POINT));
		srpRasterPointSize((fzI(&s, 12) == 0) ? 0.0f : fzF(&s, 0.5f, 5.5f));
		srpScissorTest(fzI(&s, 2) == 0);
		srpScissorOptions((size_t) fzI(&s, (int) W), (size_t) fzI(&s, (int) H), (size_t) fzI(&s, (int) W + 8), (size_t) fzI(&s, (int) H + 8));
		if (fzI(&s, 2) == 0)
		{
			srpStencilTest(true);
			const SRPStencilFaceState f = fzFace(&s), b = fzFace(&s);
			srpStencilFuncSeparate(SRP_FACE_FRONT, f.func, f.ref, f.mask);
			srpStencilOpSeparate(SRP_FACE_FRONT, f.sfailOp, f.dfailOp, f.passOp);
			srpStencilWriteMaskSeparate(SRP_FACE_FRONT, f.writeMask);
			srpStencilFuncSeparate(SRP_FACE_BACK, b.func, b.ref, b.mask);
			srpStencilOpSeparate(SRP_FACE_BACK, b.sfailOp, b.dfailOp, b.passOp);
			srpStencilWriteMaskSeparate(SRP_FACE_BACK, b.writeMask);
		}
		srpDepthTest(fzI(&s, 4) != 0);
		srpDepthWrite(fzI(&s, 4) != 0);
		srpDepthCompareOp((SRPCompareOp) fzI(&s, 8));
		srpProvokingVertexMode((SRPProvokingVertexMode) fzI(&s, 2));

		const SRPPrimitive prim = (SRPPrimitive) fzI(&s, 7);
		const bool indexed = fzI(&s, 2) == 0;
		const int total = indexed ? nIdx : nVerts;
		const int start = fzI(&s, total);
		const int count = (fzI(&s, 20) == 0) ? total + 3 : fzI(&s, total - start + 1);
		if (indexed)
			srpDrawIndexBuffer(ib, vb, fb, &sp, prim, (size_t) start, (size_t) count);
		else
			srpDrawVertexBuffer(vb, fb, &sp, prim, (size_t) start, (size_t) count);
	}

	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	free(ibytes);
	free(v);

	int rc = 0;
	if (o->capacity < fb->size) rc = -2;
	else
	{
		memcpy(o->color, fb->color, fb->size * 4);
		memcpy(o->depth, fb->depth, fb->size * 4);
		memcpy(o->stencil, fb->stencil, fb->size);
	}
	o->width = fb->width;
	o->height = fb->height;
	srpFreeFramebuffer(fb);
	return rc;
}
