/* Scene programs for the parity tests and benchmark workloads.
 *
 * The first eighteen restate the inputs of the reference's image tests
 * (tests/scenes/<group>/<name>.c of kitrofimov/srp): same vertex data,
 * matrices, render state and draw calls, so their output can be compared
 * with the golden PNGs in tests/golden/references.  The `cfg_*` scenes are
 * the benchmark workloads and a randomised state/geometry fuzzer.
 *
 * Written against the public API only, so the same file links against the
 * CPU oracle, the compiled reference and the GPU library. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "shaders.h"
#include "obj_load.h"
#include "scenes.h"
#include "terrain.h"

SRPContext srpContext;

/* Shader symbols (host_shaders.c or device_shaders.hip). */
#define HOST_VS(name, fn, usize, part) void name(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
#define HOST_FS(name, fn, usize, part) void name(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
#include "shader_list.h"
#undef HOST_VS
#undef HOST_FS

#ifndef M_PI
#  define M_PI 3.14159265358979323846
#endif
/* reference examples/utility/rad.h:7 */
#define SC_RAD(x) (((float) (x) / 180.) * M_PI)

/* ------------------------------------------------------------ helpers */

static SRPFramebuffer* scBegin(const SceneArgs* a, size_t w, size_t h)
{
	(void) a;
	srpNewContext(&srpContext);
	SRPFramebuffer* fb = srpNewFramebuffer(w, h);
	/* The reference programs run in a fresh process where the stencil
	 * allocation comes from zeroed pages; state that explicitly. */
	memset(fb->stencil, 0, fb->size);
	return fb;
}

static int scFinish(SRPFramebuffer* fb, SceneOut* o)
{
	int rc = 0;
	if (o->capacity < fb->size)
		rc = -2;
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

static char* scPath(const SceneArgs* a, const char* file)
{
	static char buf[4096];
	snprintf(buf, sizeof(buf), "%s/%s", a->resDir ? a->resDir : ".", file);
	return buf;
}

static SRPVaryingInfo g_vColor[1] = {{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE}};

static SRPShaderProgram scProgram(void* uniform, SRPVertexShader* vs, SRPFragmentShader* fs)
{
	SRPShaderProgram p;
	p.uniform = (SRPUniform*) uniform;
	p.vs = vs;
	p.fs = fs;
	return p;
}

static SRPVertexShader scVS(void (*f)(SRPVertexShaderIn*, SRPVertexShaderOut*), size_t n, SRPVaryingInfo* info, size_t size)
{
	SRPVertexShader v;
	v.shader = f; v.nVaryings = n; v.varyingsInfo = info; v.varyingsSize = size;
	return v;
}

static SRPFragmentShader scFS(void (*f)(SRPFragmentShaderIn*, SRPFragmentShaderOut*), bool mayOverwriteDepth)
{
	SRPFragmentShader s;
	s.shader = f; s.mayOverwriteDepth = mayOverwriteDepth;
	return s;
}

static VtxPosColor pc(float x, float y, float z, float r, float g, float b)
{
	VtxPosColor v;
	v.position = VEC3(x, y, z);
	v.color = VEC3(r, g, b);
	return v;
}

/* The 24-vertex textured cube of the interpolation and stencil scenes. */
static const float kCubeUV[24][5] = {
	{-1, -1, -1, 0, 0}, { 1, -1, -1, 1, 0}, { 1,  1, -1, 1, 1}, {-1,  1, -1, 0, 1},
	{-1,  1, -1, 0, 0}, { 1,  1, -1, 1, 0}, { 1,  1,  1, 1, 1}, {-1,  1,  1, 0, 1},
	{ 1, -1,  1, 0, 0}, {-1, -1,  1, 1, 0}, {-1,  1,  1, 1, 1}, { 1,  1,  1, 0, 1},
	{ 1, -1,  1, 0, 0}, { 1, -1, -1, 1, 0}, { 1,  1, -1, 1, 1}, { 1,  1,  1, 0, 1},
	{-1, -1, -1, 0, 0}, {-1, -1,  1, 1, 0}, {-1,  1,  1, 1, 1}, {-1,  1, -1, 0, 1},
	{-1, -1, -1, 0, 0}, { 1, -1, -1, 1, 0}, { 1, -1,  1, 1, 1}, {-1, -1,  1, 0, 1},
};
static const uint8_t kCubeIdx[36] = {
	 0,  1,  2,   0,  2,  3,
	 4,  5,  6,   4,  6,  7,
	 8,  9, 10,   8, 10, 11,
	12, 15, 14,  12, 14, 13,
	16, 18, 17,  16, 19, 18,
	20, 23, 22,  20, 22, 21
};
/* The 8-corner wire cube and its 12 edges (lines, clipping/line, depth_test). */
static const float kCorners[8][3] = {
	{-1, -1, -1}, { 1, -1, -1}, { 1, -1,  1}, {-1, -1,  1},
	{-1,  1, -1}, { 1,  1, -1}, { 1,  1,  1}, {-1,  1,  1},
};
static const uint8_t kEdges[24] = {0, 1, 1, 2, 2, 3, 3, 0, 4, 5, 5, 6, 6, 7, 7, 4, 0, 4, 1, 5, 2, 6, 3, 7};

/* ------------------------------------------------- reference scenes */

/* misc/rasterization_rules.c */
static int scRasterizationRules(const SceneArgs* a, SceneOut* o)
{
	VtxPosColor data[6] = {
		pc(-0.5, -0.5, 0., 1., 0., 0.), pc(0.5, 0.5, 0., 0., 1., 0.), pc(-0.5, 0.5, 0., 0., 0., 1.),
		pc(0.5, 0.5, 0., 1., 1., 0.), pc(0.5, -0.5, 0., 1., 0., 1.), pc(-0.5, -0.5, 0., 0., 1., 1.),
	};
	UniMat u;
	u.m = mat4ConstructRotate(0., 0., 1.5);
	SRPVertexShader vs = scVS(sc_vsColorMat, 1, g_vColor, sizeof(VarColor));
	SRPFragmentShader fs = scFS(sc_fsColor, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), sizeof(data), data);
	srpFramebufferClear(fb);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 6);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}

/* misc/scissor_test.c */
static int scScissor(const SceneArgs* a, SceneOut* o)
{
	VtxPosColor data[3] = {pc(-0.5, -0.5, 0., 1., 0., 0.), pc(0.5, -0.5, 0., 0., 1., 0.), pc(0, 0.5, 0., 0., 0., 1.)};
	SRPVertexShader vs = scVS(sc_vsColorPass, 1, g_vColor, sizeof(VarColor));
	SRPFragmentShader fs = scFS(sc_fsColor, false);
	SRPShaderProgram sp = scProgram(NULL, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	srpScissorTest(true);
	srpScissorOptions(140, 220, 200, 150);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), sizeof(data), data);
	srpFramebufferClear(fb);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 3);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}

/* primitives/triangle_strip.c and primitives/triangle_fan.c */
static int scStripFan(const SceneArgs* a, SceneOut* o, bool fan)
{
	VtxPosColor strip[4] = {
		pc(-0.5, -0.5, 0., 1., 0., 0.), pc(0.5, 0, 0., 0., 1., 0.),
		pc(-0.5, 0., 0., 0., 0., 1.), pc(0.5, 0.5, 0., 1., 1., 1.),
	};
	VtxPosColor fanv[4] = {
		pc(-0.5, -0.5, 0., 1., 0., 0.), pc(-0.5, 0.5, 0., 0., 1., 0.),
		pc(0.5, 0.75, 0., 0., 0., 1.), pc(0.5, -0.25, 0., 1., 1., 1.),
	};
	SRPVertexShader vs = scVS(sc_vsColorPass, 1, g_vColor, sizeof(VarColor));
	SRPFragmentShader fs = scFS(sc_fsColor, false);
	SRPShaderProgram sp = scProgram(NULL, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), sizeof(strip), fan ? fanv : strip);
	srpFramebufferClear(fb);
	srpDrawVertexBuffer(vb, fb, &sp, fan ? SRP_PRIM_TRIANGLE_FAN : SRP_PRIM_TRIANGLE_STRIP, 0, 4);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}
static int scTriangleStrip(const SceneArgs* a, SceneOut* o) { return scStripFan(a, o, false); }
static int scTriangleFan(const SceneArgs* a, SceneOut* o) { return scStripFan(a, o, true); }

/* primitives/line_strip.c and primitives/line_loop.c */
static int scLineStripLoop(const SceneArgs* a, SceneOut* o, SRPPrimitive prim)
{
	VtxPosColor data[4] = {
		pc(-0.5, -0.5, 0., 0, 0, 0), pc(-0.5, 0.5, 0., 0, 0, 0), pc(0.5, 0.5, 0., 0, 0, 0), pc(0.5, -0.5, 0., 0, 0, 0),
	};
	SRPVertexShader vs = scVS(sc_vsPosPass, 0, NULL, 0);
	SRPFragmentShader fs = scFS(sc_fsWhite, false);
	SRPShaderProgram sp = scProgram(NULL, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), sizeof(data), data);
	srpFramebufferClear(fb);
	srpDrawVertexBuffer(vb, fb, &sp, prim, 0, 4);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}
static int scLineStrip(const SceneArgs* a, SceneOut* o) { return scLineStripLoop(a, o, SRP_PRIM_LINE_STRIP); }
static int scLineLoop(const SceneArgs* a, SceneOut* o) { return scLineStripLoop(a, o, SRP_PRIM_LINE_LOOP); }

/* Wire cube: primitives/lines.c (no varyings, white), clipping/line.c and
 * misc/depth_test.c (white colour varying, plus two triangles). */
static int scWireCube(const SceneArgs* a, SceneOut* o, int variant)
{
	VtxPosColor data[14];
	const float white = (variant == 0) ? 0.f : 1.f;
	for (int i = 0; i < 8; i++) data[i] = pc(kCorners[i][0], kCorners[i][1], kCorners[i][2], white, white, white);
	data[8] = pc(-1, -1, -1, 1, 0, 0);  data[9] = pc(-1, 1, 1, 0, 1, 0);  data[10] = pc(1, 1, -1, 0, 0, 1);
	data[11] = pc(-1, 1, -1, 1, 1, 0);  data[12] = pc(-1, -1, 1, 0, 1, 1); data[13] = pc(1, -1, -1, 1, 0, 1);
	const size_t nVerts = (variant == 2) ? 14 : 8;

	UniMVP u;
	if (variant == 1)
	{
		u.model = mat4ConstructTRS(0, 0, 0, 0, 0.4, 0, 0.5, 0.5, 0.5);
		u.view = mat4ConstructView(0, 0, -1.5, 0, 0, 0, 1, 1, 1);
	}
	else
	{
		u.model = mat4ConstructTRS(0, 0, 0, 0, 0.3, 0, 0.5, 0.5, 0.5);
		u.view = mat4ConstructView(0, 0, -2, 0, 0, 0, 1, 1, 1);
	}
	u.projection = mat4ConstructPerspectiveProjection(-1, 1, -1, 1, 1, 10);

	SRPVertexShader vs = (variant == 0) ? scVS(sc_vsPosMVP, 0, NULL, 0)
	                                    : scVS(sc_vsColorMVP, 1, g_vColor, sizeof(VarColor));
	SRPFragmentShader fs = scFS((variant == 0) ? sc_fsWhite : sc_fsColor, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	if (variant != 0) srpDepthTest(true);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), nVerts * sizeof(VtxPosColor), data);
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpIndexBufferCopyData(ib, SRP_UINT8, sizeof(kEdges), kEdges);
	srpFramebufferClear(fb);
	srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_LINES, 0, 24);
	if (variant == 2)
		srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 8, 6);
	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}
static int scLines(const SceneArgs* a, SceneOut* o) { return scWireCube(a, o, 0); }
static int scClipLine(const SceneArgs* a, SceneOut* o) { return scWireCube(a, o, 1); }
static int scDepthTest(const SceneArgs* a, SceneOut* o) { return scWireCube(a, o, 2); }

/* Textured cube: interpolation/perspective.c, interpolation/affine.c, misc/stencil_test.c */
static int scTexCube(const SceneArgs* a, SceneOut* o, int variant)
{
	VtxPosUV data[24];
	for (int i = 0; i < 24; i++)
	{
		data[i].position = VEC3(kCubeUV[i][0], kCubeUV[i][1], kCubeUV[i][2]);
		data[i].uv = VEC2(kCubeUV[i][3], kCubeUV[i][4]);
	}
	UniFrameMVPTex u;
	memset(&u, 0, sizeof(u));
	u.model = (variant == 2) ? mat4ConstructRotate(0.7, 0.7, 0.7) : mat4ConstructRotate(2.5, 0.7, 0.5);
	u.view = mat4ConstructView(0, 0, -3, 0, 0, 0, 1, 1, 1);
	u.projection = mat4ConstructPerspectiveProjection(-1, 1, -1, 1, 1, 50);
	u.texture = srpNewTexture(scPath(a, "stoneWall.png"), TW_REPEAT, TW_REPEAT);
	if (!u.texture) return -3;
	u.frameCount = 0;

	SRPVaryingInfo info[1] = {{2, SRP_FLOAT, (variant == 1) ? SRP_INTERPOLATION_MODE_AFFINE : SRP_INTERPOLATION_MODE_PERSPECTIVE}};
	SRPVertexShader vs = scVS(sc_vsUVMVP, 1, info, sizeof(VarUV));
	SRPFragmentShader fs = scFS(sc_fsTexture, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	srpRasterFrontFace(SRP_WINDING_CCW);
	srpRasterCullFace(SRP_FACE_BACK);
	srpDepthTest(true);
	if (variant == 2) srpStencilTest(true);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosUV), sizeof(data), data);
	srpIndexBufferCopyData(ib, SRP_UINT8, sizeof(kCubeIdx), kCubeIdx);
	srpFramebufferClear(fb);
	if (variant == 2)
	{
		srpStencilOp(SRP_STENCIL_KEEP, SRP_STENCIL_KEEP, SRP_STENCIL_REPLACE);
		srpStencilFunc(SRP_COMPARE_ALWAYS, 1, 0xFF);
		srpStencilWriteMask(0xFF);
	}
	srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 36);
	if (variant == 2)
	{
		/* Outline pass: the cube scaled by 1.05 where the stencil is not 1. */
		const mat4 s = mat4ConstructScale(1.05, 1.05, 1.05);
		u.model = mat4MultiplyMat4(&s, &u.model);
		SRPFragmentShader red = scFS(sc_fsRed, false);
		sp.fs = &red;
		srpStencilFunc(SRP_COMPARE_NOTEQUAL, 1, 0xFF);
		srpStencilWriteMask(0x00);
		srpDepthTest(false);
		srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 36);
	}
	srpFreeTexture(u.texture);
	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}
static int scPerspective(const SceneArgs* a, SceneOut* o) { return scTexCube(a, o, 0); }
static int scAffine(const SceneArgs* a, SceneOut* o) { return scTexCube(a, o, 1); }
static int scStencil(const SceneArgs* a, SceneOut* o) { return scTexCube(a, o, 2); }

/* interpolation/flat.c: a one-byte varying described (as in the reference)
 * as three flat floats; only its first byte is read back. */
static int scFlat(const SceneArgs* a, SceneOut* o)
{
	VtxPosByte data[3];
	data[0].position = VEC3(-0.5, -0.5, 0.); data[0].color = 0;
	data[1].position = VEC3(0.5, -0.5, 0.);  data[1].color = 1;
	data[2].position = VEC3(0., 0.5, 0.);    data[2].color = 2;
	UniMat u;
	memset(&u, 0, sizeof(u));
	SRPVaryingInfo info[1] = {{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_FLAT}};
	SRPVertexShader vs = scVS(sc_vsByteMat, 1, info, sizeof(VarByte));
	SRPFragmentShader fs = scFS(sc_fsByte, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	srpFramebufferClear(fb);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosByte), sizeof(data), data);
	srpProvokingVertexMode(SRP_PROVOKING_VERTEX_FIRST);
	u.m = mat4ConstructTRS(-0.5, 0, 0, 0, 0, 0, 0.5, 0.5, 0.5);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 3);
	srpProvokingVertexMode(SRP_PROVOKING_VERTEX_LAST);
	u.m = mat4ConstructTRS(0.5, 0, 0, 0, 0, 0, 0.5, 0.5, 0.5);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 3);
	srpFreeVertexBuffer(vb);
	return scFinish(fb, o);
}

/* Teapot scenes: primitives/points.c, polygon_modes/{line,point}.c,
 * clipping/{triangle,point}.c */
enum { TP_POINTS, TP_PMLINE, TP_PMPOINT, TP_CLIPTRI, TP_CLIPPOINT };

static int scTeapot(const SceneArgs* a, SceneOut* o, int variant)
{
	SRPObjMesh mesh;
	if (!srpObjLoad(scPath(a, "utah_teapot.obj"), &mesh))
		return -3;
	UniFrameMVP u;
	memset(&u, 0, sizeof(u));
	float camZ = -6.f, rotY = 0.3f;
	if (variant == TP_CLIPTRI) { camZ = -4.f; rotY = 0.2f; }
	if (variant == TP_CLIPPOINT) camZ = -2.f;
	u.model = mat4ConstructRotate(SC_RAD(-90), rotY, 0);
	u.view = mat4ConstructView(0, 1.75, camZ, 0, 0, 0, 1, 1, 1);
	u.projection = mat4ConstructPerspectiveProjection(-1, 1, -1, 1, 1, 10);
	u.frameCount = 0;

	SRPVertexShader vs = scVS(sc_vsObjMVP, 0, NULL, 0);
	SRPFragmentShader fs = scFS((variant == TP_CLIPPOINT) ? sc_fsPrimIDHalf : sc_fsPrimID, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	switch (variant)
	{
		case TP_POINTS: srpRasterPointSize(2.); break;
		case TP_PMLINE: srpRasterPolygonMode(SRP_POLYGON_MODE_LINE); break;
		case TP_PMPOINT: srpRasterPolygonMode(SRP_POLYGON_MODE_POINT); srpRasterPointSize(2.); break;
		case TP_CLIPTRI:
			srpRasterFrontFace(SRP_WINDING_CW);
			srpRasterCullFace(SRP_FACE_BACK);
			srpRasterPolygonMode(SRP_POLYGON_MODE_FILL);
			break;
		case TP_CLIPPOINT: srpRasterPolygonMode(SRP_POLYGON_MODE_LINE); srpRasterPointSize(3.); break;
	}
	srpDepthTest(true);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpVertexBufferCopyData(vb, sizeof(SRPObjVertex), mesh.vertexCount * sizeof(SRPObjVertex), mesh.vertices);
	srpIndexBufferCopyData(ib, SRP_UINT32, mesh.indexCount * sizeof(uint32_t), mesh.indices);
	srpFramebufferClear(fb);
	srpDrawIndexBuffer(ib, vb, fb, &sp, (variant == TP_POINTS) ? SRP_PRIM_POINTS : SRP_PRIM_TRIANGLES, 0, mesh.indexCount);
	if (variant == TP_CLIPPOINT)
	{
		SRPFragmentShader white = scFS(sc_fsWhite, false);
		SRPShaderProgram sp2 = sp;
		sp2.fs = &white;
		srpDrawIndexBuffer(ib, vb, fb, &sp2, SRP_PRIM_POINTS, 0, mesh.indexCount);
	}
	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	srpObjFree(&mesh);
	return scFinish(fb, o);
}
static int scPoints(const SceneArgs* a, SceneOut* o) { return scTeapot(a, o, TP_POINTS); }
static int scPMLine(const SceneArgs* a, SceneOut* o) { return scTeapot(a, o, TP_PMLINE); }
static int scPMPoint(const SceneArgs* a, SceneOut* o) { return scTeapot(a, o, TP_PMPOINT); }
static int scClipTri(const SceneArgs* a, SceneOut* o) { return scTeapot(a, o, TP_CLIPTRI); }
static int scClipPoint(const SceneArgs* a, SceneOut* o) { return scTeapot(a, o, TP_CLIPPOINT); }

/* ------------------------------------------------------- workloads */

/* splitmix64 */
static uint64_t rngNext(uint64_t* s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static double rngU(uint64_t* s) { return (double) (rngNext(s) >> 11) * (1.0 / 9007199254740992.0); }
static float rngF(uint64_t* s, float lo, float hi) { return (float) (lo + (hi - lo) * rngU(s)); }
static double rngNormal(uint64_t* s)
{
	const double u1 = rngU(s) + 1e-300, u2 = rngU(s);
	return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static void scDepthInit(SRPFramebuffer* fb, float depth)
{
	for (size_t i = 0; i < fb->size; i++) fb->depth[i] = depth;
	memset(fb->stencil, 0, fb->size);
}

/* Random coloured triangles in clip space, xyz ~ U[-1.2, 1.2], w = 1;
 * depth LESS against a +1 depth buffer.  p0 = triangle count. */
/* The random clip-space triangles of the plain-method workload. */
void srpSceneBuildTris(uint64_t seed, size_t nTri, VtxPosColor* v)
{
	uint64_t s = seed;
	for (size_t i = 0; i < 3 * nTri; i++)
	{
		const float x = rngF(&s, -1.2f, 1.2f), y = rngF(&s, -1.2f, 1.2f), z = rngF(&s, -1.2f, 1.2f);
		const float r = rngF(&s, 0, 1), g = rngF(&s, 0, 1), b = rngF(&s, 0, 1);
		v[i] = pc(x, y, z, r, g, b);
	}
}

static int scCfgTris(const SceneArgs* a, SceneOut* o)
{
	const size_t nTri = a->p0 ? (size_t) a->p0 : 2000;
	VtxPosColor* v = (VtxPosColor*) malloc(3 * nTri * sizeof(VtxPosColor));
	srpSceneBuildTris(a->seed, nTri, v);
	SRPVertexShader vs = scVS(sc_vsColorPass, 1, g_vColor, sizeof(VarColor));
	SRPFragmentShader fs = scFS(sc_fsColor, false);
	SRPShaderProgram sp = scProgram(NULL, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, a->width ? (size_t) a->width : 512, a->height ? (size_t) a->height : 512);
	srpDepthTest(true);
	srpDepthCompareOp(SRP_COMPARE_LESS);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), 3 * nTri * sizeof(VtxPosColor), v);
	srpFramebufferClear(fb);
	scDepthInit(fb, 1.0f);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 3 * nTri);
	srpFreeVertexBuffer(vb);
	free(v);
	return scFinish(fb, o);
}

/* One draw issued again and again with counts that change between calls:
 * p0 random clip-space triangles (seed) through vsColorMat, scaled about the
 * origin by each of 0.05, 1, 0.3, 1, 0.5 in turn, filled, then as lines,
 * then as points, all into one 512x512 framebuffer with depth LESS.  The
 * GPU path sizes a draw's buffers from the same draw's previous counts and
 * repeats its emitting half when they were too small; every growth here
 * takes that path.  p1 = 1: in the GPU build the framebuffer stays on the
 * device between the draws (deferred), so each draw is checked only at the
 * next call. */
static int scSpecGrow(const SceneArgs* a, SceneOut* o)
{
	const size_t nTri = a->p0 ? (size_t) a->p0 : 300;
	VtxPosColor* v = (VtxPosColor*) malloc(3 * nTri * sizeof(VtxPosColor));
	srpSceneBuildTris(a->seed, nTri, v);
	UniMat u;
	SRPVertexShader vs = scVS(sc_vsColorMat, 1, g_vColor, sizeof(VarColor));
	SRPFragmentShader fs = scFS(sc_fsColor, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, 512, 512);
	srpDepthTest(true);
	srpDepthCompareOp(SRP_COMPARE_LESS);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), 3 * nTri * sizeof(VtxPosColor), v);
	srpFramebufferClear(fb);
	scDepthInit(fb, 1.0f);
#ifdef SRP_GPU_BUILD
	if (a->p1)
		srpFramebufferSetDeferred(fb, true);
#endif
	static const float scales[5] = {0.05f, 1.0f, 0.3f, 1.0f, 0.5f};
	static const SRPPolygonMode modes[3] = {SRP_POLYGON_MODE_FILL, SRP_POLYGON_MODE_LINE, SRP_POLYGON_MODE_POINT};
	for (int m = 0; m < 3; m++)
	{
		srpRasterPolygonMode(modes[m]);
		for (int k = 0; k < 5; k++)
		{
			u.m = mat4ConstructScale(scales[k], scales[k], scales[k]);
			srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 3 * nTri);
		}
	}
#ifdef SRP_GPU_BUILD
	if (a->p1)
		srpFramebufferSetDeferred(fb, false);
#endif
	srpFreeVertexBuffer(vb);
	free(v);
	return scFinish(fb, o);
}

/* UV sphere of p0 rings x p1 segments: (p0+1)(p1+1) vertices {pos, uv},
 * 2*p0*p1 u32-indexed triangles (degenerate at the poles). */
void srpSceneBuildSphere(int rings, int segs, VtxPosUV** verts, size_t* nVerts, uint32_t** idx, size_t* nIdx)
{
	const size_t nv = (size_t) (rings + 1) * (size_t) (segs + 1);
	VtxPosUV* v = (VtxPosUV*) malloc(nv * sizeof(VtxPosUV));
	for (int i = 0; i <= rings; i++)
		for (int j = 0; j <= segs; j++)
		{
			const double th = M_PI * i / rings, ph = 2.0 * M_PI * j / segs;
			VtxPosUV* p = &v[(size_t) i * (segs + 1) + j];
			p->position = VEC3((float) (sin(th) * cos(ph)), (float) cos(th), (float) (sin(th) * sin(ph)));
			p->uv = VEC2((float) j / (float) segs, 1.0f - (float) i / (float) rings);
		}
	const size_t ni = (size_t) rings * (size_t) segs * 6;
	uint32_t* ix = (uint32_t*) malloc(ni * sizeof(uint32_t));
	size_t k = 0;
	for (int i = 0; i < rings; i++)
		for (int j = 0; j < segs; j++)
		{
			const uint32_t q = (uint32_t) (i * (segs + 1) + j), r = q + (uint32_t) (segs + 1);
			ix[k++] = q; ix[k++] = r; ix[k++] = q + 1;
			ix[k++] = r; ix[k++] = r + 1; ix[k++] = q + 1;
		}
	*verts = v; *nVerts = nv; *idx = ix; *nIdx = ni;
}

/* Camera and state of the indexed textured-mesh workload. */
void srpSceneMeshUniform(int width, int height, float spin, UniFrameMVPTex* u)
{
	const float t = (float) tan(M_PI / 6.0), n = 0.1f, f = 100.f;
	const float aspect = (float) width / (float) height;
	u->frameCount = 0;
	u->model = mat4ConstructTRS(0, 0, 2.6f, 0.35f, spin, 0.1f, 1, 1, 1);
	u->view = mat4ConstructView(0, 0, 0, 0, 0, 0, 1, 1, 1);
	u->projection = mat4ConstructPerspectiveProjection(-t * n * aspect, t * n * aspect, -t * n, t * n, n, f);
}

void srpSceneMeshState(int width, int height)
{
	srpScissorTest(true);
	srpScissorOptions((size_t) width / 8, (size_t) height / 8, (size_t) width * 3 / 4, (size_t) height * 3 / 4);
	srpStencilTest(true);
	srpStencilFunc(SRP_COMPARE_ALWAYS, 0, 0xFF);
	srpStencilOp(SRP_STENCIL_KEEP, SRP_STENCIL_KEEP, SRP_STENCIL_INCR);
	srpDepthTest(true);
	srpDepthCompareOp(SRP_COMPARE_LESS);
}

SRPVaryingInfo g_meshVaryings[2] = {
	{2, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE},
	{1, SRP_UINT32, SRP_INTERPOLATION_MODE_FLAT},
};

/* Indexed textured sphere: p0 rings, p1 segments, texPath = noise PNG. */
static int scCfgMesh(const SceneArgs* a, SceneOut* o)
{
	const int w = a->width ? a->width : 1920, h = a->height ? a->height : 1080;
	VtxPosUV* verts; uint32_t* idx; size_t nv, ni;
	srpSceneBuildSphere(a->p0 ? a->p0 : 256, a->p1 ? a->p1 : 512, &verts, &nv, &idx, &ni);
	UniFrameMVPTex u;
	memset(&u, 0, sizeof(u));
	srpSceneMeshUniform(w, h, a->f0, &u);
	u.texture = srpNewTexture(a->texPath, TW_REPEAT, TW_REPEAT);
	if (!u.texture) { free(verts); free(idx); return -3; }
	SRPVertexShader vs = scVS(sc_vsMeshTex, 2, g_meshVaryings, sizeof(VarUVTag));
	SRPFragmentShader fs = scFS(sc_fsMeshTex, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, (size_t) w, (size_t) h);
	srpSceneMeshState(w, h);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosUV), nv * sizeof(VtxPosUV), verts);
	srpIndexBufferCopyData(ib, SRP_UINT32, ni * sizeof(uint32_t), idx);
	srpFramebufferClear(fb);
	scDepthInit(fb, 1.0f);
	srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, ni);
	srpFreeTexture(u.texture);
	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	free(verts); free(idx);
	return scFinish(fb, o);
}

/* Primitive mix geometry: p0 triangles, p1 lines, p2 points. */
void srpSceneBuildMix(uint64_t seed, size_t nTri, size_t nLine, size_t nPoint,
                      VtxMix** tris, VtxMix** lines, VtxMix** points)
{
	uint64_t s = seed;
	VtxMix* t = (VtxMix*) malloc((3 * nTri + 1) * sizeof(VtxMix));
	for (size_t i = 0; i < nTri; i++)
	{
		const float cx = rngF(&s, -1, 1), cy = rngF(&s, -1, 1);
		const double len = exp(-3.0 + rngNormal(&s));
		const double a0 = 2.0 * M_PI * rngU(&s);
		const double kind = rngU(&s);
		const float z0 = rngF(&s, -0.9f, 0.9f);
		for (int k = 0; k < 3; k++)
		{
			VtxMix* v = &t[3 * i + k];
			const double ang = a0 + k * (2.0 * M_PI / 3.0) + 0.3 * rngU(&s);
			float x = (float) (cx + len * cos(ang)), y = (float) (cy + len * sin(ang));
			float z = z0 + rngF(&s, -0.05f, 0.05f);
			const float w = rngF(&s, 0.6f, 1.6f);
			if (kind < 0.3 && k < 1 + (int) (kind * 10.0) % 2)
				z = rngF(&s, -3.0f, -1.2f);       /* behind the near plane */
			if (kind >= 0.3 && kind < 0.4)
				x += 3.5f;                        /* fully outside (right) */
			v->clip = VEC4(x * w, y * w, z * w, w);
			v->st = VEC2(rngF(&s, 0, 1), rngF(&s, 0, 1));
			v->level = rngF(&s, 0, 1);
			v->tag = (uint32_t) rngNext(&s);
		}
	}
	VtxMix* l = (VtxMix*) malloc((2 * nLine + 1) * sizeof(VtxMix));
	for (size_t i = 0; i < nLine; i++)
	{
		const float cx = rngF(&s, -1.1f, 1.1f), cy = rngF(&s, -1.1f, 1.1f);
		const double len = exp(-2.5 + 0.7 * rngNormal(&s));
		const double ang = 2.0 * M_PI * rngU(&s);
		for (int k = 0; k < 2; k++)
		{
			VtxMix* v = &l[2 * i + k];
			const double sgn = k ? 0.5 : -0.5;
			const float w = rngF(&s, 0.8f, 1.2f);
			const float x = (float) (cx + sgn * len * cos(ang)), y = (float) (cy + sgn * len * sin(ang));
			v->clip = VEC4(x * w, y * w, rngF(&s, -1.1f, 1.1f) * w, w);
			v->st = VEC2(rngF(&s, 0, 1), rngF(&s, 0, 1));
			v->level = rngF(&s, 0, 1);
			v->tag = (uint32_t) rngNext(&s);
		}
	}
	VtxMix* p = (VtxMix*) malloc((nPoint + 1) * sizeof(VtxMix));
	for (size_t i = 0; i < nPoint; i++)
	{
		VtxMix* v = &p[i];
		const float w = rngF(&s, 0.8f, 1.2f);
		v->clip = VEC4(rngF(&s, -1.05f, 1.05f) * w, rngF(&s, -1.05f, 1.05f) * w, rngF(&s, -1.0f, 1.0f) * w, w);
		v->st = VEC2(rngF(&s, 0, 1), rngF(&s, 0, 1));
		v->level = rngF(&s, 0, 1);
		v->tag = (uint32_t) rngNext(&s);
	}
	*tris = t; *lines = l; *points = p;
}

SRPVaryingInfo g_mixVaryings[3] = {
	{2, SRP_FLOAT, SRP_INTERPOLATION_MODE_AFFINE},
	{1, SRP_FLOAT, SRP_INTERPOLATION_MODE_FLAT},
	{1, SRP_UINT32, SRP_INTERPOLATION_MODE_FLAT},
};

void srpSceneMixState(void)
{
	srpDepthTest(true);
	srpDepthCompareOp(SRP_COMPARE_LEQUAL);
	srpRasterCullFace(SRP_FACE_BACK);
}

/* Draw sequence of the primitive-mix workload: triangles, lines, then
 * points in four draws of sizes 1..4. */
void srpSceneMixDraw(const SRPFramebuffer* fb, const SRPShaderProgram* sp, const SRPVertexBuffer* vbT,
                     const SRPVertexBuffer* vbL, const SRPVertexBuffer* vbP, size_t nTri, size_t nLine, size_t nPoint)
{
	if (nTri) srpDrawVertexBuffer(vbT, fb, sp, SRP_PRIM_TRIANGLES, 0, 3 * nTri);
	if (nLine) srpDrawVertexBuffer(vbL, fb, sp, SRP_PRIM_LINES, 0, 2 * nLine);
	const size_t q = nPoint / 4;
	for (int k = 0; k < 4 && q; k++)
	{
		srpRasterPointSize((float) (k + 1));
		srpDrawVertexBuffer(vbP, fb, sp, SRP_PRIM_POINTS, (size_t) k * q, q);
	}
}

static int scCfgMix(const SceneArgs* a, SceneOut* o)
{
	const size_t nT = a->p0 >= 0 ? (size_t) a->p0 : 500000;
	const size_t nL = a->p1 >= 0 ? (size_t) a->p1 : 100000;
	const size_t nP = a->p2 >= 0 ? (size_t) a->p2 : 100000;
	VtxMix *t, *l, *p;
	srpSceneBuildMix(a->seed, nT, nL, nP, &t, &l, &p);
	SRPVertexShader vs = scVS(sc_vsMix, 3, g_mixVaryings, sizeof(VarMix));
	SRPFragmentShader fs = scFS(sc_fsMix, false);
	SRPShaderProgram sp = scProgram(NULL, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, a->width ? (size_t) a->width : 1920, a->height ? (size_t) a->height : 1080);
	srpSceneMixState();
	SRPVertexBuffer* vbT = srpNewVertexBuffer();
	SRPVertexBuffer* vbL = srpNewVertexBuffer();
	SRPVertexBuffer* vbP = srpNewVertexBuffer();
	srpVertexBufferCopyData(vbT, sizeof(VtxMix), (3 * nT + 1) * sizeof(VtxMix), t);
	srpVertexBufferCopyData(vbL, sizeof(VtxMix), (2 * nL + 1) * sizeof(VtxMix), l);
	srpVertexBufferCopyData(vbP, sizeof(VtxMix), (nP + 1) * sizeof(VtxMix), p);
	srpFramebufferClear(fb);
	scDepthInit(fb, 1.0f);
	srpSceneMixDraw(fb, &sp, vbT, vbL, vbP, nT, nL, nP);
	srpFreeVertexBuffer(vbT); srpFreeVertexBuffer(vbL); srpFreeVertexBuffer(vbP);
	free(t); free(l); free(p);
	return scFinish(fb, o);
}

/* Terrain heightfield of p0 x p1 cells (default 1000 x 500) from pose p2;
 * seed selects the fragment shader: 0 the procedural material, 1+k the
 * G-buffer capture pass k. */
static int scCfgTerrain(const SceneArgs* a, SceneOut* o)
{
	const int w = a->width ? a->width : 1920, h = a->height ? a->height : 1080;
	VtxTerrain* verts; uint32_t* idx; size_t nv, ni;
	UniTerrain u;
	memset(&u, 0, sizeof(u));
	srpTerrainBuild(a->p0 > 0 ? a->p0 : 1000, a->p1 > 0 ? a->p1 : 500, &verts, &nv, &idx, &ni, &u.bounds);
	const TerrainBounds b = u.bounds;
	srpTerrainPose(a->p2 > 0 ? a->p2 : 0, w, h, &u);
	u.bounds = b;
	if (a->seed > 4) { free(verts); free(idx); return -1; }
	u.pass = a->seed ? (uint32_t) a->seed - 1u : 0u;
	SRPVertexShader vs = scVS(sc_vsTerrain, 3, g_terrainVaryings, sizeof(VarTerrain));
	SRPFragmentShader fs = scFS(a->seed ? sc_fsTerrainCapture : sc_fsTerrain, false);
	SRPShaderProgram sp = scProgram(&u, &vs, &fs);
	SRPFramebuffer* fb = scBegin(a, (size_t) w, (size_t) h);
	srpTerrainState();
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxTerrain), nv * sizeof(VtxTerrain), verts);
	srpIndexBufferCopyData(ib, SRP_UINT32, ni * sizeof(uint32_t), idx);
	srpFramebufferClear(fb);
	scDepthInit(fb, 1.0f);
	srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, ni);
	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	free(verts); free(idx);
	return scFinish(fb, o);
}

/* ----------------------------------------------------------- registry */

typedef struct SceneEntry { const char* name; int (*fn)(const SceneArgs*, SceneOut*); } SceneEntry;

static const SceneEntry kScenes[] = {
	{"primitives/points", scPoints},
	{"primitives/lines", scLines},
	{"primitives/line_strip", scLineStrip},
	{"primitives/line_loop", scLineLoop},
	{"primitives/triangle_strip", scTriangleStrip},
	{"primitives/triangle_fan", scTriangleFan},
	{"polygon_modes/line", scPMLine},
	{"polygon_modes/point", scPMPoint},
	{"clipping/triangle", scClipTri},
	{"clipping/line", scClipLine},
	{"clipping/point", scClipPoint},
	{"interpolation/perspective", scPerspective},
	{"interpolation/affine", scAffine},
	{"interpolation/flat", scFlat},
	{"misc/depth_test", scDepthTest},
	{"misc/stencil_test", scStencil},
	{"misc/scissor_test", scScissor},
	{"misc/rasterization_rules", scRasterizationRules},
	{"cfg/tris", scCfgTris},
	{"spec/grow", scSpecGrow},
	{"cfg/mesh", scCfgMesh},
	{"cfg/mix", scCfgMix},
	{"cfg/fuzz", srpSceneFuzz},
	{"cfg/terrain", scCfgTerrain},
};

int srpSceneCount(void) { return (int) (sizeof(kScenes) / sizeof(kScenes[0])); }
const char* srpSceneName(int i) { return (i >= 0 && i < srpSceneCount()) ? kScenes[i].name : NULL; }

int srpSceneRun(const char* name, const SceneArgs* args, SceneOut* out)
{
	for (int i = 0; i < srpSceneCount(); i++)
		if (strcmp(kScenes[i].name, name) == 0)
			return kScenes[i].fn(args, out);
	return -1;
}
