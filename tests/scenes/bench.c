/* Frame drivers of the benchmark workloads, compiled into the GPU scene
 * module (SRP_GPU_BUILD: device-resident framebuffer, device fills) and
 * into the oracle module (CPU baseline: host arrays).  Setup (geometry,
 * texture upload) happens once; a frame is clear + depth/stencil init +
 * the workload's draws, ending with the buffers final in the renderer's
 * memory. */
#include <stdlib.h>
#include <string.h>

#include "shaders.h"
#include "scenes.h"
#include "terrain.h"
#ifdef SRP_GPU_BUILD
#include <srp/srp_hip.h>
#include <srp/srp_neural.h>
#endif

void sc_vsMeshTex(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsMeshTex(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_vsMix(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsMix(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_vsColorPass(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_vsTerrain(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsTerrain(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_fsTerrainCapture(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_fsColor(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);

void srpSceneBuildSphere(int rings, int segs, VtxPosUV** verts, size_t* nVerts, uint32_t** idx, size_t* nIdx);
void srpSceneMeshUniform(int width, int height, float spin, UniFrameMVPTex* u);
void srpSceneMeshState(int width, int height);
void srpSceneBuildMix(uint64_t seed, size_t nTri, size_t nLine, size_t nPoint, VtxMix** tris, VtxMix** lines, VtxMix** points);
void srpSceneMixState(void);
void srpSceneMixDraw(const SRPFramebuffer* fb, const SRPShaderProgram* sp, const SRPVertexBuffer* vbT,
                     const SRPVertexBuffer* vbL, const SRPVertexBuffer* vbP, size_t nTri, size_t nLine, size_t nPoint);
extern SRPVaryingInfo g_meshVaryings[2];
extern SRPVaryingInfo g_mixVaryings[3];
void srpSceneBuildTris(uint64_t seed, size_t nTri, VtxPosColor* v);
static SRPVaryingInfo g_trisVaryings[1] = {{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE}};

typedef struct Bench
{
	int kind;  /* 0 mesh, 1 mix, 2 terrain, 3 random triangles */
	SRPFramebuffer* fb;       /* current target: fbs[0] or fbs[1] */
	SRPFramebuffer* fbs[2];
	SRPVertexBuffer *vb, *vbL, *vbP;
	SRPIndexBuffer* ib;
	size_t nIdx, nTri, nLine, nPoint;
	UniFrameMVPTex u;
	UniTerrain ut;
	SRPVertexShader vs;
	SRPFragmentShader fs;
	SRPShaderProgram sp;
	int width, height;
} Bench;

static Bench g_b;

void srpBenchFrame(void);

static void benchFbInit(int w, int h)
{
	g_b.width = w;
	g_b.height = h;
	g_b.fb = srpNewFramebuffer((size_t) w, (size_t) h);
	g_b.fbs[0] = g_b.fb;
	memset(g_b.fb->stencil, 0, g_b.fb->size);
#ifdef SRP_GPU_BUILD
	srpFramebufferSetDeferred(g_b.fb, true);
#endif
}

/* Indexed textured UV sphere (rings x segs) at w x h with the noise texture at texPath. */
int srpBenchMeshSetup(int w, int h, int rings, int segs, const char* texPath)
{
	memset(&g_b, 0, sizeof(g_b));
	g_b.kind = 0;
	srpNewContext(&srpContext);
	VtxPosUV* verts; uint32_t* idx; size_t nv, ni;
	srpSceneBuildSphere(rings, segs, &verts, &nv, &idx, &ni);
	g_b.vb = srpNewVertexBuffer();
	g_b.ib = srpNewIndexBuffer();
	srpVertexBufferCopyData(g_b.vb, sizeof(VtxPosUV), nv * sizeof(VtxPosUV), verts);
	srpIndexBufferCopyData(g_b.ib, SRP_UINT32, ni * sizeof(uint32_t), idx);
	free(verts); free(idx);
	g_b.nIdx = ni;
	srpSceneMeshUniform(w, h, 0.0f, &g_b.u);
	g_b.u.texture = srpNewTexture(texPath, TW_REPEAT, TW_REPEAT);
	if (!g_b.u.texture) return -3;
	g_b.vs.shader = sc_vsMeshTex; g_b.vs.nVaryings = 2; g_b.vs.varyingsInfo = g_meshVaryings; g_b.vs.varyingsSize = sizeof(VarUVTag);
	g_b.fs.shader = sc_fsMeshTex; g_b.fs.mayOverwriteDepth = false;
	g_b.sp.uniform = (SRPUniform*) &g_b.u; g_b.sp.vs = &g_b.vs; g_b.sp.fs = &g_b.fs;
	benchFbInit(w, h);
	srpSceneMeshState(w, h);
	return 0;
}

/* Clip-space triangles, lines and points (the primitive-mix workload). */
int srpBenchMixSetup(int w, int h, int nTri, int nLine, int nPoint, uint64_t seed)
{
	memset(&g_b, 0, sizeof(g_b));
	g_b.kind = 1;
	srpNewContext(&srpContext);
	VtxMix *t, *l, *p;
	srpSceneBuildMix(seed, (size_t) nTri, (size_t) nLine, (size_t) nPoint, &t, &l, &p);
	g_b.vb = srpNewVertexBuffer(); g_b.vbL = srpNewVertexBuffer(); g_b.vbP = srpNewVertexBuffer();
	srpVertexBufferCopyData(g_b.vb, sizeof(VtxMix), (3 * (size_t) nTri + 1) * sizeof(VtxMix), t);
	srpVertexBufferCopyData(g_b.vbL, sizeof(VtxMix), (2 * (size_t) nLine + 1) * sizeof(VtxMix), l);
	srpVertexBufferCopyData(g_b.vbP, sizeof(VtxMix), ((size_t) nPoint + 1) * sizeof(VtxMix), p);
	free(t); free(l); free(p);
	g_b.nTri = (size_t) nTri; g_b.nLine = (size_t) nLine; g_b.nPoint = (size_t) nPoint;
	g_b.vs.shader = sc_vsMix; g_b.vs.nVaryings = 3; g_b.vs.varyingsInfo = g_mixVaryings; g_b.vs.varyingsSize = sizeof(VarMix);
	g_b.fs.shader = sc_fsMix; g_b.fs.mayOverwriteDepth = false;
	g_b.sp.uniform = NULL; g_b.sp.vs = &g_b.vs; g_b.sp.fs = &g_b.fs;
	benchFbInit(w, h);
	srpSceneMixState();
	return 0;
}

/* Random clip-space triangles (the plain-method workload): nTri
 * non-indexed triangles, xyz uniform in [-1.2, 1.2]^3 with w = 1, RGB per
 * vertex, passthrough VS, perspective colour, depth LESS, no culling. */
int srpBenchTrisSetup(int w, int h, int nTri, uint64_t seed)
{
	memset(&g_b, 0, sizeof(g_b));
	g_b.kind = 3;
	srpNewContext(&srpContext);
	VtxPosColor* v = (VtxPosColor*) malloc(3 * (size_t) nTri * sizeof(VtxPosColor));
	srpSceneBuildTris(seed, (size_t) nTri, v);
	g_b.vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(g_b.vb, sizeof(VtxPosColor), 3 * (size_t) nTri * sizeof(VtxPosColor), v);
	free(v);
	g_b.nTri = (size_t) nTri;
	g_b.vs.shader = sc_vsColorPass; g_b.vs.nVaryings = 1; g_b.vs.varyingsInfo = g_trisVaryings; g_b.vs.varyingsSize = sizeof(VarColor);
	g_b.fs.shader = sc_fsColor; g_b.fs.mayOverwriteDepth = false;
	g_b.sp.uniform = NULL; g_b.sp.vs = &g_b.vs; g_b.sp.fs = &g_b.fs;
	benchFbInit(w, h);
	srpDepthTest(true);
	srpDepthCompareOp(SRP_COMPARE_LESS);
	return 0;
}

/* Terrain heightfield of nx x nz cells seen from `pose`, with the
 * procedural material (capture = 0) or G-buffer capture pass capture-1. */
int srpBenchTerrainSetup(int w, int h, int nx, int nz, int pose, int capture)
{
	memset(&g_b, 0, sizeof(g_b));
	g_b.kind = 2;
	if (capture < 0 || capture > 4) return -1;
	srpNewContext(&srpContext);
	VtxTerrain* verts; uint32_t* idx; size_t nv, ni;
	srpTerrainBuild(nx, nz, &verts, &nv, &idx, &ni, &g_b.ut.bounds);
	g_b.vb = srpNewVertexBuffer();
	g_b.ib = srpNewIndexBuffer();
	srpVertexBufferCopyData(g_b.vb, sizeof(VtxTerrain), nv * sizeof(VtxTerrain), verts);
	srpIndexBufferCopyData(g_b.ib, SRP_UINT32, ni * sizeof(uint32_t), idx);
	free(verts); free(idx);
	g_b.nIdx = ni;
	g_b.width = w; g_b.height = h;
	const TerrainBounds b = g_b.ut.bounds;
	srpTerrainPose(pose, w, h, &g_b.ut);
	g_b.ut.bounds = b;
	g_b.ut.pass = capture ? (uint32_t) capture - 1u : 0u;
	g_b.vs.shader = sc_vsTerrain; g_b.vs.nVaryings = 3; g_b.vs.varyingsInfo = g_terrainVaryings; g_b.vs.varyingsSize = sizeof(VarTerrain);
	g_b.fs.shader = capture ? sc_fsTerrainCapture : sc_fsTerrain; g_b.fs.mayOverwriteDepth = false;
	g_b.sp.uniform = (SRPUniform*) &g_b.ut; g_b.sp.vs = &g_b.vs; g_b.sp.fs = &g_b.fs;
	benchFbInit(w, h);
	srpTerrainState();
	return 0;
}

/* Switch the terrain workload's camera pose and fragment shader:
 * fs = 0 procedural, 1+k capture pass k, or an explicit shader function
 * with its uniform network pointer (neural shading). */
void srpBenchTerrainPose(int pose)
{
	const TerrainBounds b = g_b.ut.bounds;
	const uint32_t pass = g_b.ut.pass;
	const void* net = g_b.ut.network;
	srpTerrainPose(pose, g_b.width, g_b.height, &g_b.ut);
	g_b.ut.bounds = b;
	g_b.ut.pass = pass;
	g_b.ut.network = net;
}

void srpBenchTerrainShader(int capture, void (*fs)(SRPFragmentShaderIn*, SRPFragmentShaderOut*), const void* network)
{
	if (fs)
		g_b.fs.shader = fs;
	else
		g_b.fs.shader = capture ? sc_fsTerrainCapture : sc_fsTerrain;
	g_b.ut.pass = capture ? (uint32_t) capture - 1u : 0u;
	g_b.ut.network = network;
}

const SRPFramebuffer* srpBenchFramebuffer(void) { return g_b.fb; }

/* Select the framebuffer later frames render into: 0 the workload's, 1 a
 * second one of the same size (created on first use). */
void srpBenchUseFramebuffer(int which)
{
	if (which == 1 && !g_b.fbs[1])
	{
		g_b.fbs[1] = srpNewFramebuffer((size_t) g_b.width, (size_t) g_b.height);
		memset(g_b.fbs[1]->stencil, 0, g_b.fbs[1]->size);
#ifdef SRP_GPU_BUILD
		srpFramebufferSetDeferred(g_b.fbs[1], true);
#endif
	}
	g_b.fb = g_b.fbs[which ? 1 : 0];
}

#ifdef SRP_GPU_BUILD
/* Training data of the neural material: for each pose, the four G-buffer
 * capture passes and the procedural shading through the ordinary pipeline,
 * then the covered pixels kept with probability `keep`.  Returns the
 * samples appended.  Leaves the procedural shader selected. */
size_t srpBenchTerrainCollect(SRPNeuralDataset* ds, int firstPose, int nPoses, float keep, uint64_t seed)
{
	size_t total = 0;
	for (int p = firstPose; p < firstPose + nPoses; p++)
	{
		srpBenchTerrainPose(p);
		for (int k = 0; k < 4; k++)
		{
			srpBenchTerrainShader(k + 1, NULL, NULL);
			srpBenchFrame();
			srpNeuralDatasetStage(ds, g_b.fb, k);
		}
		srpBenchTerrainShader(0, NULL, NULL);
		srpBenchFrame();
		srpNeuralDatasetStage(ds, g_b.fb, 4);
		total += srpNeuralDatasetAppend(ds, g_b.fb, keep, seed + (uint64_t) p);
	}
	return total;
}
#endif

void srpBenchTerrainBounds(float out[6])
{
	for (int i = 0; i < 3; i++)
	{
		out[i] = g_b.ut.bounds.lo[i];
		out[3 + i] = g_b.ut.bounds.inv[i];
	}
}

void srpBenchFrame(void)
{
	srpFramebufferClear(g_b.fb);
#ifdef SRP_GPU_BUILD
	srpFramebufferFillDepth(g_b.fb, 1.0f);
	srpFramebufferFillStencil(g_b.fb, 0);
#else
	for (size_t i = 0; i < g_b.fb->size; i++) g_b.fb->depth[i] = 1.0f;
	memset(g_b.fb->stencil, 0, g_b.fb->size);
#endif
	if (g_b.kind == 0 || g_b.kind == 2)
		srpDrawIndexBuffer(g_b.ib, g_b.vb, g_b.fb, &g_b.sp, SRP_PRIM_TRIANGLES, 0, g_b.nIdx);
	else if (g_b.kind == 3)
		srpDrawVertexBuffer(g_b.vb, g_b.fb, &g_b.sp, SRP_PRIM_TRIANGLES, 0, 3 * g_b.nTri);
	else
		srpSceneMixDraw(g_b.fb, &g_b.sp, g_b.vb, g_b.vbL, g_b.vbP, g_b.nTri, g_b.nLine, g_b.nPoint);
}

/* Copies the final buffers out (GPU: downloads the device copy). */
void srpBenchRead(uint32_t* color, float* depth, uint8_t* stencil)
{
#ifdef SRP_GPU_BUILD
	srpFramebufferDownload(g_b.fb);
#endif
	memcpy(color, g_b.fb->color, g_b.fb->size * 4);
	memcpy(depth, g_b.fb->depth, g_b.fb->size * 4);
	memcpy(stencil, g_b.fb->stencil, g_b.fb->size);
}

void srpBenchFinish(void)
{
	if (g_b.u.texture) srpFreeTexture(g_b.u.texture);
	if (g_b.ib) srpFreeIndexBuffer(g_b.ib);
	if (g_b.vb) srpFreeVertexBuffer(g_b.vb);
	if (g_b.vbL) srpFreeVertexBuffer(g_b.vbL);
	if (g_b.vbP) srpFreeVertexBuffer(g_b.vbP);
	if (g_b.fbs[0]) srpFreeFramebuffer(g_b.fbs[0]);
	if (g_b.fbs[1]) srpFreeFramebuffer(g_b.fbs[1]);
	memset(&g_b, 0, sizeof(g_b));
}
