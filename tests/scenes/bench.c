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
#ifdef SRP_GPU_BUILD
#include <srp/srp_hip.h>
#endif

void sc_vsMeshTex(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsMeshTex(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_vsMix(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsMix(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);
void sc_vsColorPass(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
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

typedef struct Bench
{
	int kind;  /* 0 mesh, 1 mix */
	SRPFramebuffer* fb;
	SRPVertexBuffer *vb, *vbL, *vbP;
	SRPIndexBuffer* ib;
	size_t nIdx, nTri, nLine, nPoint;
	UniFrameMVPTex u;
	SRPVertexShader vs;
	SRPFragmentShader fs;
	SRPShaderProgram sp;
	int width, height;
} Bench;

static Bench g_b;

static void benchFbInit(int w, int h)
{
	g_b.width = w;
	g_b.height = h;
	g_b.fb = srpNewFramebuffer((size_t) w, (size_t) h);
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
	if (g_b.kind == 0)
		srpDrawIndexBuffer(g_b.ib, g_b.vb, g_b.fb, &g_b.sp, SRP_PRIM_TRIANGLES, 0, g_b.nIdx);
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
	if (g_b.fb) srpFreeFramebuffer(g_b.fb);
	memset(&g_b, 0, sizeof(g_b));
}
