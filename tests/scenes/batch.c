/* The batched-scenes workload: independent scenes of uneven size, each a
 * seeded triangle soup seen by its own perspective camera, with its own
 * framebuffer and depth / stencil / cull / scissor state.
 *
 *   side   uniform in {128, 256, 512, 1024, 2048}
 *   tris   min(2M, round(lognormal(ln 20000, 1))) * scale (at least 1)
 *   seed   0x5EED0004 (scene k uses a stream derived from it)
 *
 * The GPU build renders the scenes as one batch (srpBatchRender, optionally
 * a captured HIP graph) or one draw at a time; the CPU builds (oracle,
 * compiled reference) draw them one at a time.  Both orders give the same
 * frames: the scenes are independent. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "shaders.h"
#include "scenes.h"
#ifdef SRP_GPU_BUILD
#include <srp/srp_hip.h>
#endif

#ifndef M_PI
#  define M_PI 3.14159265358979323846
#endif

void sc_vsColorMVP(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsColor(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);

typedef struct BScene
{
	int side;
	size_t nTri;
	UniMVP u;
	SRPContext ctx;
	float clearDepth;
	uint8_t clearStencil;
	SRPFramebuffer* fb;
	SRPVertexBuffer* vb;
} BScene;

static struct
{
	int n;
	BScene* s;
	SRPVertexShader vs;
	SRPFragmentShader fs;
	SRPShaderProgram* sp;   /* one per scene (its own uniform) */
#ifdef SRP_GPU_BUILD
	SRPBatch* batch;
#endif
} g_bb;

static SRPVaryingInfo g_bbVaryings[1] = {{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE}};

static uint64_t bbNext(uint64_t* s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static double bbU(uint64_t* s) { return (double) (bbNext(s) >> 11) * (1.0 / 9007199254740992.0); }
static double bbNormal(uint64_t* s)
{
	const double u1 = bbU(s) + 1e-300, u2 = bbU(s);
	return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}
static int bbPick(uint64_t* s, int n) { return (int) (bbU(s) * n) % n; }

/* Perspective camera at distance `dist` from the origin along a random
 * direction, looking at the origin (view space: x right, y up, z forward). */
static void bbCamera(uint64_t* s, int side, UniMVP* u)
{
	double dir[3];
	for (;;)
	{
		for (int i = 0; i < 3; i++) dir[i] = 2.0 * bbU(s) - 1.0;
		const double l = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
		if (l > 0.2 && l <= 1.0 && fabs(dir[1]) < 0.9 * l)
		{
			for (int i = 0; i < 3; i++) dir[i] /= l;
			break;
		}
	}
	const double dist = 2.6 + 1.6 * bbU(s);
	const double e[3] = {dir[0] * dist, dir[1] * dist, dir[2] * dist};
	const double f[3] = {-dir[0], -dir[1], -dir[2]};
	double r[3] = {f[2], 0.0, -f[0]};
	const double rl = sqrt(r[0] * r[0] + r[2] * r[2]);
	r[0] /= rl; r[2] /= rl;
	const double up[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
	const double* rows[3] = {r, up, f};
	memset(&u->view, 0, sizeof(u->view));
	for (int i = 0; i < 3; i++)
	{
		for (int j = 0; j < 3; j++) u->view.data[i][j] = (float) rows[i][j];
		u->view.data[i][3] = (float) -(rows[i][0] * e[0] + rows[i][1] * e[1] + rows[i][2] * e[2]);
	}
	u->view.data[3][3] = 1.0f;
	u->model = mat4ConstructIdentity();
	const double fov = (50.0 + 20.0 * bbU(s)) * M_PI / 180.0;
	const float t = (float) tan(fov / 2.0), n = 0.1f, fa = 20.0f;
	(void) side;
	u->projection = mat4ConstructPerspectiveProjection(-t * n, t * n, -t * n, t * n, n, fa);
}

/* Render state of a scene, drawn from the API's enum sets. */
static void bbState(uint64_t* s, BScene* sc)
{
	srpNewContext(&srpContext);
	const SRPCompareOp ops[] = {SRP_COMPARE_NEVER, SRP_COMPARE_LESS, SRP_COMPARE_EQUAL, SRP_COMPARE_LEQUAL,
	                            SRP_COMPARE_GREATER, SRP_COMPARE_NOTEQUAL, SRP_COMPARE_GEQUAL, SRP_COMPARE_ALWAYS};
	const SRPStencilOp sops[] = {SRP_STENCIL_KEEP, SRP_STENCIL_ZERO, SRP_STENCIL_REPLACE, SRP_STENCIL_INCR,
	                             SRP_STENCIL_INCR_WRAP, SRP_STENCIL_DECR, SRP_STENCIL_DECR_WRAP, SRP_STENCIL_INVERT};
	const SRPCompareOp depthOps[] = {SRP_COMPARE_LESS, SRP_COMPARE_LEQUAL, SRP_COMPARE_GREATER, SRP_COMPARE_GEQUAL,
	                                 SRP_COMPARE_NOTEQUAL, SRP_COMPARE_ALWAYS};
	sc->clearDepth = 1.0f;
	if (bbU(s) < 0.85)
	{
		const SRPCompareOp op = depthOps[bbPick(s, 6)];
		srpDepthTest(true);
		srpDepthCompareOp(op);
		srpDepthWrite(bbU(s) < 0.9);
		sc->clearDepth = (op == SRP_COMPARE_GREATER || op == SRP_COMPARE_GEQUAL) ? -1.0f : 1.0f;
	}
	sc->clearStencil = 0;
	if (bbU(s) < 0.5)
	{
		srpStencilTest(true);
		sc->clearStencil = (uint8_t) bbPick(s, 4);
		srpStencilFunc(ops[bbPick(s, 8)], (uint8_t) bbPick(s, 4), (uint8_t) (bbU(s) < 0.5 ? 0xFF : 0x0F));
		srpStencilOp(sops[bbPick(s, 8)], sops[bbPick(s, 8)], sops[bbPick(s, 8)]);
		srpStencilWriteMask((uint8_t) (bbU(s) < 0.7 ? 0xFF : 0x3C));
	}
	const SRPFace faces[] = {SRP_FACE_NONE, SRP_FACE_BACK, SRP_FACE_FRONT};
	srpRasterCullFace(faces[bbPick(s, 3)]);
	srpRasterFrontFace(bbU(s) < 0.5 ? SRP_WINDING_CCW : SRP_WINDING_CW);
	if (bbU(s) < 0.2)
	{
		const size_t x = (size_t) bbPick(s, sc->side / 2), y = (size_t) bbPick(s, sc->side / 2);
		srpScissorTest(true);
		srpScissorOptions(x, y, (size_t) (sc->side / 4 + bbPick(s, sc->side / 2)), (size_t) (sc->side / 4 + bbPick(s, sc->side / 2)));
	}
	sc->ctx = srpContext;
}

static void bbClear(BScene* sc)
{
	srpFramebufferClear(sc->fb);
#ifdef SRP_GPU_BUILD
	srpFramebufferFillDepth(sc->fb, sc->clearDepth);
	srpFramebufferFillStencil(sc->fb, sc->clearStencil);
#else
	for (size_t i = 0; i < sc->fb->size; i++) sc->fb->depth[i] = sc->clearDepth;
	memset(sc->fb->stencil, sc->clearStencil, sc->fb->size);
#endif
}

/* nScenes scenes, triangle counts scaled by `scale`; batched (GPU build):
 * also make the batch.  Returns 0, or -1 when a batch cannot be made. */
int srpBenchBatchSetup(int nScenes, double scale, uint64_t seed, int batched)
{
	memset(&g_bb, 0, sizeof(g_bb));
	g_bb.n = nScenes;
	g_bb.s = (BScene*) calloc((size_t) nScenes, sizeof(BScene));
	g_bb.sp = (SRPShaderProgram*) calloc((size_t) nScenes, sizeof(SRPShaderProgram));
	g_bb.vs.shader = sc_vsColorMVP;
	g_bb.vs.nVaryings = 1;
	g_bb.vs.varyingsInfo = g_bbVaryings;
	g_bb.vs.varyingsSize = sizeof(VarColor);
	g_bb.fs.shader = sc_fsColor;
	g_bb.fs.mayOverwriteDepth = false;
	const int sides[] = {128, 256, 512, 1024, 2048};
	for (int k = 0; k < nScenes; k++)
	{
		BScene* sc = &g_bb.s[k];
		uint64_t s = seed + 0x1000003ull * (uint64_t) (k + 1);
		sc->side = sides[bbPick(&s, 5)];
		double t = floor(exp(log(20000.0) + bbNormal(&s)) + 0.5);
		if (t > 2000000.0) t = 2000000.0;
		t = floor(t * scale + 0.5);
		sc->nTri = t < 1.0 ? 1 : (size_t) t;
		bbCamera(&s, sc->side, &sc->u);
		bbState(&s, sc);
		VtxPosColor* v = (VtxPosColor*) malloc(3 * sc->nTri * sizeof(VtxPosColor));
		for (size_t i = 0; i < sc->nTri; i++)
		{
			const double cx = 2.0 * bbU(&s) - 1.0, cy = 2.0 * bbU(&s) - 1.0, cz = 2.0 * bbU(&s) - 1.0;
			const double len = exp(log(0.05) + 0.6 * bbNormal(&s));
			for (int j = 0; j < 3; j++)
			{
				VtxPosColor* p = &v[3 * i + j];
				p->position = VEC3((float) (cx + len * (2.0 * bbU(&s) - 1.0)), (float) (cy + len * (2.0 * bbU(&s) - 1.0)),
				                   (float) (cz + len * (2.0 * bbU(&s) - 1.0)));
				p->color = VEC3((float) bbU(&s), (float) bbU(&s), (float) bbU(&s));
			}
		}
		sc->vb = srpNewVertexBuffer();
		srpVertexBufferCopyData(sc->vb, sizeof(VtxPosColor), 3 * sc->nTri * sizeof(VtxPosColor), v);
		free(v);
		sc->fb = srpNewFramebuffer((size_t) sc->side, (size_t) sc->side);
		memset(sc->fb->stencil, 0, sc->fb->size);
#ifdef SRP_GPU_BUILD
		srpFramebufferSetDeferred(sc->fb, true);
#endif
		g_bb.sp[k].uniform = (SRPUniform*) &sc->u;
		g_bb.sp[k].vs = &g_bb.vs;
		g_bb.sp[k].fs = &g_bb.fs;
	}
#ifdef SRP_GPU_BUILD
	if (batched)
	{
		SRPDrawCommand* cmds = (SRPDrawCommand*) calloc((size_t) nScenes, sizeof(SRPDrawCommand));
		for (int k = 0; k < nScenes; k++)
		{
			BScene* sc = &g_bb.s[k];
			cmds[k].vb = sc->vb;
			cmds[k].fb = sc->fb;
			cmds[k].sp = &g_bb.sp[k];
			cmds[k].primitive = SRP_PRIM_TRIANGLES;
			cmds[k].startIndex = 0;
			cmds[k].count = 3 * sc->nTri;
			cmds[k].context = &sc->ctx;
			cmds[k].clear = 1;
			cmds[k].clearDepth = sc->clearDepth;
			cmds[k].clearStencil = sc->clearStencil;
		}
		g_bb.batch = srpNewBatch(cmds, (size_t) nScenes);
		free(cmds);
		if (!g_bb.batch)
			return -1;
	}
#else
	(void) batched;
#endif
	return 0;
}

/* One frame of every scene: batched (GPU build with a batch) or one clear
 * and one draw per scene. */
void srpBenchBatchFrame(int batched)
{
#ifdef SRP_GPU_BUILD
	if (batched && g_bb.batch)
	{
		srpBatchRender(g_bb.batch);
		return;
	}
#else
	(void) batched;
#endif
	for (int k = 0; k < g_bb.n; k++)
	{
		BScene* sc = &g_bb.s[k];
		srpContext = sc->ctx;
		bbClear(sc);
		srpDrawVertexBuffer(sc->vb, sc->fb, &g_bb.sp[k], SRP_PRIM_TRIANGLES, 0, 3 * sc->nTri);
	}
}

int srpBenchBatchScene(int k, int* side, uint64_t* nTri)
{
	if (k < 0 || k >= g_bb.n) return -1;
	*side = g_bb.s[k].side;
	*nTri = g_bb.s[k].nTri;
	return 0;
}

void srpBenchBatchRead(int k, uint32_t* color, float* depth, uint8_t* stencil)
{
	BScene* sc = &g_bb.s[k];
#ifdef SRP_GPU_BUILD
	srpFramebufferDownload(sc->fb);
#endif
	memcpy(color, sc->fb->color, sc->fb->size * 4);
	memcpy(depth, sc->fb->depth, sc->fb->size * 4);
	memcpy(stencil, sc->fb->stencil, sc->fb->size);
}

#ifdef SRP_GPU_BUILD
int srpBenchBatchCapture(void) { return g_bb.batch ? srpBatchCapture(g_bb.batch) : -1; }

double srpBenchBatchAutotune(int trials, int out[5])
{
	SRPBatchTuning t;
	const double ms = srpBatchAutotune(g_bb.batch, trials, &t);
	out[0] = t.wavesPerCU;
	out[1] = t.maskBox;
	out[2] = t.xcdQueues;
	out[3] = t.pixelRuns;
	out[4] = t.tileSize;
	return ms;
}

void srpBenchBatchSetTuning(int waves, int box, int xcd, int runs, int tile)
{
	SRPBatchTuning t = {waves, box, xcd, runs, 1, tile};
	srpBatchSetTuning(g_bb.batch, &t);
}

void srpBenchBatchInfo(uint64_t out[4]) { srpBatchInfo(g_bb.batch, out); }

/* Free scene k's vertex buffer while the batch still names it (the batch's
 * frames must then draw nothing and say so); the scene keeps no buffer. */
void srpBenchBatchDropScene(int k)
{
	srpFreeVertexBuffer(g_bb.s[k].vb);
	g_bb.s[k].vb = NULL;
}
#endif

/* New vertex data for scene k: the same number of triangles drawn from
 * another seed, in a larger allocation (the device buffer moves), so a batch
 * made earlier must be prepared again before its next frame. */
void srpBenchBatchRegenerate(int k, uint64_t seed)
{
	BScene* sc = &g_bb.s[k];
	uint64_t s = seed;
	const size_t nv = 3 * sc->nTri, spare = 3 * sc->nTri + 64;
	VtxPosColor* v = (VtxPosColor*) calloc(nv + spare, sizeof(VtxPosColor));
	for (size_t i = 0; i < sc->nTri; i++)
	{
		const double cx = 2.0 * bbU(&s) - 1.0, cy = 2.0 * bbU(&s) - 1.0, cz = 2.0 * bbU(&s) - 1.0;
		const double len = exp(log(0.08) + 0.6 * bbNormal(&s));
		for (int j = 0; j < 3; j++)
		{
			VtxPosColor* p = &v[3 * i + j];
			p->position = VEC3((float) (cx + len * (2.0 * bbU(&s) - 1.0)), (float) (cy + len * (2.0 * bbU(&s) - 1.0)),
			                   (float) (cz + len * (2.0 * bbU(&s) - 1.0)));
			p->color = VEC3((float) bbU(&s), (float) bbU(&s), (float) bbU(&s));
		}
	}
	srpVertexBufferCopyData(sc->vb, sizeof(VtxPosColor), (nv + spare) * sizeof(VtxPosColor), v);
	free(v);
}

void srpBenchBatchFinish(void)
{
#ifdef SRP_GPU_BUILD
	if (g_bb.batch) srpFreeBatch(g_bb.batch);
#endif
	for (int k = 0; k < g_bb.n; k++)
	{
		if (g_bb.s[k].vb) srpFreeVertexBuffer(g_bb.s[k].vb);
		srpFreeFramebuffer(g_bb.s[k].fb);
	}
	free(g_bb.s);
	free(g_bb.sp);
	memset(&g_bb, 0, sizeof(g_bb));
}
