/* Host side of libsrp_amd: the srp C API (include/srp) over the HIP
 * pipeline.  Entry points keep the reference's names, argument meaning and
 * error behaviour (message callback, early return; kitrofimov/srp
 * core/ and pipeline/draw.c); the work of a draw runs on the GPU:
 *
 *   vertex shader (shader module)  ->  assembly/clip/setup COUNT pass
 *   -> scans -> EMIT pass, sized from the same draw's previous counts and
 *      checked against its counts on the device (the host reads them at its
 *      next call; a new draw, or one that outgrew its buffers, waits)
 *   -> filled triangles: stable radix sort of (tile, primitive) pairs,
 *      per-tile rasteriser (shader module)
 *   -> lines/points: stable radix sort of fragments by pixel, per-pixel
 *      ordered fragment kernel (shader module). */
#include <hip/hip_runtime.h>

#include <sched.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#define SRP_INCLUDE_VEC
#define SRP_INCLUDE_MAT
#include "srp/srp.h"
#include "srp/srp_hip.h"
#include "srp/device/abi.h"
#include "kernels.h"
#include "sizing.h"
#include "../host/image_decode.h"
#include "../host/png_encode.h"

using namespace srp::lib;

/* The user program defines srpContext (reference include/srp/context.h:194);
 * this weak definition only serves programs that never do. */
extern "C" {
__attribute__((weak)) SRPContext srpContext;
}

namespace {

#define HIP_CHECK(x)                                                                        \
	do {                                                                                    \
		hipError_t e_ = (x);                                                                \
		if (e_ != hipSuccess)                                                               \
		{                                                                                   \
			fprintf(stderr, "libsrp_amd: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
			        __FILE__, __LINE__);                                                    \
			abort();                                                                        \
		}                                                                                   \
	} while (0)

SRPContext* g_ctx = nullptr;

SRPContext* ctx() { return g_ctx ? g_ctx : &srpContext; }

void resolvePending();

/* utils/message_callback.c:18-35.  A draw still waiting for its counts
 * reports first, so messages keep the order of the calls that caused them. */
void message(SRPMessageType type, SRPMessageSeverity sev, const char* fn, const char* fmt, ...)
{
	resolvePending();
	SRPContext* c = ctx();
	if (c->messageCallback.func == nullptr)
		return;
	char text[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(text, sizeof(text), fmt, ap);
	va_end(ap);
	c->messageCallback.func(type, sev, fn, text, c->messageCallback.userParameter);
}

/* Grow-only device allocation. */
struct DevBuf
{
	void* p = nullptr;
	size_t cap = 0;
	template <typename T>
	T* get(size_t bytes)
	{
		if (bytes > cap)
		{
			/* in-flight kernels of earlier draws may still use the old block */
			if (p)
			{
				HIP_CHECK(hipDeviceSynchronize());
				HIP_CHECK(hipFree(p));
			}
			cap = bytes + bytes / 4 + 256;
			HIP_CHECK(hipMalloc(&p, cap));
		}
		return static_cast<T*>(p);
	}
};

/* Parameter block: SrpDrawParams, then the uniform snapshot. */
/* Parameter block of a draw: SrpDrawParams, the draw's assembly record,
 * then (256-aligned) the inline uniform. */
constexpr size_t kAsmDrawOff = (sizeof(SrpDrawParams) + 63) & ~(size_t) 63;
/* the draw's device counters ride in the block's padding, so the parameter
 * upload zeroes them: eight u32 (assembly totals mod 2^32, error flags, clip
 * count), then the three exact u64 totals of the counting pass */
constexpr size_t kTotalsOff = (kAsmDrawOff + sizeof(AsmDraw) + 15) & ~(size_t) 15;
constexpr size_t kTotalsBytes = 8 * sizeof(uint32_t) + 4 * sizeof(uint64_t);
constexpr size_t kParamsHead = (kTotalsOff + kTotalsBytes + 255) & ~(size_t) 255;
constexpr size_t kMaxInlineUniform = 16384;
constexpr size_t kStatWords = (size_t) SRP_STAT_SLOTS * SRP_STAT_STRIDE;
constexpr size_t kParamsBlock = kParamsHead + kMaxInlineUniform;

/* A single draw as the counts of its emitting half depend on it: the
 * buffers and range it reads and the framebuffer's size. */
struct DrawKey
{
	const void* ib;
	const void* vb;
	size_t start, count;
	uint32_t prim, width, height;
	bool operator==(const DrawKey& o) const
	{
		return ib == o.ib && vb == o.vb && start == o.start && count == o.count && prim == o.prim &&
		       width == o.width && height == o.height;
	}
};
struct DrawKeyHash
{
	size_t operator()(const DrawKey& k) const
	{
		size_t h = std::hash<const void*>()(k.ib) * 31u ^ std::hash<const void*>()(k.vb);
		h = h * 1000003u ^ k.start;
		h = h * 1000003u ^ k.count;
		h = h * 1000003u ^ ((size_t) k.prim << 40 ^ (size_t) k.width << 20 ^ k.height);
		return h;
	}
};

struct Runtime
{
	hipStream_t stream = nullptr;
	hipStream_t side = nullptr;       /* a single draw's clipped-triangle emitting kernel */
	hipEvent_t forkEv = nullptr, joinEv = nullptr;
	SrpDrawParams* hParams = nullptr;   /* pinned staging */
	SrpDrawParams* dParams = nullptr;
	uint8_t* hUniform = nullptr;        /* pinned staging */
	size_t uniformCap = 0;
	uint32_t* hTotals = nullptr;        /* pinned, mapped: the draw's counters (kTotalsBytes) */
	uint32_t* dTotals = nullptr;        /* its device address */
	unsigned long long* dStats = nullptr;
	uint64_t hostPrims = 0, hostDraws = 0;
	int tile = 8;
	int batched = 1;  /* three-phase batched rasteriser (tile 8) */
	uint32_t maskBox = SRP_SMALL_BBOX;  /* exact binning masks up to this box side */
	uint32_t bigPixel = 1;  /* larger triangles shaded pixel-parallel in runs */
	uint32_t depthCull = 1; /* tile-level depth culling in runs */
	uint32_t tileOrder = 8; /* single draws: runs of this many tiles dealt to the XCDs in turn (0: ranges) */
	uint32_t deferShading = 0;  /* wave shaders shade each pixel's visible fragment only, after the rasteriser */
	uint32_t lastClipped = 0;   /* triangles the previous draw clipped (grid hint of a new draw) */

	/* A single draw whose emitting half was enqueued before its counts
	 * reached the host (see PendingDraw). */
	struct Pending
	{
		bool active = false;
		int kind = 0, backend = 0;
		uint32_t nIn = 0, nTiles = 0, cacheSize = 0, varyingsSize = 0, lateZ = 0, words = 0;
		uint32_t clipHint = 0;
		uint32_t fragWords = 0;    /* 4-byte words of a fragment's interpolated varyings */
		bool clipSkipped = false;  /* counted without the clipping kernel (it clipped nothing last time) */
		uint32_t* tileStart = nullptr;   /* tile ranges (filled triangles), zeroed by the counting pass */
		int fragBits = 0;
		bool emitted = false;      /* the emitting half is enqueued (else it waits for the counts) */
		DrawKey key;
		AsmArgs A;                 /* counting-pass arguments: offsets, lists, counters */
		SrpRasterLaunch rl;        /* fields known before the counts */
		const SrpFsEntry* fse = nullptr;
		hipEvent_t evDraw = nullptr;
	} pending;
	uint32_t countsSeq = 0;   /* publication number of the pending draw's counters (hTotals[16]) */
	/* each draw's counts at its last call (clipped: triangles it clipped),
	 * which size its next emitting half */
	std::unordered_map<DrawKey, FitCaps, DrawKeyHash> drawCounts;

	DevBuf uniform, clip, varyings, genVaryings, counts;
	DevBuf tris, lines, points, pairKeys, pairVals, pairRec, pairMask, frags, fragKeys, order, tileCount, spans;
	DevBuf deferList;   /* (pixel, record) of each pixel a deferred-shading draw writes */
	DevBuf deferCounts; /* pixels each tile lists there, then their offsets */
	DevBuf deferFrags;  /* those pixels' fragments, rebuilt for shading */
	ScanScratch scan;
	SortScratch sort;

	std::unordered_map<const void*, const SrpVsEntry*> vs;
	std::unordered_map<const void*, const SrpFsEntry*> fs;

	/* Profiling: event pairs recorded around launches, read back lazily so
	 * that timing adds no host synchronisation to the draws. */
	struct ProfRec { hipEvent_t a, b; int slot; };
	bool profile = false;
	uint32_t profileSlots = 0xFu;   /* slots that record events while profiling */
	bool prof(int slot) const { return profile && ((profileSlots >> slot) & 1u); }
	double profMs[4] = {0, 0, 0, 0};
	uint64_t profLaunches[4] = {0, 0, 0, 0};
	std::vector<ProfRec> profPending;
	std::vector<hipEvent_t> eventPool;
	/* A draw's staging block is free again once its counters have been
	 * published (they follow its uploads on the stream); a parameter patch
	 * after that (a larger clip-vertex buffer) is waited for with this. */
	hipEvent_t paramsCopied = nullptr;
	/* What the device parameter block holds outside its counters (bytes
	 * [0, kTotalsOff) and the inline uniform from kParamsHead), so that a
	 * draw with the same parameters need not send them again; size 0: not
	 * known. */
	std::vector<uint8_t> sentParams;
	bool paramsPatched = false;

	hipEvent_t mark()
	{
		hipEvent_t e;
		if (eventPool.empty()) HIP_CHECK(hipEventCreate(&e));
		else { e = eventPool.back(); eventPool.pop_back(); }
		HIP_CHECK(hipEventRecord(e, stream));
		return e;
	}
	void span(int slot, hipEvent_t a, hipEvent_t b) { profPending.push_back({a, b, slot}); }
	void collect()
	{
		HIP_CHECK(hipStreamSynchronize(stream));
		/* one event may bound several spans (a draw's start); each goes
		 * back to the pool once */
		std::vector<hipEvent_t> done;
		for (const ProfRec& r : profPending)
		{
			float ms = 0.0f;
			HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
			profMs[r.slot] += ms;
			profLaunches[r.slot]++;
			done.push_back(r.a);
			done.push_back(r.b);
		}
		std::sort(done.begin(), done.end());
		done.erase(std::unique(done.begin(), done.end()), done.end());
		eventPool.insert(eventPool.end(), done.begin(), done.end());
		profPending.clear();
	}

	Runtime()
	{
		HIP_CHECK(hipSetDevice(0));
		HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
		HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
		HIP_CHECK(hipEventCreateWithFlags(&forkEv, hipEventDisableTiming));
		HIP_CHECK(hipEventCreateWithFlags(&joinEv, hipEventDisableTiming));
		HIP_CHECK(hipHostMalloc((void**) &hParams, kParamsBlock, hipHostMallocDefault));
		HIP_CHECK(hipMalloc((void**) &dParams, kParamsBlock));
		HIP_CHECK(hipHostMalloc((void**) &hTotals, kTotalsBytes + 64, hipHostMallocMapped));
		memset(hTotals, 0, kTotalsBytes + 64);
		HIP_CHECK(hipHostGetDevicePointer((void**) &dTotals, hTotals, 0));
		HIP_CHECK(hipMalloc((void**) &dStats, kStatWords * sizeof(unsigned long long)));
		HIP_CHECK(hipMemset(dStats, 0, kStatWords * sizeof(unsigned long long)));
		HIP_CHECK(hipEventCreateWithFlags(&paramsCopied, hipEventDisableTiming));
		const char* t = getenv("SRP_TILE");
		if (t && (atoi(t) == 8 || atoi(t) == 16)) tile = atoi(t);
		const char* b = getenv("SRP_RASTER");
		if (b && strcmp(b, "simple") == 0) batched = 0;
		if (b && strcmp(b, "lists") == 0) bigPixel = 0;
	}
};

Runtime* g_rt = nullptr;   /* set once the runtime exists */

Runtime& rt()
{
	static Runtime* r = (g_rt = new Runtime());
	return *r;
}

/* Registries live outside the runtime so that shader modules loaded before
 * the first API call register without touching the GPU. */
std::mutex g_regLock;
std::unordered_map<const void*, const SrpVsEntry*>& vsRegistry()
{
	static auto* m = new std::unordered_map<const void*, const SrpVsEntry*>();
	return *m;
}
std::unordered_map<const void*, const SrpFsEntry*>& fsRegistry()
{
	static auto* m = new std::unordered_map<const void*, const SrpFsEntry*>();
	return *m;
}

/* ---------------------------------------------------------------- objects */

struct FbImpl
{
	SRPFramebuffer pub;       /* first member: SRPFramebuffer* == FbImpl* */
	uint32_t* dColor = nullptr;
	float* dDepth = nullptr;
	uint8_t* dStencil = nullptr;
	bool deferred = false;
	/* Deferred mode: clears and fills are recorded and run as one fused
	 * kernel before the next use of the device copy. */
	bool pendColor = false, pendDepth = false, pendStencil = false;
	uint32_t fillColor = 0, fillDepthBits = 0;
	uint8_t fillStencil = 0;
};

FbImpl* fbImpl(const SRPFramebuffer* fb) { return (FbImpl*) fb; }

void fbFlushFills(FbImpl* F)
{
	if (!(F->pendColor || F->pendDepth || F->pendStencil))
		return;
	fillFramebuffer(F->pendColor ? F->dColor : nullptr, F->fillColor,
	                F->pendDepth ? (uint32_t*) F->dDepth : nullptr, F->fillDepthBits,
	                F->pendStencil ? F->dStencil : nullptr, F->fillStencil, F->pub.size, rt().stream);
	F->pendColor = F->pendDepth = F->pendStencil = false;
}

void fbUpload(FbImpl* F)
{
	Runtime& R = rt();
	F->pendColor = F->pendDepth = F->pendStencil = false;  /* the host arrays win */
	const size_t n = F->pub.size;
	HIP_CHECK(hipMemcpyAsync(F->dColor, F->pub.color, n * 4, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(F->dDepth, F->pub.depth, n * 4, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(F->dStencil, F->pub.stencil, n, hipMemcpyHostToDevice, R.stream));
}

void fbDownload(FbImpl* F)
{
	Runtime& R = rt();
	fbFlushFills(F);
	const size_t n = F->pub.size;
	HIP_CHECK(hipMemcpyAsync(F->pub.color, F->dColor, n * 4, hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipMemcpyAsync(F->pub.depth, F->dDepth, n * 4, hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipMemcpyAsync(F->pub.stencil, F->dStencil, n, hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
}

struct TexImpl
{
	SrpTexDesc* dDesc = nullptr;
	uint32_t* dTexels = nullptr;
	std::vector<uint8_t> rgb;
	int w = 0, h = 0;
	SrpTexDesc desc;
};

std::unordered_map<const void*, TexImpl*>& textures()
{
	static auto* m = new std::unordered_map<const void*, TexImpl*>();
	return *m;
}

TexImpl* texImpl(const SRPTexture* t)
{
	auto& m = textures();
	auto it = m.find((const void*) t);
	return it == m.end() ? nullptr : it->second;
}

SRPTexture* makeTexture(uint8_t* rgb, int w, int h, SRPTextureWrappingMode wx, SRPTextureWrappingMode wy)
{
	Runtime& R = rt();
	TexImpl* T = new TexImpl();
	T->w = w;
	T->h = h;
	T->rgb.assign(rgb, rgb + (size_t) w * h * 3);
	std::vector<uint32_t> rgba((size_t) w * h);
	for (size_t i = 0; i < rgba.size(); i++)
		rgba[i] = (uint32_t) rgb[3 * i] | ((uint32_t) rgb[3 * i + 1] << 8) | ((uint32_t) rgb[3 * i + 2] << 16) | 0xFF000000u;
	HIP_CHECK(hipMalloc((void**) &T->dTexels, rgba.size() * 4 + 16));
	HIP_CHECK(hipMalloc((void**) &T->dDesc, sizeof(SrpTexDesc)));
	T->desc.texels = T->dTexels;
	T->desc.w = w; T->desc.h = h;
	T->desc.wM1 = w - 1; T->desc.hM1 = h - 1;
	T->desc.wrapX = wx; T->desc.wrapY = wy;
	HIP_CHECK(hipMemcpyAsync(T->dTexels, rgba.data(), rgba.size() * 4, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(T->dDesc, &T->desc, sizeof(SrpTexDesc), hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
	textures()[(const void*) T->dDesc] = T;
	return (SRPTexture*) T->dDesc;
}

void texPush(TexImpl* T)
{
	Runtime& R = rt();
	HIP_CHECK(hipMemcpyAsync(T->dDesc, &T->desc, sizeof(SrpTexDesc), hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
}

}  // namespace

namespace srp {
namespace lib {

hipStream_t libStream()
{
	resolvePending();
	return rt().stream;
}

void libMessage(SRPMessageType type, SRPMessageSeverity sev, const char* fn, const char* fmt, ...)
{
	char text[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(text, sizeof(text), fmt, ap);
	va_end(ap);
	message(type, sev, fn, "%s", text);
}

void framebufferDevice(const SRPFramebuffer* fb, uint32_t** color, float** depth, uint8_t** stencil)
{
	resolvePending();
	FbImpl* F = fbImpl(fb);
	if (F->deferred)
		fbFlushFills(F);
	else
		fbUpload(F);
	if (color) *color = F->dColor;
	if (depth) *depth = F->dDepth;
	if (stencil) *stencil = F->dStencil;
}

}  // namespace lib
}  // namespace srp

struct SRPVertexBuffer
{
	size_t nBytesPerVertex = 0, nVertices = 0, nBytesAllocated = 0;
	uint8_t* dData = nullptr;
	uint64_t version = 0;   /* bumped by every CopyData (batches re-prepare) */
};

struct SRPIndexBuffer
{
	SRPType type = SRP_UINT8;
	size_t elem = 1, nIndices = 0, nBytesAllocated = 0;
	uint8_t* dData = nullptr;
	std::vector<uint8_t> host;
	/* min/max vertex index per block of 4096 indices, for range queries */
	std::vector<uint64_t> blkMin, blkMax;
	uint64_t version = 0;   /* bumped by every CopyData (batches re-prepare) */
};

namespace {

constexpr size_t kIdxBlock = 4096;

/* Buffers and framebuffers not yet freed (a batch checks the ones it uses
 * before each frame). */
std::unordered_set<const void*>& liveObjects()
{
	static auto* s = new std::unordered_set<const void*>();
	return *s;
}

uint64_t hostIndex(const SRPIndexBuffer* b, size_t i)
{
	const uint8_t* p = b->host.data() + i * b->elem;
	switch (b->elem)
	{
		case 1: return *p;
		case 2: { uint16_t v; memcpy(&v, p, 2); return v; }
		case 4: { uint32_t v; memcpy(&v, p, 4); return v; }
		default: { uint64_t v; memcpy(&v, p, 8); return v; }
	}
}

/* vertex_processing.c:90-112: smallest and largest vertex index of the draw. */
void indexRange(const SRPIndexBuffer* b, size_t start, size_t count, uint64_t& lo, uint64_t& hi)
{
	lo = UINT64_MAX;
	hi = 0;
	size_t i = start;
	const size_t end = start + count;
	while (i < end)
	{
		if (i % kIdxBlock == 0 && i + kIdxBlock <= end)
		{
			const size_t k = i / kIdxBlock;
			lo = b->blkMin[k] < lo ? b->blkMin[k] : lo;
			hi = b->blkMax[k] > hi ? b->blkMax[k] : hi;
			i += kIdxBlock;
			continue;
		}
		const uint64_t v = hostIndex(b, i);
		lo = v < lo ? v : lo;
		hi = v > hi ? v : hi;
		i++;
	}
}

}  // namespace

/* ======================================================================= */
extern "C" {

/* ---------------------------------------------------------------- context */

/* core/context.c:15-51 */
void srpNewContext(SRPContext* c)
{
	resolvePending();
	g_ctx = c;
	c->messageCallback.func = nullptr;
	c->messageCallback.userParameter = nullptr;
	c->provokingVertexMode = SRP_PROVOKING_VERTEX_LAST;
	c->raster.frontFace = SRP_WINDING_CCW;
	c->raster.cullFace = SRP_FACE_NONE;
	c->raster.polygonMode = SRP_POLYGON_MODE_FILL;
	c->raster.pointSize = 1.0f;
	memset(&c->scissor, 0, sizeof(c->scissor));
	c->depth.testEnable = false;
	c->depth.writeEnable = true;
	c->depth.compareOp = SRP_COMPARE_GREATER;
	SRPStencilFaceState f;
	f.func = SRP_COMPARE_ALWAYS;
	f.ref = 0; f.mask = 0xFF; f.writeMask = 0xFF;
	f.sfailOp = SRP_STENCIL_KEEP; f.dfailOp = SRP_STENCIL_KEEP; f.passOp = SRP_STENCIL_KEEP;
	c->stencil.enabled = false;
	c->stencil.front = f;
	c->stencil.back = f;
	c->arena = (SRPArena*) &rt();
}

/* core/context.c:53-180 */
void srpSetMessageCallback(SRPMessageCallback cb)
{
	resolvePending();   /* a pending draw reports to the callback of its call */
	ctx()->messageCallback = cb;
}
void srpProvokingVertexMode(SRPProvokingVertexMode m) { ctx()->provokingVertexMode = m; }
void srpRasterCullFace(SRPFace f) { ctx()->raster.cullFace = f; }
void srpRasterFrontFace(SRPWinding w) { ctx()->raster.frontFace = w; }
void srpRasterPolygonMode(SRPPolygonMode m) { ctx()->raster.polygonMode = m; }
void srpRasterPointSize(float s) { ctx()->raster.pointSize = s; }
void srpScissorTest(bool e) { ctx()->scissor.enabled = e; }
void srpScissorOptions(size_t x, size_t y, size_t w, size_t h)
{
	SRPContext* c = ctx();
	c->scissor.x = x; c->scissor.y = y; c->scissor.width = w; c->scissor.height = h;
}
void srpStencilTest(bool) { ctx()->stencil.enabled = true; }
void srpStencilFunc(SRPCompareOp func, uint8_t ref, uint8_t mask)
{
	SRPContext* c = ctx();
	c->stencil.front.func = func; c->stencil.front.ref = ref; c->stencil.front.mask = mask;
	c->stencil.back.func = func; c->stencil.back.ref = ref; c->stencil.back.mask = mask;
}
static SRPStencilFaceState* faceState(SRPFace f)
{
	return (f == SRP_FACE_FRONT) ? &ctx()->stencil.front : &ctx()->stencil.back;
}
void srpStencilFuncSeparate(SRPFace face, SRPCompareOp func, uint8_t ref, uint8_t mask)
{
	if (face == SRP_FACE_NONE) return;
	if (face == SRP_FACE_FRONT_AND_BACK) srpStencilFunc(func, ref, mask);
	SRPStencilFaceState* s = faceState(face);
	s->func = func; s->ref = ref; s->mask = mask;
}
void srpStencilOp(SRPStencilOp sfail, SRPStencilOp dfail, SRPStencilOp pass)
{
	SRPContext* c = ctx();
	c->stencil.front.sfailOp = sfail; c->stencil.front.dfailOp = dfail; c->stencil.front.passOp = pass;
	c->stencil.back.sfailOp = sfail; c->stencil.back.dfailOp = dfail; c->stencil.back.passOp = pass;
}
void srpStencilOpSeparate(SRPFace face, SRPStencilOp sfail, SRPStencilOp dfail, SRPStencilOp pass)
{
	if (face == SRP_FACE_NONE) return;
	if (face == SRP_FACE_FRONT_AND_BACK) srpStencilOp(sfail, dfail, pass);
	SRPStencilFaceState* s = faceState(face);
	s->sfailOp = sfail; s->dfailOp = dfail; s->passOp = pass;
}
void srpStencilWriteMask(uint8_t m)
{
	ctx()->stencil.front.writeMask = m;
	ctx()->stencil.back.writeMask = m;
}
void srpStencilWriteMaskSeparate(SRPFace face, uint8_t m)
{
	if (face == SRP_FACE_NONE) return;
	if (face == SRP_FACE_FRONT_AND_BACK) srpStencilWriteMask(m);
	faceState(face)->writeMask = m;
}
void srpDepthTest(bool e) { ctx()->depth.testEnable = e; }
void srpDepthWrite(bool e) { ctx()->depth.writeEnable = e; }
void srpDepthCompareOp(SRPCompareOp op) { ctx()->depth.compareOp = op; }

/* ---------------------------------------------------------------- buffers */

SRPVertexBuffer* srpNewVertexBuffer(void)
{
	SRPVertexBuffer* b = new SRPVertexBuffer();
	liveObjects().insert(b);
	return b;
}

void srpFreeVertexBuffer(SRPVertexBuffer* b)
{
	resolvePending();
	if (!b) return;
	liveObjects().erase(b);
	if (b->dData)
	{
		HIP_CHECK(hipStreamSynchronize(rt().stream));
		HIP_CHECK(hipFree(b->dData));
	}
	delete b;
}

/* core/buffer.c:30-45 (the copy lands in device memory) */
void srpVertexBufferCopyData(SRPVertexBuffer* b, size_t stride, size_t nBytes, const void* data)
{
	resolvePending();
	Runtime& R = rt();
	if (nBytes > b->nBytesAllocated)
	{
		if (b->dData)
		{
			HIP_CHECK(hipStreamSynchronize(R.stream));
			HIP_CHECK(hipFree(b->dData));
		}
		/* slack: shaders may read a little past the last vertex struct */
		HIP_CHECK(hipMalloc((void**) &b->dData, nBytes + 256));
		b->nBytesAllocated = nBytes;
	}
	b->nBytesPerVertex = stride;
	b->nVertices = stride ? nBytes / stride : 0;
	b->version++;
	HIP_CHECK(hipMemcpyAsync(b->dData, data, nBytes, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
}

SRPIndexBuffer* srpNewIndexBuffer(void)
{
	SRPIndexBuffer* b = new SRPIndexBuffer();
	liveObjects().insert(b);
	return b;
}

void srpFreeIndexBuffer(SRPIndexBuffer* b)
{
	resolvePending();
	if (!b) return;
	liveObjects().erase(b);
	if (b->dData)
	{
		HIP_CHECK(hipStreamSynchronize(rt().stream));
		HIP_CHECK(hipFree(b->dData));
	}
	delete b;
}

/* utils/type.c:15-40 */
static size_t sizeofIndexType(SRPType t)
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
			message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpSizeofType", "Unknown type (%i)", t);
			return 0;
	}
}

/* core/buffer.c:77-94 */
void srpIndexBufferCopyData(SRPIndexBuffer* b, SRPType type, size_t nBytes, const void* data)
{
	resolvePending();
	Runtime& R = rt();
	if (nBytes > b->nBytesAllocated)
	{
		if (b->dData)
		{
			HIP_CHECK(hipStreamSynchronize(R.stream));
			HIP_CHECK(hipFree(b->dData));
		}
		HIP_CHECK(hipMalloc((void**) &b->dData, nBytes + 16));
		b->nBytesAllocated = nBytes;
	}
	b->type = type;
	b->elem = sizeofIndexType(type);
	b->version++;
	b->nIndices = b->elem ? nBytes / b->elem : 0;
	b->host.assign((const uint8_t*) data, (const uint8_t*) data + nBytes);
	const size_t nb = (b->nIndices + kIdxBlock - 1) / kIdxBlock;
	b->blkMin.assign(nb, UINT64_MAX);
	b->blkMax.assign(nb, 0);
	if (b->elem == 1 || b->elem == 2 || b->elem == 4 || b->elem == 8)
		for (size_t i = 0; i < b->nIndices; i++)
		{
			const uint64_t v = hostIndex(b, i);
			uint64_t& lo = b->blkMin[i / kIdxBlock];
			uint64_t& hi = b->blkMax[i / kIdxBlock];
			lo = v < lo ? v : lo;
			hi = v > hi ? v : hi;
		}
	HIP_CHECK(hipMemcpyAsync(b->dData, data, nBytes, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
}

/* ------------------------------------------------------------ framebuffer */

SRPFramebuffer* srpNewFramebuffer(size_t w, size_t h)
{
	Runtime& R = rt();
	FbImpl* F = new FbImpl();
	F->pub.width = w;
	F->pub.height = h;
	F->pub.size = w * h;
	const size_t n = F->pub.size;
	HIP_CHECK(hipHostMalloc((void**) &F->pub.color, n * 4 + 16, hipHostMallocDefault));
	HIP_CHECK(hipHostMalloc((void**) &F->pub.depth, n * 4 + 16, hipHostMallocDefault));
	HIP_CHECK(hipHostMalloc((void**) &F->pub.stencil, n + 16, hipHostMallocDefault));
	HIP_CHECK(hipMalloc((void**) &F->dColor, n * 4 + 16));
	HIP_CHECK(hipMalloc((void**) &F->dDepth, n * 4 + 16));
	HIP_CHECK(hipMalloc((void**) &F->dStencil, n + 16));
	/* Stencil starts zeroed (the reference's tests rely on fresh zero pages). */
	memset(F->pub.stencil, 0, n);
	memset(F->pub.color, 0, n * 4);
	memset(F->pub.depth, 0, n * 4);
	fillU8(F->dStencil, 0, n, R.stream);
	fillU32(F->dColor, 0, n, R.stream);
	fillU32((uint32_t*) F->dDepth, 0, n, R.stream);
	HIP_CHECK(hipStreamSynchronize(R.stream));
	liveObjects().insert(&F->pub);
	return &F->pub;
}

void srpFreeFramebuffer(SRPFramebuffer* fb)
{
	resolvePending();
	if (!fb) return;
	liveObjects().erase(fb);
	FbImpl* F = fbImpl(fb);
	HIP_CHECK(hipStreamSynchronize(rt().stream));
	HIP_CHECK(hipHostFree(F->pub.color));
	HIP_CHECK(hipHostFree(F->pub.depth));
	HIP_CHECK(hipHostFree(F->pub.stencil));
	HIP_CHECK(hipFree(F->dColor));
	HIP_CHECK(hipFree(F->dDepth));
	HIP_CHECK(hipFree(F->dStencil));
	delete F;
}

/* core/framebuffer.c:57-62: colour 0, depth -1, stencil untouched */
void srpFramebufferClear(const SRPFramebuffer* fb)
{
	resolvePending();
	FbImpl* F = fbImpl(fb);
	if (F->deferred)
	{
		const float m1 = -1.0f;
		F->pendColor = true;
		F->fillColor = 0u;
		F->pendDepth = true;
		memcpy(&F->fillDepthBits, &m1, 4);
		return;
	}
	memset(fb->color, 0, fb->size * 4);
	for (size_t i = 0; i < fb->size; i++) fb->depth[i] = -1.0f;
}

void srpFramebufferSetDeferred(const SRPFramebuffer* fb, bool deferred)
{
	resolvePending();
	if (!fb)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid framebuffer");
		return;
	}
	FbImpl* F = fbImpl(fb);
	if (deferred && !F->deferred)
		fbUpload(F);
	if (!deferred && F->deferred)
		fbDownload(F);
	F->deferred = deferred;
}

void srpFramebufferDownload(const SRPFramebuffer* fb)
{
	resolvePending();
	if (!fb)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid framebuffer");
		return;
	}
	fbDownload(fbImpl(fb));
}

/* tests/utils/save.c:4-33 of the reference (its test harness): the colour
 * buffer as an RGBA PNG with alpha 255.  A deferred framebuffer's device
 * copy is downloaded first. */
int srpFramebufferSaveImage(const SRPFramebuffer* fb, const char* path)
{
	resolvePending();
	if (!fb || !path)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid framebuffer or path");
		return -1;
	}
	FbImpl* F = fbImpl(fb);
	if (F->deferred)
		fbDownload(F);
	if (fb->width > (1u << 24) || fb->height > (1u << 24) ||
	    srpPngSaveColorBuffer(path, (int) fb->width, (int) fb->height, fb->color) != 0)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Could not write %s", path);
		return -1;
	}
	return 0;
}

void srpFramebufferUpload(const SRPFramebuffer* fb)
{
	resolvePending();
	if (!fb)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid framebuffer");
		return;
	}
	fbUpload(fbImpl(fb));
	HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void srpFramebufferFillDepth(const SRPFramebuffer* fb, float v)
{
	resolvePending();
	if (!fb)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid framebuffer");
		return;
	}
	FbImpl* F = fbImpl(fb);
	if (F->deferred)
	{
		F->pendDepth = true;
		memcpy(&F->fillDepthBits, &v, 4);
		return;
	}
	for (size_t i = 0; i < fb->size; i++) fb->depth[i] = v;
}

void srpFramebufferFillStencil(const SRPFramebuffer* fb, uint8_t v)
{
	resolvePending();
	if (!fb)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Invalid framebuffer");
		return;
	}
	FbImpl* F = fbImpl(fb);
	if (F->deferred)
	{
		F->pendStencil = true;
		F->fillStencil = v;
		return;
	}
	memset(fb->stencil, v, fb->size);
}

/* --------------------------------------------------------------- textures */

/* core/texture.c:27-48 */
SRPTexture* srpNewTexture(const char* path, SRPTextureWrappingMode wx, SRPTextureWrappingMode wy)
{
	int w = 0, h = 0;
	const char* err = "unknown error";
	uint8_t* rgb = srpImageLoadRGB(path, &w, &h, &err);
	if (!rgb)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpNewTexture", "Failed to load image `%s`: %s", path, err);
		return nullptr;
	}
	SRPTexture* t = makeTexture(rgb, w, h, wx, wy);
	free(rgb);
	return t;
}

SRPTexture* srpNewTextureFromMemory(const uint8_t* rgb, int w, int h, SRPTextureWrappingMode wx, SRPTextureWrappingMode wy)
{
	if (!rgb || w <= 0 || h <= 0)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpNewTextureFromMemory", "Invalid texture data");
		return nullptr;
	}
	return makeTexture((uint8_t*) rgb, w, h, wx, wy);
}

void srpFreeTexture(SRPTexture* t)
{
	resolvePending();
	TexImpl* T = texImpl(t);
	if (!T) return;
	HIP_CHECK(hipStreamSynchronize(rt().stream));
	HIP_CHECK(hipFree(T->dTexels));
	HIP_CHECK(hipFree(T->dDesc));
	textures().erase((const void*) t);
	delete T;
}

/* core/texture.c:56-81 */
void srpTextureGetFilteredColorHost(const SRPTexture* t, float u, float v, float out[4])
{
	const TexImpl* T = texImpl(t);
	if (!T)
	{
		out[0] = out[1] = out[2] = 0.0f;
		out[3] = 1.0f;
		return;
	}
	if (u < 0 || u > 1)
		u = (T->desc.wrapX == TW_REPEAT) ? (float) ((double) u - floor((double) u)) : (float) fmax(0.0, fmin(1.0, (double) u));
	if (v < 0 || v > 1)
		v = (T->desc.wrapY == TW_REPEAT) ? (float) ((double) v - floor((double) v)) : (float) fmax(0.0, fmin(1.0, (double) v));
	const float x = (float) T->desc.wM1 * u;
	const float y = (float) T->desc.hM1 * (1 - v);
	const double xd = (double) x + 0.5, yd = (double) y + 0.5;
	int64_t xi = (xd >= 0.0 && xd < 9.0e18) ? (int64_t) xd : 0;
	int64_t yi = (yd >= 0.0 && yd < 9.0e18) ? (int64_t) yd : 0;
	if (xi > T->desc.wM1) xi = T->desc.wM1;
	if (yi > T->desc.hM1) yi = T->desc.hM1;
	const uint8_t* p = T->rgb.data() + (size_t) (yi * T->w + xi) * 3;
	const float inv255 = (float) (1. / 255.);
	out[0] = p[0] * inv255;
	out[1] = p[1] * inv255;
	out[2] = p[2] * inv255;
	out[3] = 1.0f;
}

/* core/texture.c:83-137 */
int srpTextureGet(SRPTexture* t, SRPTextureParameter p)
{
	TexImpl* T = texImpl(t);
	if (T && p == SRP_TEXTURE_WRAPPING_MODE_X) return T->desc.wrapX;
	if (T && p == SRP_TEXTURE_WRAPPING_MODE_Y) return T->desc.wrapY;
	message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpTextureGet", "Unknown texture parameter (%i)", p);
	return -1;
}

void srpTextureSet(SRPTexture* t, SRPTextureParameter p, int v)
{
	resolvePending();
	TexImpl* T = texImpl(t);
	if (!T || (p != SRP_TEXTURE_WRAPPING_MODE_X && p != SRP_TEXTURE_WRAPPING_MODE_Y))
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpTextureSet", "Unknown texture parameter (%i)", p);
		return;
	}
	if (p == SRP_TEXTURE_WRAPPING_MODE_X) T->desc.wrapX = v;
	else T->desc.wrapY = v;
	texPush(T);
	if (v != TW_REPEAT && v != TW_CLAMP_TO_EDGE)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpTextureSet",
		        "Unknown texture wrapping mode (%i). Falling back to TW_REPEAT", v);
}

/* ----------------------------------------------------- registration, misc */

void srpHipRegisterVertexShader(const void* fn, const void* entry)
{
	std::lock_guard<std::mutex> g(g_regLock);
	const SrpVsEntry* e = (const SrpVsEntry*) entry;
	if (e->abiVersion != SRP_ABI_VERSION)
	{
		fprintf(stderr, "libsrp_amd: vertex shader module built for ABI %u, library is %u\n", e->abiVersion, SRP_ABI_VERSION);
		return;
	}
	vsRegistry()[fn] = e;
}

void srpHipRegisterFragmentShader(const void* fn, const void* entry)
{
	std::lock_guard<std::mutex> g(g_regLock);
	const SrpFsEntry* e = (const SrpFsEntry*) entry;
	if (e->abiVersion != SRP_ABI_VERSION)
	{
		fprintf(stderr, "libsrp_amd: fragment shader module built for ABI %u, library is %u\n", e->abiVersion, SRP_ABI_VERSION);
		return;
	}
	fsRegistry()[fn] = e;
}

void srpHipShaderCalledOnHost(const char* name)
{
	message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, name,
	        "device shader `%s` was called from host code; it only runs inside srpDraw*", name);
}

void srpHipGetStats(uint64_t out[4])
{
	resolvePending();
	Runtime& R = rt();
	std::vector<unsigned long long> d(kStatWords);
	HIP_CHECK(hipStreamSynchronize(R.stream));
	HIP_CHECK(hipMemcpy(d.data(), R.dStats, kStatWords * sizeof(unsigned long long), hipMemcpyDeviceToHost));
	out[0] = out[1] = 0;
	for (size_t s = 0; s < SRP_STAT_SLOTS; s++)
	{
		out[0] += d[s * SRP_STAT_STRIDE];
		out[1] += d[s * SRP_STAT_STRIDE + 1];
	}
	out[2] = R.hostPrims;
	out[3] = R.hostDraws;
}

void srpHipResetStats(void)
{
	resolvePending();
	Runtime& R = rt();
	HIP_CHECK(hipStreamSynchronize(R.stream));
	HIP_CHECK(hipMemset(R.dStats, 0, kStatWords * sizeof(unsigned long long)));
	R.hostPrims = 0;
	R.hostDraws = 0;
}

void* srpHipGetStream(void)
{
	resolvePending();   /* work the caller adds follows the pending draw's, repeated or not */
	return (void*) rt().stream;
}
void srpHipSynchronize(void)
{
	resolvePending();
	HIP_CHECK(hipStreamSynchronize(rt().stream));
}

void srpHipProfileEnable(bool enable)
{
	resolvePending();
	Runtime& R = rt();
	R.collect();
	R.profile = enable;
	for (int i = 0; i < 4; i++) { R.profMs[i] = 0; R.profLaunches[i] = 0; }
}

void srpHipProfileSlots(unsigned mask)
{
	resolvePending();
	rt().profileSlots = mask & 0xFu;
}

void srpHipProfileRead(double ms[4], uint64_t launches[4])
{
	resolvePending();
	Runtime& R = rt();
	R.collect();
	for (int i = 0; i < 4; i++) { ms[i] = R.profMs[i]; launches[i] = R.profLaunches[i]; }
}

void srpHipSetTileSize(int t)
{
	resolvePending();
	if (t == 8 || t == 16) rt().tile = t;
}
int srpHipGetTileSize(void) { return rt().tile; }

void srpHipSetRasterOptions(int batched, int pixelRuns, int maskBox)
{
	resolvePending();
	Runtime& R = rt();
	if (batched >= 0) R.batched = batched ? 1 : 0;
	if (pixelRuns >= 0) R.bigPixel = pixelRuns ? 1u : 0u;
	if (maskBox >= 0) R.maskBox = (uint32_t) (maskBox > SRP_SMALL_BBOX ? SRP_SMALL_BBOX : maskBox);
}

/* Test hook (not part of the public headers): the multi-array exclusive
 * scan the draws use, on host data.  counts holds nArrays arrays of n u32;
 * offsets receives their scans (mod 2^32), totals32 the sums mod 2^32 and
 * totals64 the exact sums. */
void srp_lib_scan_exact(const uint32_t* counts, size_t n, int nArrays, uint32_t* offsets, uint32_t* totals32,
                        uint64_t* totals64)
{
	resolvePending();
	Runtime& R = rt();
	uint32_t* d = nullptr;
	uint32_t* t32 = nullptr;
	unsigned long long* t64 = nullptr;
	const size_t bytes = n * (size_t) nArrays * sizeof(uint32_t);
	HIP_CHECK(hipMalloc((void**) &d, bytes + 16));
	HIP_CHECK(hipMalloc((void**) &t32, 16 * sizeof(uint32_t)));
	HIP_CHECK(hipMalloc((void**) &t64, 16 * sizeof(unsigned long long)));
	HIP_CHECK(hipMemcpyAsync(d, counts, bytes, hipMemcpyHostToDevice, R.stream));
	ScanScratch sc;
	scanExclusiveMulti(d, n, nArrays, n, t32, sc, R.stream, t64);
	HIP_CHECK(hipMemcpyAsync(offsets, d, bytes, hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipMemcpyAsync(totals32, t32, (size_t) nArrays * sizeof(uint32_t), hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipMemcpyAsync(totals64, t64, (size_t) nArrays * sizeof(uint64_t), hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
	HIP_CHECK(hipFree(d));
	HIP_CHECK(hipFree(t32));
	HIP_CHECK(hipFree(t64));
	if (sc.buf) HIP_CHECK(hipFree(sc.buf));
	if (sc.buf64) HIP_CHECK(hipFree(sc.buf64));
}

/* Test hook (not part of the public headers): the stable radix sort the
 * draws use, on host data, with the runtime's own sort scratch.  keys and
 * vals hold cap items; the first count of them (count <= cap, read on the
 * device as the draws pass it) are sorted on the low `bits` bits of the
 * keys, in place. */
void srp_lib_radix_sort(uint32_t* keys, uint32_t* vals, size_t cap, size_t count, int bits)
{
	resolvePending();
	if (count > cap)
		count = cap;
	Runtime& R = rt();
	uint32_t* dk = nullptr;
	uint32_t* dv = nullptr;
	uint32_t* dn = nullptr;
	const size_t bytes = cap * sizeof(uint32_t);
	HIP_CHECK(hipMalloc((void**) &dk, bytes + 16));
	HIP_CHECK(hipMalloc((void**) &dv, bytes + 16));
	HIP_CHECK(hipMalloc((void**) &dn, sizeof(uint32_t)));
	const uint32_t count32 = (uint32_t) count;
	HIP_CHECK(hipMemcpyAsync(dk, keys, bytes, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(dv, vals, bytes, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(dn, &count32, sizeof(uint32_t), hipMemcpyHostToDevice, R.stream));
	uint32_t* ok = nullptr;
	uint32_t* ov = nullptr;
	radixSort(dk, dv, cap, bits, R.sort, &ok, &ov, R.stream, dn);
	HIP_CHECK(hipMemcpyAsync(keys, ok, count * sizeof(uint32_t), hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipMemcpyAsync(vals, ov, count * sizeof(uint32_t), hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
	HIP_CHECK(hipFree(dk));
	HIP_CHECK(hipFree(dv));
	HIP_CHECK(hipFree(dn));
}

void srpHipSetTileDepthCull(int enable)
{
	resolvePending();
	rt().depthCull = enable ? 1u : 0u;
}

void srpHipSetDeferredShading(int enable)
{
	resolvePending();
	rt().deferShading = enable ? 1u : 0u;
}
int srpHipGetDeferredShading(void) { return (int) rt().deferShading; }

void srpHipSetTileOrder(int run)
{
	resolvePending();
	if (run < 0 || run > 64)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, __func__, "Tile run %d is not in [0, 64]", run);
		return;
	}
	rt().tileOrder = (uint32_t) run;
}
int srpHipGetTileOrder(void) { return (int) rt().tileOrder; }
int srpHipGetTileDepthCull(void) { return (int) rt().depthCull; }

}  // extern "C"

/* =================================================================== draw */

namespace {

size_t elemSizeOf(uint32_t type)
{
	switch (type)
	{
		case SRP_FLOAT: case SRP_INT32: case SRP_UINT32: return 4;
		case SRP_DOUBLE: case SRP_INT64: case SRP_UINT64: return 8;
		case SRP_INT8: case SRP_UINT8: return 1;
		case SRP_INT16: case SRP_UINT16: return 2;
		default: return 0;
	}
}

int bitsFor(uint64_t n)
{
	int b = 0;
	while ((1ull << b) < n) b++;
	return b;
}

/* Varying layout of a vertex shader (interpolation.c:93-163): offsets,
 * the 4-byte word path when every element allows it.  False (with a
 * message) when the layout exceeds the device limits. */
bool varyingLayout(const SRPVertexShader* V, SrpDrawParams& H)
{
	uint32_t off = 0;
	bool words = (V->varyingsSize % 4) == 0;
	uint32_t nWords = 0;
	for (size_t a = 0; a < V->nVaryings; a++)
	{
		const SRPVaryingInfo& vi = V->varyingsInfo[a];
		const size_t es = elemSizeOf(vi.type);
		if (es == 0)
			message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "interpolateAttributes", "Unexpected type (%i)", vi.type);
		H.var[a].nItems = (uint32_t) vi.nItems;
		H.var[a].type = vi.type;
		H.var[a].mode = vi.interpolationMode;
		H.var[a].offset = off;
		const bool fourByte = vi.type == SRP_FLOAT || vi.type == SRP_INT32 || vi.type == SRP_UINT32;
		if (!fourByte || (off % 4) != 0)
			words = false;
		for (size_t e = 0; e < vi.nItems && words; e++)
		{
			if (nWords >= SRP_MAX_WORDS) { words = false; break; }
			uint8_t k = SRP_WORD_PROVOKING;
			if (vi.type == SRP_FLOAT && vi.interpolationMode == SRP_INTERPOLATION_MODE_PERSPECTIVE) k = SRP_WORD_PERSPECTIVE;
			else if (vi.type == SRP_FLOAT && vi.interpolationMode == SRP_INTERPOLATION_MODE_AFFINE) k = SRP_WORD_AFFINE;
			H.wordKind[nWords++] = k;
		}
		off += (uint32_t) (es * vi.nItems);
	}
	if (off > SRP_MAX_VARYING_BYTES)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "Varyings larger than %d bytes are not supported",
		        SRP_MAX_VARYING_BYTES);
		return false;
	}
	H.nVaryings = (uint32_t) V->nVaryings;
	H.describedSize = off;
	H.varyingsSize = (uint32_t) V->varyingsSize;
	H.wordPath = words ? 1u : 0u;
	H.nWords = words ? nWords : 0u;
	return true;
}

/* Render state captured when a draw is issued. */
bool snapshotState(const SRPContext* C, const SRPShaderProgram* sp, int kind, SrpState& S)
{
	S.provokingLast = C->provokingVertexMode == SRP_PROVOKING_VERTEX_LAST;
	S.cullFace = C->raster.cullFace;
	S.frontFace = C->raster.frontFace;
	S.polygonMode = C->raster.polygonMode;
	if (kind == 0 && S.polygonMode > SRP_POLYGON_MODE_POINT)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "assembleTrianglesGeneric",
		        "Unexpected srpContext.raster.polygonMode (%i)", C->raster.polygonMode);
		return false;
	}
	S.pointSize = C->raster.pointSize;
	S.scissorEnabled = C->scissor.enabled;
	S.sx0 = C->scissor.x;
	S.sx1 = C->scissor.x + C->scissor.width;
	S.sy0 = C->scissor.y;
	S.sy1 = C->scissor.y + C->scissor.height;
	S.stencilEnabled = C->stencil.enabled;
	S.stencil[0] = C->stencil.front;
	S.stencil[1] = C->stencil.back;
	S.depthTest = C->depth.testEnable;
	S.depthWrite = C->depth.writeEnable;
	S.depthOp = C->depth.compareOp;
	S.lateZ = sp->fs->mayOverwriteDepth;
	return true;
}

/* Buffers of a single draw's emitting half for `want` (the counts, or an
 * estimate of them), the fit check of the counts against those buffers,
 * and the half itself: emitting pass, span expansion, sort, tile ranges,
 * rasteriser or fragment lists.  Speculative (the counts are not known
 * yet): the fit check also copies the counters to the host, for
 * resolvePending, and want.spans or want.clipped of 0 leave out the span
 * and clipped-triangle kernels (the check then requires none). */
void enqueueEmit(Runtime& R, const FitCaps& want, bool speculative)
{
	Runtime::Pending& D = R.pending;
	AsmArgs A = D.A;
	uint32_t* totals = reinterpret_cast<uint32_t*>((uint8_t*) R.dParams + kTotalsOff);
	FitCaps caps = {kMaxDrawRecords, kMaxDrawVertexSlots - D.cacheSize, kMaxDrawItems, 0xFFFFFFFFull, want.clipped};
	auto take = [](DevBuf& b, unsigned long long n, size_t elem, size_t pad, unsigned long long& cap) {
		uint8_t* p = b.get<uint8_t>((size_t) n * elem + pad);
		cap = std::min<unsigned long long>(cap, (b.cap - pad) / elem);
		return p;
	};
	/* clip-generated vertices: the device parameters hold the buffer's
	 * address, so a new buffer is sent there */
	const size_t vs = D.varyingsSize;
	uint8_t* genBuf = R.genVaryings.get<uint8_t>((size_t) want.generated * vs + 512);
	if (vs)
		caps.generated = std::min<unsigned long long>(caps.generated, (R.genVaryings.cap - 512) / vs);
	if (genBuf != R.hParams->genVaryings)
	{
		/* the block's upload itself may still be queued */
		HIP_CHECK(hipStreamSynchronize(R.stream));
		R.hParams->genVaryings = genBuf;
		HIP_CHECK(hipMemcpyAsync(&R.dParams->genVaryings, &R.hParams->genVaryings, sizeof(genBuf),
		                         hipMemcpyHostToDevice, R.stream));
		R.sentParams.clear();
		HIP_CHECK(hipEventRecord(R.paramsCopied, R.stream));
		R.paramsPatched = true;
	}
	A.genVaryings = genBuf;
	uint32_t* tileStart = D.tileStart;   /* zeroed by the counting pass */
	uint32_t* tileEnd = tileStart ? tileStart + D.nTiles : nullptr;
	if (D.backend == 0)
	{
		A.tris = (SrpTri*) take(R.tris, want.records, sizeof(SrpTri), 0, caps.records);
		A.pairKeys = (uint32_t*) take(R.pairKeys, want.items, 4, 16, caps.items);
		A.pairVals = (uint32_t*) take(R.pairVals, want.items, 4, 16, caps.items);
		A.pairRec = (uint32_t*) take(R.pairRec, want.items, 4, 16, caps.items);
		A.pairMask = (unsigned long long*) take(R.pairMask, want.items, 8, 16, caps.items);
		A.spans = (SrpSpan*) take(R.spans, want.spans, sizeof(SrpSpan), 64, caps.spans);
	}
	else
	{
		if (D.backend == 1) A.lines = (SrpLine*) take(R.lines, want.records, sizeof(SrpLine), 0, caps.records);
		else A.points = (SrpPoint*) take(R.points, want.records, sizeof(SrpPoint), 0, caps.records);
		A.frags = (SrpFrag*) take(R.frags, want.items, sizeof(SrpFrag), 16, caps.items);
		A.fragKeys = (uint32_t*) take(R.fragKeys, want.items, 4, 16, caps.items);
		caps.spans = 0;
	}
	uint32_t* order = (D.backend != 0) ? (uint32_t*) take(R.order, want.items, 4, 16, caps.items) : nullptr;
	/* the sort, tile ranges, span and fragment kernels size their grids by these */
	caps.items = std::min(caps.items, want.items);
	caps.spans = std::min(caps.spans, want.spans);
	A.recCap = (uint32_t) caps.records;
	A.itemCap = (uint32_t) caps.items;
	A.genCap = (uint32_t) std::min<unsigned long long>(caps.generated, 0xFFFFFFFFull);
	A.spanCap = (uint32_t) caps.spans;
	A.spanCount = totals + 5;   /* emitting pass: append cursor (zeroed by the upload) */
	/* the emitting kernels check the totals against these capacities
	 * themselves; the leading one records the outcome and, for a
	 * speculative half, publishes the counters to the host */
	A.fitTotals = totals;
	A.fitCaps = caps;
	A.fitHost = speculative ? R.dTotals : nullptr;
	A.fitSeq = speculative ? ++R.countsSeq : 0u;
	uint32_t* items = totals + 2;   /* the item count, or 0 when skipped (emitGo) */

	const uint32_t clipped = speculative ? (caps.clipped ? kClipUnknown : 0u) : R.hTotals[4];
	if (D.kind == 0 && clipped != 0u)
	{
		/* the clipped triangles' emitting kernel runs beside the others'
		 * on a second stream (disjoint outputs; its few groups are latency
		 * bound), joined before the sort */
		HIP_CHECK(hipEventRecord(R.forkEv, R.stream));
		HIP_CHECK(hipStreamWaitEvent(R.side, R.forkEv, 0));
		launchClipped(A, true, R.side, clipped, D.clipHint);
		HIP_CHECK(hipEventRecord(R.joinEv, R.side));
		launchAssemble(A, D.kind, true, R.stream, 0u, 0u);
		HIP_CHECK(hipStreamWaitEvent(R.stream, R.joinEv, 0));
	}
	else
		launchAssemble(A, D.kind, true, R.stream, clipped, D.clipHint);
	if (D.backend == 0)
	{
		/* spans: the grid strides over the device-side count */
		launchExpandSpans(A, (uint32_t) std::min<unsigned long long>((caps.spans + 3) / 4, 2048ull), R.stream);
		uint32_t *keys, *vals;
		radixSort(A.pairKeys, A.pairVals, caps.items, bitsFor(D.nTiles), R.sort, &keys, &vals, R.stream, items);
		tileBounds(keys, caps.items, tileStart, tileEnd, R.stream, items);
		SrpRasterLaunch rl = D.rl;
		rl.tris = A.tris;
		rl.tileList = vals;
		rl.tileStart = tileStart;
		rl.tileEnd = tileEnd;
		rl.pairRec = A.pairRec;
		rl.pairMask = A.pairMask;
		if (D.fse->resolve && !D.lateZ && rl.batched && R.deferShading)
		{
			/* deferred shading: a list entry per pixel at most; the counter
			 * was zeroed by the parameter upload */
			rl.deferList = R.deferList.get<uint32_t>((size_t) D.nTiles * 64u * 8u);
			rl.deferCounts = R.deferCounts.get<uint32_t>(((size_t) D.nTiles + 1u) * 4u);
			/* the rebuilt fragments: 7 words plus the varyings' words */
			const size_t words = 7u + D.fragWords;
			rl.deferStride = D.nTiles * 64u;
			rl.deferFrags = R.deferFrags.get<uint32_t>(words * rl.deferStride * 4u);
		}
		hipEvent_t e0 = R.prof(0) ? R.mark() : nullptr;
		D.fse->raster(&rl, (void*) R.stream);
		if (rl.deferList)
		{
			scanExclusive(rl.deferCounts, D.nTiles, rl.deferCounts + D.nTiles, R.scan, R.stream);
			D.fse->resolve(&rl, (void*) R.stream);
		}
		if (e0) R.span(0, e0, R.mark());
	}
	else
	{
		iota(order, caps.items, R.stream);
		uint32_t *keys, *vals;
		radixSort(A.fragKeys, order, caps.items, D.fragBits, R.sort, &keys, &vals, R.stream, items);
		SrpFragLaunch fl;
		memset(&fl, 0, sizeof(fl));
		fl.params = R.dParams;
		fl.frags = A.frags;
		fl.keys = keys;
		fl.order = vals;
		fl.lines = A.lines;
		fl.points = A.points;
		fl.count = (uint32_t) caps.items;
		fl.countDev = items;
		fl.kind = (D.backend == 1) ? 0u : 1u;
		fl.lateZ = D.lateZ;
		fl.words = D.words;
		hipEvent_t e0 = R.prof(2) ? R.mark() : nullptr;
		D.fse->frags(&fl, (void*) R.stream);
		if (e0) R.span(2, e0, R.mark());
	}
	if (R.prof(1) && D.evDraw)
		R.span(1, D.evDraw, R.mark());
}

/* The counting pass of the pending draw (triangles: with its clipping
 * kernel unless clipSkipped) and the scans of its counts. */
void enqueueCount(Runtime& R, bool clipSkipped)
{
	Runtime::Pending& D = R.pending;
	uint32_t* totals = reinterpret_cast<uint32_t*>((uint8_t*) R.dParams + kTotalsOff);
	AsmArgs A = D.A;
	A.zero = D.tileStart;
	A.zeroCount = D.tileStart ? 2u * D.nTiles : 0u;
	launchAssemble(A, D.kind, false, R.stream, clipSkipped ? 0u : kClipUnknown, D.clipHint);
	scanExclusiveMulti(A.recOff, (size_t) D.nIn + 1, 3, D.nIn, totals, R.scan, R.stream,
	                   reinterpret_cast<unsigned long long*>(totals + 8));
}

/* Publish the pending draw's counters to the host behind the work enqueued
 * so far (waitCounts waits for them). */
void copyCounts(Runtime& R)
{
	uint32_t* totals = reinterpret_cast<uint32_t*>((uint8_t*) R.dParams + kTotalsOff);
	launchPublish(totals, R.dTotals, ++R.countsSeq, R.stream);
}

/* Wait until the counters of publication R.countsSeq are on the host: the
 * publishing kernel writes them, then the number (no event, so the stream
 * runs on without a barrier).  Should the stream drain without it, which
 * would be a library fault, the process stops with a message rather than
 * waiting for ever. */
void waitCounts(Runtime& R)
{
	volatile uint32_t* seq = R.hTotals + 16;
	for (uint32_t spin = 0;; spin++)
	{
		if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == R.countsSeq)
			return;
		if ((spin & 63u) == 63u)
		{
			const hipError_t q = hipStreamQuery(R.stream);
			if (q == hipSuccess && __atomic_load_n(seq, __ATOMIC_ACQUIRE) != R.countsSeq)
			{
				fprintf(stderr, "libsrp_amd: draw counters %u were not published (last %u)\n", R.countsSeq, *seq);
				abort();
			}
			if (q != hipSuccess && q != hipErrorNotReady)
				HIP_CHECK(q);
			sched_yield();
		}
	}
}

/* Settle the pending draw, if any: wait for its counts (not for the draw),
 * report what they show as the synchronous path would have (clip overflow,
 * a draw over the limits: refused, nothing drawn), and when the buffers it
 * was given were too small, enqueue its emitting half again with buffers
 * for its counts (the first attempt was skipped on the device).  A draw
 * whose clipping kernel was left out but that clips triangles is counted
 * again, with it, first. */
void resolvePending()
{
	if (!g_rt || !g_rt->pending.active)
		return;
	Runtime& R = *g_rt;
	Runtime::Pending& D = R.pending;
	D.active = false;
	waitCounts(R);
	const uint32_t* h = R.hTotals;
	const uint64_t* h64 = reinterpret_cast<const uint64_t*>(h + 8);
	if (D.clipSkipped && h[4] != 0u)
	{
		/* the emitting half did nothing (emitGo): count again, with the
		 * clipping kernel, and wait for those counts */
		uint32_t* totals = reinterpret_cast<uint32_t*>((uint8_t*) R.dParams + kTotalsOff);
		fillU32(totals + 4, 0u, 3, R.stream);   /* clip list, span cursor, span count */
		D.clipHint = h[4];
		D.clipSkipped = false;
		D.emitted = false;
		enqueueCount(R, false);
		copyCounts(R);
		waitCounts(R);
	}
	if (h[3] & 1u)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "clipTriangle",
		        "clipped polygon exceeded %d vertices; the primitive was dropped", SRP_MAX_POLY);
	const DrawTotals dt = {h64[0], h64[1], h64[2], D.cacheSize};
	const DrawLimit lim = checkDrawTotals(dt);
	if (lim != DRAW_FITS)
	{
		char text[256];
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "%s", drawLimitText(lim, dt, text, sizeof(text)));
		return;   /* skipped on the device: nothing drawn */
	}
	if (D.kind == 0)
		R.lastClipped = h[4];
	R.hostDraws++;
	R.hostPrims += h64[0];
	const FitCaps need = {h64[0], h64[1], h64[2], h[6], h[4]};
	if (R.drawCounts.size() >= 4096)
		R.drawCounts.clear();
	R.drawCounts[D.key] = need;
	if (h64[0] != 0u && (!D.emitted || h[7] == 1u))
		enqueueEmit(R, need, false);
}

void drawBuffer(const SRPIndexBuffer* ib, const SRPVertexBuffer* vb, const SRPFramebuffer* fb,
                const SRPShaderProgram* sp, SRPPrimitive prim, size_t start, size_t count)
{
	resolvePending();
	SRPContext* C = ctx();
	/* pipeline/draw.c:63-82, :167-189 */
	if (count == 0)
		return;
	const size_t end = start + count - 1;
	const size_t size = ib ? ib->nIndices : vb->nVertices;
	if (end >= size)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "checkOOB",
		        ib ? "Attempt to OOB access index buffer (read) at indices %zu-%zu (size: %zu)\n"
		           : "Attempt to OOB access vertex buffer (read) at indices %zu-%zu (size: %zu)\n",
		        start, end, size);
		return;
	}
	int kind;
	if (prim == SRP_PRIM_TRIANGLES || prim == SRP_PRIM_TRIANGLE_STRIP || prim == SRP_PRIM_TRIANGLE_FAN) kind = 0;
	else if (prim == SRP_PRIM_LINES || prim == SRP_PRIM_LINE_STRIP || prim == SRP_PRIM_LINE_LOOP) kind = 1;
	else if (prim == SRP_PRIM_POINTS) kind = 2;
	else
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "Unknown primitive type: %i", prim);
		return;
	}

	size_t nIn = 0;
	if (kind == 0)
	{
		/* draw.c:89-90 */
		if (C->raster.cullFace == SRP_FACE_FRONT_AND_BACK)
			return;
		if (prim == SRP_PRIM_TRIANGLES && count % 3 != 0)
			message(SRP_MESSAGE_WARNING, SRP_MESSAGE_SEVERITY_LOW, "warnOnExcessVertexCount",
			        "Vertex count not divisible by 3 when drawing SRP_PRIM_TRIANGLES. The last %i vertex/vertices will be ignored\n",
			        (int) (count % 3));
		nIn = (prim == SRP_PRIM_TRIANGLES) ? count / 3 : (count >= 3 ? count - 2 : 0);
	}
	else if (kind == 1)
	{
		if (prim == SRP_PRIM_LINES && count % 2 != 0)
			message(SRP_MESSAGE_WARNING, SRP_MESSAGE_SEVERITY_LOW, "warnOnExcessVertexCount",
			        "Odd vertex count when drawing SRP_PRIM_LINES. The last vertex will be ignored\n");
		if (prim == SRP_PRIM_LINES) nIn = count / 2;
		else if (prim == SRP_PRIM_LINE_STRIP) nIn = (count != 1) ? count - 1 : 0;
		else nIn = (count != 1) ? count : 0;
	}
	else
	{
		if (C->raster.pointSize <= 0.)
			return;
		nIn = count;
	}
	if (nIn == 0)
		return;

	const SrpVsEntry* vse = nullptr;
	const SrpFsEntry* fse = nullptr;
	{
		std::lock_guard<std::mutex> g(g_regLock);
		auto iv = vsRegistry().find((const void*) sp->vs->shader);
		auto ifs = fsRegistry().find((const void*) sp->fs->shader);
		if (iv != vsRegistry().end()) vse = iv->second;
		if (ifs != fsRegistry().end()) fse = ifs->second;
	}
	if (!vse || !fse)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer",
		        "%s shader is not a registered device shader (see SRP_HIP_*_SHADER in srp/srp_device.h)",
		        vse ? "fragment" : "vertex");
		return;
	}

	Runtime& R = rt();
	FbImpl* F = fbImpl(fb);
	if (fb->width > 65535 || fb->height > 65535 || fb->size >= 0xFFFFFFFFull)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "Framebuffer too large for the GPU path");
		return;
	}

	/* Varying layout (interpolation.c:93-163).  The staging block is free:
	 * the previous draw's counters, published after its uploads, have been
	 * waited for (resolvePending), unless a patch followed them. */
	if (R.paramsPatched)
	{
		HIP_CHECK(hipEventSynchronize(R.paramsCopied));
		R.paramsPatched = false;
	}
	SrpDrawParams& H = *R.hParams;
	memset(&H, 0, sizeof(H));
	const SRPVertexShader* V = sp->vs;
	if (V->nVaryings > SRP_MAX_VARYINGS)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "At most %d varyings are supported", SRP_MAX_VARYINGS);
		return;
	}
	if (!varyingLayout(V, H))
		return;

	/* Post-VS cache range (vertex_processing.c:23-36). */
	uint64_t lo, hi;
	if (ib)
	{
		indexRange(ib, start, count, lo, hi);
		if (hi >= vb->nVertices)
		{
			message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer",
			        "Index %llu exceeds the vertex buffer (%zu vertices)", (unsigned long long) hi, vb->nVertices);
			return;
		}
	}
	else
	{
		lo = start;
		hi = start + count - 1;
	}
	if (hi - lo + 1 > 0x7FFFFFFFull)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "drawBuffer", "Vertex index range too large");
		return;
	}
	const uint32_t cacheSize = (uint32_t) (hi - lo + 1);

	/* State snapshot. */
	SrpState& S = H.st;
	if (!snapshotState(C, sp, kind, S))
		return;

	H.width = (uint32_t) fb->width;
	H.height = (uint32_t) fb->height;
	H.color = F->dColor;
	H.depth = F->dDepth;
	H.stencil = F->dStencil;

	const size_t usize = vse->uniformSize > fse->uniformSize ? vse->uniformSize : fse->uniformSize;
	size_t paramsBytes = kTotalsOff + kTotalsBytes;
	memset((uint8_t*) R.hParams + kTotalsOff, 0, kTotalsBytes);
	if (sp->uniform && usize)
	{
		if (usize <= kMaxInlineUniform)
		{
			memcpy((uint8_t*) R.hParams + kParamsHead, sp->uniform, usize);
			H.uniform = (const uint8_t*) R.dParams + kParamsHead;
			paramsBytes = kParamsHead + usize;
		}
		else
		{
			if (usize > R.uniformCap)
			{
				if (R.hUniform) HIP_CHECK(hipHostFree(R.hUniform));
				R.uniformCap = usize * 2;
				HIP_CHECK(hipHostMalloc((void**) &R.hUniform, R.uniformCap, hipHostMallocDefault));
			}
			memcpy(R.hUniform, sp->uniform, usize);
			uint8_t* dU = R.uniform.get<uint8_t>(usize + 16);
			HIP_CHECK(hipMemcpyAsync(dU, R.hUniform, usize, hipMemcpyHostToDevice, R.stream));
			H.uniform = dU;
		}
	}
	else
		H.uniform = nullptr;

	H.vertices = vb->dData;
	H.vertexStride = vb->nBytesPerVertex;
	H.cacheBase = lo;
	H.cacheSize = cacheSize;
	H.clip = R.clip.get<float>((size_t) cacheSize * 16);
	H.varyings = R.varyings.get<uint8_t>((size_t) cacheSize * V->varyingsSize + 512);
	/* the buffer of clip-generated vertices as it stands; re-sent after the
	 * counting pass only if that pass needs it larger */
	H.genVaryings = R.genVaryings.get<uint8_t>(512);
	H.tileShift = (R.tile == 16) ? 4u : 3u;
	H.maskBox = R.maskBox;
	H.tilesX = (H.width + R.tile - 1) / R.tile;
	H.tilesY = (H.height + R.tile - 1) / R.tile;
	H.stats = R.dStats;

	/* a deferred framebuffer's pending fill rides along with the vertex
	 * stage below (nothing before the rasteriser touches the framebuffer) */
	if (!F->deferred)
		fbUpload(F);
	hipEvent_t evDraw = (R.prof(1) || R.prof(3)) ? R.mark() : nullptr;
	{
		AsmDraw& AD = *reinterpret_cast<AsmDraw*>((uint8_t*) R.hParams + kAsmDrawOff);
		AD.P = R.dParams;
		AD.indices = ib ? ib->dData : nullptr;
		AD.start = start;
		AD.count = count;
		AD.prim = (uint32_t) prim;
		AD.nIn = (uint32_t) nIn;
		AD.indexElem = ib ? (uint32_t) ib->elem : 0u;
		AD.primBase = 0;
		AD.tileBase = 0;
		AD.firstBlock = 0;
	}
	/* The parameters go to the device unless it holds the same already (a
	 * draw repeated with the same state, uniform, buffers and framebuffer);
	 * the counters the upload would have zeroed are then cleared by the
	 * vertex stage. */
	const uint8_t* hp = reinterpret_cast<const uint8_t*>(R.hParams);
	const size_t uniformBytes = paramsBytes > kParamsHead ? paramsBytes - kParamsHead : 0;
	const bool resend = R.sentParams.size() != kTotalsOff + uniformBytes ||
	                    memcmp(R.sentParams.data(), hp, kTotalsOff) != 0 ||
	                    (uniformBytes && memcmp(R.sentParams.data() + kTotalsOff, hp + kParamsHead, uniformBytes) != 0);
	if (resend)
	{
		HIP_CHECK(hipMemcpyAsync(R.dParams, &H, paramsBytes, hipMemcpyHostToDevice, R.stream));
		R.sentParams.assign(hp, hp + kTotalsOff);
		R.sentParams.insert(R.sentParams.end(), hp + kParamsHead, hp + kParamsHead + uniformBytes);
	}

	/* Vertex stage (single draw: the multi-draw fields stay zero). */
	SrpVsLaunch vl;
	memset(&vl, 0, sizeof(vl));
	vl.params = R.dParams;
	vl.count = cacheSize;
	if (!resend)
	{
		vl.zero = reinterpret_cast<uint32_t*>((uint8_t*) R.dParams + kTotalsOff);
		vl.zeroCount = (uint32_t) (kTotalsBytes / sizeof(uint32_t));
	}
	if (F->deferred && (F->pendColor || F->pendDepth || F->pendStencil))
	{
		vl.fillColor = F->pendColor ? F->dColor : nullptr;
		vl.fillDepth = F->pendDepth ? (uint32_t*) F->dDepth : nullptr;
		vl.fillStencil = F->pendStencil ? F->dStencil : nullptr;
		vl.fillColorValue = F->fillColor;
		vl.fillDepthValue = F->fillDepthBits;
		vl.fillStencilValue = F->fillStencil;
		vl.fillCount = F->pub.size;
		vl.fillBlocks = (uint32_t) std::min<size_t>((F->pub.size + 1023) / 1024, 2048);
		F->pendColor = F->pendDepth = F->pendStencil = false;
	}
	vse->launch(&vl, (void*) R.stream);

	/* Assembly COUNT pass + scans. */
	uint32_t* counts = R.counts.get<uint32_t>((nIn + 1) * 4 * sizeof(uint32_t));
	uint32_t* totals = reinterpret_cast<uint32_t*>((uint8_t*) R.dParams + kTotalsOff);   /* zeroed by the upload or the vertex stage */
	Runtime::Pending& D = R.pending;
	AsmArgs& A = D.A;
	memset(&A, 0, sizeof(A));
	A.draws = reinterpret_cast<const AsmDraw*>((const uint8_t*) R.dParams + kAsmDrawOff);
	A.nDraws = 1;
	A.nBlocks = (uint32_t) ((nIn + 127) / 128);
	A.recOff = counts;
	A.genOff = counts + (nIn + 1);
	A.itemOff = counts + 2 * (nIn + 1);
	A.clipList = counts + 3 * (nIn + 1);
	A.error = totals + 3;
	A.clipCount = totals + 4;
	A.spanCount = totals + 6;   /* counting pass: how many spans */
	A.recCap = A.itemCap = A.genCap = 0xFFFFFFFFu;
	D.kind = kind;
	D.backend = (kind == 0) ? (int) S.polygonMode : (kind == 1 ? 1 : 2);  /* 0 tiles, 1 lines, 2 points */
	D.nIn = (uint32_t) nIn;
	D.nTiles = H.tilesX * H.tilesY;
	D.cacheSize = cacheSize;
	D.varyingsSize = V->varyingsSize;
	D.lateZ = S.lateZ;
	D.words = H.wordPath;
	D.fragWords = H.wordPath ? H.nWords : (H.describedSize + 3u) / 4u;
	D.fragBits = bitsFor((uint64_t) H.width * H.height + 1);
	D.fse = fse;
	D.evDraw = evDraw;
	D.tileStart = (D.backend == 0) ? R.tileCount.get<uint32_t>(((size_t) 2 * D.nTiles + 2) * 4) : nullptr;
	/* A draw seen before: its emitting half goes in behind the counts at
	 * once, with buffers for a quarter more than its last counts, and
	 * without the span or clipping kernels when it had no spans or clipped
	 * nothing; the host checks the counts at its next library call (here,
	 * for a synchronous framebuffer), and in the rare draw that did not fit
	 * (the device skipped the attempt) repeats the half for its counts.  A
	 * new draw waits for its counts. */
	D.key = {ib, vb, start, count, (uint32_t) prim, H.width, H.height};
	auto seen = R.drawCounts.find(D.key);
	D.emitted = seen != R.drawCounts.end();
	FitCaps last = {0, 0, 0, 0, 0};
	if (D.emitted)
		last = seen->second;
	D.clipHint = D.emitted ? (uint32_t) last.clipped : R.lastClipped;
	D.clipSkipped = D.emitted && kind == 0 && last.clipped == 0;
	enqueueCount(R, D.clipSkipped);
	if (R.prof(3) && evDraw) R.span(3, evDraw, R.mark());

	memset(&D.rl, 0, sizeof(D.rl));
	D.rl.params = R.dParams;
	D.rl.nTiles = D.nTiles;
	D.rl.lateZ = S.lateZ;
	D.rl.words = H.wordPath;
	D.rl.tile = (uint32_t) R.tile;
	D.rl.batched = (R.batched && R.tile == 8) ? 1u : 0u;
	D.rl.bigPixel = R.bigPixel;
	/* culling kernels only where culling can act (the kernel checks the
	 * same conditions; the other kernels need fewer registers) */
	const bool monotone = S.depthOp == SRP_COMPARE_LESS || S.depthOp == SRP_COMPARE_LEQUAL ||
	                      S.depthOp == SRP_COMPARE_GREATER || S.depthOp == SRP_COMPARE_GEQUAL ||
	                      S.depthOp == SRP_COMPARE_NEVER;
	D.rl.depthCull = (R.depthCull && S.depthTest && monotone && !S.stencilEnabled && !S.lateZ) ? 1u : 0u;
	const uint32_t perBlock = D.rl.batched ? 1u : ((R.tile == 8) ? 4u : 1u);
	D.rl.nBlocks = (D.nTiles + perBlock - 1) / perBlock;
	D.rl.tileOrder = R.tileOrder;
	if (D.emitted)
	{
		typedef unsigned long long u64;
		const FitCaps want = {std::min<u64>(last.records + last.records / 4 + 64, kMaxDrawRecords),
		                      std::min<u64>(last.generated + last.generated / 4 + 64, kMaxDrawVertexSlots - cacheSize),
		                      std::min<u64>(last.items + last.items / 4 + 4096, kMaxDrawItems),
		                      last.spans ? last.spans + last.spans / 4 + 64 : 0ull,
		                      last.clipped ? ~0ull : 0ull};
		enqueueEmit(R, want, true);
	}
	else
		copyCounts(R);
	D.active = true;
	if (!F->deferred || !D.emitted)
		resolvePending();
	if (!F->deferred)
		fbDownload(F);
}

}  // namespace

extern "C" {

/* core/buffer.c:53-59 */
void srpDrawVertexBuffer(const SRPVertexBuffer* vb, const SRPFramebuffer* fb, const SRPShaderProgram* sp,
                         SRPPrimitive prim, size_t start, size_t count)
{
	drawBuffer(nullptr, vb, fb, sp, prim, start, count);
}

/* core/buffer.c:130-136 */
void srpDrawIndexBuffer(const SRPIndexBuffer* ib, const SRPVertexBuffer* vb, const SRPFramebuffer* fb,
                        const SRPShaderProgram* sp, SRPPrimitive prim, size_t start, size_t count)
{
	drawBuffer(ib, vb, fb, sp, prim, start, count);
}

}  // extern "C"

/* ====================================================================== batches
 *
 * A batch renders many independent triangle draws as one frame: one vertex
 * launch over every draw's cache slots, one counting and one emitting
 * assembly launch over every draw's primitives (blocks split by draw), one
 * multi-array scan, one radix sort of the (tile, triangle) pairs of all
 * framebuffers (tiles numbered globally), and one persistent rasteriser
 * launch whose waves take tiles from device queues.  Buffers are sized by
 * a counting pass when the batch is made, so a frame needs no host
 * synchronisation and can be captured into a HIP graph. */

struct SRPBatch
{
	std::vector<SRPDrawCommand> src;         /* every command as given; contexts point into ctxCopy */
	std::vector<SRPContext> ctxCopy;         /* each draw's render state when the batch was made */
	/* objects the batch reads: checked live (and unchanged, else the batch
	 * is prepared again) before each frame */
	struct Watch { const void* obj; int kind; size_t draw; uint64_t version; };
	std::vector<Watch> watch;
	uint32_t* hErr = nullptr;                /* host-mapped: error bits of the last frame */
	uint32_t* dErr = nullptr;                /* its device address */
	uint32_t reported = 0;                   /* error bits already reported */
	std::vector<SRPDrawCommand> cmds;        /* drawn (not culled away) */
	struct Clear { FbImpl* F; uint32_t depthBits; uint8_t stencil; };
	std::vector<Clear> clears;
	std::vector<FbImpl*> fbs;                /* every framebuffer touched */
	const SrpVsEntry* vse = nullptr;
	const SrpFsEntry* fse = nullptr;
	const SRPShaderProgram* sp = nullptr;
	size_t usize = 0, stride = 0;
	uint8_t* hBlocks = nullptr;              /* pinned parameter blocks */
	uint8_t* dBlocks = nullptr;
	std::vector<AsmDraw> hDraws;
	AsmDraw* dDraws = nullptr;
	std::vector<SrpSeg> hVsSeg, hRasterSeg;
	SrpSeg* dVsSeg = nullptr;
	SrpSeg* dRasterSeg = nullptr;
	uint32_t nIn = 0, inBlocks = 0, vsBlocks = 0, totalTiles = 0;
	uint32_t lateZ = 0, words = 0;
	float* clip = nullptr;
	uint8_t* varyings = nullptr;
	uint32_t* counts = nullptr;
	uint32_t* totals = nullptr;
	uint32_t* hTotals = nullptr;             /* pinned */
	SrpTri* tris = nullptr;
	SrpSpan* spans = nullptr;                /* recCap entries */
	uint32_t *pairKeys = nullptr, *pairVals = nullptr, *pairRec = nullptr;
	unsigned long long* pairMask = nullptr;
	uint8_t* genVaryings = nullptr;
	uint32_t recCap = 0, itemCap = 0, genCap = 0;
	uint32_t nRec = 0, nItems = 0, nGen = 0;
	uint32_t nClipped = 0;                   /* triangles clipped at sizing (grid hint) */
	uint32_t* tileRange = nullptr;           /* start[totalTiles], end[totalTiles] */
	uint32_t* queues = nullptr;
	FillDesc* dFills = nullptr;              /* the clears as one fill launch */
	uint32_t fillBlocks = 0;
	ScanScratch scan;
	SortScratch sort;
	SRPBatchTuning tuning = {8, SRP_SMALL_BBOX, 1, 1, 1, 8};
	int cus = 256;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

namespace {

void batchFreeDev(void*& p)
{
	if (p)
	{
		HIP_CHECK(hipFree(p));
		p = nullptr;
	}
}

void batchDropGraph(SRPBatch* B)
{
	if (B->exec) HIP_CHECK(hipGraphExecDestroy(B->exec));
	if (B->graph) HIP_CHECK(hipGraphDestroy(B->graph));
	B->exec = nullptr;
	B->graph = nullptr;
}

SrpDrawParams* batchParams(SRPBatch* B, size_t d) { return (SrpDrawParams*) (B->hBlocks + d * B->stride); }

void batchUpload(SRPBatch* B)
{
	Runtime& R = rt();
	const size_t nd = B->cmds.size();
	for (size_t d = 0; d < nd; d++)
	{
		batchParams(B, d)->maskBox = (uint32_t) B->tuning.maskBox;
		batchParams(B, d)->genVaryings = B->genVaryings;
	}
	HIP_CHECK(hipMemcpyAsync(B->dBlocks, B->hBlocks, nd * B->stride, hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(B->dDraws, B->hDraws.data(), nd * sizeof(AsmDraw), hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(B->dVsSeg, B->hVsSeg.data(), nd * sizeof(SrpSeg), hipMemcpyHostToDevice, R.stream));
	HIP_CHECK(hipMemcpyAsync(B->dRasterSeg, B->hRasterSeg.data(), nd * sizeof(SrpSeg), hipMemcpyHostToDevice, R.stream));
	/* the clears, as one fill launch over every cleared framebuffer */
	std::vector<FillDesc> fills;
	uint32_t blocks = 0;
	for (const SRPBatch::Clear& c : B->clears)
	{
		FillDesc f;
		f.color = c.F->dColor;
		f.depth = (uint32_t*) c.F->dDepth;
		f.stencil = c.F->dStencil;
		f.colorValue = 0u;
		f.depthValue = c.depthBits;
		f.stencilValue = c.stencil;
		f.firstBlock = blocks;
		f.n = c.F->pub.size;
		blocks += fillBlocksFor(f.n);
		fills.push_back(f);
	}
	if (!B->dFills && !fills.empty())
		HIP_CHECK(hipMalloc((void**) &B->dFills, B->clears.size() * sizeof(FillDesc)));
	if (!fills.empty())
		HIP_CHECK(hipMemcpyAsync(B->dFills, fills.data(), fills.size() * sizeof(FillDesc), hipMemcpyHostToDevice,
		                         R.stream));
	B->fillBlocks = blocks;
	HIP_CHECK(hipStreamSynchronize(R.stream));
}

/* Tile side of the batch's draws from its tuning: the draws' tile grids,
 * their ranges of the batch's tiles and the tile-range buffer (the caller
 * sizes and uploads the batch afterwards). */
void batchRetile(SRPBatch* B)
{
	const uint32_t side = (B->tuning.tileSize == 16) ? 16u : 8u;
	uint32_t tileBase = 0;
	for (size_t d = 0; d < B->cmds.size(); d++)
	{
		SrpDrawParams* H = batchParams(B, d);
		H->tileShift = (side == 16u) ? 4u : 3u;
		H->tilesX = (H->width + side - 1u) / side;
		H->tilesY = (H->height + side - 1u) / side;
		const uint32_t nTiles = H->tilesX * H->tilesY;
		B->hDraws[d].tileBase = tileBase;
		B->hRasterSeg[d].first = tileBase;
		B->hRasterSeg[d].count = nTiles;
		tileBase += nTiles;
	}
	if (tileBase != B->totalTiles)
	{
		HIP_CHECK(hipDeviceSynchronize());
		if (B->tileRange) HIP_CHECK(hipFree(B->tileRange));
		HIP_CHECK(hipMalloc((void**) &B->tileRange, ((size_t) 2 * tileBase + 2) * sizeof(uint32_t)));
	}
	B->totalTiles = tileBase;
}

AsmArgs batchAsmArgs(SRPBatch* B, bool emit)
{
	AsmArgs A;
	memset(&A, 0, sizeof(A));
	A.draws = B->dDraws;
	A.nDraws = (uint32_t) B->cmds.size();
	A.nBlocks = B->inBlocks;
	A.recOff = B->counts;
	A.genOff = B->counts + (B->nIn + 1);
	A.itemOff = B->counts + 2 * ((size_t) B->nIn + 1);
	A.clipList = B->counts + 3 * ((size_t) B->nIn + 1);
	A.clipCount = B->totals + 4;
	A.tris = B->tris;
	A.pairKeys = B->pairKeys;
	A.pairVals = B->pairVals;
	A.pairRec = B->pairRec;
	A.pairMask = B->pairMask;
	A.genVaryings = B->genVaryings;
	A.error = B->totals + 3;
	A.recCap = B->recCap;
	A.itemCap = B->itemCap;
	A.genCap = B->genCap;
	A.spans = B->spans;
	A.spanCap = B->recCap;
	A.spanCount = B->totals + (emit ? 5 : 6);
	return A;
}

/* Vertex stage + counting pass + scans (device totals in B->totals). */
void batchFront(SRPBatch* B, hipStream_t s)
{
	SrpVsLaunch vl;
	memset(&vl, 0, sizeof(vl));
	vl.seg = B->dVsSeg;
	vl.nSeg = (uint32_t) B->cmds.size();
	vl.nBlocks = B->vsBlocks;
	B->vse->launch(&vl, (void*) s);
	fillU32(B->totals, 0u, kTotalsBytes / 4, s);
	AsmArgs A = batchAsmArgs(B, false);
	launchAssemble(A, 0, false, s, kClipUnknown, B->nClipped);
	scanExclusiveMulti(A.recOff, (size_t) B->nIn + 1, 3, B->nIn, B->totals, B->scan, s,
	                   reinterpret_cast<unsigned long long*>(B->totals + 8));
}

/* Counting pass on the device, totals to the host; buffers (re)sized.
 * False (with a message) when the batch exceeds what one frame can hold. */
bool batchSize(SRPBatch* B, const char* caller)
{
	Runtime& R = rt();
	batchUpload(B);
	batchFront(B, R.stream);
	HIP_CHECK(hipMemcpyAsync(B->hTotals, B->totals, kTotalsBytes, hipMemcpyDeviceToHost, R.stream));
	HIP_CHECK(hipStreamSynchronize(R.stream));
	const uint64_t* t64 = reinterpret_cast<const uint64_t*>(B->hTotals + 8);
	if (B->hTotals[3] & 1u)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "clipTriangle",
		        "clipped polygon exceeded %d vertices; the primitive was dropped", SRP_MAX_POLY);
	uint64_t slots = 0;
	for (const SrpSeg& sg : B->hVsSeg)
		slots += sg.count;
	/* with headroom for the capacities below */
	const DrawTotals dt = {t64[0] + 64, t64[1] + 64, t64[2] + 256 + (t64[2] + 256) / 8, slots};
	const DrawLimit lim = checkDrawTotals(dt);
	if (lim != DRAW_FITS)
	{
		char text[256];
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, caller, "%s", drawLimitText(lim, dt, text, sizeof(text)));
		return false;
	}
	B->nRec = (uint32_t) t64[0];
	B->nGen = (uint32_t) t64[1];
	B->nItems = (uint32_t) t64[2];
	B->nClipped = B->hTotals[4];
	const uint32_t recCap = B->nRec + 64, genCap = B->nGen + 64, itemCap = B->nItems + 256;
	/* the scan's and sort's scratch too: frames allocate nothing */
	scanReserve(B->scan, B->nIn, 3);
	if (recCap > B->recCap || genCap > B->genCap || itemCap > B->itemCap)
	{
		HIP_CHECK(hipDeviceSynchronize());
		void* bufs[] = {B->tris, B->spans, B->pairKeys, B->pairVals, B->pairRec, B->pairMask, B->genVaryings};
		for (void* p : bufs)
			if (p) HIP_CHECK(hipFree(p));
		/* headroom for frames whose data grow; never past the per-call limits */
		auto grow = [](uint64_t c, uint64_t lim) { return (uint32_t) std::min<uint64_t>(c + c / 8, lim); };
		B->recCap = grow(recCap, kMaxDrawRecords);
		B->genCap = grow(genCap, kMaxDrawVertexSlots);
		B->itemCap = grow(itemCap, kMaxDrawItems);
		HIP_CHECK(hipMalloc((void**) &B->tris, (size_t) B->recCap * sizeof(SrpTri)));
		HIP_CHECK(hipMalloc((void**) &B->spans, (size_t) B->recCap * sizeof(SrpSpan)));
		HIP_CHECK(hipMalloc((void**) &B->pairKeys, (size_t) B->itemCap * 4));
		HIP_CHECK(hipMalloc((void**) &B->pairVals, (size_t) B->itemCap * 4));
		HIP_CHECK(hipMalloc((void**) &B->pairRec, (size_t) B->itemCap * 4));
		HIP_CHECK(hipMalloc((void**) &B->pairMask, (size_t) B->itemCap * 8));
		HIP_CHECK(hipMalloc((void**) &B->genVaryings, (size_t) B->genCap * B->sp->vs->varyingsSize + 512));
		batchUpload(B);
	}
	radixSortReserve(B->sort, B->itemCap);
	return true;
}

/* One frame of launches (capturable: fixed grids, no allocation, no host
 * synchronisation, fills by kernels). */
void batchLaunch(SRPBatch* B, hipStream_t s)
{
	fillFramebuffers(B->dFills, (uint32_t) B->clears.size(), B->fillBlocks, s);
	batchFront(B, s);
	AsmArgs A = batchAsmArgs(B, true);
	launchAssemble(A, 0, true, s, kClipUnknown, B->nClipped);
	launchExpandSpans(A, std::min<uint32_t>((B->recCap + 3u) / 4u, 2048u), s);
	uint32_t* tileStart = B->tileRange;
	uint32_t* tileEnd = B->tileRange + B->totalTiles;
	fillU32(B->tileRange, 0u, (size_t) 2 * B->totalTiles, s);
	uint32_t *keys, *vals;
	radixSort(B->pairKeys, B->pairVals, B->itemCap, bitsFor(B->totalTiles), B->sort, &keys, &vals, s, B->totals + 2);
	tileBounds(keys, B->itemCap, tileStart, tileEnd, s, B->totals + 2);
	fillU32(B->queues, 0u, 8, s);
	SrpRasterLaunch rl;
	memset(&rl, 0, sizeof(rl));
	rl.params = nullptr;
	rl.tris = B->tris;
	rl.tileList = vals;
	rl.tileStart = tileStart;
	rl.tileEnd = tileEnd;
	rl.pairRec = B->pairRec;
	rl.pairMask = B->pairMask;
	rl.nTiles = B->totalTiles;
	rl.lateZ = B->lateZ;
	rl.words = B->words;
	rl.tile = (B->tuning.tileSize == 16) ? 16u : 8u;
	rl.batched = (rl.tile == 8u) ? 1u : 0u;   /* 8x8: a wave a tile; 16x16: four (a block) */
	rl.bigPixel = B->tuning.pixelRuns ? 1u : 0u;
	rl.depthCull = B->tuning.depthCull ? 1u : 0u;
	const uint32_t perTile = (rl.tile == 16u) ? 4u : 1u;
	uint32_t workers = (uint32_t) (B->cus * B->tuning.wavesPerCU) / perTile;
	if (workers > B->totalTiles) workers = B->totalTiles;
	rl.nBlocks = workers ? workers : 1u;
	rl.seg = B->dRasterSeg;
	rl.nSeg = (uint32_t) B->cmds.size();
	rl.xcdQueues = B->tuning.xcdQueues ? 1u : 0u;
	rl.queues = B->queues;
	B->fse->raster(&rl, (void*) s);
	/* the frame's error bits, where the host reads them without waiting */
	copyU32(B->totals + 3, B->dErr, s);
}

/* Report error bits raised by frames since the last report. */
void batchReport(SRPBatch* B)
{
	const uint32_t e = __atomic_load_n(B->hErr, __ATOMIC_ACQUIRE);
	const uint32_t fresh = e & ~B->reported;
	if (fresh & 1u)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "clipTriangle",
		        "clipped polygon exceeded %d vertices; the primitive was dropped", SRP_MAX_POLY);
	if (fresh & 2u)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpBatchRender",
		        "A frame exceeded the batch's record, tile-entry or clip-vertex capacity; its excess primitives "
		        "were dropped");
	if (fresh & 4u)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpBatchRender",
		        "A frame met vertex indices outside the range the batch was prepared for; those primitives were "
		        "dropped");
	B->reported = e;
}

/* Apply deferred fills recorded on the batch's framebuffers outside it. */
void batchFlushPending(SRPBatch* B)
{
	for (FbImpl* F : B->fbs)
		fbFlushFills(F);
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

/* Validate and prepare a batch of n commands (srpNewBatch; also when a
 * batch's buffers changed).  Each command's render state is copied (its
 * context, or the current one), so the batch can be prepared again later
 * with the state it was made with. */
SRPBatch* buildBatch(const SRPDrawCommand* cmds, size_t n, const char* caller)
{
	Runtime& R = rt();
	SRPBatch* B = new SRPBatch();
	auto fail = [&](const char* fmt, size_t i) {
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, caller, fmt, i);
		delete B;
		return (SRPBatch*) nullptr;
	};
	B->ctxCopy.resize(n);
	B->src.assign(cmds, cmds + n);
	for (size_t i = 0; i < n; i++)
	{
		B->ctxCopy[i] = *(cmds[i].context ? cmds[i].context : ctx());
		B->src[i].context = &B->ctxCopy[i];
	}
	struct Prep { uint64_t lo; uint32_t cacheSize, nIn; const SRPContext* C; };
	std::vector<Prep> prep;
	for (size_t i = 0; i < n; i++)
	{
		const SRPDrawCommand& c = B->src[i];
		if (!c.vb || !c.fb || !c.sp || !c.sp->vs || !c.sp->fs)
			return fail("Draw %zu: missing vertex buffer, framebuffer or shader program", i);
		if (!liveObjects().count(c.vb) || !liveObjects().count(c.fb) || (c.ib && !liveObjects().count(c.ib)))
			return fail("Draw %zu: a buffer or framebuffer it names was freed", i);
		B->watch.push_back({c.vb, 0, i, c.vb->version});
		if (c.ib)
			B->watch.push_back({c.ib, 1, i, c.ib->version});
		B->watch.push_back({c.fb, 2, i, 0});
		FbImpl* F = fbImpl(c.fb);
		if (!F->deferred)
			return fail("Draw %zu: batches draw into deferred framebuffers (srpFramebufferSetDeferred)", i);
		if (c.fb->width > 65535 || c.fb->height > 65535 || c.fb->size >= 0xFFFFFFFFull)
			return fail("Draw %zu: framebuffer too large for the GPU path", i);
		if (std::find(B->fbs.begin(), B->fbs.end(), F) == B->fbs.end())
			B->fbs.push_back(F);
		if (c.clear)
		{
			uint32_t bits;
			memcpy(&bits, &c.clearDepth, 4);
			B->clears.push_back({F, bits, c.clearStencil});
		}
		if (c.primitive != SRP_PRIM_TRIANGLES && c.primitive != SRP_PRIM_TRIANGLE_STRIP &&
		    c.primitive != SRP_PRIM_TRIANGLE_FAN)
			return fail("Draw %zu: batches draw triangles only", i);
		const SRPContext* C = c.context ? c.context : ctx();
		if (C->raster.polygonMode != SRP_POLYGON_MODE_FILL)
			return fail("Draw %zu: batches draw filled triangles only", i);
		if (!B->sp)
			B->sp = c.sp;
		else if (c.sp->vs->shader != B->sp->vs->shader || c.sp->fs->shader != B->sp->fs->shader ||
		         c.sp->fs->mayOverwriteDepth != B->sp->fs->mayOverwriteDepth ||
		         c.sp->vs->varyingsSize != B->sp->vs->varyingsSize || c.sp->vs->nVaryings != B->sp->vs->nVaryings)
			return fail("Draw %zu: the draws of a batch share one vertex and one fragment shader", i);
		if (c.count == 0 || C->raster.cullFace == SRP_FACE_FRONT_AND_BACK)
			continue;   /* draws nothing (pipeline/draw.c:63-65, :89-90) */
		const size_t size = c.ib ? c.ib->nIndices : c.vb->nVertices;
		if (c.startIndex + c.count - 1 >= size)
			return fail("Draw %zu: indices out of bounds", i);
		const size_t nIn = (c.primitive == SRP_PRIM_TRIANGLES) ? c.count / 3 : (c.count >= 3 ? c.count - 2 : 0);
		if (nIn == 0)
			continue;
		uint64_t lo, hi;
		if (c.ib)
		{
			indexRange(c.ib, c.startIndex, c.count, lo, hi);
			if (hi >= c.vb->nVertices)
				return fail("Draw %zu: an index exceeds the vertex buffer", i);
		}
		else
		{
			lo = c.startIndex;
			hi = c.startIndex + c.count - 1;
		}
		if (hi - lo + 1 > 0x7FFFFFFFull || (uint64_t) B->nIn + nIn > 0x7FFFFFFFull)
			return fail("Draw %zu: too many vertices or primitives", i);
		B->cmds.push_back(c);
		prep.push_back({lo, (uint32_t) (hi - lo + 1), (uint32_t) nIn, C});
		B->nIn += (uint32_t) nIn;
	}
	{
		std::lock_guard<std::mutex> g(g_regLock);
		auto iv = vsRegistry().find((const void*) B->sp->vs->shader);
		auto ifs = fsRegistry().find((const void*) B->sp->fs->shader);
		if (iv != vsRegistry().end()) B->vse = iv->second;
		if (ifs != fsRegistry().end()) B->fse = ifs->second;
	}
	if (!B->vse || !B->fse)
		return fail("Draw %zu: a shader is not a registered device shader (see SRP_HIP_*_SHADER in srp/srp_device.h)",
		            (size_t) 0);
	const SRPVertexShader* V = B->sp->vs;
	if (V->nVaryings > SRP_MAX_VARYINGS)
		return fail("At most %zu varyings are supported", (size_t) SRP_MAX_VARYINGS);
	HIP_CHECK(hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, 0));

	const size_t nd = B->cmds.size();
	B->usize = B->vse->uniformSize > B->fse->uniformSize ? B->vse->uniformSize : B->fse->uniformSize;
	B->stride = kParamsHead + ((B->usize + 255) & ~(size_t) 255);
	HIP_CHECK(hipHostMalloc((void**) &B->hTotals, kTotalsBytes, hipHostMallocDefault));
	HIP_CHECK(hipMalloc((void**) &B->totals, kTotalsBytes));
	HIP_CHECK(hipHostMalloc((void**) &B->hErr, sizeof(uint32_t), hipHostMallocMapped));
	*B->hErr = 0;
	HIP_CHECK(hipHostGetDevicePointer((void**) &B->dErr, B->hErr, 0));
	if (nd == 0)
		return B;   /* only clears */
	HIP_CHECK(hipHostMalloc((void**) &B->hBlocks, nd * B->stride, hipHostMallocDefault));
	memset(B->hBlocks, 0, nd * B->stride);
	HIP_CHECK(hipMalloc((void**) &B->dBlocks, nd * B->stride));
	size_t clipSlots = 0, varBytes = 0;
	for (size_t d = 0; d < nd; d++)
	{
		clipSlots += prep[d].cacheSize;
		varBytes += (size_t) prep[d].cacheSize * V->varyingsSize;
	}
	HIP_CHECK(hipMalloc((void**) &B->clip, clipSlots * 16 + 16));
	HIP_CHECK(hipMalloc((void**) &B->varyings, varBytes + 512));
	HIP_CHECK(hipMalloc((void**) &B->counts, ((size_t) B->nIn + 1) * 4 * sizeof(uint32_t)));
	B->hDraws.resize(nd);
	B->hVsSeg.resize(nd);
	B->hRasterSeg.resize(nd);
	size_t clipOff = 0, varOff = 0;
	uint32_t primBase = 0, tileBase = 0, inBlock = 0, vsBlock = 0;
	for (size_t d = 0; d < nd; d++)
	{
		const SRPDrawCommand& c = B->cmds[d];
		SrpDrawParams& H = *batchParams(B, d);
		const SrpDrawParams* dP = (const SrpDrawParams*) (B->dBlocks + d * B->stride);
		if (!varyingLayout(V, H) || !snapshotState(prep[d].C, c.sp, 0, H.st))
		{
			srpFreeBatch(B);
			return nullptr;
		}
		FbImpl* F = fbImpl(c.fb);
		H.width = (uint32_t) c.fb->width;
		H.height = (uint32_t) c.fb->height;
		H.color = F->dColor;
		H.depth = F->dDepth;
		H.stencil = F->dStencil;
		if (c.sp->uniform && B->usize)
		{
			memcpy((uint8_t*) &H + kParamsHead, c.sp->uniform, B->usize);
			H.uniform = (const uint8_t*) dP + kParamsHead;
		}
		else
			H.uniform = nullptr;
		H.vertices = c.vb->dData;
		H.vertexStride = c.vb->nBytesPerVertex;
		H.cacheBase = prep[d].lo;
		H.cacheSize = prep[d].cacheSize;
		H.clip = B->clip + clipOff * 4;
		H.varyings = B->varyings + varOff;
		const uint32_t side = (B->tuning.tileSize == 16) ? 16u : 8u;
		H.tileShift = (side == 16u) ? 4u : 3u;
		H.tilesX = (H.width + side - 1u) / side;
		H.tilesY = (H.height + side - 1u) / side;
		H.maskBox = (uint32_t) B->tuning.maskBox;
		H.stats = R.dStats;
		clipOff += prep[d].cacheSize;
		varOff += (size_t) prep[d].cacheSize * V->varyingsSize;
		B->lateZ = H.st.lateZ;
		B->words = H.wordPath;

		AsmDraw& AD = B->hDraws[d];
		AD.P = dP;
		AD.indices = c.ib ? c.ib->dData : nullptr;
		AD.start = c.startIndex;
		AD.count = c.count;
		AD.prim = (uint32_t) c.primitive;
		AD.nIn = prep[d].nIn;
		AD.indexElem = c.ib ? (uint32_t) c.ib->elem : 0u;
		AD.primBase = primBase;
		AD.tileBase = tileBase;
		AD.firstBlock = inBlock;
		const uint32_t nTiles = H.tilesX * H.tilesY;
		B->hVsSeg[d] = SrpSeg{dP, 0u, prep[d].cacheSize, vsBlock, 0u};
		B->hRasterSeg[d] = SrpSeg{dP, tileBase, nTiles, 0u, 0u};
		primBase += prep[d].nIn;
		tileBase += nTiles;
		inBlock += (prep[d].nIn + 127u) / 128u;
		vsBlock += (prep[d].cacheSize + 255u) / 256u;
	}
	B->inBlocks = inBlock;
	B->vsBlocks = vsBlock;
	B->totalTiles = tileBase;
	HIP_CHECK(hipMalloc((void**) &B->dDraws, nd * sizeof(AsmDraw)));
	HIP_CHECK(hipMalloc((void**) &B->dVsSeg, nd * sizeof(SrpSeg)));
	HIP_CHECK(hipMalloc((void**) &B->dRasterSeg, nd * sizeof(SrpSeg)));
	HIP_CHECK(hipMalloc((void**) &B->tileRange, ((size_t) 2 * B->totalTiles + 2) * sizeof(uint32_t)));
	HIP_CHECK(hipMalloc((void**) &B->queues, 8 * sizeof(uint32_t)));
	if (!batchSize(B, caller))
	{
		srpFreeBatch(B);
		return nullptr;
	}
	return B;
}

/* Before a frame: the objects the batch reads must still exist, and a
 * batch whose vertex or index buffers were given new data (which may move
 * them, or change the vertex range the draws need) is prepared again from
 * its commands and their recorded state, keeping its tuning and its graph
 * (captured again).  False, with a message, when the frame cannot run. */
bool batchValidate(SRPBatch* B, const char* caller)
{
	bool changed = false;
	for (const SRPBatch::Watch& w : B->watch)
	{
		if (!liveObjects().count(w.obj))
		{
			message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, caller,
			        "Draw %zu: its %s was freed after the batch was made; nothing is drawn", w.draw,
			        w.kind == 0 ? "vertex buffer" : (w.kind == 1 ? "index buffer" : "framebuffer"));
			return false;
		}
		if (w.kind == 0 && ((const SRPVertexBuffer*) w.obj)->version != w.version)
			changed = true;
		if (w.kind == 1 && ((const SRPIndexBuffer*) w.obj)->version != w.version)
			changed = true;
	}
	if (!changed)
		return true;
	HIP_CHECK(hipStreamSynchronize(rt().stream));
	const bool captured = B->exec != nullptr;
	SRPBatch* fresh = buildBatch(B->src.data(), B->src.size(), caller);
	if (!fresh)
		return false;
	if (memcmp(&fresh->tuning, &B->tuning, sizeof(SRPBatchTuning)) != 0)
		srpBatchSetTuning(fresh, &B->tuning);
	fresh->reported = B->reported;
	std::swap(*B, *fresh);
	srpFreeBatch(fresh);   /* the old preparation */
	if (captured && srpBatchCapture(B) != 0)
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, caller, "The batch's graph could not be captured again");
	return true;
}

}  // namespace

extern "C" {

SRPBatch* srpNewBatch(const SRPDrawCommand* cmds, size_t n)
{
	resolvePending();
	if (!cmds || n == 0)
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpNewBatch", "No draws");
		return nullptr;
	}
	return buildBatch(cmds, n, "srpNewBatch");
}

void srpFreeBatch(SRPBatch* B)
{
	resolvePending();
	if (!B)
		return;
	HIP_CHECK(hipStreamSynchronize(rt().stream));
	batchDropGraph(B);
	void* bufs[] = {B->dBlocks, B->dDraws, B->dVsSeg, B->dRasterSeg, B->clip, B->varyings, B->counts, B->totals,
	                B->tris, B->spans, B->pairKeys, B->pairVals, B->pairRec, B->pairMask, B->genVaryings, B->tileRange,
	                B->queues, B->scan.buf, B->scan.buf64, B->sort.keys, B->sort.vals, B->sort.hist, B->sort.grp, B->dFills,
	                B->sort.scan.buf, B->sort.scan.buf64};
	for (void* p : bufs)
		batchFreeDev(p);
	if (B->hErr)
	{
		batchReport(B);
		HIP_CHECK(hipHostFree(B->hErr));
	}
	if (B->hBlocks) HIP_CHECK(hipHostFree(B->hBlocks));
	if (B->hTotals) HIP_CHECK(hipHostFree(B->hTotals));
	delete B;
}

void srpBatchRender(SRPBatch* B)
{
	resolvePending();
	if (!B)
		return;
	Runtime& R = rt();
	batchReport(B);
	if (!batchValidate(B, "srpBatchRender"))
		return;
	batchFlushPending(B);
	if (B->exec)
	{
		HIP_CHECK(hipGraphLaunch(B->exec, R.stream));
		return;
	}
	if (B->cmds.empty())
	{
		for (const SRPBatch::Clear& c : B->clears)
			fillFramebuffer(c.F->dColor, 0u, (uint32_t*) c.F->dDepth, c.depthBits, c.F->dStencil, c.stencil,
			                c.F->pub.size, R.stream);
		return;
	}
	hipEvent_t e0 = R.prof(1) ? R.mark() : nullptr;
	batchLaunch(B, R.stream);
	if (e0) R.span(1, e0, R.mark());
}

int srpBatchCapture(SRPBatch* B)
{
	resolvePending();
	if (!B || B->cmds.empty())
		return -1;
	Runtime& R = rt();
	batchReport(B);
	batchDropGraph(B);
	if (!batchValidate(B, "srpBatchCapture"))
		return -1;
	batchFlushPending(B);
	/* every scratch buffer was sized when the batch was prepared, so the
	 * captured frame allocates nothing; nothing is drawn now */
	HIP_CHECK(hipStreamSynchronize(R.stream));
	if (hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
		return -2;
	batchLaunch(B, R.stream);
	hipGraph_t g = nullptr;
	if (hipStreamEndCapture(R.stream, &g) != hipSuccess || !g)
		return -3;
	if (hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0) != hipSuccess)
	{
		HIP_CHECK(hipGraphDestroy(g));
		B->exec = nullptr;
		return -4;
	}
	B->graph = g;
	return 0;
}

void srpBatchSetTuning(SRPBatch* B, const SRPBatchTuning* t)
{
	resolvePending();
	if (!B || !t)
		return;
	if (t->wavesPerCU < 1 || t->wavesPerCU > 32 || t->maskBox < 0 || t->maskBox > SRP_SMALL_BBOX ||
	    (t->tileSize != 8 && t->tileSize != 16))
	{
		message(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "srpBatchSetTuning",
		        "Waves per CU must lie in 1..32, the mask box in 0..%d and the tile size be 8 or 16", SRP_SMALL_BBOX);
		return;
	}
	HIP_CHECK(hipStreamSynchronize(rt().stream));
	batchDropGraph(B);
	const bool retile = t->tileSize != B->tuning.tileSize;
	const bool resize = retile || t->maskBox != B->tuning.maskBox;
	B->tuning = *t;
	if (!B->cmds.empty())
	{
		if (retile)
			batchRetile(B);
		if (resize)
			batchSize(B, "srpBatchSetTuning");   /* tile size and mask box change the number of tile pairs */
		else
			batchUpload(B);
	}
}

void srpBatchGetTuning(const SRPBatch* B, SRPBatchTuning* t)
{
	if (B && t)
		*t = B->tuning;
}

double srpBatchAutotune(SRPBatch* B, int trials, SRPBatchTuning* best)
{
	resolvePending();
	if (!B || B->cmds.empty())
		return 0.0;
	batchReport(B);
	if (!batchValidate(B, "srpBatchAutotune"))
		return 0.0;
	if (trials < 1) trials = 1;
	Runtime& R = rt();
	/* 8x8 tiles, a wave each (the batched kernel): waves per CU x mask box x
	 * queues x pixel runs; 16x16 tiles, four waves each (a lane a pixel, no
	 * binning masks or runs): waves per CU x queues */
	const int waves[] = {4, 8, 16, 28};
	const int boxes[] = {0, 16, 24};
	const int queues[] = {0, 1};
	const int runs[] = {0, 1};
	std::vector<SRPBatchTuning> cands;
	for (int box : boxes)
		for (int w : waves)
			for (int q : queues)
				for (int pr : runs)
					cands.push_back({w, box, q, pr, B->tuning.depthCull, 8});
	for (int w : waves)
		for (int q : queues)
			cands.push_back({w, SRP_SMALL_BBOX, q, 1, B->tuning.depthCull, 16});
	SRPBatchTuning bestT = B->tuning;
	double bestMs = 1e30;
	hipEvent_t a, b;
	HIP_CHECK(hipEventCreate(&a));
	HIP_CHECK(hipEventCreate(&b));
	for (const SRPBatchTuning& t : cands)
	{
		srpBatchSetTuning(B, &t);
		batchFlushPending(B);
		batchLaunch(B, R.stream);   /* warm-up */
		HIP_CHECK(hipEventRecord(a, R.stream));
		for (int k = 0; k < trials; k++)
			batchLaunch(B, R.stream);
		HIP_CHECK(hipEventRecord(b, R.stream));
		HIP_CHECK(hipEventSynchronize(b));
		float ms = 0.0f;
		HIP_CHECK(hipEventElapsedTime(&ms, a, b));
		const double per = (double) ms / trials;
		if (per < bestMs)
		{
			bestMs = per;
			bestT = t;
		}
	}
	HIP_CHECK(hipEventDestroy(a));
	HIP_CHECK(hipEventDestroy(b));
	srpBatchSetTuning(B, &bestT);
	if (best)
		*best = bestT;
	return bestMs;
}

void srpBatchInfo(const SRPBatch* B, uint64_t out[4])
{
	resolvePending();
	if (!B || !out)
		return;
	batchReport(const_cast<SRPBatch*>(B));
	out[0] = B->cmds.size();
	out[1] = B->nRec;
	out[2] = B->nItems;
	out[3] = B->totalTiles;
}

}  // extern "C"
