/* Kernels that call user shaders, instantiated inside each HIP shader
 * module by the SRP_HIP_*_SHADER macros (srp/srp_device.h):
 *
 *   srp_vs_kernel      vertex shader over the draw's post-VS cache slots
 *   srp_raster_kernel  sort-middle tile rasteriser for triangles
 *   srp_frag_kernel    per-pixel ordered fragment lists (lines, points)
 *
 * Numerics follow the reference renderer bit for bit (kitrofimov/srp):
 * FP contraction is off for this whole header, float division and sqrt
 * are correctly rounded, the barycentric walk replays the reference's
 * sequence of float additions, and every per-fragment test runs in the
 * reference's order per pixel and per primitive. */
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "srp/shaders.h"
#include "srp/device/abi.h"
#include "srp/device/replay.h"
#include "srp/device/depth_cull.h"

#pragma clang fp contract(off)

namespace srp {
namespace dev {

/* reference math/utils.h:41-45 (comparison in double precision) */
__device__ __forceinline__ bool roughlyZero(float x) { return fabsf(x) <= SRP_EPSILON_F; }

/* Per-edge thresholds of the inside test: l[e] > thr[e] for every edge
 * is (l[e] > 0) || (roughlyZero(l[e]) && topLeft(e)) (SRP_EDGE_TL_THRESHOLD). */
__device__ __forceinline__ void edgeThresholds(uint32_t flags, float thr[3])
{
#pragma unroll
	for (int e = 0; e < 3; e++)
		thr[e] = ((flags >> e) & 1u) ? SRP_EDGE_TL_THRESHOLD : 0.0f;
}
__device__ __forceinline__ bool insideEdges(const float* l, const float* thr)
{
	return (l[0] > thr[0]) & (l[1] > thr[1]) & (l[2] > thr[2]);
}
/* The reference's coverage test of a pixel with barycentrics l. */
__device__ __forceinline__ bool insideTri(const SrpTri& T, const float* l)
{
	float thr[3];
	edgeThresholds(T.flags, thr);
	return insideEdges(l, thr);
}

/* reference raster/fragment.c:156-178, :225-245 */
template <typename T>
__device__ __forceinline__ bool compare(uint32_t op, T lhs, T rhs)
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

/* reference raster/fragment.c:180-223: op result merged under the write mask */
__device__ __forceinline__ uint32_t stencilApply(uint32_t op, uint32_t s, const SRPStencilFaceState& st)
{
	uint32_t v;
	switch (op)
	{
		case SRP_STENCIL_KEEP: v = s; break;
		case SRP_STENCIL_ZERO: v = 0; break;
		case SRP_STENCIL_REPLACE: v = st.ref; break;
		case SRP_STENCIL_INCR: v = (s < 255u) ? s + 1u : 255u; break;
		case SRP_STENCIL_INCR_WRAP: v = (s + 1u) & 255u; break;
		case SRP_STENCIL_DECR: v = (s > 0u) ? s - 1u : 0u; break;
		case SRP_STENCIL_DECR_WRAP: v = (s - 1u) & 255u; break;
		case SRP_STENCIL_INVERT: v = (~s) & 255u; break;
		default: v = s; break;
	}
	const uint32_t m = st.writeMask;
	return (s & ~m & 255u) | (v & m);
}

/* reference core/color.c:14-23: clamp(c*255) to [0,255], truncate; R in the top byte. */
__device__ __forceinline__ uint32_t packColor(const float c[4])
{
	uint32_t r = 0;
#pragma unroll
	for (int i = 0; i < 4; i++)
	{
		const float v = c[i] * 255.0f;
		const float cl = (v < 0.0f) ? 0.0f : ((v > 255.0f) ? 255.0f : v);
		const uint32_t b = (cl != cl) ? 0u : (uint32_t) cl;
		r |= b << (24 - 8 * i);
	}
	return r;
}

/* The reference walks the bounding box adding dldx per pixel and dldy per
 * row (raster/triangle.c:101-109); `n` additions of `d` in that order. */
__device__ __forceinline__ float replay(float a, float d, int n)
{
	return srpReplay(a, d, n);
}

/* ---------------------------------------------------------------- varyings */

/* Varying bytes of a vertex slot: post-VS cache or clip-generated vertex. */
__device__ __forceinline__ const uint8_t* varyingPtr(const SrpDrawParams* P, uint32_t vref)
{
	return (vref < P->cacheSize) ? P->varyings + (size_t) vref * P->varyingsSize
	                             : P->genVaryings + (size_t) (vref - P->cacheSize) * P->varyingsSize;
}

__device__ __forceinline__ float loadF(const uint8_t* p)
{
	float v;
	__builtin_memcpy(&v, p, 4);
	return v;
}

__device__ __forceinline__ double loadD(const uint8_t* p)
{
	double v;
	__builtin_memcpy(&v, p, 8);
	return v;
}

__device__ __forceinline__ uint32_t elemSize(uint32_t type)
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

/* Fast path: every element is a 4-byte word at a 4-byte aligned offset.
 * reference pipeline/interpolation.c:63-91 per word: perspective
 * ((A*invW)*w summed, times 1/sum(invW*w)), affine (A*w summed) or a copy
 * from the provoking vertex.  `out` stays in registers when the shader's
 * reads have constant offsets. */
template <int N>
__device__ __forceinline__ void interpWords(const SrpDrawParams* P, const uint32_t* const* v, const float* w,
                                            const float* invW, float rec, uint32_t pv, uint32_t* out)
{
#pragma unroll
	for (int k = 0; k < SRP_MAX_WORDS; k++)
	{
		if (k >= (int) P->nWords)
		{
			out[k] = 0u;
			continue;
		}
		const uint32_t kind = P->wordKind[k];
		if (kind == SRP_WORD_PERSPECTIVE)
		{
			float val = 0.0f;
#pragma unroll
			for (int i = 0; i < N; i++)
				val += __uint_as_float(v[i][k]) * invW[i] * w[i];
			val *= rec;
			out[k] = __float_as_uint(val);
		}
		else if (kind == SRP_WORD_AFFINE)
		{
			float val = 0.0f;
#pragma unroll
			for (int i = 0; i < N; i++)
				val += __uint_as_float(v[i][k]) * w[i];
			out[k] = __float_as_uint(val);
		}
		else if (kind == SRP_WORD_PROVOKING)
			out[k] = v[pv][k];
		else
			out[k] = 0u;
	}
}

/* General path: any layout, byte addressed (interpolation.c:93-163). */
template <int N>
__device__ inline void interpBytes(const SrpDrawParams* P, const uint8_t* const* v, const float* w,
                                   const float* invW, float rec, uint32_t pv, uint8_t* out)
{
	for (uint32_t a = 0; a < P->nVaryings; a++)
	{
		const SrpVaryingDesc d = P->var[a];
		const bool persp = d.mode == SRP_INTERPOLATION_MODE_PERSPECTIVE;
		const bool affine = d.mode == SRP_INTERPOLATION_MODE_AFFINE;
		if (d.type == SRP_FLOAT)
		{
			for (uint32_t e = 0; e < d.nItems; e++)
			{
				const uint32_t off = d.offset + 4 * e;
				float val = 0.0f;
				if (persp)
				{
					for (int i = 0; i < N; i++) val += loadF(v[i] + off) * invW[i] * w[i];
					val *= rec;
				}
				else if (affine)
					for (int i = 0; i < N; i++) val += loadF(v[i] + off) * w[i];
				else
					val = loadF(v[pv] + off);
				__builtin_memcpy(out + off, &val, 4);
			}
		}
		else if (d.type == SRP_DOUBLE)
		{
			for (uint32_t e = 0; e < d.nItems; e++)
			{
				const uint32_t off = d.offset + 8 * e;
				double val = 0.0;
				if (persp)
				{
					for (int i = 0; i < N; i++) val += loadD(v[i] + off) * (double) invW[i] * (double) w[i];
					val *= (double) rec;
				}
				else if (affine)
					for (int i = 0; i < N; i++) val += loadD(v[i] + off) * (double) w[i];
				else
					val = loadD(v[pv] + off);
				__builtin_memcpy(out + off, &val, 8);
			}
		}
		else
		{
			const uint32_t n = elemSize(d.type) * d.nItems;
			for (uint32_t b = 0; b < n; b++) out[d.offset + b] = v[pv][d.offset + b];
		}
	}
}

/* --------------------------------------------------------- fragment ops */

struct PixelState
{
	uint32_t color;
	float depth;
	uint32_t stencil;
	uint32_t covered;
	uint32_t shaded;
};

/* reference raster/fragment.c:63-125, for one fragment of pixel (px, py)
 * whose fb values live in `ps`.  `interp(buf)` fills the varying buffer and
 * returns the pointer handed to the shader; it runs only when the fragment
 * reaches the shader (interpolation has no side effects). */
template <class FS, bool LATE, class Interp>
__device__ __forceinline__ void fragment(const SrpDrawParams* P, int px, int py, bool front, float depth,
                                         float fc3, uint32_t primID, PixelState& ps, Interp&& interp)
{
	const SrpState& st = P->st;
	ps.covered++;
	if (st.scissorEnabled)
	{
		const uint64_t ux = (uint64_t) (int64_t) px, uy = (uint64_t) (int64_t) py;
		if (ux < st.sx0 || ux >= st.sx1 || uy < st.sy0 || uy >= st.sy1)
			return;
	}
	const SRPStencilFaceState& sf = st.stencil[front ? 0 : 1];
	const float stored = ps.depth;
	if (st.stencilEnabled && !compare<uint32_t>(sf.func, (uint32_t) (sf.ref & sf.mask), ps.stencil & sf.mask))
	{
		ps.stencil = stencilApply(sf.sfailOp, ps.stencil, sf);
		return;
	}
	if (!LATE && st.depthTest && !compare<float>(st.depthOp, depth, stored))
	{
		if (st.stencilEnabled) ps.stencil = stencilApply(sf.dfailOp, ps.stencil, sf);
		return;
	}
	SRPFragmentShaderIn in;
	in.uniform = (SRPUniform*) P->uniform;
	in.fragCoord[0] = (float) ((double) px + 0.5);
	in.fragCoord[1] = (float) ((double) py + 0.5);
	in.fragCoord[2] = depth;
	in.fragCoord[3] = fc3;
	in.frontFacing = front;
	in.primitiveID = primID;
	in.varyings = interp();
	SRPFragmentShaderOut out;
	out.color[0] = 0.0f; out.color[1] = 0.0f; out.color[2] = 0.0f; out.color[3] = 0.0f;
	out.fragDepth = __builtin_nanf("");
	ps.shaded++;
	FS()(&in, &out);
	if (LATE)
	{
		if (!isnan(out.fragDepth)) depth = out.fragDepth;
		if (st.depthTest && !compare<float>(st.depthOp, depth, stored))
		{
			if (st.stencilEnabled) ps.stencil = stencilApply(sf.dfailOp, ps.stencil, sf);
			return;
		}
	}
	if (st.stencilEnabled) ps.stencil = stencilApply(sf.passOp, ps.stencil, sf);
	ps.color = packColor(out.color);
	if (st.depthTest && st.depthWrite)
		ps.depth = depth;
}

__device__ __forceinline__ void flushStats(const SrpDrawParams* P, uint32_t covered, uint32_t shaded)
{
	/* wave-level reduction, one atomic per wave and counter */
	for (int off = 32; off > 0; off >>= 1)
	{
		covered += __shfl_xor(covered, off);
		shaded += __shfl_xor(shaded, off);
	}
	if ((threadIdx.x & 63) == 0 && (covered | shaded))
	{
		unsigned long long* c = P->stats + (size_t) (blockIdx.x % SRP_STAT_SLOTS) * SRP_STAT_STRIDE;
		atomicAdd(&c[0], (unsigned long long) covered);
		atomicAdd(&c[1], (unsigned long long) shaded);
	}
}

/* ------------------------------------------------------------- kernels */

/* Block b of a vertex launch's framebuffer fill (SrpVsLaunch.fill*). */
__device__ __forceinline__ void fillBlock(const SrpVsLaunch& L, uint32_t b)
{
	const size_t n = L.fillCount;
	const size_t i0 = (size_t) b * 256u + threadIdx.x;
	const size_t stride = (size_t) L.fillBlocks * 256u;
	for (size_t i = i0; i < n; i += stride)
	{
		if (L.fillColor) L.fillColor[i] = L.fillColorValue;
		if (L.fillDepth) L.fillDepth[i] = L.fillDepthValue;
	}
	if (L.fillStencil)
	{
		/* four bytes at a time when aligned, then the rest */
		const uint32_t v = (L.fillStencilValue & 255u) * 0x01010101u;
		size_t done = 0;
		if (((uintptr_t) L.fillStencil & 3u) == 0)
		{
			const size_t n4 = n / 4;
			for (size_t i = i0; i < n4; i += stride)
				reinterpret_cast<uint32_t*>(L.fillStencil)[i] = v;
			done = n4 * 4;
		}
		for (size_t i = done + i0; i < n; i += stride)
			L.fillStencil[i] = (uint8_t) v;
	}
}

/* Vertex stage: one thread per post-VS cache slot (vertex index
 * cacheBase + i).  reference pipeline/vertex_processing.c:55-74.  Blocks
 * past the vertex blocks fill the framebuffer (fillBlock). */
template <class VS>
__global__ __launch_bounds__(256) void srp_vs_kernel(SrpVsLaunch L)
{
	if (L.zero && blockIdx.x == 0 && threadIdx.x < L.zeroCount)
		L.zero[threadIdx.x] = 0u;
	const uint32_t vsBlocks = L.nSeg ? L.nBlocks : (L.count + 255u) / 256u;
	if (blockIdx.x >= vsBlocks)
	{
		fillBlock(L, blockIdx.x - vsBlocks);
		return;
	}
	const SrpDrawParams* P = L.params;
	uint32_t i = blockIdx.x * 256u + threadIdx.x;
	uint32_t count = L.count;
	if (L.nSeg)
	{
		uint32_t a = 0, b = L.nSeg;
		while (b - a > 1)
		{
			const uint32_t m = (a + b) >> 1;
			if (L.seg[m].firstBlock <= blockIdx.x) a = m;
			else b = m;
		}
		P = L.seg[a].P;
		i = (blockIdx.x - L.seg[a].firstBlock) * 256u + threadIdx.x;
		count = L.seg[a].count;
	}
	if (i >= count)
		return;
	const uint64_t vi = P->cacheBase + i;
	SRPVertexShaderIn in;
	in.uniform = (SRPUniform*) P->uniform;
	in.vertex = (SRPVertex*) (P->vertices + vi * P->vertexStride);
	in.vertexID = (size_t) vi;
	SRPVertexShaderOut out;
	out.clipPosition[0] = 0.0f; out.clipPosition[1] = 0.0f;
	out.clipPosition[2] = 0.0f; out.clipPosition[3] = 0.0f;
	out.varyings = (SRPVarying*) (P->varyings + (size_t) i * P->varyingsSize);
	VS()(&in, &out);
	float4 c;
	c.x = out.clipPosition[0]; c.y = out.clipPosition[1];
	c.z = out.clipPosition[2]; c.w = out.clipPosition[3];
	reinterpret_cast<float4*>(P->clip)[i] = c;
}

/* Blocks are issued round-robin over the 8 XCDs (block b to XCD b % 8).
 * The tile a block renders:
 *   run == 0   each XCD gets one contiguous range of tiles, so primitives
 *              shared by neighbouring tiles hit the same L2; a frame whose
 *              busy tiles gather in a band of rows then loads a few XCDs;
 *   run == n   runs of n consecutive tiles are dealt to the XCDs in turn
 *              (groups of 8n blocks; a last partial group keeps block
 *              order), so every XCD gets a share of any region of the
 *              frame and neighbouring tiles within a run share an L2.
 * Both are permutations of [0, nBlocks). */
__device__ __forceinline__ uint32_t srpTileOfBlock(uint32_t b, uint32_t nBlocks, uint32_t run)
{
	if (run == 0u)
	{
		const uint32_t q = nBlocks / 8u, r = nBlocks % 8u;
		const uint32_t x = b % 8u, k = b / 8u;
		return x * q + (x < r ? x : r) + k;
	}
	const uint32_t group = 8u * run;
	const uint32_t m = b / group;
	if ((m + 1u) * group > nBlocks)
		return b;
	const uint32_t r = b - m * group;
	return m * group + (r % 8u) * run + r / 8u;
}

/* Triangle rasteriser: each lane owns one pixel of a TILE x TILE tile and
 * keeps its colour/depth/stencil in registers while it walks the tile's
 * primitive list in primitive order (so per-pixel API order holds).  The
 * barycentrics of a lane are replayed from the bbox origin exactly as the
 * reference's scan would have produced them. */
/* One tile of that kernel: `tile` is the tile of draw P, `gtile` its entry
 * in the tile ranges, `local` this lane's pixel of the tile; the counts
 * accumulate into `covered` / `shaded`. */
template <class FS, bool LATE, bool WORDS, int TILE>
__device__ __forceinline__ void rasterTileSimple(const SrpRasterLaunch& L, const SrpDrawParams* P, uint32_t tile,
                                                 uint32_t gtile, uint32_t local, uint32_t& covered, uint32_t& shaded)
{
	const uint32_t begin = L.tileStart[gtile], end = L.tileEnd[gtile];
	if (begin == end)
		return;
	const uint32_t tx = tile % P->tilesX, ty = tile / P->tilesX;
	const int px = (int) (tx * TILE + local % TILE);
	const int py = (int) (ty * TILE + local / TILE);
	const bool inFb = px < (int) P->width && py < (int) P->height;
	const size_t idx = (size_t) py * P->width + (size_t) px;

	PixelState ps;
	ps.color = 0; ps.depth = 0.0f; ps.stencil = 0; ps.covered = 0; ps.shaded = 0;
	if (inFb)
	{
		ps.color = P->color[idx];
		ps.depth = P->depth[idx];
		ps.stencil = P->stencil[idx];
	}
	const uint32_t pv = P->st.provokingLast ? 2u : 0u;

	for (uint32_t i = begin; i < end; i++)
	{
		const uint32_t pr = L.pairRec[L.tileList[i]];
		if (pr == 0xFFFFFFFFu)   /* a pair without a record (a batch's capacity exceeded) covers nothing */
			continue;
		const SrpTri* T = &L.tris[pr & 0x7FFFFFFFu];
		if (!(px >= (int) T->x0 && px < (int) T->x1 && py >= (int) T->y0 && py < (int) T->y1))
			continue;
		const int ny = py - (int) T->y0, nx = px - (int) T->x0;
		float l[3];
#pragma unroll
		for (int e = 0; e < 3; e++)
			l[e] = replay(replay(T->l0[e], T->dy[e], ny), T->dx[e], nx);
		if (!insideTri(*T, l))
			continue;
		const float iw[3] = {T->invW[0], T->invW[1], T->invW[2]};
		const float rec = 1.0f / (iw[0] * l[0] + iw[1] * l[1] + iw[2] * l[2]);
		const float depth = T->zw[0] * l[0] + T->zw[1] * l[1] + T->zw[2] * l[2];
		const bool front = (T->flags >> 3) & 1u;
		if (WORDS)
		{
			uint32_t buf[SRP_MAX_WORDS];
			fragment<FS, LATE>(P, px, py, front, depth, rec, T->id, ps, [&]() {
				const uint32_t* v[3];
#pragma unroll
				for (int k = 0; k < 3; k++)
					v[k] = reinterpret_cast<const uint32_t*>(varyingPtr(P, T->vref[k]));
				interpWords<3>(P, v, l, iw, rec, pv, buf);
				return (SRPInterpolated*) buf;
			});
		}
		else
		{
			uint8_t buf[SRP_MAX_VARYING_BYTES];
			fragment<FS, LATE>(P, px, py, front, depth, rec, T->id, ps, [&]() {
				const uint8_t* v[3];
#pragma unroll
				for (int k = 0; k < 3; k++) v[k] = varyingPtr(P, T->vref[k]);
				for (uint32_t b = 0; b < P->describedSize; b++) buf[b] = 0;
				interpBytes<3>(P, v, l, iw, rec, pv, buf);
				return (SRPInterpolated*) buf;
			});
		}
	}
	if (inFb)
	{
		P->color[idx] = ps.color;
		P->depth[idx] = ps.depth;
		P->stencil[idx] = (uint8_t) ps.stencil;
	}
	covered += ps.covered;
	shaded += ps.shaded;
}

template <class FS, bool LATE, bool WORDS, int TILE>
__global__ __launch_bounds__(256) void srp_raster_kernel(SrpRasterLaunch L)
{
	constexpr int TILES_PER_BLOCK = (TILE == 8) ? 4 : 1;
	const uint32_t t = threadIdx.x;
	const uint32_t local = (TILE == 8) ? (t & 63u) : t;
	const uint32_t tile = srpTileOfBlock(blockIdx.x, L.nBlocks, L.tileOrder) * TILES_PER_BLOCK + ((TILE == 8) ? (t >> 6) : 0u);
	if (tile >= L.nTiles)
		return;
	uint32_t covered = 0, shaded = 0;
	rasterTileSimple<FS, LATE, WORDS, TILE>(L, L.params, tile, tile, local, covered, shaded);
	flushStats(L.params, covered, shaded);
}

/* Persistent multi-draw form of srp_raster_kernel for 16x16 tiles (four
 * waves a tile, a lane a pixel): blocks take tiles of all the draws from
 * the device queues as srp_raster_persistent does, the whole block one
 * tile at a time. */
template <class FS, bool LATE, bool WORDS>
__global__ __launch_bounds__(256) void srp_raster_persistent16(SrpRasterLaunch L)
{
	__shared__ uint32_t taken;
	const uint32_t nq = L.xcdQueues ? 8u : 1u;
	const uint32_t q0 = blockIdx.x % nq;
	uint32_t covered = 0, shaded = 0;
	const SrpDrawParams* last = nullptr;
	for (uint32_t k = 0; k < nq; k++)
	{
		const uint32_t q = (q0 + k) % nq;
		const uint32_t lo = (uint32_t) (((uint64_t) L.nTiles * q) / nq);
		const uint32_t hi = (uint32_t) (((uint64_t) L.nTiles * (q + 1u)) / nq);
		for (;;)
		{
			if (threadIdx.x == 0)
				taken = atomicAdd(&L.queues[q], 1u) + lo;
			__syncthreads();
			const uint32_t t = taken;
			__syncthreads();   /* every lane has read it before the next take */
			if (t >= hi)
				break;
			uint32_t a = 0, b = L.nSeg;
			while (b - a > 1)
			{
				const uint32_t m = (a + b) >> 1;
				if (L.seg[m].first <= t) a = m;
				else b = m;
			}
			const SrpSeg sg = L.seg[a];
			rasterTileSimple<FS, LATE, WORDS, 16>(L, sg.P, t - sg.first, t, threadIdx.x, covered, shaded);
			last = sg.P;
		}
	}
	/* every draw of a batch shares one counter array */
	if (last)
		flushStats(last, covered, shaded);
}

/* Batched tile rasteriser (one 64-lane wave per 8x8 tile).  A long tile
 * list (thousands of slivers around a mesh pole, say) would serialise one
 * shader call, and its memory latency, per primitive.  Instead the list is
 * consumed in batches of up to 64 primitives / RF fragments:
 *   A  coverage: lane b holds primitive b's 64-bit mask of the tile, exact
 *      from binning for small triangles, else found here pixel-parallel by
 *      the reference's barycentric replay; a wave scan of the mask
 *      popcounts places the batch's fragments in LDS in (primitive, pixel)
 *      order;
 *   B  shading, lane = fragment: interpolation and the fragment shader for
 *      64 fragments at a time.  The shader is a pure function of its
 *      inputs, so it may run before the ordered tests; a fragment is
 *      skipped only when it is certain to fail (scissor, or an early depth
 *      test with a monotone compare op against the depth the pixel had
 *      when the batch began: stored depth only ever moves towards values
 *      that fail more);
 *   C  resolve, lane = pixel: stencil, depth and colour applied in
 *      primitive order exactly as fragment.c:63-125 does, counting a
 *      fragment as shaded only where the reference invokes the shader. */
constexpr int RF = 128;

/* Waves per SIMD the batched rasteriser kernels are compiled for (register
 * budget 512 / waves, so 72 VGPRs at 7).  Tiles are latency-bound and a
 * 1080p frame has a few thousand busy tiles, so the more of them resident
 * at once the better; 128-fragment batches keep LDS at 4.7 KB per wave.
 * Measured on the 1080p mesh: 5 waves 71.9 us, 6 71.6, 7 62.8, 8 70.1 (the
 * 64-register budget spills).  A wave-cooperative shader (a network on the
 * matrix cores) needs more than 128 registers and gets 2. */
#ifndef SRP_RASTER_WAVES
#define SRP_RASTER_WAVES 7
#endif

/* Fragment shaders that evaluate a whole wave of fragments together (all 64
 * lanes call FS().wave(in, out, active) convergently) declare
 * `static constexpr bool kWave = true`; the batched rasteriser then calls
 * that entry, the other kernels the per-fragment one. */
template <class FS, class = void>
struct FsIsWave
{
	static constexpr bool value = false;
};
template <class FS>
struct FsIsWave<FS, decltype((void) FS::kWave)>
{
	static constexpr bool value = FS::kWave;
};

/* Interpolated-varying buffer of a fragment: words or bytes. */
template <bool WORDS>
struct VaryingBuf
{
	typedef uint32_t type;
	static constexpr uint32_t n = SRP_MAX_WORDS;
};
template <>
struct VaryingBuf<false>
{
	typedef uint32_t type;
	static constexpr uint32_t n = SRP_MAX_VARYING_BYTES / 4;
};

/* Per-fragment state kept in LDS between shading and resolve; FRAG_FRONT
 * is or'ed in for fragments of front-facing triangles, so the resolve
 * reads no triangle record. */
enum : uint8_t { FRAG_SCISSORED = 1, FRAG_CULLED = 2, FRAG_SHADED = 3, FRAG_FRONT = 0x80 };

__device__ __forceinline__ bool monotoneOp(uint32_t op)
{
	return op == SRP_COMPARE_LESS || op == SRP_COMPARE_LEQUAL || op == SRP_COMPARE_GREATER ||
	       op == SRP_COMPARE_GEQUAL || op == SRP_COMPARE_NEVER;
}

/* Barycentrics of tile pixel p (lane index layout) for triangle T. */
__device__ __forceinline__ void baryAt(const SrpTri& T, int x, int y, float* l)
{
	const int ny = y - (int) T.y0, nx = x - (int) T.x0;
#pragma unroll
	for (int e = 0; e < 3; e++)
		l[e] = replay(replay(T.l0[e], T.dy[e], ny), T.dx[e], nx);
}

/* Row bases of up to 8 triangles of a tile (the set bits of `round`, in
 * order; bit b is the record in slot (head + b) % 64).  Lane (j = lane % 8, i = lane / 8) walks triangle j's
 * barycentrics down to row i of the tile and along that row to the tile's
 * first column inside the triangle's box, as the reference's scan would
 * reach them; a pixel of the row then continues with at most 7 additions.
 * (A walk split in two is the same sequence of additions.) */
__device__ __forceinline__ void rowBases(const SrpTri* tris, const uint32_t* recIdx, uint32_t head, uint64_t round,
                                         uint32_t tx, uint32_t ty, uint32_t lane, float (*base)[8][3])
{
	const uint32_t j = lane & 7u, i = lane >> 3;
	uint64_t r = round;
	for (uint32_t k = 0; k < j && r; k++)
		r &= r - 1ull;
	if (!r)
		return;
	const SrpTri& T = tris[recIdx[(head + (uint32_t) __builtin_ctzll(r)) & 63u]];
	const int y = (int) (ty * 8u + i);
	if (y < (int) T.y0 || y >= (int) T.y1)
		return;
	const int cs = ((int) (tx * 8u) > (int) T.x0) ? (int) (tx * 8u) : (int) T.x0;
#pragma unroll
	for (int e = 0; e < 3; e++)
		base[j][i][e] = replay(replay(T.l0[e], T.dy[e], y - (int) T.y0), T.dx[e], cs - (int) T.x0);
}

/* Barycentrics of pixel (px, py) inside triangle slot j's box from its row base. */
__device__ __forceinline__ void baryFromBase(const SrpTri& T, const float (*base)[8][3], uint32_t j, int px, int py,
                                             uint32_t tx, float* l)
{
	const int cs = ((int) (tx * 8u) > (int) T.x0) ? (int) (tx * 8u) : (int) T.x0;
	const int n = px - cs;
#pragma unroll
	for (int e = 0; e < 3; e++)
	{
		float v = base[j][py & 7][e];
		for (int k = 0; k < n; k++)
			v = v + T.dx[e];
		l[e] = v;
	}
}

/* The lowest (up to) 8 set bits of *rest, removed from it. */
__device__ __forceinline__ uint64_t takeRound(uint64_t& rest)
{
	uint64_t round = 0, u = rest;
	for (int k = 0; k < 8 && u; k++)
	{
		round |= u & (~u + 1ull);
		u &= u - 1ull;
	}
	rest = u;
	return round;
}

/* DEFER (a wave shader without depth output, single draws): fragments go
 * through every test but are not shaded; a pixel's colour register holds
 * the triangle record of the last fragment that wrote it, and the tile's
 * written pixels are appended to L.deferList for srp_resolve_kernel. */
template <class FS, bool LATE, bool WORDS, bool CULL, bool DEFER = false>
__device__ __forceinline__ void rasterTileBatched(const SrpRasterLaunch& L, const SrpDrawParams* P, uint32_t tile,
                                                  uint32_t gtile, uint32_t& covered, uint32_t& shaded)
{
	__shared__ uint32_t recIdx[64];   /* record of each window slot (records stay in L2) */
	__shared__ float base[8][8][3];
	__shared__ unsigned long long triMask[64];
	__shared__ uint32_t fragBase[64];
	__shared__ float fragL[3][RF];
	__shared__ uint8_t fragPix[RF];
	__shared__ uint8_t fragTri[RF];
	__shared__ uint32_t fragCol[RF];
	__shared__ float fragDep[RF];
	__shared__ uint8_t fragState[RF];

	const uint32_t lane = threadIdx.x;
	const uint32_t begin = L.tileStart[gtile], end = L.tileEnd[gtile];
	if (begin == end)
	{
		if (DEFER && lane == 0)
			L.deferCounts[gtile] = 0;
		return;
	}
	const uint32_t tx = tile % P->tilesX, ty = tile / P->tilesX;
	const int px = (int) (tx * 8 + (lane & 7u));
	const int py = (int) (ty * 8 + (lane >> 3));
	const bool inFb = px < (int) P->width && py < (int) P->height;
	const size_t idx = (size_t) py * P->width + (size_t) px;
	const uint64_t ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64u - lane));
	const SrpState& st = P->st;

	uint32_t col = 0, sten = 0;
	float dep = 0.0f;
	bool wrote = false;   /* DEFER: the draw wrote this pixel's colour */
	if (inFb)
	{
		col = P->color[idx];
		dep = P->depth[idx];
		sten = P->stencil[idx];
	}
	const uint32_t pv = st.provokingLast ? 2u : 0u;
	const bool precull = !LATE && st.depthTest && monotoneOp(st.depthOp);
	/* whole-tile depth culling of pixel-parallel runs: a culled fragment
	 * must have no effect at all, so no stencil (its dfail op) and no
	 * shader-written depth */
	const bool cull = CULL && L.depthCull && precull && !st.stencilEnabled;

	/* The tile's list is consumed through a window of 64 slots: slot s
	 * (lane s keeps its mask, recIdx[s] its record) holds list entry
	 * i + ((s - head) mod 64).  Each entry is loaded and its mask found once,
	 * however many batches it waits before its fragments fit. */
	uint32_t head = 0, filled = 0;
	uint64_t smask = 0;
	bool sknown = true;
	for (uint32_t i = begin; i < end;)
	{
		/* A: masks of the entries new to the window. */
		const uint32_t avail = (end - i < 64u) ? end - i : 64u;
		const uint32_t mine = (lane - head) & 63u;   /* this slot's position in the window */
		const bool fresh = mine >= filled && mine < avail;
		if (fresh)
		{
			const uint32_t pi = L.tileList[i + mine];
			const uint32_t pr = L.pairRec[pi];
			sknown = (pr >> 31) != 0;
			smask = sknown ? L.pairMask[pi] : 0ull;
			if (pr != 0xFFFFFFFFu)   /* a pair without a record (capacity exceeded) covers nothing */
				recIdx[lane] = pr & 0x7FFFFFFFu;
		}
		__syncthreads();
		/* Triangles without an exact mask: pixel-parallel coverage, 8
		 * triangles per round from shared row bases (bits are slots). */
		uint64_t unknown = L.bigPixel ? 0ull : __ballot(fresh && !sknown);
		while (unknown)
		{
			const uint64_t round = takeRound(unknown);
			rowBases(L.tris, recIdx, 0u, round, tx, ty, lane, base);
			__syncthreads();
			uint64_t r = round;
			for (uint32_t j = 0; r; j++)
			{
				const uint32_t b = (uint32_t) __builtin_ctzll(r);
				r &= r - 1ull;
				const SrpTri& T = L.tris[recIdx[b]];
				bool cov = px >= (int) T.x0 && px < (int) T.x1 && py >= (int) T.y0 && py < (int) T.y1;
				if (cov)
				{
					float l[3];
					baryFromBase(T, base, j, px, py, tx, l);
					cov = insideTri(T, l);
				}
				const uint64_t m = __ballot(cov);
				if (lane == b)
					smask = m;
			}
			__syncthreads();
		}
		filled = avail;
		/* Window order: lane b sees entry i + b. */
		const uint32_t src = (head + lane) & 63u;
		const uint32_t mlo = (uint32_t) __shfl((uint32_t) smask, (int) src);
		const uint32_t mhi = (uint32_t) __shfl((uint32_t) (smask >> 32), (int) src);
		const uint64_t mask = ((uint64_t) mhi << 32) | mlo;
		const bool known = __shfl((int) sknown, (int) src) != 0;
		uint32_t nb;
		const uint64_t unkWin = __ballot(lane < avail && !known);
		if (L.bigPixel && (unkWin & 1ull))
		{
			/* A run of triangles without exact masks heads the window:
			 * lane = pixel, the triangles in order, each fragment tested
			 * and shaded as the reference does, with no fragment list
			 * (these triangles cover much of the tile).  Triangles whose
			 * every fragment here must fail the depth test are dropped
			 * first (srpTileDepthCull); the run then takes window entries
			 * until it holds 8 triangles that remain. */
			const uint64_t notUnk = ~unkWin;
			const uint32_t runMax = notUnk ? (uint32_t) __builtin_ctzll(notUnk) : 64u;
			uint64_t keep = (runMax >= 64u) ? ~0ull : ((1ull << runMax) - 1ull);
			if (cull)
			{
				/* the tile's stored depth range (pixels inside the
				 * framebuffer); a run's culls use its value at the run's
				 * start, which is conservative because a monotone compare
				 * only writes depths that make later fragments fail more */
				float zMax = inFb ? dep : -__builtin_inff();
				float zMin = inFb ? dep : __builtin_inff();
				for (int off = 32; off > 0; off >>= 1)
				{
					zMax = fmaxf(zMax, __shfl_xor(zMax, off));
					zMin = fminf(zMin, __shfl_xor(zMin, off));
				}
				bool drop = false;
				if (lane < runMax)
					drop = srpTileDepthCull(&L.tris[recIdx[(head + lane) & 63u]], tx, ty, st.depthOp, zMin, zMax);
				keep &= ~__ballot(drop);
			}
			uint32_t run = runMax;
			{
				uint64_t k = keep;
				for (int c = 0; c < 8 && k; c++)
					k &= k - 1ull;
				if (k)
					run = (uint32_t) __builtin_ctzll(k);   /* the 9th remaining entry starts the next run */
			}
			const uint64_t keepRun = keep & ((run >= 64u) ? ~0ull : ((1ull << run) - 1ull));
			if (keepRun)
			{
				rowBases(L.tris, recIdx, head, keepRun, tx, ty, lane, base);
				__syncthreads();
			}
			uint32_t jb = 0;   /* row-base slot: rank of the triangle among those kept */
			for (uint64_t rest = keepRun; rest; rest &= rest - 1ull, jb++)
			{
				const uint32_t j = (uint32_t) __builtin_ctzll(rest);
				const SrpTri& T = L.tris[recIdx[(head + j) & 63u]];
				bool cov = px >= (int) T.x0 && px < (int) T.x1 && py >= (int) T.y0 && py < (int) T.y1;
				float l[3] = {0.0f, 0.0f, 0.0f};
				if (cov)
				{
					baryFromBase(T, base, jb, px, py, tx, l);
					cov = insideTri(T, l);
				}
				const bool front = (T.flags >> 3) & 1u;
				const SRPStencilFaceState& sf = st.stencil[front ? 0 : 1];
				const float iw[3] = {T.invW[0], T.invW[1], T.invW[2]};
				float depth = 0.0f, rec = 0.0f;
				bool active = false;
				if (cov)
				{
					covered++;
					rec = 1.0f / (iw[0] * l[0] + iw[1] * l[1] + iw[2] * l[2]);
					depth = T.zw[0] * l[0] + T.zw[1] * l[1] + T.zw[2] * l[2];
					active = true;
					if (st.scissorEnabled)
					{
						const uint64_t ux = (uint64_t) (int64_t) px, uy = (uint64_t) (int64_t) py;
						if (ux < st.sx0 || ux >= st.sx1 || uy < st.sy0 || uy >= st.sy1)
							active = false;
					}
					if (active && st.stencilEnabled &&
					    !compare<uint32_t>(sf.func, (uint32_t) (sf.ref & sf.mask), sten & sf.mask))
					{
						sten = stencilApply(sf.sfailOp, sten, sf);
						active = false;
					}
					if (active && !LATE && st.depthTest && !compare<float>(st.depthOp, depth, dep))
					{
						if (st.stencilEnabled) sten = stencilApply(sf.dfailOp, sten, sf);
						active = false;
					}
				}
				SRPFragmentShaderIn in;
				SRPFragmentShaderOut out;
				out.color[0] = 0.0f; out.color[1] = 0.0f; out.color[2] = 0.0f; out.color[3] = 0.0f;
				out.fragDepth = __builtin_nanf("");
				typedef typename VaryingBuf<WORDS>::type BufT;
				BufT buf[VaryingBuf<WORDS>::n];
				in.varyings = (SRPInterpolated*) buf;
				in.uniform = (SRPUniform*) P->uniform;
				in.fragCoord[0] = (float) ((double) px + 0.5);
				in.fragCoord[1] = (float) ((double) py + 0.5);
				in.fragCoord[2] = depth;
				in.fragCoord[3] = rec;
				in.frontFacing = front;
				in.primitiveID = T.id;
				if (active && !DEFER)
				{
					if (WORDS)
					{
						const uint32_t* v[3];
#pragma unroll
						for (int k = 0; k < 3; k++) v[k] = reinterpret_cast<const uint32_t*>(varyingPtr(P, T.vref[k]));
						interpWords<3>(P, v, l, iw, rec, pv, reinterpret_cast<uint32_t*>(buf));
					}
					else
					{
						const uint8_t* v[3];
#pragma unroll
						for (int k = 0; k < 3; k++) v[k] = varyingPtr(P, T.vref[k]);
						uint8_t* bb = reinterpret_cast<uint8_t*>(buf);
						for (uint32_t q = 0; q < P->describedSize; q++) bb[q] = 0;
						interpBytes<3>(P, v, l, iw, rec, pv, bb);
					}
				}
				if constexpr (DEFER)
				{
				}
				else if constexpr (FsIsWave<FS>::value)
				{
					if (!active)
						for (uint32_t q = 0; q < VaryingBuf<WORDS>::n; q++) buf[q] = 0;
					FS().wave(&in, &out, active);
				}
				else if (active)
					FS()(&in, &out);
				if (active)
				{
					shaded++;
					float d = depth;
					bool ok = true;
					if (LATE)
					{
						if (!isnan(out.fragDepth)) d = out.fragDepth;
						if (st.depthTest && !compare<float>(st.depthOp, d, dep))
						{
							if (st.stencilEnabled) sten = stencilApply(sf.dfailOp, sten, sf);
							ok = false;
						}
					}
					if (ok)
					{
						if (st.stencilEnabled) sten = stencilApply(sf.passOp, sten, sf);
						col = DEFER ? recIdx[(head + j) & 63u] : packColor(out.color);
						wrote = true;
						if (st.depthTest && st.depthWrite)
							dep = d;
					}
				}
			}
			__syncthreads();
			nb = run;
		}
		else
		{
		/* entries taken by the fragment-list path this time */
		const uint32_t lim = (L.bigPixel && unkWin) ? (uint32_t) __builtin_ctzll(unkWin) : avail;
		/* Inclusive scan of fragment counts; keep the longest prefix that fits. */
		const uint32_t cnt = (lane < lim) ? (uint32_t) __popcll(mask) : 0u;
		uint32_t incl = cnt;
		for (int off = 1; off < 64; off <<= 1)
		{
			const uint32_t v = __shfl_up(incl, off);
			if ((int) lane >= off) incl += v;
		}
		nb = (uint32_t) __popcll(__ballot(lane < lim && incl <= (uint32_t) RF));
		const uint32_t nFrag = __shfl(incl, (int) nb - 1);
		if (lane < nb)
		{
			triMask[lane] = mask;
			fragBase[lane] = incl - cnt;
		}
		__syncthreads();
		/* Fragment list: exact-mask triangles write their own entries;
		 * barycentrics of these are replayed in B. */
		if (lane < nb && known)
		{
			uint64_t m = mask;
			uint32_t pos = incl - cnt;
			while (m)
			{
				const uint32_t p = (uint32_t) __builtin_ctzll(m);
				m &= m - 1ull;
				fragPix[pos] = (uint8_t) p;
				fragTri[pos] = (uint8_t) lane;
				pos++;
			}
		}
		/* ...other triangles pixel-parallel, storing their barycentrics. */
		uint64_t unk = __ballot(lane < nb && !known);
		while (unk)
		{
			const uint64_t round = takeRound(unk);
			rowBases(L.tris, recIdx, head, round, tx, ty, lane, base);
			__syncthreads();
			uint64_t r = round;
			for (uint32_t j = 0; r; j++)
			{
			const uint32_t b = (uint32_t) __builtin_ctzll(r);
			r &= r - 1ull;
			const uint64_t m = triMask[b];
			if ((m >> lane) & 1ull)
			{
				const SrpTri& T = L.tris[recIdx[(head + b) & 63u]];
				float l[3];
				baryFromBase(T, base, j, px, py, tx, l);
				const uint32_t pos = fragBase[b] + (uint32_t) __popcll(m & ltMask);
				fragL[0][pos] = l[0];
				fragL[1][pos] = l[1];
				fragL[2][pos] = l[2];
				fragPix[pos] = (uint8_t) lane;
				fragTri[pos] = (uint8_t) b;
			}
			}
			__syncthreads();
		}
		const uint64_t knownLanes = __ballot(known);
		__syncthreads();

		/* B: shading */
		for (uint32_t f0 = 0; f0 < nFrag; f0 += 64)
		{
			const uint32_t f = f0 + lane;
			const uint32_t p = (f < nFrag) ? fragPix[f] : 0u;
			const float pixDepth = __shfl(dep, (int) p);
			uint8_t state = FRAG_CULLED;
			uint8_t facing = 0;   /* FRAG_FRONT when the fragment's triangle is front facing */
			float depth = 0.0f;
			SRPFragmentShaderIn in;
			SRPFragmentShaderOut out;
			out.color[0] = 0.0f; out.color[1] = 0.0f; out.color[2] = 0.0f; out.color[3] = 0.0f;
			out.fragDepth = __builtin_nanf("");
			typedef typename VaryingBuf<WORDS>::type BufT;
			BufT buf[VaryingBuf<WORDS>::n];
			in.varyings = (SRPInterpolated*) buf;
			uint32_t fragRec = 0;   /* the fragment's triangle record */
			if (f < nFrag)
			{
				const uint32_t b = fragTri[f];
				fragRec = recIdx[(head + b) & 63u];
				const SrpTri& T = L.tris[fragRec];
				facing = ((T.flags >> 3) & 1u) ? FRAG_FRONT : 0;
				const int fx = (int) (tx * 8 + (p & 7u)), fy = (int) (ty * 8 + (p >> 3));
				float l[3];
				if ((knownLanes >> b) & 1ull)
					baryAt(T, fx, fy, l);
				else
				{
					l[0] = fragL[0][f];
					l[1] = fragL[1][f];
					l[2] = fragL[2][f];
				}
				const float iw[3] = {T.invW[0], T.invW[1], T.invW[2]};
				const float rec = 1.0f / (iw[0] * l[0] + iw[1] * l[1] + iw[2] * l[2]);
				depth = T.zw[0] * l[0] + T.zw[1] * l[1] + T.zw[2] * l[2];
				state = FRAG_SHADED;
				if (st.scissorEnabled)
				{
					const uint64_t ux = (uint64_t) (int64_t) fx, uy = (uint64_t) (int64_t) fy;
					if (ux < st.sx0 || ux >= st.sx1 || uy < st.sy0 || uy >= st.sy1)
						state = FRAG_SCISSORED;
				}
				if (state == FRAG_SHADED && precull && !compare<float>(st.depthOp, depth, pixDepth))
					state = FRAG_CULLED;
				if (state == FRAG_SHADED && !DEFER)
				{
					in.uniform = (SRPUniform*) P->uniform;
					in.fragCoord[0] = (float) ((double) fx + 0.5);
					in.fragCoord[1] = (float) ((double) fy + 0.5);
					in.fragCoord[2] = depth;
					in.fragCoord[3] = rec;
					in.frontFacing = (T.flags >> 3) & 1u;
					in.primitiveID = T.id;
					if (WORDS)
					{
						const uint32_t* v[3];
#pragma unroll
						for (int k = 0; k < 3; k++) v[k] = reinterpret_cast<const uint32_t*>(varyingPtr(P, T.vref[k]));
						interpWords<3>(P, v, l, iw, rec, pv, reinterpret_cast<uint32_t*>(buf));
					}
					else
					{
						const uint8_t* v[3];
#pragma unroll
						for (int k = 0; k < 3; k++) v[k] = varyingPtr(P, T.vref[k]);
						uint8_t* bb = reinterpret_cast<uint8_t*>(buf);
						for (uint32_t q = 0; q < P->describedSize; q++) bb[q] = 0;
						interpBytes<3>(P, v, l, iw, rec, pv, bb);
					}
				}
			}
			const bool active = state == FRAG_SHADED;
			if constexpr (DEFER)
			{
			}
			else if constexpr (FsIsWave<FS>::value)
			{
				/* every lane calls; lanes without a fragment pass zeroed inputs */
				if (!active)
				{
					in.uniform = (SRPUniform*) P->uniform;
					in.fragCoord[0] = 0.0f; in.fragCoord[1] = 0.0f; in.fragCoord[2] = 0.0f; in.fragCoord[3] = 0.0f;
					in.frontFacing = 0;
					in.primitiveID = 0;
					for (uint32_t q = 0; q < VaryingBuf<WORDS>::n; q++) buf[q] = 0;
				}
				FS().wave(&in, &out, active);
			}
			else if (active)
				FS()(&in, &out);
			if (f < nFrag)
			{
				if (active)
				{
					if (LATE && !isnan(out.fragDepth))
						depth = out.fragDepth;
					fragCol[f] = DEFER ? fragRec : packColor(out.color);
				}
				fragDep[f] = depth;
				fragState[f] = state | facing;
			}
		}
		__syncthreads();

		/* C: ordered resolve.  The pixel's covering primitives, in order. */
		uint64_t pm = 0;
		for (uint32_t b = 0; b < nb; b++)
			pm |= ((triMask[b] >> lane) & 1ull) << b;
		while (pm)
		{
			const uint32_t b = (uint32_t) __builtin_ctzll(pm);
			pm &= pm - 1ull;
			const uint32_t f = fragBase[b] + (uint32_t) __popcll(triMask[b] & ltMask);
			covered++;
			const uint8_t state = fragState[f] & (uint8_t) ~FRAG_FRONT;
			if (state == FRAG_SCISSORED)
				continue;
			const bool front = (fragState[f] & FRAG_FRONT) != 0;
			const SRPStencilFaceState& sf = st.stencil[front ? 0 : 1];
			if (st.stencilEnabled && !compare<uint32_t>(sf.func, (uint32_t) (sf.ref & sf.mask), sten & sf.mask))
			{
				sten = stencilApply(sf.sfailOp, sten, sf);
				continue;
			}
			const float d = fragDep[f];
			if (LATE)
				shaded++;
			if (st.depthTest && !compare<float>(st.depthOp, d, dep))
			{
				if (st.stencilEnabled) sten = stencilApply(sf.dfailOp, sten, sf);
				continue;
			}
			if (!LATE)
				shaded++;
			if (st.stencilEnabled) sten = stencilApply(sf.passOp, sten, sf);
			col = fragCol[f];
			wrote = true;
			if (st.depthTest && st.depthWrite)
				dep = d;
		}
		__syncthreads();
		}
		i += nb;
		head = (head + nb) & 63u;
		filled -= nb;
	}
	if (inFb)
	{
		if (!DEFER)
			P->color[idx] = col;
		P->depth[idx] = dep;
		P->stencil[idx] = (uint8_t) sten;
	}
	if (DEFER)
	{
		/* the tile's written pixels in its own slots, and their number */
		const bool app = inFb && wrote;
		const uint64_t m = __ballot(app);
		if (lane == 0)
			L.deferCounts[gtile] = (uint32_t) __popcll(m);
		if (app)
		{
			const size_t at = (size_t) gtile * 64u + (size_t) __popcll(m & ltMask);
			L.deferList[2u * at] = (uint32_t) idx;
			L.deferList[2u * at + 1u] = col;
		}
	}
}

template <class FS, bool LATE, bool WORDS, int WAVES, bool CULL, bool DEFER = false>
__global__ __launch_bounds__(64, WAVES) void srp_raster_batched(SrpRasterLaunch L)
{
	const uint32_t tile = srpTileOfBlock(blockIdx.x, L.nBlocks, L.tileOrder);
	if (tile >= L.nTiles)
		return;
	uint32_t covered = 0, shaded = 0;
	rasterTileBatched<FS, LATE, WORDS, CULL, DEFER>(L, L.params, tile, tile, covered, shaded);
	flushStats(L.params, covered, shaded);
}

/* Deferred shading of a draw's written pixels (the tiles' slots of
 * L.deferList, pixel index and triangle record, counted and then offset by
 * L.deferCounts), in two kernels so that each runs at the occupancy it
 * needs:
 *   srp_defer_gather  rebuilds each listed pixel's fragment as the batched
 *                     rasteriser built it (barycentrics by the reference's
 *                     additions, depth, 1/w, the same interpolation) into
 *                     the word arrays of L.deferFrags (many waves, few
 *                     registers: the dependent loads overlap);
 *   srp_defer_shade   shades them with the wave entry, 64 per wave, and
 *                     writes the colour as the rasteriser would have. */
__device__ __forceinline__ uint32_t deferWords(const SrpDrawParams* P, bool words)
{
	return words ? P->nWords : (P->describedSize + 3u) / 4u;
}

template <bool WORDS>
__global__ __launch_bounds__(256) void srp_defer_gather(SrpRasterLaunch L)
{
	const SrpDrawParams* P = L.params;
	const uint32_t pv = P->st.provokingLast ? 2u : 0u;
	const uint32_t width = P->width;
	const size_t S = L.deferStride;
	const uint32_t nw = deferWords(P, WORDS);
	uint32_t* F = L.deferFrags;
	const uint32_t slots = L.nTiles * 64u;
	for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < slots; s += gridDim.x * 256u)
	{
		/* slot r of tile t holds entry off[t] + r of the compacted list */
		const uint32_t t = s >> 6, r = s & 63u;
		const uint32_t off = L.deferCounts[t];
		if (r >= L.deferCounts[t + 1] - off)
			continue;
		const uint32_t k = off + r;
		const uint32_t pix = L.deferList[2u * s];
		const SrpTri& T = L.tris[L.deferList[2u * s + 1u]];
		const int px = (int) (pix % width), py = (int) (pix / width);
		float l[3];
		baryAt(T, px, py, l);
		const float iw[3] = {T.invW[0], T.invW[1], T.invW[2]};
		const float rec = 1.0f / (iw[0] * l[0] + iw[1] * l[1] + iw[2] * l[2]);
		F[0 * S + k] = __float_as_uint((float) ((double) px + 0.5));
		F[1 * S + k] = __float_as_uint((float) ((double) py + 0.5));
		F[2 * S + k] = __float_as_uint(T.zw[0] * l[0] + T.zw[1] * l[1] + T.zw[2] * l[2]);
		F[3 * S + k] = __float_as_uint(rec);
		F[4 * S + k] = (T.flags >> 3) & 1u;
		F[5 * S + k] = T.id;
		F[6 * S + k] = pix;
		if (WORDS)
		{
			const uint32_t* v[3];
#pragma unroll
			for (int q = 0; q < 3; q++) v[q] = reinterpret_cast<const uint32_t*>(varyingPtr(P, T.vref[q]));
			uint32_t buf[SRP_MAX_WORDS];
			interpWords<3>(P, v, l, iw, rec, pv, buf);
#pragma unroll
			for (uint32_t w = 0; w < SRP_MAX_WORDS; w++)
				if (w < nw) F[(7 + w) * S + k] = buf[w];
		}
		else
		{
			const uint8_t* v[3];
#pragma unroll
			for (int q = 0; q < 3; q++) v[q] = varyingPtr(P, T.vref[q]);
			uint32_t buf[SRP_MAX_VARYING_BYTES / 4];
			for (uint32_t w = 0; w < nw; w++) buf[w] = 0;
			interpBytes<3>(P, v, l, iw, rec, pv, reinterpret_cast<uint8_t*>(buf));
			for (uint32_t w = 0; w < nw; w++) F[(7 + w) * S + k] = buf[w];
		}
	}
}

template <class FS, bool WORDS>
__global__ __launch_bounds__(256, 2) void srp_defer_shade(SrpRasterLaunch L)
{
	const SrpDrawParams* P = L.params;
	const uint32_t n = L.deferCounts[L.nTiles];
	const uint32_t lane = threadIdx.x & 63u;
	const size_t S = L.deferStride;
	const uint32_t nw = deferWords(P, WORDS);
	const uint32_t* F = L.deferFrags;
	for (uint32_t base = (blockIdx.x * 4u + (threadIdx.x >> 6)) * 64u; base < n; base += gridDim.x * 256u)
	{
		const uint32_t k = base + lane;
		const bool active = k < n;
		SRPFragmentShaderIn in;
		SRPFragmentShaderOut out;
		out.color[0] = 0.0f; out.color[1] = 0.0f; out.color[2] = 0.0f; out.color[3] = 0.0f;
		out.fragDepth = __builtin_nanf("");
		typedef typename VaryingBuf<WORDS>::type BufT;
		BufT buf[VaryingBuf<WORDS>::n];
		for (uint32_t q = 0; q < VaryingBuf<WORDS>::n; q++)
			buf[q] = (active && q < nw) ? F[(7 + q) * S + k] : 0u;
		in.varyings = (SRPInterpolated*) buf;
		in.uniform = (SRPUniform*) P->uniform;
		in.fragCoord[0] = active ? __uint_as_float(F[0 * S + k]) : 0.0f;
		in.fragCoord[1] = active ? __uint_as_float(F[1 * S + k]) : 0.0f;
		in.fragCoord[2] = active ? __uint_as_float(F[2 * S + k]) : 0.0f;
		in.fragCoord[3] = active ? __uint_as_float(F[3 * S + k]) : 0.0f;
		in.frontFacing = active ? F[4 * S + k] : 0u;
		in.primitiveID = active ? F[5 * S + k] : 0u;
		FS().wave(&in, &out, active);
		if (active)
			P->color[F[6 * S + k]] = packColor(out.color);
	}
}

/* Persistent multi-draw rasteriser: L.nBlocks waves take tiles of all the
 * draws' framebuffers from device queues until none is left (every wave
 * ends when it finds all queues drained).  With xcdQueues the tiles are
 * split into 8 contiguous ranges, one per XCD (blocks are dispatched to
 * XCDs round-robin), and a wave that drains its own range moves on to the
 * next; otherwise one queue serves all. */
template <class FS, bool LATE, bool WORDS, int WAVES, bool CULL>
__global__ __launch_bounds__(64, WAVES) void srp_raster_persistent(SrpRasterLaunch L)
{
	const uint32_t nq = L.xcdQueues ? 8u : 1u;
	const uint32_t q0 = blockIdx.x % nq;
	uint32_t covered = 0, shaded = 0;
	const SrpDrawParams* last = nullptr;
	for (uint32_t k = 0; k < nq; k++)
	{
		const uint32_t q = (q0 + k) % nq;
		const uint32_t lo = (uint32_t) (((uint64_t) L.nTiles * q) / nq);
		const uint32_t hi = (uint32_t) (((uint64_t) L.nTiles * (q + 1u)) / nq);
		for (;;)
		{
			uint32_t t = 0;
			if (threadIdx.x == 0)
				t = atomicAdd(&L.queues[q], 1u);
			t = __shfl(t, 0) + lo;
			if (t >= hi)
				break;
			/* the draw owning global tile t */
			uint32_t a = 0, b = L.nSeg;
			while (b - a > 1)
			{
				const uint32_t m = (a + b) >> 1;
				if (L.seg[m].first <= t) a = m;
				else b = m;
			}
			const SrpSeg sg = L.seg[a];
			rasterTileBatched<FS, LATE, WORDS, CULL>(L, sg.P, t - sg.first, t, covered, shaded);
			last = sg.P;
		}
	}
	/* every draw of a batch shares one counter array */
	if (last)
		flushStats(last, covered, shaded);
}

/* Lines and points: fragments sorted (stably) by pixel; the first thread
 * of each pixel's run applies them in primitive order. */
template <class FS, bool LATE, bool WORDS>
__global__ __launch_bounds__(256) void srp_frag_kernel(SrpFragLaunch L)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	const SrpDrawParams* P = L.params;
	PixelState ps;
	ps.color = 0; ps.depth = 0.0f; ps.stencil = 0; ps.covered = 0; ps.shaded = 0;
	const uint32_t npix = P->width * P->height;
	uint32_t count = L.count;
	if (L.countDev && *L.countDev < count)
		count = *L.countDev;
	uint32_t key = 0xFFFFFFFFu;
	bool head = false;
	if (i < count)
	{
		key = L.keys[i];
		head = key < npix && (i == 0 || L.keys[i - 1] != key);
	}
	if (head)
	{
		ps.color = P->color[key];
		ps.depth = P->depth[key];
		ps.stencil = P->stencil[key];
		const uint32_t pv = P->st.provokingLast ? 1u : 0u;
		for (uint32_t j = i; j < count && L.keys[j] == key; j++)
		{
			const SrpFrag f = L.frags[L.order[j]];
			if (L.kind == 0)
			{
				const SrpLine* Ln = &L.lines[f.prim];
				const float w[2] = {1.0f - f.t, f.t};
				const float iw[2] = {Ln->invW[0], Ln->invW[1]};
				const float rec = 1.0f / (iw[0] * w[0] + iw[1] * w[1]);
				const float depth = Ln->zw[0] * w[0] + Ln->zw[1] * w[1];
				if (WORDS)
				{
					uint32_t buf[SRP_MAX_WORDS];
					fragment<FS, LATE>(P, f.px, f.py, true, depth, rec, Ln->id, ps, [&]() {
						const uint32_t* v[2];
						for (int k = 0; k < 2; k++)
							v[k] = reinterpret_cast<const uint32_t*>(varyingPtr(P, Ln->vref[k]));
						interpWords<2>(P, v, w, iw, rec, pv, buf);
						return (SRPInterpolated*) buf;
					});
				}
				else
				{
					uint8_t buf[SRP_MAX_VARYING_BYTES];
					fragment<FS, LATE>(P, f.px, f.py, true, depth, rec, Ln->id, ps, [&]() {
						const uint8_t* v[2];
						for (int k = 0; k < 2; k++) v[k] = varyingPtr(P, Ln->vref[k]);
						for (uint32_t b = 0; b < P->describedSize; b++) buf[b] = 0;
						interpBytes<2>(P, v, w, iw, rec, pv, buf);
						return (SRPInterpolated*) buf;
					});
				}
			}
			else
			{
				/* reference raster/point.c:64-71: varyings passed through */
				const SrpPoint* Pt = &L.points[f.prim];
				fragment<FS, LATE>(P, f.px, f.py, true, Pt->z, Pt->w, Pt->id, ps, [&]() {
					return (SRPInterpolated*) varyingPtr(P, Pt->vref);
				});
			}
		}
		P->color[key] = ps.color;
		P->depth[key] = ps.depth;
		P->stencil[key] = (uint8_t) ps.stencil;
	}
	flushStats(P, ps.covered, ps.shaded);
}

/* ------------------------------------------------------------ launchers */

template <class VS>
void launchVS(const SrpVsLaunch* L, void* stream)
{
	const uint32_t blocks = (L->nSeg ? L->nBlocks : (L->count + 255u) / 256u) + L->fillBlocks;
	if (blocks)
		hipLaunchKernelGGL(srp_vs_kernel<VS>, dim3(blocks), dim3(256), 0, (hipStream_t) stream, *L);
}

template <class FS, bool LATE, bool WORDS>
void launchRasterT(const SrpRasterLaunch* L, hipStream_t s)
{
	if (L->nSeg)
	{
		/* batches: 16x16 tiles (8x8 ones use the batched kernel) */
		hipLaunchKernelGGL((srp_raster_persistent16<FS, LATE, WORDS>), dim3(L->nBlocks), dim3(256), 0, s, *L);
		return;
	}
	if (L->tile == 16)
		hipLaunchKernelGGL((srp_raster_kernel<FS, LATE, WORDS, 16>), dim3(L->nBlocks), dim3(256), 0, s, *L);
	else
		hipLaunchKernelGGL((srp_raster_kernel<FS, LATE, WORDS, 8>), dim3(L->nBlocks), dim3(256), 0, s, *L);
}

/* Depth culling has its own kernels (CULL), so that draws it cannot help
 * (stencil on, late depth, culling off) run kernels without its registers. */
template <class FS, bool LATE, bool WORDS>
void launchBatchedT(const SrpRasterLaunch* L, hipStream_t s)
{
	constexpr int W = FsIsWave<FS>::value ? 2 : SRP_RASTER_WAVES;
	const bool cull = !LATE && L->depthCull;
	if constexpr (FsIsWave<FS>::value && !LATE)
		if (L->deferList && !L->nSeg)
		{
			/* no shading here: the registers of a plain shader's kernel */
			constexpr int WD = SRP_RASTER_WAVES;
			if (cull)
				hipLaunchKernelGGL((srp_raster_batched<FS, false, WORDS, WD, true, true>), dim3(L->nBlocks), dim3(64), 0, s, *L);
			else
				hipLaunchKernelGGL((srp_raster_batched<FS, false, WORDS, WD, false, true>), dim3(L->nBlocks), dim3(64), 0, s, *L);
			return;
		}
	if (L->nSeg)
	{
		if (cull)
			hipLaunchKernelGGL((srp_raster_persistent<FS, LATE, WORDS, W, !LATE>), dim3(L->nBlocks), dim3(64), 0, s, *L);
		else
			hipLaunchKernelGGL((srp_raster_persistent<FS, LATE, WORDS, W, false>), dim3(L->nBlocks), dim3(64), 0, s, *L);
	}
	else
	{
		if (cull)
			hipLaunchKernelGGL((srp_raster_batched<FS, LATE, WORDS, W, !LATE>), dim3(L->nBlocks), dim3(64), 0, s, *L);
		else
			hipLaunchKernelGGL((srp_raster_batched<FS, LATE, WORDS, W, false>), dim3(L->nBlocks), dim3(64), 0, s, *L);
	}
}

template <class FS>
void launchRaster(const SrpRasterLaunch* L, void* stream)
{
	hipStream_t s = (hipStream_t) stream;
	if (L->nBlocks == 0) return;
	if (L->batched)
	{
		if (L->lateZ)
		{
			if (L->words) launchBatchedT<FS, true, true>(L, s);
			else launchBatchedT<FS, true, false>(L, s);
		}
		else
		{
			if (L->words) launchBatchedT<FS, false, true>(L, s);
			else launchBatchedT<FS, false, false>(L, s);
		}
		return;
	}
	if (L->lateZ)
	{
		if (L->words) launchRasterT<FS, true, true>(L, s);
		else launchRasterT<FS, true, false>(L, s);
	}
	else
	{
		if (L->words) launchRasterT<FS, false, true>(L, s);
		else launchRasterT<FS, false, false>(L, s);
	}
}

/* Deferred shading after a raster launch with a list (wave shaders only):
 * grids sized for every pixel of the framebuffer listed (the kernels stride
 * over the device-side count). */
template <class FS>
void launchResolve(const SrpRasterLaunch* L, void* stream)
{
	hipStream_t s = (hipStream_t) stream;
	if (!L->deferList || !L->deferCounts || !L->deferFrags || L->nTiles == 0)
		return;
	const uint32_t gather = L->nTiles / 4u + 1u < 4096u ? L->nTiles / 4u + 1u : 4096u;
	const uint32_t shade = L->nBlocks < 1024u ? L->nBlocks : 1024u;
	if (L->words)
	{
		hipLaunchKernelGGL((srp_defer_gather<true>), dim3(gather), dim3(256), 0, s, *L);
		hipLaunchKernelGGL((srp_defer_shade<FS, true>), dim3(shade), dim3(256), 0, s, *L);
	}
	else
	{
		hipLaunchKernelGGL((srp_defer_gather<false>), dim3(gather), dim3(256), 0, s, *L);
		hipLaunchKernelGGL((srp_defer_shade<FS, false>), dim3(shade), dim3(256), 0, s, *L);
	}
}

template <class FS>
void launchFrags(const SrpFragLaunch* L, void* stream)
{
	hipStream_t s = (hipStream_t) stream;
	const uint32_t blocks = (L->count + 255u) / 256u;
	if (!blocks) return;
	if (L->lateZ)
	{
		if (L->words) hipLaunchKernelGGL((srp_frag_kernel<FS, true, true>), dim3(blocks), dim3(256), 0, s, *L);
		else hipLaunchKernelGGL((srp_frag_kernel<FS, true, false>), dim3(blocks), dim3(256), 0, s, *L);
	}
	else
	{
		if (L->words) hipLaunchKernelGGL((srp_frag_kernel<FS, false, true>), dim3(blocks), dim3(256), 0, s, *L);
		else hipLaunchKernelGGL((srp_frag_kernel<FS, false, false>), dim3(blocks), dim3(256), 0, s, *L);
	}
}

}  // namespace dev
}  // namespace srp
