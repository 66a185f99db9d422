/* Primitive assembly on the GPU: topology, index fetch, clipping, setup.
 *
 * One thread per input primitive, in two passes over the same code:
 *   COUNT: records, clip-generated vertices and binning items (tile pairs
 *          for filled triangles, fragments for lines and points) per input;
 *   EMIT:  after exclusive scans of those counts, writes the records at
 *          their final positions (so primitive IDs equal record indices,
 *          matching the reference's sequential numbering after clipping
 *          and culling, pipeline/primitive_assembly.c:87-120), materialises
 *          the varyings of clip-generated vertices and writes the items.
 *
 * Arithmetic restates the reference (kitrofimov/srp) operation by operation:
 * clipping pipeline/clipping.c:68-277, perspective divide
 * pipeline/vertex_processing.c:76-88, triangle setup raster/triangle.c:113-238,
 * line setup and DDA raster/line.c:34-86, points raster/point.c:32-118,
 * NDC to screen core/framebuffer.c:48-55. */
#include "kernels.h"

#pragma clang fp contract(off)

namespace srp {
namespace lib {

using srp::dev::roughlyZero;

/* x86 cvttsd2si semantics: NaN and out-of-range give INT_MIN. */
__device__ __forceinline__ int cvtI32(double d)
{
	if (!(d >= -2147483648.0 && d < 2147483648.0))
		return (int) 0x80000000u;
	return (int) d;
}

/* clipping.c:260-277 */
__device__ __forceinline__ float planeDist(const float* c, int p)
{
	switch (p)
	{
		case 0: return c[0] + c[3];
		case 1: return c[3] - c[0];
		case 2: return c[1] + c[3];
		case 3: return c[3] - c[1];
		case 4: return c[2] + c[3];
		default: return c[3] - c[2];
	}
}

__device__ __forceinline__ uint32_t outcode(const float* c)
{
	uint32_t code = 0;
#pragma unroll
	for (int p = 0; p < 6; p++)
		code |= (planeDist(c, p) < 0.0f) ? (1u << p) : 0u;
	return code;
}

/* framebuffer.c:48-55 */
__device__ __forceinline__ void toScreen(const SrpDrawParams* P, float nx, float ny, float& sx, float& sy)
{
	sx = (float) (((double) (float) P->width / 2.0) * ((double) nx + 1.0));
	sy = (float) (((double) (float) P->height / 2.0) * (1.0 - (double) ny));
}

/* vertex_processing.c:76-88 */
__device__ __forceinline__ float divide(float c[4])
{
	const float invW = (float) (1.0 / (double) c[3]);
	c[0] *= invW;
	c[1] *= invW;
	c[2] *= invW;
	c[3] = 1.0f;
	return invW;
}

__device__ __forceinline__ uint64_t loadIndex(const uint8_t* ib, uint32_t elem, uint64_t i)
{
	switch (elem)
	{
		case 1: return ib[i];
		case 2: return reinterpret_cast<const uint16_t*>(ib)[i];
		case 4: return reinterpret_cast<const uint32_t*>(ib)[i];
		default: return reinterpret_cast<const uint64_t*>(ib)[i];
	}
}

__device__ __forceinline__ uint32_t slotOf(const AsmDraw& W, uint64_t stream)
{
	const uint64_t vi = W.indexElem ? loadIndex(W.indices, W.indexElem, stream) : stream;
	const uint64_t slot = vi - W.P->cacheBase;
	return slot < (uint64_t) W.P->cacheSize ? (uint32_t) slot : 0xFFFFFFFFu;
}

/* A vertex index outside the range the draw's post-VS cache was prepared
 * for (only possible when a batch's index data changed after it was made):
 * the primitive is dropped and error bit 2 is raised, never read out of
 * bounds. */
__device__ __forceinline__ bool slotsValid(const AsmArgs& A, const SrpDrawParams* P, const uint32_t* slot, int n)
{
	bool ok = true;
	for (int i = 0; i < n; i++)
		ok = ok && slot[i] < P->cacheSize;
	if (!ok)
		atomicOr(A.error, 4u);
	return ok;
}

/* The draw a block of a (possibly multi-draw) assembly launch works on:
 * the last draw whose first block is at or before it. */
__device__ __forceinline__ uint32_t drawOfBlock(const AsmDraw* D, uint32_t n, uint32_t block)
{
	uint32_t lo = 0, hi = n;
	while (hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (D[mid].firstBlock <= block) lo = mid;
		else hi = mid;
	}
	return lo;
}

__device__ __forceinline__ void loadClip(const SrpDrawParams* P, uint32_t slot, float* c)
{
	const float4 v = reinterpret_cast<const float4*>(P->clip)[slot];
	c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
}

/* ------------------------------------------------------- clip varyings */

/* Varyings of a vertex generated at parameter t between vertices whose
 * varyings are va and vb, by the clip rule (interpolateAttributes with no
 * 1/w, interpolation.c:116-120, weights {1-t, t}, clipping.c:244-258):
 * every float and double element blended affinely, FLAT ones included;
 * other types copied from the provoking end (vb for LAST, va for FIRST). */
__device__ void lerpVaryings(const SrpDrawParams* P, const uint8_t* va, const uint8_t* vb, float t, uint8_t* dst)
{
	const float w0 = 1 - t;
	for (uint32_t a = 0; a < P->nVaryings; a++)
	{
		const SrpVaryingDesc d = P->var[a];
		if (d.type == SRP_FLOAT)
			for (uint32_t e = 0; e < d.nItems; e++)
			{
				const uint32_t off = d.offset + 4 * e;
				float v = 0.0f;
				v += srp::dev::loadF(va + off) * w0;
				v += srp::dev::loadF(vb + off) * t;
				__builtin_memcpy(dst + off, &v, 4);
			}
		else if (d.type == SRP_DOUBLE)
			for (uint32_t e = 0; e < d.nItems; e++)
			{
				const uint32_t off = d.offset + 8 * e;
				double v = 0.0;
				v += srp::dev::loadD(va + off) * (double) w0;
				v += srp::dev::loadD(vb + off) * (double) t;
				__builtin_memcpy(dst + off, &v, 8);
			}
		else
		{
			const uint32_t n = srp::dev::elemSize(d.type) * d.nItems;
			const uint8_t* src = (P->st.provokingLast ? vb : va) + d.offset;
			for (uint32_t b = 0; b < n; b++) dst[d.offset + b] = src[b];
		}
	}
}

/* ------------------------------------------------------------------ setup */

struct Vtx
{
	float c[4];      /* ndc after the divide */
	float invW;
	uint32_t vref;
};

__device__ __forceinline__ float minf_(float a, float b) { return a > b ? b : a; }
__device__ __forceinline__ float maxf_(float a, float b) { return a > b ? a : b; }

/* triangle.c:113-238.  Returns false when culled or degenerate. */
__device__ bool setupTriangle(const SrpDrawParams* P, Vtx v0, Vtx v1, Vtx v2, uint32_t id, SrpTri& T)
{
	const SrpState& st = P->st;
	const float e0x = v1.c[0] - v0.c[0], e0y = v1.c[1] - v0.c[1];
	const float e1x = v2.c[0] - v0.c[0], e1y = v2.c[1] - v0.c[1];
	const float area = e0x * e1y - e0y * e1x;
	const bool ccw = area > 0.0f;
	if (st.cullFace == SRP_FACE_FRONT_AND_BACK)
		return false;
	const bool front = ((area > 0.0f) && st.frontFace == SRP_WINDING_CCW) || ((area < 0.0f) && st.frontFace == SRP_WINDING_CW);
	const bool cull = (front && st.cullFace == SRP_FACE_FRONT) || (!front && st.cullFace == SRP_FACE_BACK);
	if (cull)
		return false;
	if (!ccw)
	{
		const Vtx t = v1; v1 = v2; v2 = t;
	}
	const Vtx v[3] = {v0, v1, v2};
	float sx[3], sy[3];
	for (int i = 0; i < 3; i++) toScreen(P, v[i].c[0], v[i].c[1], sx[i], sy[i]);
	float ex[3], ey[3];
	for (int i = 0; i < 3; i++)
	{
		ex[i] = sx[(i + 1) % 3] - sx[i];
		ey[i] = sy[(i + 1) % 3] - sy[i];
	}
	const float areaX2 = fabsf(ex[0] * ey[2] - ey[0] * ex[2]);
	if (roughlyZero(areaX2))
		return false;

	const float mnx = minf_(sx[0], minf_(sx[1], sx[2])), mny = minf_(sy[0], minf_(sy[1], sy[2]));
	const float mxx = maxf_(sx[0], maxf_(sx[1], sx[2])), mxy = maxf_(sy[0], maxf_(sy[1], sy[2]));
	const double fx = floor((double) mnx), fy = floor((double) mny);
	const double cx = ceil((double) mxx), cy = ceil((double) mxy);
	const float minX = (float) (fx > 0.0 ? fx : 0.0);
	const float minY = (float) (fy > 0.0 ? fy : 0.0);
	const float maxX = (float) (cx > (double) P->width ? (double) P->width : cx);
	const float maxY = (float) (cy > (double) P->height ? (double) P->height : cy);

	const float px = (float) ((double) minX + 0.5), py = (float) ((double) minY + 0.5);
	for (int i = 0; i < 3; i++)
	{
		const int j = (i + 1) % 3;
		const float rx = px - sx[j], ry = py - sy[j];
		T.l0[i] = (rx * ey[j] - ry * ex[j]) / areaX2;
		T.dx[i] = ey[j] / areaX2;
		T.dy[i] = -ex[j] / areaX2;
	}
	uint32_t flags = front ? 8u : 0u;
	for (int i = 0; i < 3; i++)
		if (((ex[i] > 0.0f) && roughlyZero(ey[i])) || (ey[i] < 0.0f))
			flags |= 1u << i;
	/* The reference's scan runs while (float) x < maxX from (size_t) minX;
	 * a NaN or inverted box scans nothing. */
	const int x0 = (int) minX, y0 = (int) minY;
	const int x1 = (maxX > minX) ? (int) maxX : x0;
	const int y1 = (maxY > minY) ? (int) maxY : y0;
	T.x0 = (uint16_t) x0; T.y0 = (uint16_t) y0;
	T.x1 = (uint16_t) x1; T.y1 = (uint16_t) y1;
	T.flags = flags;
	T.id = id;
	for (int i = 0; i < 3; i++)
	{
		T.invW[i] = v[i].invW;
		T.zw[i] = v[i].c[2] * v[i].invW;
		T.vref[i] = v[i].vref;
	}
	T.pad = 0;
	return true;
}

/* Tiles a triangle is binned to.  A triangle whose bounding box is at
 * most P->maskBox (<= SRP_SMALL_BBOX) pixels on each side is scanned here exactly as the
 * reference scans it (raster/triangle.c:78-110: rows from the box origin,
 * dldy per row, dldx per pixel), giving one 64-bit coverage mask per 8x8
 * tile; tiles it does not cover get no entry (an empty entry would cost the
 * rasteriser a window slot, and sliver fans such as a sphere's poles pile
 * hundreds of them onto a few tiles).  Larger triangles are binned to every
 * tile of their box and their coverage is found by the rasteriser.  Calls
 * emit(tileKey, mask, known) in row-major tile order.  The counting pass
 * (EMIT false) needs only which tiles are covered: once every tile of a
 * band has a pixel, the band's remaining rows are stepped but not scanned. */
template <bool EMIT, class F>
__device__ __forceinline__ uint32_t binTriangle(const SrpDrawParams* P, const SrpTri& T, uint32_t tileBase, F&& emit)
{
	if (T.x1 <= T.x0 || T.y1 <= T.y0)
		return 0;
	const uint32_t sh = P->tileShift;
	const uint32_t tx0 = T.x0 >> sh, tx1 = (T.x1 - 1u) >> sh;
	const uint32_t ty0 = T.y0 >> sh, ty1 = (T.y1 - 1u) >> sh;
	const int bw = (int) T.x1 - (int) T.x0, bh = (int) T.y1 - (int) T.y0;
	uint32_t n = 0;
	if (sh == 3 && bw <= (int) P->maskBox && bh <= (int) P->maskBox)
	{
		/* One band of tile rows at a time; the (at most four) tiles of a
		 * band keep their masks in registers.  The walk itself is the
		 * reference's row-major scan of the whole box. */
		float row[3] = {T.l0[0], T.l0[1], T.l0[2]};
		float thr[3];
		srp::dev::edgeThresholds(T.flags, thr);
		int y = T.y0;
		const uint32_t nCols = tx1 - tx0 + 1u;
		const uint32_t allCols = (1u << nCols) - 1u;
		for (uint32_t ty = ty0; ty <= ty1; ty++)
		{
			uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
			const int yEnd = ((int) (ty + 1u) << 3) < (int) T.y1 ? (int) (ty + 1u) << 3 : (int) T.y1;
			for (; y < yEnd; y++)
			{
				const uint32_t found = (m0 != 0ull ? 1u : 0u) | (m1 != 0ull ? 2u : 0u) | (m2 != 0ull ? 4u : 0u) |
				                       (m3 != 0ull ? 8u : 0u);
				const int xEnd = (!EMIT && found == allCols) ? (int) T.x0 : (int) T.x1;
				float l[3] = {row[0], row[1], row[2]};
				const uint32_t rowShift = ((uint32_t) y & 7u) << 3;
				int x = T.x0;
				/* the row tile by tile: a byte of bits per tile, placed in
				 * that tile's mask once the tile's part of the row is done */
				for (uint32_t c = 0; c < nCols && x < xEnd; c++)
				{
					const int segEnd = ((int) ((tx0 + c + 1u) << 3) < xEnd) ? (int) ((tx0 + c + 1u) << 3) : xEnd;
					uint32_t bits = 0;
					for (; x < segEnd; x++)
					{
						bits |= (srp::dev::insideEdges(l, thr) ? 1u : 0u) << ((uint32_t) x & 7u);
						for (int e = 0; e < 3; e++) l[e] += T.dx[e];
					}
					const uint64_t add = (uint64_t) bits << rowShift;
					m0 |= c == 0 ? add : 0ull;
					m1 |= c == 1 ? add : 0ull;
					m2 |= c == 2 ? add : 0ull;
					m3 |= c == 3 ? add : 0ull;
				}
				for (int e = 0; e < 3; e++) row[e] += T.dy[e];
			}
			const uint64_t ms[4] = {m0, m1, m2, m3};
#pragma unroll
			for (uint32_t c = 0; c < 4; c++)
				if (tx0 + c <= tx1 && ms[c])
				{
					emit(tileBase + ty * P->tilesX + tx0 + c, ms[c], true);
					n++;
				}
		}
		return n;
	}
	/* every tile of the box, without a mask: written by k_expand_spans */
	return (tx1 - tx0 + 1u) * (ty1 - ty0 + 1u);
}

/* True when binTriangle gives T one unmasked entry per tile of its box
 * (a span) rather than exact masks. */
__device__ __forceinline__ bool isSpan(const SrpDrawParams* P, const SrpTri& T)
{
	if (T.x1 <= T.x0 || T.y1 <= T.y0)
		return false;
	const int bw = (int) T.x1 - (int) T.x0, bh = (int) T.y1 - (int) T.y0;
	return !(P->tileShift == 3 && bw <= (int) P->maskBox && bh <= (int) P->maskBox);
}

/* Reserve a span-list entry (one atomic per group of lanes calling
 * together) and fill it: the pairs of T's box tiles, items item.. in
 * row-major tile order, record `rec` (0xFFFFFFFF: none). */
__device__ void appendSpan(const AsmArgs& A, const SrpDrawParams* P, const SrpTri& T, uint32_t tileBase, uint32_t item,
                           uint32_t rec)
{
	const uint64_t act = __ballot(1);
	const uint32_t lane = __lane_id();
	const uint32_t leader = (uint32_t) __builtin_ctzll(act);
	uint32_t at = 0;
	if (lane == leader)
		at = atomicAdd(A.spanCount, (uint32_t) __popcll(act));
	at = (uint32_t) __shfl((int) at, (int) leader) + (uint32_t) __popcll(act & ((1ull << lane) - 1ull));
	if (at >= A.spanCap)
	{
		atomicOr(A.error, 2u);
		return;
	}
	const uint32_t sh = P->tileShift;
	const uint32_t tx0 = T.x0 >> sh, tx1 = (T.x1 - 1u) >> sh;
	const uint32_t ty0 = T.y0 >> sh, ty1 = (T.y1 - 1u) >> sh;
	SrpSpan sp;
	sp.item = item;
	sp.rec = rec;
	sp.keyBase = tileBase + ty0 * P->tilesX + tx0;
	sp.stride = P->tilesX;
	sp.nx = tx1 - tx0 + 1u;
	sp.ny = ty1 - ty0 + 1u;
	A.spans[at] = sp;
}

/* line.c:79-86 */
__device__ void setupLine(const SrpDrawParams* P, const Vtx& a, const Vtx& b, uint32_t id, SrpLine& L)
{
	toScreen(P, a.c[0], a.c[1], L.ssx[0], L.ssy[0]);
	toScreen(P, b.c[0], b.c[1], L.ssx[1], L.ssy[1]);
	L.invW[0] = a.invW; L.invW[1] = b.invW;
	L.zw[0] = a.c[2] * a.invW; L.zw[1] = b.c[2] * b.invW;
	L.vref[0] = a.vref; L.vref[1] = b.vref;
	L.id = id;
	L.pad = 0;
}

/* line.c:41-46: number of DDA samples (steps + 1); 0 when steps is not finite. */
__device__ __forceinline__ int lineSteps(const SrpLine& L)
{
	const float dx = L.ssx[1] - L.ssx[0], dy = L.ssy[1] - L.ssy[0];
	int steps = cvtI32(ceil(fmax(fabs((double) dx), fabs((double) dy))));
	if (steps == 0) steps = 1;
	return steps;
}

__device__ __forceinline__ uint32_t lineFragCount(const SrpLine& L)
{
	const int steps = lineSteps(L);
	return steps >= 0 ? (uint32_t) steps + 1u : 0u;
}

__device__ __forceinline__ uint32_t pixelKey(const SrpDrawParams* P, int px, int py)
{
	if (px < 0 || py < 0)
		return 0xFFFFFFFFu;
	const uint64_t idx = (uint64_t) py * P->width + (uint64_t) px;
	return idx < (uint64_t) P->width * P->height ? (uint32_t) idx : 0xFFFFFFFFu;
}

/* line.c:48-76: the DDA walks x, y, t by repeated float addition. */
__device__ void emitLineFrags(const AsmArgs& A, const SrpDrawParams* P, const SrpLine& L, uint32_t rec, uint32_t base)
{
	const int steps = lineSteps(L);
	if (steps < 0) return;
	const float dx = L.ssx[1] - L.ssx[0], dy = L.ssy[1] - L.ssy[0];
	const float xInc = dx / (float) steps, yInc = dy / (float) steps;
	const float tInc = (float) (1.0 / (double) steps);
	float x = L.ssx[0], y = L.ssy[0], t = 0.0f;
	uint32_t invalid = 0;
	for (int i = 0; i <= steps; i++)
	{
		const int px = cvtI32(round((double) x)), py = cvtI32(round((double) y));
		SrpFrag f;
		f.px = px; f.py = py; f.prim = rec; f.t = t;
		const uint32_t key = pixelKey(P, px, py);
		invalid += key == 0xFFFFFFFFu;
		A.frags[base + i] = f;
		A.fragKeys[base + i] = key;
		x += xInc; y += yInc; t += tInc;
	}
	if (invalid)
		atomicAdd(&P->stats[(size_t) (blockIdx.x % SRP_STAT_SLOTS) * SRP_STAT_STRIDE], (unsigned long long) invalid);
}

/* point.c:81-118: pixel bounds of a point; false when fully outside. */
struct PointBox { float mnx, mny, mxx, mxy; int x0, x1, y0, y1; };

__device__ bool pointBox(const SrpDrawParams* P, const SrpPoint& pt, PointBox& b)
{
	const float half = (float) ((double) P->st.pointSize * 0.5);
	b.mnx = pt.ssx - half; b.mny = pt.ssy - half;
	b.mxx = pt.ssx + half; b.mxy = pt.ssy + half;
	int x0 = cvtI32(floor((double) b.mnx)), x1 = cvtI32(floor((double) b.mxx));
	int y0 = cvtI32(floor((double) b.mny)), y1 = cvtI32(floor((double) b.mxy));
	if (x1 < 0 || y1 < 0 || x0 >= (int) P->width || y0 >= (int) P->height)
		return false;
	if (x0 < 0) x0 = 0;
	if (y0 < 0) y0 = 0;
	if (x1 >= (int) P->width) x1 = (int) P->width - 1;
	if (y1 >= (int) P->height) y1 = (int) P->height - 1;
	b.x0 = x0; b.x1 = x1; b.y0 = y0; b.y1 = y1;
	return true;
}

__device__ __forceinline__ bool pointCovers(const PointBox& b, int x, int y)
{
	const float px = (float) ((double) x + 0.5), py = (float) ((double) y + 0.5);
	return !(px < b.mnx || px >= b.mxx || py < b.mny || py >= b.mxy);
}

template <bool EMIT>
__device__ uint32_t pointFrags(const AsmArgs& A, const SrpDrawParams* P, const SrpPoint& pt, uint32_t rec, uint32_t base)
{
	PointBox b;
	if (!pointBox(P, pt, b))
		return 0;
	uint32_t n = 0;
	for (int y = b.y0; y <= b.y1; y++)
		for (int x = b.x0; x <= b.x1; x++)
			if (pointCovers(b, x, y))
			{
				if (EMIT)
				{
					SrpFrag f;
					f.px = x; f.py = y; f.prim = rec; f.t = 0.0f;
					A.frags[base + n] = f;
					A.fragKeys[base + n] = (uint32_t) y * P->width + (uint32_t) x;
				}
				n++;
			}
	return n;
}

__device__ __forceinline__ void makePoint(const SrpDrawParams* P, const Vtx& v, uint32_t id, SrpPoint& pt)
{
	toScreen(P, v.c[0], v.c[1], pt.ssx, pt.ssy);
	pt.z = v.c[2];
	pt.w = v.c[3];
	pt.vref = v.vref;
	pt.id = id;
	pt.pad[0] = pt.pad[1] = 0;
}

/* ---------------------------------------------------------------- kernels */

/* topology.c:24-48 */
__device__ __forceinline__ void triStream(uint32_t prim, uint64_t base, uint64_t k, uint64_t* s)
{
	if (prim == SRP_PRIM_TRIANGLES) { s[0] = base + 3 * k; s[1] = base + 3 * k + 1; s[2] = base + 3 * k + 2; }
	else if (prim == SRP_PRIM_TRIANGLE_STRIP)
	{
		const bool odd = (k & 1) != 0;
		s[0] = base + k + (odd ? 1 : 0);
		s[1] = base + k + (odd ? 0 : 1);
		s[2] = base + k + 2;
	}
	else { s[0] = base; s[1] = base + k + 1; s[2] = base + k + 2; }
}

/* The primitives of one fan triangle (a, b, c) per polygon mode
 * (primitive_assembly.c:80-128): records from recBase + nRec, items from
 * itemBase + nItems, both counters advanced.  COUNT_SPANS (counting pass):
 * spans are also counted into *A.spanCount. */
template <bool EMIT, bool COUNT_SPANS>
__device__ __forceinline__ void fanTriangle(const AsmArgs& A, const AsmDraw& W, const Vtx& a, const Vtx& b, const Vtx& c,
                                            uint32_t recBase, uint32_t idBase, uint32_t itemBase, uint32_t& nRec,
                                            uint32_t& nItems, bool drop)
{
	const SrpDrawParams* P = W.P;
	const uint32_t mode = P->st.polygonMode;
	if (mode == SRP_POLYGON_MODE_FILL)
	{
		SrpTri T;
		if (!setupTriangle(P, a, b, c, recBase + nRec - idBase, T))
			return;
		const uint32_t r = recBase + nRec;
		uint32_t np;
		if (EMIT)
		{
			/* A triangle beyond the record capacity (or with a clip vertex
			 * beyond its capacity) keeps its counted pairs, written as
			 * empty pairs without a record. */
			const bool fits = r < A.recCap && !drop;
			if (fits)
				A.tris[r] = T;
			else
				atomicOr(A.error, 2u);
			uint32_t o = itemBase + nItems;
			np = binTriangle<true>(P, T, W.tileBase, [&](uint32_t key, uint64_t m, bool known) {
				if (o < A.itemCap)
				{
					A.pairKeys[o] = key;
					A.pairVals[o] = o;
					A.pairRec[o] = fits ? (r | (known ? 0x80000000u : 0u)) : 0xFFFFFFFFu;
					A.pairMask[o] = fits ? m : 0ull;
				}
				o++;
			});
			if (isSpan(P, T))
			{
				appendSpan(A, P, T, W.tileBase, o, fits ? r : 0xFFFFFFFFu);
				o += np;
			}
			if (o > A.itemCap)
				atomicOr(A.error, 2u);
		}
		else
		{
			np = binTriangle<false>(P, T, W.tileBase, [](uint32_t, uint64_t, bool) {});
			if (COUNT_SPANS && isSpan(P, T))
			{
				/* spans of the draw (sizes the span list) */
				const uint64_t act = __ballot(1);
				if (__lane_id() == (uint32_t) __builtin_ctzll(act))
					atomicAdd(A.spanCount, (uint32_t) __popcll(act));
			}
		}
		nRec++;
		nItems += np;
	}
	else if (mode == SRP_POLYGON_MODE_LINE)
	{
		/* edges a-b, b-c, c-a; called by name, not through an array of
		 * vertex pointers, which would put the vertices of every
		 * triangle of the launch in scratch memory */
		auto edge = [&](const Vtx& p, const Vtx& q) {
			SrpLine L;
			setupLine(P, p, q, recBase + nRec - idBase, L);
			const uint32_t nf = lineFragCount(L);
			if (EMIT)
			{
				A.lines[recBase + nRec] = L;
				emitLineFrags(A, P, L, recBase + nRec, itemBase + nItems);
			}
			nRec++;
			nItems += nf;
		};
		edge(a, b);
		edge(b, c);
		edge(c, a);
	}
	else
	{
		auto corner = [&](const Vtx& p) {
			SrpPoint pt;
			makePoint(P, p, recBase + nRec - idBase, pt);
			if (EMIT) A.points[recBase + nRec] = pt;
			/* up to 42 points of up to W*H fragments each: the count
			 * saturates rather than wraps, and the host's check of the
			 * exact total then refuses the draw before any emit */
			const uint32_t nf = pointFrags<EMIT>(A, P, pt, recBase + nRec, itemBase + nItems);
			nItems = (nItems + nf < nItems) ? 0xFFFFFFFFu : nItems + nf;
			nRec++;
		};
		corner(a);
		corner(b);
		corner(c);
	}
}

/* Start of an emitting kernel of a single draw (AsmArgs.fitTotals): false
 * when it has nothing to do, because the counted totals do not fit the
 * capacities it was given or nothing was assembled.  Block 0 of the
 * leading kernel records the outcome and publishes the counters. */
template <bool EMIT>
__device__ __forceinline__ bool emitGo(const AsmArgs& A, bool leader)
{
	if (!EMIT || !A.fitTotals)
		return true;
	uint32_t* t = A.fitTotals;
	const unsigned long long* t64 = reinterpret_cast<const unsigned long long*>(t + 8);
	const FitCaps& c = A.fitCaps;
	const bool fits = t64[0] <= c.records && t64[1] <= c.generated && t64[2] <= c.items &&
	                  (unsigned long long) t[6] <= c.spans && (unsigned long long) t[4] <= c.clipped;
	const bool any = t64[0] != 0ull;
	if (leader && blockIdx.x == 0 && threadIdx.x == 0)
	{
		t[7] = !fits ? 1u : (any ? 0u : 2u);
		t[2] = fits ? (uint32_t) t64[2] : 0u;
		if (A.fitHost)
			publish(t, A.fitHost, A.fitSeq);
	}
	return fits && any;
}

/* Triangle draws (all polygon modes): primitive_assembly.c:37-135.  The
 * common case — all three vertices inside the frustum, trivially accepted
 * by clipping.c:74-78 — stays in registers. */
template <bool EMIT>
__global__ __launch_bounds__(128) void k_assemble_tris(AsmArgs A)
{
	if (!emitGo<EMIT>(A, true))   /* the counted totals do not fit, or nothing to emit */
		return;
	if (!EMIT && A.zero)
		for (uint32_t z = blockIdx.x * 128u + threadIdx.x; z < A.zeroCount; z += gridDim.x * 128u)
			A.zero[z] = 0u;
	const uint32_t d = A.nDraws > 1 ? drawOfBlock(A.draws, A.nDraws, blockIdx.x) : 0u;
	const AsmDraw& W = A.draws[d];
	const uint32_t k = (blockIdx.x - W.firstBlock) * 128u + threadIdx.x;
	if (k >= W.nIn)
		return;
	const uint32_t g = W.primBase + k;   /* slot of this input in the count arrays */
	const SrpDrawParams* P = W.P;
	uint64_t s[3];
	triStream(W.prim, W.start, k, s);
	uint32_t slot[3];
	Vtx v[3];
	uint32_t codeOr = 0, codeAnd = 0x3Fu;
	for (int i = 0; i < 3; i++)
		slot[i] = slotOf(W, s[i]);
	if (!slotsValid(A, P, slot, 3))
	{
		if (!EMIT)
		{
			A.recOff[g] = 0;
			A.genOff[g] = 0;
			A.itemOff[g] = 0;
		}
		return;
	}
	for (int i = 0; i < 3; i++)
	{
		loadClip(P, slot[i], v[i].c);
		const uint32_t c = outcode(v[i].c);
		codeOr |= c;
		codeAnd &= c;
		v[i].vref = slot[i];
	}
	/* Triangles that need clipping are listed here and assembled by
	 * k_assemble_clipped: the clipper's registers and stack would otherwise
	 * be reserved by every wave of this kernel (148 against 59 VGPRs). */
	const bool clip = codeOr != 0 && codeAnd == 0;
	if (!EMIT)
	{
		const uint64_t want = __ballot(clip);
		if (want)
		{
			const uint32_t lane = __lane_id();
			const uint32_t leader = (uint32_t) __builtin_ctzll(want);
			uint32_t at = 0;
			if (lane == leader)
				at = atomicAdd(A.clipCount, (uint32_t) __builtin_popcountll(want));
			at = (uint32_t) __shfl((int) at, (int) leader);
			if (clip)
				A.clipList[at + (uint32_t) __builtin_popcountll(want & ((1ull << lane) - 1ull))] = g;
		}
	}
	if (clip)
		return;
	uint32_t recBase = 0, itemBase = 0, idBase = 0;
	if (EMIT)
	{
		recBase = A.recOff[g];
		itemBase = A.itemOff[g];
		idBase = A.recOff[W.primBase];
	}
	uint32_t nRec = 0, nItems = 0;
	if (codeOr == 0)
	{
		for (int i = 0; i < 3; i++) v[i].invW = divide(v[i].c);
		fanTriangle<EMIT, !EMIT>(A, W, v[0], v[1], v[2], recBase, idBase, itemBase, nRec, nItems, false);
	}
	if (!EMIT)
	{
		A.recOff[g] = nRec;
		A.genOff[g] = 0;
		A.itemOff[g] = nItems;
	}
}

/* The draw holding input g of a (possibly multi-draw) launch: the last
 * draw whose first input is at or before it. */
__device__ __forceinline__ uint32_t drawOfInput(const AsmDraw* D, uint32_t n, uint32_t g)
{
	uint32_t lo = 0, hi = n;
	while (hi - lo > 1)
	{
		const uint32_t mid = (lo + hi) >> 1;
		if (D[mid].primBase <= g) lo = mid;
		else hi = mid;
	}
	return lo;
}

/* ------------------------------------------------------------- clipping */

/* Lanes that clip one triangle together.  A triangle clipped by the six
 * planes has at most 9 vertices (a polygon gains at most one per plane);
 * the limits below only guard numerically degenerate input. */
constexpr uint32_t CLIP_GROUP = 16;

/* The emitting pass of the filled fan triangles of one clipped triangle
 * (polygon mode FILL): set up by their fan lanes, binned by the whole
 * group.  binTriangle scans a box with one lane; here the group's lanes
 * take work items in turn: each tile of a fan triangle's box (exact-mask
 * triangles) or the whole span (larger ones).  A tile's mask is found by
 * the reference's walk restricted to that tile: the row value after
 * (y - y0) additions of dy (replayed exactly, then added row by row as
 * triangle.c:101-109 does) and, along the row, (x - x0) additions of dx,
 * so every bit equals binTriangle's.  Items are numbered fan by fan and,
 * within a fan, in row-major tile order, and pairs are placed in that
 * order skipping empty tiles: the same records, pairs and positions as
 * fan lanes calling fanTriangle one after the other. */
__device__ __forceinline__ void clipFillFans(const AsmArgs& A, const AsmDraw& W, uint32_t gl, const Vtx& a,
                                             const Vtx& b, const Vtx& c, bool fan, uint32_t recBase, uint32_t idBase,
                                             uint32_t itemBase, bool drop)
{
	const SrpDrawParams* P = W.P;
	SrpTri T = {};
	const bool ok = fan && setupTriangle(P, a, b, c, 0u, T);
	const bool span = ok && isSpan(P, T);
	const bool boxed = ok && T.x1 > T.x0 && T.y1 > T.y0;
	const uint32_t sh = P->tileShift;
	uint32_t nItem = 0;
	if (span)
		nItem = 1u;
	else if (boxed)
		nItem = ((((uint32_t) T.x1 - 1u) >> sh) - ((uint32_t) T.x0 >> sh) + 1u) *
		        ((((uint32_t) T.y1 - 1u) >> sh) - ((uint32_t) T.y0 >> sh) + 1u);
	/* exclusive scans over the group: work items and records */
	uint32_t iIncl = nItem, rIncl = ok ? 1u : 0u;
#pragma unroll
	for (uint32_t off = 1; off < CLIP_GROUP; off <<= 1)
	{
		const uint32_t vi = __shfl_up(iIncl, off, CLIP_GROUP);
		const uint32_t vr = __shfl_up(rIncl, off, CLIP_GROUP);
		if (gl >= off)
		{
			iIncl += vi;
			rIncl += vr;
		}
	}
	const uint32_t totI = __shfl(iIncl, (int) CLIP_GROUP - 1, CLIP_GROUP);
	const uint32_t iEx = iIncl - nItem;
	const uint32_t r = recBase + rIncl - (ok ? 1u : 0u);
	bool fits = false;
	if (ok)
	{
		/* a triangle beyond the record capacity (or with a clip vertex
		 * beyond its capacity) keeps its counted pairs, written as empty
		 * pairs without a record */
		fits = r < A.recCap && !drop;
		T.id = r - idBase;
		if (fits)
			A.tris[r] = T;
		else
			atomicOr(A.error, 2u);
	}
	unsigned long long pairs = 0;   /* pairs of the items before this round */
	const uint32_t box0 = (uint32_t) T.x0 | ((uint32_t) T.y0 << 16);
	const uint32_t box1 = (uint32_t) T.x1 | ((uint32_t) T.y1 << 16);
	for (uint32_t base = 0; base < totI; base += CLIP_GROUP)
	{
		const uint32_t item = base + gl;
		const bool valid = item < totI;
		/* the fan lane whose items include this one: the first lane whose
		 * inclusive item count passes it (the counts never decrease) */
		uint32_t owner = 0;
#pragma unroll
		for (uint32_t step = CLIP_GROUP / 2; step > 0; step >>= 1)
		{
			const uint32_t v = __shfl(iIncl, (int) (owner + step - 1u), CLIP_GROUP);
			if (v <= item) owner += step;
		}
		const int o = valid ? (int) owner : 0;
		float l0[3], dx[3], dy[3];
#pragma unroll
		for (int e = 0; e < 3; e++)
		{
			l0[e] = __shfl(T.l0[e], o, CLIP_GROUP);
			dx[e] = __shfl(T.dx[e], o, CLIP_GROUP);
			dy[e] = __shfl(T.dy[e], o, CLIP_GROUP);
		}
		const uint32_t ob0 = __shfl(box0, o, CLIP_GROUP), ob1 = __shfl(box1, o, CLIP_GROUP);
		const uint32_t oflags = __shfl(T.flags, o, CLIP_GROUP);
		const uint32_t oEx = __shfl(iEx, o, CLIP_GROUP);
		const bool oSpan = __shfl((int) span, o, CLIP_GROUP) != 0;
		const uint32_t oRec = __shfl(r, o, CLIP_GROUP);
		const bool oFits = __shfl((int) fits, o, CLIP_GROUP) != 0;
		const int x0 = (int) (ob0 & 0xFFFFu), y0 = (int) (ob0 >> 16);
		const int x1 = (int) (ob1 & 0xFFFFu), y1 = (int) (ob1 >> 16);
		uint64_t m = 0;
		uint32_t weight = 0, key = 0;
		if (valid)
		{
			const uint32_t tx0 = (uint32_t) x0 >> sh, ty0 = (uint32_t) y0 >> sh;
			const uint32_t ntx = (((uint32_t) x1 - 1u) >> sh) - tx0 + 1u;
			if (oSpan)
				weight = ntx * ((((uint32_t) y1 - 1u) >> sh) - ty0 + 1u);
			else
			{
				/* exact masks: tileShift 3 */
				const uint32_t j = item - oEx;
				const uint32_t ty = ty0 + j / ntx, tx = tx0 + j % ntx;
				key = W.tileBase + ty * P->tilesX + tx;
				const int xs = ((int) (tx * 8u) > x0) ? (int) (tx * 8u) : x0;
				const int xe = ((int) (tx * 8u + 8u) < x1) ? (int) (tx * 8u + 8u) : x1;
				const int ys = ((int) (ty * 8u) > y0) ? (int) (ty * 8u) : y0;
				const int ye = ((int) (ty * 8u + 8u) < y1) ? (int) (ty * 8u + 8u) : y1;
				float row[3], thr[3];
				srp::dev::edgeThresholds(oflags, thr);
#pragma unroll
				for (int e = 0; e < 3; e++)
					row[e] = srp::dev::replay(l0[e], dy[e], ys - y0);
				for (int y = ys; y < ye; y++)
				{
					float l[3];
#pragma unroll
					for (int e = 0; e < 3; e++)
						l[e] = srp::dev::replay(row[e], dx[e], xs - x0);
					uint32_t bits = 0;
					for (int x = xs; x < xe; x++)
					{
						bits |= (srp::dev::insideEdges(l, thr) ? 1u : 0u) << ((uint32_t) x & 7u);
#pragma unroll
						for (int e = 0; e < 3; e++) l[e] += dx[e];
					}
					m |= (uint64_t) bits << (((uint32_t) y & 7u) << 3);
#pragma unroll
					for (int e = 0; e < 3; e++) row[e] += dy[e];
				}
				weight = m ? 1u : 0u;
			}
		}
		/* this round's pairs, in item order */
		uint32_t wIncl = weight;
#pragma unroll
		for (uint32_t off = 1; off < CLIP_GROUP; off <<= 1)
		{
			const uint32_t v = __shfl_up(wIncl, off, CLIP_GROUP);
			if (gl >= off) wIncl += v;
		}
		const uint32_t roundPairs = __shfl(wIncl, (int) CLIP_GROUP - 1, CLIP_GROUP);
		if (weight)
		{
			const unsigned long long at = (unsigned long long) itemBase + pairs + (wIncl - weight);
			if (oSpan)
			{
				SrpTri S = {};
				S.x0 = (uint16_t) x0; S.y0 = (uint16_t) y0;
				S.x1 = (uint16_t) x1; S.y1 = (uint16_t) y1;
				appendSpan(A, P, S, W.tileBase, (uint32_t) at, oFits ? oRec : 0xFFFFFFFFu);
			}
			else if (at < A.itemCap)
			{
				const uint32_t q = (uint32_t) at;
				A.pairKeys[q] = key;
				A.pairVals[q] = q;
				A.pairRec[q] = oFits ? (oRec | 0x80000000u) : 0xFFFFFFFFu;
				A.pairMask[q] = oFits ? m : 0ull;
			}
			if (at + weight > A.itemCap)
				atomicOr(A.error, 2u);
		}
		pairs += roundPairs;
	}
}

/* One triangle's Sutherland-Hodgman clip (clipping.c:68-119, :199-258) by
 * a group of CLIP_GROUP lanes, lane j holding polygon vertex j, so nothing
 * lives in per-lane scratch memory.  Each of the six passes runs as the
 * reference's loop does (an all-inside pass rotates the polygon by one):
 * lane i tests edge (i, i+1) and yields 0, 1 or 2 output vertices; a scan
 * of those counts places them, and output lane j gathers its source edge's
 * two vertices and either copies the edge's end or computes the crossing
 * a*(1-t) + b*t with the reference's operation order.  The recipe of each
 * pass (source edge, crossing or copy, t) is kept so that the varyings of
 * generated vertices can be replayed afterwards, element by element, with
 * the clip rule of interpolation.c:116-120 (floats and doubles blended by
 * {1-t, t} from their two parents' values, other types from the provoking
 * parent) - the same values a vertex-by-vertex clip produces.
 *
 * A pass that yields SRP_MAX_POLY - 1 or more vertices, or a clip that
 * generates more than SRP_POOL - 3 vertices, overflows: the triangle is
 * dropped and error bit 0 raised.
 *
 * Then lane f (1 <= f <= n-2) assembles fan triangle (0, f, f+1)
 * (clipping.c:107-116); its records and items follow those of the lanes
 * before it (a scan over the group). */
template <bool EMIT>
__device__ __forceinline__ void clipGroup(const AsmArgs& A, uint32_t g, bool has, uint32_t gl)
{
	const uint32_t d = (has && A.nDraws > 1) ? drawOfInput(A.draws, A.nDraws, g) : 0u;
	const AsmDraw& W = A.draws[d];
	const SrpDrawParams* P = W.P;
	const uint32_t gshift = __lane_id() & ~(CLIP_GROUP - 1u);
	const uint64_t below = (1ull << gl) - 1ull;

	/* lanes 0..2: the triangle's vertices (slots validated by the counting
	 * pass of k_assemble_tris before it listed the triangle) */
	uint32_t mySlot = 0;
	float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	if (has && gl < 3u)
	{
		uint64_t st[3];
		triStream(W.prim, W.start, g - W.primBase, st);
		mySlot = slotOf(W, gl == 0u ? st[0] : (gl == 1u ? st[1] : st[2]));
		loadClip(P, mySlot, c);
	}
	int np = has ? 3 : 0;
	uint32_t orig = gl;   /* original vertex (0..2) this lane's vertex is, 3: generated */
	uint32_t genTotal = 0;
	bool overflow = false;
	uint32_t rSrc[6], rNxt[6];
	bool rLerp[6];
	float rT[6];
#pragma unroll
	for (int p = 0; p < 6; p++)
	{
		const bool live = (int) gl < np;
		const uint32_t nxtL = ((int) gl + 1 < np) ? gl + 1u : 0u;
		const float da = planeDist(c, p);
		/* Group-uniform shortcuts: with every vertex inside the plane each
		 * edge yields its end, which rotates the polygon by one; with every
		 * vertex outside nothing is left.  (The lanes of one group branch
		 * together, so their shuffles stay within active lanes.) */
		const uint64_t liveMask = (np >= (int) CLIP_GROUP) ? 0xFFFFull : ((1ull << np) - 1ull);
		const uint64_t inMask = (__ballot(live && da >= 0.0f) >> gshift) & 0xFFFFull;
		if (inMask == liveMask)
		{
#pragma unroll
			for (int q = 0; q < 4; q++) c[q] = __shfl(c[q], (int) nxtL, CLIP_GROUP);
			orig = __shfl(orig, (int) nxtL, CLIP_GROUP);
			rSrc[p] = gl;
			rNxt[p] = nxtL;
			rLerp[p] = false;
			rT[p] = 0.0f;
			continue;
		}
		if (inMask == 0ull)
		{
			np = 0;
			rSrc[p] = gl;
			rNxt[p] = gl;
			rLerp[p] = false;
			rT[p] = 0.0f;
			continue;
		}
		float nc[4];
#pragma unroll
		for (int q = 0; q < 4; q++) nc[q] = __shfl(c[q], (int) nxtL, CLIP_GROUP);
		const float db = planeDist(nc, p);
		const bool ci = da >= 0.0f, ni = db >= 0.0f;
		const bool cross = live && (ci != ni);
		const bool gen = cross && !roughlyZero(da - db);
		const float t = gen ? da / (da - db) : 0.0f;
		const uint32_t cnt = !live ? 0u : ((ci && ni) ? 1u : (!gen ? 0u : ((!ci && ni) ? 2u : 1u)));
		uint32_t incl = cnt;
#pragma unroll
		for (uint32_t off = 1; off < CLIP_GROUP; off <<= 1)
		{
			const uint32_t v = __shfl_up(incl, off, CLIP_GROUP);
			if (gl >= off) incl += v;
		}
		const uint32_t m = __shfl(incl, (int) CLIP_GROUP - 1, CLIP_GROUP);
		genTotal += (uint32_t) __popcll((__ballot(gen) >> gshift) & 0xFFFFull);
		/* output j comes from the edge i with incl[i-1] <= j < incl[i] */
		uint32_t src = 0;
#pragma unroll
		for (uint32_t step = CLIP_GROUP / 2; step > 0; step >>= 1)
		{
			const uint32_t v = __shfl(incl, (int) (src + step - 1u), CLIP_GROUP);
			if (v <= gl) src += step;
		}
		const uint32_t sIncl = __shfl(incl, (int) src, CLIP_GROUP);
		const uint32_t sCnt = __shfl(cnt, (int) src, CLIP_GROUP);
		const bool sGen = __shfl((int) gen, (int) src, CLIP_GROUP) != 0;
		const float sT = __shfl(t, (int) src, CLIP_GROUP);
		const uint32_t sNxt = __shfl(nxtL, (int) src, CLIP_GROUP);
		float a[4], b[4];
#pragma unroll
		for (int q = 0; q < 4; q++)
		{
			a[q] = __shfl(c[q], (int) src, CLIP_GROUP);
			b[q] = __shfl(c[q], (int) sNxt, CLIP_GROUP);
		}
		/* an edge's crossing is its first output; its second (entering
		 * edges) and an all-inside edge's only output are its end */
		const bool lerp = sGen && gl + sCnt == sIncl;
#pragma unroll
		for (int q = 0; q < 4; q++)
			c[q] = lerp ? a[q] * (1 - sT) + b[q] * sT : b[q];
		const uint32_t oNxt = __shfl(orig, (int) sNxt, CLIP_GROUP);
		orig = lerp ? 3u : oNxt;
		rSrc[p] = src;
		rNxt[p] = sNxt;
		rLerp[p] = lerp;
		rT[p] = sT;
		np = (int) m;
		if (m + 2u > (uint32_t) SRP_MAX_POLY || genTotal > (uint32_t) SRP_POOL - 3u)
		{
			overflow = true;
			np = 0;
		}
	}
	if (overflow && gl == 0u)
		atomicOr(A.error, 1u);

	/* the polygon's vertices after the divide; generated ones numbered in
	 * polygon order */
	const bool live = (int) gl < np;
	Vtx v;
#pragma unroll
	for (int q = 0; q < 4; q++) v.c[q] = c[q];
	v.invW = live ? divide(v.c) : 0.0f;
	const uint64_t genLanes = (__ballot(live && orig == 3u) >> gshift) & 0xFFFFull;
	const uint32_t nGen = (uint32_t) __popcll(genLanes);
	const uint32_t rank = (uint32_t) __popcll(genLanes & below);
	const uint32_t origSlot = __shfl(mySlot, (int) (orig < 3u ? orig : 0u), CLIP_GROUP);
	uint32_t recBase = 0, genBase = 0, itemBase = 0, idBase = 0;
	if (EMIT && has)
	{
		recBase = A.recOff[g];
		genBase = A.genOff[g];
		itemBase = A.itemOff[g];
		idBase = A.recOff[W.primBase];
	}
	v.vref = (orig < 3u) ? origSlot : P->cacheSize + genBase + rank;

	bool drop = false;
	if (EMIT)
	{
		const bool mine = live && orig == 3u;
		const bool over = mine && genBase + rank >= A.genCap;
		drop = ((__ballot(over) >> gshift) & 0xFFFFull) != 0ull;
		uint8_t* dst = (mine && !over) ? A.genVaryings + (size_t) (genBase + rank) * P->varyingsSize : nullptr;
		const uint8_t* own = (has && gl < 3u) ? srp::dev::varyingPtr(P, mySlot) : nullptr;
		/* integer elements: the original each vertex takes them from */
		uint32_t root = gl < 3u ? gl : 0u;
#pragma unroll
		for (int p = 0; p < 6; p++)
		{
			const uint32_t cu = __shfl(root, (int) rSrc[p], CLIP_GROUP), nx = __shfl(root, (int) rNxt[p], CLIP_GROUP);
			root = rLerp[p] ? (P->st.provokingLast ? nx : cu) : nx;
		}
		const uint32_t rootSlot = __shfl(mySlot, (int) root, CLIP_GROUP);
		for (uint32_t at = 0; at < P->nVaryings; at++)
		{
			const SrpVaryingDesc dsc = P->var[at];
			if (dsc.type == SRP_FLOAT)
				for (uint32_t e = 0; e < dsc.nItems; e++)
				{
					const uint32_t off = dsc.offset + 4 * e;
					float val = own ? srp::dev::loadF(own + off) : 0.0f;
#pragma unroll
					for (int p = 0; p < 6; p++)
					{
						const float cu = __shfl(val, (int) rSrc[p], CLIP_GROUP);
						const float nx = __shfl(val, (int) rNxt[p], CLIP_GROUP);
						const float w0 = 1 - rT[p];
						float lv = 0.0f;
						lv += cu * w0;
						lv += nx * rT[p];
						val = rLerp[p] ? lv : nx;
					}
					if (dst)
						__builtin_memcpy(dst + off, &val, 4);
				}
			else if (dsc.type == SRP_DOUBLE)
				for (uint32_t e = 0; e < dsc.nItems; e++)
				{
					const uint32_t off = dsc.offset + 8 * e;
					double val = own ? srp::dev::loadD(own + off) : 0.0;
#pragma unroll
					for (int p = 0; p < 6; p++)
					{
						const double cu = __shfl(val, (int) rSrc[p], CLIP_GROUP);
						const double nx = __shfl(val, (int) rNxt[p], CLIP_GROUP);
						const float w0 = 1 - rT[p];
						double lv = 0.0;
						lv += cu * (double) w0;
						lv += nx * (double) rT[p];
						val = rLerp[p] ? lv : nx;
					}
					if (dst)
						__builtin_memcpy(dst + off, &val, 8);
				}
			else if (dst)
			{
				const uint32_t n = srp::dev::elemSize(dsc.type) * dsc.nItems;
				const uint8_t* from = srp::dev::varyingPtr(P, rootSlot) + dsc.offset;
				for (uint32_t b = 0; b < n; b++) dst[dsc.offset + b] = from[b];
			}
		}
	}

	/* fan triangle (0, gl, gl + 1) */
	Vtx a, cc;
	const uint32_t nextL = gl + 1u < CLIP_GROUP ? gl + 1u : gl;
#pragma unroll
	for (int q = 0; q < 4; q++)
	{
		a.c[q] = __shfl(v.c[q], 0, CLIP_GROUP);
		cc.c[q] = __shfl(v.c[q], (int) nextL, CLIP_GROUP);
	}
	a.invW = __shfl(v.invW, 0, CLIP_GROUP);
	cc.invW = __shfl(v.invW, (int) nextL, CLIP_GROUP);
	a.vref = __shfl(v.vref, 0, CLIP_GROUP);
	cc.vref = __shfl(v.vref, (int) nextL, CLIP_GROUP);
	const bool fan = gl >= 1u && (int) gl + 1 < np;
	if (EMIT && P->st.polygonMode == SRP_POLYGON_MODE_FILL)
	{
		/* (the counting pass keeps one lane per fan triangle: its scan
		 * stops at a tile's first covered pixel, which for the small
		 * triangles clipping mostly makes costs less than the group's
		 * shuffles) */
		clipFillFans(A, W, gl, a, v, cc, fan, recBase, idBase, itemBase, drop);
		return;
	}
	uint32_t myRec = 0, myItems = 0;
	if (fan)
		fanTriangle<false, !EMIT>(A, W, a, v, cc, 0u, 0u, 0u, myRec, myItems, false);
	/* exclusive scans over the group (items saturate: see fanTriangle) */
	uint32_t exRec = myRec;
	unsigned long long exItems = myItems;
#pragma unroll
	for (uint32_t off = 1; off < CLIP_GROUP; off <<= 1)
	{
		const uint32_t r = __shfl_up(exRec, off, CLIP_GROUP);
		const unsigned long long it = __shfl_up(exItems, off, CLIP_GROUP);
		if (gl >= off)
		{
			exRec += r;
			exItems += it;
		}
	}
	const uint32_t totRec = __shfl(exRec, (int) CLIP_GROUP - 1, CLIP_GROUP);
	const unsigned long long totItems = __shfl(exItems, (int) CLIP_GROUP - 1, CLIP_GROUP);
	exRec -= myRec;
	exItems -= myItems;
	if (!EMIT)
	{
		if (has && gl == 0u)
		{
			A.recOff[g] = totRec;
			A.genOff[g] = nGen;
			A.itemOff[g] = totItems > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t) totItems;
		}
	}
	else if (fan)
	{
		uint32_t r = 0, it = 0;
		fanTriangle<true, false>(A, W, a, v, cc, recBase + exRec, idBase, itemBase + (uint32_t) exItems, r, it, drop);
	}
}

/* The triangles k_assemble_tris listed for clipping, in either pass (the
 * list stays valid from the counting pass to the emitting one): one group
 * of CLIP_GROUP lanes per triangle, the grid striding over the list (every
 * lane of a block runs the same number of rounds, so the group shuffles
 * stay convergent).  Each triangle writes only at its own scanned offsets,
 * so list order does not matter. */
template <bool EMIT>
__global__ __launch_bounds__(128) void k_assemble_clipped(AsmArgs A)
{
	if (!emitGo<EMIT>(A, false))   /* the counted totals do not fit, or nothing to emit */
		return;
	const uint32_t n = *A.clipCount;
	constexpr uint32_t perBlock = 128u / CLIP_GROUP;
	const uint32_t gl = threadIdx.x & (CLIP_GROUP - 1u);
	for (uint32_t base = blockIdx.x * perBlock; base < n; base += gridDim.x * perBlock)
	{
		const uint32_t i = base + threadIdx.x / CLIP_GROUP;
		const bool has = i < n;
		clipGroup<EMIT>(A, has ? A.clipList[i] : 0u, has, gl);
	}
}

/* topology.c:62-83 */
__device__ __forceinline__ void lineStream(uint32_t prim, uint64_t base, uint64_t k, uint64_t count, uint64_t* s)
{
	if (prim == SRP_PRIM_LINES) { s[0] = base + 2 * k; s[1] = base + 2 * k + 1; }
	else if (prim == SRP_PRIM_LINE_STRIP) { s[0] = base + k; s[1] = base + k + 1; }
	else { s[0] = base + k; s[1] = base + (k + 1) % count; }
}

/* Line draws: primitive_assembly.c:137-178 with clipping.c:139-183. */
template <bool EMIT>
__global__ __launch_bounds__(128) void k_assemble_lines(AsmArgs A)
{
	if (!emitGo<EMIT>(A, true))   /* the counted totals do not fit, or nothing to emit */
		return;
	const uint32_t d = A.nDraws > 1 ? drawOfBlock(A.draws, A.nDraws, blockIdx.x) : 0u;
	const AsmDraw& W = A.draws[d];
	const uint32_t k = (blockIdx.x - W.firstBlock) * 128u + threadIdx.x;
	if (k >= W.nIn)
		return;
	const uint32_t g = W.primBase + k;   /* slot of this input in the count arrays */
	const SrpDrawParams* P = W.P;
	uint64_t s[2];
	lineStream(W.prim, W.start, k, W.count, s);
	uint32_t slot[2];
	for (int i = 0; i < 2; i++)
		slot[i] = slotOf(W, s[i]);
	if (!slotsValid(A, P, slot, 2))
	{
		if (!EMIT)
		{
			A.recOff[g] = 0;
			A.genOff[g] = 0;
			A.itemOff[g] = 0;
		}
		return;
	}
	float c[2][4];
	for (int i = 0; i < 2; i++)
		loadClip(P, slot[i], c[i]);

	/* Liang-Barsky style clip (clipping.c:139-183): each clipped end is
	 * re-derived from the two originals at t0 / t1. */
	float t0 = 0.0f, t1 = 1.0f;
	bool rejected = false;
	const uint32_t c0 = outcode(c[0]), c1 = outcode(c[1]);
	if ((c0 | c1) != 0)
	{
		if ((c0 & c1) != 0)
			rejected = true;
		else
			for (int p = 0; p < 6 && !rejected; p++)
			{
				const float da = planeDist(c[0], p), db = planeDist(c[1], p);
				if (da < 0.0f && db < 0.0f) { rejected = true; break; }
				if (da < 0.0f || db < 0.0f)
				{
					if (roughlyZero(da - db)) continue;
					const float t = da / (da - db);
					if (da < 0.0f) t0 = (t0 > t) ? t0 : t;
					else t1 = (t1 > t) ? t : t1;
					if (t0 > t1) rejected = true;
				}
			}
	}

	uint32_t nRec = 0, nGen = 0, nItems = 0;
	if (!rejected)
	{
		const uint32_t recBase = EMIT ? A.recOff[g] : 0u;
		const uint32_t genBase = EMIT ? A.genOff[g] : 0u;
		const uint32_t itemBase = EMIT ? A.itemOff[g] : 0u;
		const uint32_t idBase = EMIT ? A.recOff[W.primBase] : 0u;
		Vtx e[2];
		const bool gen[2] = {t0 > 0.0f, t1 < 1.0f};
		const float tt[2] = {t0, t1};
		for (int i = 0; i < 2; i++)
		{
			if (gen[i])
			{
				/* clipping.c:244-258 */
				for (int q = 0; q < 4; q++) e[i].c[q] = c[0][q] * (1 - tt[i]) + c[1][q] * tt[i];
				e[i].vref = P->cacheSize + genBase + nGen;
				if (EMIT)
					lerpVaryings(P, srp::dev::varyingPtr(P, slot[0]), srp::dev::varyingPtr(P, slot[1]), tt[i],
					             A.genVaryings + (size_t) (genBase + nGen) * P->varyingsSize);
				nGen++;
			}
			else
			{
				for (int q = 0; q < 4; q++) e[i].c[q] = c[i][q];
				e[i].vref = slot[i];
			}
			e[i].invW = divide(e[i].c);
		}
		SrpLine L;
		setupLine(P, e[0], e[1], recBase - idBase, L);
		nItems = lineFragCount(L);
		if (EMIT)
		{
			A.lines[recBase] = L;
			emitLineFrags(A, P, L, recBase, itemBase);
		}
		nRec = 1;
	}
	if (!EMIT)
	{
		A.recOff[g] = nRec;
		A.genOff[g] = nGen;
		A.itemOff[g] = nItems;
	}
}

/* Point draws: primitive_assembly.c:221-254, clipping.c:185-197. */
template <bool EMIT>
__global__ __launch_bounds__(128) void k_assemble_points(AsmArgs A)
{
	if (!emitGo<EMIT>(A, true))   /* the counted totals do not fit, or nothing to emit */
		return;
	const uint32_t d = A.nDraws > 1 ? drawOfBlock(A.draws, A.nDraws, blockIdx.x) : 0u;
	const AsmDraw& W = A.draws[d];
	const uint32_t k = (blockIdx.x - W.firstBlock) * 128u + threadIdx.x;
	if (k >= W.nIn)
		return;
	const uint32_t g = W.primBase + k;   /* slot of this input in the count arrays */
	const SrpDrawParams* P = W.P;
	const uint32_t slot = slotOf(W, W.start + k);
	if (!slotsValid(A, P, &slot, 1))
	{
		if (!EMIT)
		{
			A.recOff[g] = 0;
			A.genOff[g] = 0;
			A.itemOff[g] = 0;
		}
		return;
	}
	float c[4];
	loadClip(P, slot, c);
	const float x = c[0], y = c[1], z = c[2], w = c[3];
	const bool clipped = (x < -w || x > w) || (y < -w || y > w) || (z < -w || z > w);
	uint32_t nRec = 0, nItems = 0;
	if (!clipped)
	{
		const uint32_t recBase = EMIT ? A.recOff[g] : 0u;
		const uint32_t itemBase = EMIT ? A.itemOff[g] : 0u;
		const uint32_t idBase = EMIT ? A.recOff[W.primBase] : 0u;
		Vtx v;
		for (int i = 0; i < 4; i++) v.c[i] = c[i];
		v.invW = divide(v.c);
		v.vref = slot;
		SrpPoint pt;
		makePoint(P, v, recBase - idBase, pt);
		if (EMIT) A.points[recBase] = pt;
		nItems = pointFrags<EMIT>(A, P, pt, recBase, itemBase);
		nRec = 1;
	}
	if (!EMIT)
	{
		A.recOff[g] = nRec;
		A.genOff[g] = 0;
		A.itemOff[g] = nItems;
	}
}

/* The tile pairs of the spans listed by the emitting pass: one wave per
 * span at a time, lanes over its tiles (consecutive items, so the writes
 * coalesce).  The pairs equal what a per-tile loop in the emitting thread
 * would write: item item + k is tile (k % nx, k / nx) of the box. */
__global__ __launch_bounds__(256) void k_expand_spans(const SrpSpan* spans, const uint32_t* count, uint32_t cap,
                                                      uint32_t* pairKeys, uint32_t* pairVals, uint32_t* pairRec,
                                                      unsigned long long* pairMask, uint32_t itemCap)
{
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
	const uint32_t nWaves = gridDim.x * 4u;
	uint32_t n = *count;
	if (n > cap)
		n = cap;
	for (uint32_t i = wave; i < n; i += nWaves)
	{
		const SrpSpan sp = spans[i];
		const uint32_t total = sp.nx * sp.ny;
		const uint64_t m = (sp.rec == 0xFFFFFFFFu) ? 0ull : ~0ull;
		for (uint32_t k = lane; k < total; k += 64u)
		{
			const uint32_t o = sp.item + k;
			if (o >= itemCap)
				break;
			const uint32_t dy = k / sp.nx, dx = k - dy * sp.nx;
			pairKeys[o] = sp.keyBase + dy * sp.stride + dx;
			pairVals[o] = o;
			pairRec[o] = sp.rec;
			pairMask[o] = m;
		}
	}
}

void launchExpandSpans(const AsmArgs& A, uint32_t blocks, hipStream_t s)
{
	if (blocks)
		hipLaunchKernelGGL(k_expand_spans, dim3(blocks), dim3(256), 0, s, A.spans, A.spanCount, A.spanCap, A.pairKeys,
		                   A.pairVals, A.pairRec, A.pairMask, A.itemCap);
}

void launchClipped(const AsmArgs& A, bool emit, hipStream_t s, uint32_t clipped, uint32_t clipHint)
{
	const uint32_t blocks = A.nBlocks;
	if (!blocks) return;
	uint32_t clipBlocks;
	if (clipped != kClipUnknown)
	{
		const uint32_t need = (uint32_t) (((uint64_t) clipped + 7u) / 8u);
		clipBlocks = need < 2048u ? need : 2048u;
	}
	else
	{
		const uint64_t want = ((uint64_t) clipHint + clipHint / 4u + 7u) / 8u;
		clipBlocks = (uint32_t) (want < 256u ? 256u : (want > 2048u ? 2048u : want));
		clipBlocks = clipBlocks < blocks ? clipBlocks : blocks;
	}
	if (!clipBlocks)
		return;
	if (emit) hipLaunchKernelGGL(k_assemble_clipped<true>, dim3(clipBlocks), dim3(128), 0, s, A);
	else hipLaunchKernelGGL(k_assemble_clipped<false>, dim3(clipBlocks), dim3(128), 0, s, A);
}

void launchAssemble(const AsmArgs& A, int kind, bool emit, hipStream_t s, uint32_t clipped, uint32_t clipHint)
{
	const uint32_t blocks = A.nBlocks;
	if (!blocks) return;
	switch (kind)
	{
		case 0:
		{
			/* The clip list's length is on the device; a grid strides over
			 * it, 8 triangles per block (groups of CLIP_GROUP lanes).  A
			 * known count gets a group per triangle (up to 2048 blocks); an
			 * unknown one is sized from the caller's hint (the count of a
			 * similar earlier draw, a quarter more), at least 256 blocks
			 * (the empty grid of a draw that clips nothing still costs its
			 * dispatch), and never more blocks than the draw's main
			 * assembly grid. */
			if (emit) hipLaunchKernelGGL(k_assemble_tris<true>, dim3(blocks), dim3(128), 0, s, A);
			else hipLaunchKernelGGL(k_assemble_tris<false>, dim3(blocks), dim3(128), 0, s, A);
			launchClipped(A, emit, s, clipped, clipHint);
			break;
		}
		case 1:
			if (emit) hipLaunchKernelGGL(k_assemble_lines<true>, dim3(blocks), dim3(128), 0, s, A);
			else hipLaunchKernelGGL(k_assemble_lines<false>, dim3(blocks), dim3(128), 0, s, A);
			break;
		default:
			if (emit) hipLaunchKernelGGL(k_assemble_points<true>, dim3(blocks), dim3(128), 0, s, A);
			else hipLaunchKernelGGL(k_assemble_points<false>, dim3(blocks), dim3(128), 0, s, A);
			break;
	}
}

}  // namespace lib
}  // namespace srp
