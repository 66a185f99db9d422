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
	return (uint32_t) (vi - W.P->cacheBase);
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

/* ------------------------------------------------------------ clip pool */

/* Vertices produced while clipping one primitive: the originals first, then
 * each generated vertex as (parent a, parent b, t) with position
 * a*(1-t) + b*t (clipping.c:244-258). */
struct Pool
{
	float c[SRP_POOL][4];
	uint8_t a[SRP_POOL], b[SRP_POOL];
	float t[SRP_POOL];
	uint32_t slot[3];
	int n;
};

__device__ __forceinline__ int poolLerp(Pool& pl, int ia, int ib, float t)
{
	const int k = pl.n;
	if (k >= SRP_POOL)
		return -1;
	for (int i = 0; i < 4; i++)
		pl.c[k][i] = pl.c[ia][i] * (1 - t) + pl.c[ib][i] * t;
	pl.a[k] = (uint8_t) ia;
	pl.b[k] = (uint8_t) ib;
	pl.t[k] = t;
	pl.n = k + 1;
	return k;
}

/* One Sutherland-Hodgman pass (clipping.c:199-242).  An edge that crosses
 * the plane with |da-db| <= 1e-9 contributes nothing, not even its inside
 * endpoint (clipping.c:226-227).  Returns the new count, -1 on overflow. */
__device__ int clipPass(Pool& pl, const uint8_t* in, int n, int plane, uint8_t* out)
{
	int m = 0;
	for (int i = 0; i < n; i++)
	{
		const int cur = in[i], nxt = in[(i + 1) % n];
		const float da = planeDist(pl.c[cur], plane);
		const float db = planeDist(pl.c[nxt], plane);
		const bool ci = da >= 0.0f, ni = db >= 0.0f;
		if (ci && ni)
			out[m++] = (uint8_t) nxt;
		else if (ci || ni)
		{
			if (roughlyZero(da - db))
				continue;
			const float t = da / (da - db);
			const int k = poolLerp(pl, cur, nxt, t);
			if (k < 0)
				return -1;
			out[m++] = (uint8_t) k;
			if (!ci && ni)
				out[m++] = (uint8_t) nxt;
		}
		if (m + 2 > SRP_MAX_POLY)
			return -1;
	}
	return m;
}

/* clipping.c:68-119: the clipped polygon as pool indices; -1 on overflow. */
__device__ int clipTriangle(Pool& pl, uint8_t* poly)
{
	const uint32_t c0 = outcode(pl.c[0]), c1 = outcode(pl.c[1]), c2 = outcode(pl.c[2]);
	poly[0] = 0; poly[1] = 1; poly[2] = 2;
	if ((c0 | c1 | c2) == 0)
		return 3;
	if ((c0 & c1 & c2) != 0)
		return 0;
	uint8_t tmp[SRP_MAX_POLY];
	uint8_t* src = poly;
	uint8_t* dst = tmp;
	int n = 3;
	for (int p = 0; p < 6; p++)
	{
		n = clipPass(pl, src, n, p, dst);
		if (n <= 0)
			return n;
		uint8_t* s = src; src = dst; dst = s;
	}
	if (src != poly)
		for (int i = 0; i < n; i++) poly[i] = src[i];
	return n;
}

/* ----------------------------------------------------- generated varyings */

/* Value of one float element for pool vertex k under the clip rule
 * (interpolation with invW == NULL: every float varying is blended
 * affinely, interpolation.c:116-120, weights {1-t, t}).  Parents always
 * precede their children in the pool, so one forward sweep suffices. */
__device__ float genFloat(const SrpDrawParams* P, const Pool& pl, int k, uint32_t off)
{
	float val[SRP_POOL];
	for (int q = 0; q <= k; q++)
	{
		if (q < 3)
			val[q] = srp::dev::loadF(srp::dev::varyingPtr(P, pl.slot[q]) + off);
		else
		{
			const float t = pl.t[q];
			const float w0 = 1 - t;
			float v = 0.0f;
			v += val[pl.a[q]] * w0;
			v += val[pl.b[q]] * t;
			val[q] = v;
		}
	}
	return val[k];
}

__device__ double genDouble(const SrpDrawParams* P, const Pool& pl, int k, uint32_t off)
{
	double val[SRP_POOL];
	for (int q = 0; q <= k; q++)
	{
		if (q < 3)
			val[q] = srp::dev::loadD(srp::dev::varyingPtr(P, pl.slot[q]) + off);
		else
		{
			const float t = pl.t[q];
			const float w0 = 1 - t;
			double v = 0.0;
			v += val[pl.a[q]] * (double) w0;
			v += val[pl.b[q]] * (double) t;
			val[q] = v;
		}
	}
	return val[k];
}

/* Integers come from the pair's provoking vertex: b for LAST, a for FIRST. */
__device__ __forceinline__ int genIntSource(const SrpDrawParams* P, const Pool& pl, int k)
{
	while (k >= 3)
		k = P->st.provokingLast ? pl.b[k] : pl.a[k];
	return k;
}

__device__ void materialize(const SrpDrawParams* P, const Pool& pl, int k, uint8_t* dst)
{
	for (uint32_t a = 0; a < P->nVaryings; a++)
	{
		const SrpVaryingDesc d = P->var[a];
		if (d.type == SRP_FLOAT)
			for (uint32_t e = 0; e < d.nItems; e++)
			{
				const float v = genFloat(P, pl, k, d.offset + 4 * e);
				__builtin_memcpy(dst + d.offset + 4 * e, &v, 4);
			}
		else if (d.type == SRP_DOUBLE)
			for (uint32_t e = 0; e < d.nItems; e++)
			{
				const double v = genDouble(P, pl, k, d.offset + 8 * e);
				__builtin_memcpy(dst + d.offset + 8 * e, &v, 8);
			}
		else
		{
			const uint32_t n = srp::dev::elemSize(d.type) * d.nItems;
			const uint8_t* src = srp::dev::varyingPtr(P, pl.slot[genIntSource(P, pl, k)]) + d.offset;
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
		int y = T.y0;
		const uint32_t allCols = (1u << (tx1 - tx0 + 1u)) - 1u;
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
				for (int x = T.x0; x < xEnd; x++)
				{
					bool inside = true;
					for (int e = 0; e < 3; e++)
					{
						const bool tl = (T.flags >> e) & 1u;
						inside = inside && ((l[e] > 0.0f) || (roughlyZero(l[e]) && tl));
					}
					if (inside)
					{
						const uint32_t c = ((uint32_t) x >> 3) - tx0;
						const uint64_t bit = 1ull << ((((uint32_t) y & 7u) << 3) | ((uint32_t) x & 7u));
						m0 |= c == 0 ? bit : 0ull;
						m1 |= c == 1 ? bit : 0ull;
						m2 |= c == 2 ? bit : 0ull;
						m3 |= c == 3 ? bit : 0ull;
					}
					for (int e = 0; e < 3; e++) l[e] += T.dx[e];
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
	for (uint32_t ty = ty0; ty <= ty1; ty++)
		for (uint32_t tx = tx0; tx <= tx1; tx++)
		{
			emit(tileBase + ty * P->tilesX + tx, ~0ull, false);
			n++;
		}
	return n;
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

/* Assembled primitives of one clipped polygon v[0..n-1] (fan from v[0],
 * clipping.c:107-116), per polygon mode (primitive_assembly.c:80-128). */
template <bool EMIT, int NMAX>
__device__ __forceinline__ void emitFan(const AsmArgs& A, const AsmDraw& W, const Vtx* v, int n, uint32_t recBase,
                                        uint32_t idBase, uint32_t itemBase, uint32_t& nRec, uint32_t& nItems,
                                        bool drop = false)
{
	const SrpDrawParams* P = W.P;
	const uint32_t mode = P->st.polygonMode;
#pragma unroll
	for (int f = 1; f + 1 < NMAX; f++)
	{
		if (f + 1 >= n)
			break;
		const Vtx& a = v[0];
		const Vtx& b = v[f];
		const Vtx& c = v[f + 1];
		if (mode == SRP_POLYGON_MODE_FILL)
		{
			SrpTri T;
			if (!setupTriangle(P, a, b, c, recBase + nRec - idBase, T))
				continue;
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
				if (o > A.itemCap)
					atomicOr(A.error, 2u);
			}
			else
				np = binTriangle<false>(P, T, W.tileBase, [](uint32_t, uint64_t, bool) {});
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
				nItems += pointFrags<EMIT>(A, P, pt, recBase + nRec, itemBase + nItems);
				nRec++;
			};
			corner(a);
			corner(b);
			corner(c);
		}
	}
}

/* A triangle that needs clipping: Sutherland-Hodgman over the pool, then
 * the varyings of the generated vertices that survive. */
template <bool EMIT>
__device__ __forceinline__ void assembleClipped(const AsmArgs& A, const AsmDraw& W, const uint32_t* slot, uint32_t genBase,
                                             uint32_t recBase, uint32_t idBase, uint32_t itemBase,
                                             uint32_t& nRec, uint32_t& nGen, uint32_t& nItems)
{
	const SrpDrawParams* P = W.P;
	Pool pl;
	for (int i = 0; i < 3; i++)
	{
		pl.slot[i] = slot[i];
		loadClip(P, slot[i], pl.c[i]);
	}
	pl.n = 3;
	uint8_t poly[SRP_MAX_POLY];
	int n = clipTriangle(pl, poly);
	if (n < 0)
	{
		atomicOr(A.error, 1u);
		n = 0;
	}
	Vtx v[SRP_MAX_POLY];
	bool drop = false;
	for (int i = 0; i < n; i++)
	{
		const int q = poly[i];
		for (int c = 0; c < 4; c++) v[i].c[c] = pl.c[q][c];
		v[i].invW = divide(v[i].c);
		if (q < 3)
			v[i].vref = pl.slot[q];
		else
		{
			v[i].vref = P->cacheSize + genBase + nGen;
			if (EMIT)
			{
				if (genBase + nGen < A.genCap)
					materialize(P, pl, q, A.genVaryings + (size_t) (genBase + nGen) * P->varyingsSize);
				else
					drop = true;
			}
			nGen++;
		}
	}
	emitFan<EMIT, SRP_MAX_POLY>(A, W, v, n, recBase, idBase, itemBase, nRec, nItems, drop);
}

/* Triangle draws (all polygon modes): primitive_assembly.c:37-135.  The
 * common case — all three vertices inside the frustum, trivially accepted
 * by clipping.c:74-78 — stays in registers. */
template <bool EMIT>
__global__ __launch_bounds__(128) void k_assemble_tris(AsmArgs A)
{
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
	{
		slot[i] = slotOf(W, s[i]);
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
		emitFan<EMIT, 3>(A, W, v, 3, recBase, idBase, itemBase, nRec, nItems);
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

/* The triangles k_assemble_tris listed for clipping, in either pass (the
 * list stays valid from the counting pass to the emitting one).  Each
 * writes only at its own scanned offsets, so list order does not matter. */
template <bool EMIT>
__global__ __launch_bounds__(128) void k_assemble_clipped(AsmArgs A)
{
	const uint32_t n = *A.clipCount;
	for (uint32_t i = blockIdx.x * 128u + threadIdx.x; i < n; i += gridDim.x * 128u)
	{
		const uint32_t g = A.clipList[i];
		const uint32_t d = A.nDraws > 1 ? drawOfInput(A.draws, A.nDraws, g) : 0u;
		const AsmDraw& W = A.draws[d];
		const uint32_t k = g - W.primBase;
		uint64_t s[3];
		triStream(W.prim, W.start, k, s);
		uint32_t slot[3];
		for (int j = 0; j < 3; j++)
			slot[j] = slotOf(W, s[j]);
		uint32_t recBase = 0, genBase = 0, itemBase = 0, idBase = 0;
		if (EMIT)
		{
			recBase = A.recOff[g];
			genBase = A.genOff[g];
			itemBase = A.itemOff[g];
			idBase = A.recOff[W.primBase];
		}
		uint32_t nRec = 0, nGen = 0, nItems = 0;
		assembleClipped<EMIT>(A, W, slot, genBase, recBase, idBase, itemBase, nRec, nGen, nItems);
		if (!EMIT)
		{
			A.recOff[g] = nRec;
			A.genOff[g] = nGen;
			A.itemOff[g] = nItems;
		}
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
	const uint32_t d = A.nDraws > 1 ? drawOfBlock(A.draws, A.nDraws, blockIdx.x) : 0u;
	const AsmDraw& W = A.draws[d];
	const uint32_t k = (blockIdx.x - W.firstBlock) * 128u + threadIdx.x;
	if (k >= W.nIn)
		return;
	const uint32_t g = W.primBase + k;   /* slot of this input in the count arrays */
	const SrpDrawParams* P = W.P;
	Pool pl;
	uint64_t s[2];
	lineStream(W.prim, W.start, k, W.count, s);
	for (int i = 0; i < 2; i++)
	{
		pl.slot[i] = slotOf(W, s[i]);
		loadClip(P, pl.slot[i], pl.c[i]);
	}
	pl.slot[2] = pl.slot[1];
	pl.n = 2;

	/* Liang-Barsky style clip: both endpoints re-derived from the originals. */
	int ia = 0, ib = 1;
	bool rejected = false;
	const uint32_t c0 = outcode(pl.c[0]), c1 = outcode(pl.c[1]);
	if ((c0 | c1) != 0)
	{
		if ((c0 & c1) != 0)
			rejected = true;
		else
		{
			float t0 = 0.0f, t1 = 1.0f;
			for (int p = 0; p < 6 && !rejected; p++)
			{
				const float da = planeDist(pl.c[0], p), db = planeDist(pl.c[1], p);
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
			if (!rejected)
			{
				/* pool slot 2 is unused for lines; generated vertices start at 3 */
				pl.n = 3;
				if (t0 > 0.0f) ia = poolLerp(pl, 0, 1, t0);
				if (t1 < 1.0f) ib = poolLerp(pl, 0, 1, t1);
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
		const int q[2] = {ia, ib};
		for (int i = 0; i < 2; i++)
		{
			for (int c = 0; c < 4; c++) e[i].c[c] = pl.c[q[i]][c];
			e[i].invW = divide(e[i].c);
			if (q[i] < 3)
				e[i].vref = pl.slot[q[i]];
			else
			{
				e[i].vref = P->cacheSize + genBase + nGen;
				if (EMIT)
					materialize(P, pl, q[i], A.genVaryings + (size_t) (genBase + nGen) * P->varyingsSize);
				nGen++;
			}
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
	const uint32_t d = A.nDraws > 1 ? drawOfBlock(A.draws, A.nDraws, blockIdx.x) : 0u;
	const AsmDraw& W = A.draws[d];
	const uint32_t k = (blockIdx.x - W.firstBlock) * 128u + threadIdx.x;
	if (k >= W.nIn)
		return;
	const uint32_t g = W.primBase + k;   /* slot of this input in the count arrays */
	const SrpDrawParams* P = W.P;
	const uint32_t slot = slotOf(W, W.start + k);
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

void launchAssemble(const AsmArgs& A, int kind, bool emit, hipStream_t s, uint32_t clipped)
{
	const uint32_t blocks = A.nBlocks;
	if (!blocks) return;
	switch (kind)
	{
		case 0:
		{
			/* The clip list's length is on the device; a grid strides over
			 * it.  Each clipped triangle is a long serial job, so a known
			 * count gets one thread per triangle (up to 1024 blocks); an
			 * unknown one gets 256 blocks, a compromise between a frame that
			 * clips tens of thousands of triangles and one that clips none
			 * (the empty grid still costs its dispatch: 1024 blocks 8.8 us,
			 * 128 blocks 4.6 us, on the 1080p mesh). */
			uint32_t clipBlocks;
			if (clipped != kClipUnknown)
			{
				const uint32_t need = (uint32_t) (((uint64_t) clipped + 127u) / 128u);
				clipBlocks = need < 1024u ? need : 1024u;
			}
			else
				clipBlocks = 256u;
			clipBlocks = clipBlocks < blocks ? clipBlocks : blocks;
			if (emit)
			{
				hipLaunchKernelGGL(k_assemble_tris<true>, dim3(blocks), dim3(128), 0, s, A);
				if (clipBlocks)
					hipLaunchKernelGGL(k_assemble_clipped<true>, dim3(clipBlocks), dim3(128), 0, s, A);
			}
			else
			{
				hipLaunchKernelGGL(k_assemble_tris<false>, dim3(blocks), dim3(128), 0, s, A);
				if (clipBlocks)
					hipLaunchKernelGGL(k_assemble_clipped<false>, dim3(clipBlocks), dim3(128), 0, s, A);
			}
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
