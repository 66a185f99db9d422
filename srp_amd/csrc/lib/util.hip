/* Data-parallel building blocks of the pipeline: exclusive scan, a stable
 * LSD radix sort used for ordered binning (tile lists, per-pixel fragment
 * lists), fills. */
#include "kernels.h"

namespace srp {
namespace lib {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;
static_assert(SCAN_ITEMS == 8, "k_scan_blocks loads a thread's values as two 16-byte vectors");

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_local(uint32_t* data, size_t n, uint32_t* sums)
{
	__shared__ uint32_t sh[SCAN_THREADS];
	const uint32_t t = threadIdx.x;
	const size_t base = (size_t) blockIdx.x * SCAN_CHUNK + (size_t) t * SCAN_ITEMS;
	uint32_t v[SCAN_ITEMS];
	uint32_t sum = 0;
#pragma unroll
	for (int j = 0; j < SCAN_ITEMS; j++)
	{
		const size_t i = base + j;
		v[j] = (i < n) ? data[i] : 0u;
		sum += v[j];
	}
	sh[t] = sum;
	__syncthreads();
	for (uint32_t off = 1; off < SCAN_THREADS; off <<= 1)
	{
		const uint32_t x = (t >= off) ? sh[t - off] : 0u;
		__syncthreads();
		sh[t] += x;
		__syncthreads();
	}
	uint32_t run = sh[t] - sum;
#pragma unroll
	for (int j = 0; j < SCAN_ITEMS; j++)
	{
		const size_t i = base + j;
		if (i < n) data[i] = run;
		run += v[j];
	}
	if (t == SCAN_THREADS - 1)
		sums[blockIdx.x] = sh[t];
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(uint32_t* data, size_t n, const uint32_t* sums)
{
	const uint32_t add = sums[blockIdx.x];
	const size_t base = (size_t) blockIdx.x * SCAN_CHUNK;
	for (int j = 0; j < SCAN_ITEMS; j++)
	{
		const size_t i = base + (size_t) j * SCAN_THREADS + threadIdx.x;
		if (i < n) data[i] += add;
	}
}

__global__ void k_copy1(const uint32_t* src, uint32_t* dst) { *dst = *src; }

void copyU32(const uint32_t* src, uint32_t* dst, hipStream_t s)
{
	hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, s, src, dst);
}

static void scanLevel(uint32_t* data, size_t n, uint32_t* total, uint32_t* scratch, hipStream_t s)
{
	const size_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
	uint32_t* sums = scratch;
	hipLaunchKernelGGL(k_scan_local, dim3((uint32_t) nb), dim3(SCAN_THREADS), 0, s, data, n, sums);
	if (nb == 1)
	{
		hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, s, sums, total);
		return;
	}
	scanLevel(sums, nb, total, scratch + nb, s);
	hipLaunchKernelGGL(k_scan_add, dim3((uint32_t) nb), dim3(SCAN_THREADS), 0, s, data, n, sums);
}

void scanExclusive(uint32_t* data, size_t n, uint32_t* total, ScanScratch& scratch, hipStream_t s)
{
	if (n == 0)
	{
		fillU32(total, 0u, 1, s);
		return;
	}
	size_t need = 0;
	for (size_t m = n; m > 1; m = (m + SCAN_CHUNK - 1) / SCAN_CHUNK)
		need += (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
	need += 2;
	if (need > scratch.cap)
	{
		if (scratch.buf) hipFree(scratch.buf);
		scratch.cap = need * 2;
		hipMalloc(&scratch.buf, scratch.cap * sizeof(uint32_t));
	}
	scanLevel(data, n, total, scratch.buf, s);
}

/* Two-launch exclusive scan of up to a few equal-length arrays (grid.y
 * selects the array): block-local scans and block totals, then each block
 * adds the sum of the totals before it (summed redundantly per block, which
 * is cheaper than a third launch while there are at most a few thousand
 * blocks).  The last block writes the array's grand total. */
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_blocks(uint32_t* base, size_t stride, size_t n, uint32_t* sums,
                                                              unsigned long long* sums64, uint32_t nb)
{
	uint32_t* data = base + (size_t) blockIdx.y * stride;
	__shared__ uint32_t waveTot[SCAN_THREADS / 64];
	__shared__ unsigned long long waveSum[SCAN_THREADS / 64];
	const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
	const size_t b0 = (size_t) blockIdx.x * SCAN_CHUNK + (size_t) t * SCAN_ITEMS;
	/* this thread's SCAN_ITEMS consecutive values: two 16-byte loads when
	 * they are all there and aligned (the arrays of one launch start at
	 * arbitrary word offsets) */
	const bool vec = b0 + SCAN_ITEMS <= n && (reinterpret_cast<uintptr_t>(data + b0) & 15u) == 0;
	uint32_t v[SCAN_ITEMS];
	if (vec)
	{
		const uint4 a = *reinterpret_cast<const uint4*>(data + b0);
		const uint4 b = *reinterpret_cast<const uint4*>(data + b0 + 4);
		v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
		v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
	}
	else
	{
#pragma unroll
		for (int j = 0; j < SCAN_ITEMS; j++)
			v[j] = (b0 + j < n) ? data[b0 + j] : 0u;
	}
	uint32_t sum = 0;
	unsigned long long wide = 0;
#pragma unroll
	for (int j = 0; j < SCAN_ITEMS; j++)
	{
		sum += v[j];
		wide += v[j];
	}
	/* inclusive scan of the threads' sums within the wave, then across the
	 * block's waves.  (The block's sum may wrap: the offsets below are
	 * exact only while the array's total fits 32 bits, which the caller
	 * checks; the 64-bit sums are exact.) */
	uint32_t incl = sum;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1)
	{
		const uint32_t x = __shfl_up(incl, off);
		if ((int) lane >= off) incl += x;
	}
	for (int off = 32; off > 0; off >>= 1)
		wide += __shfl_xor(wide, off);
	if (lane == 63u)
	{
		waveTot[wave] = incl;
		waveSum[wave] = wide;
	}
	__syncthreads();
	uint32_t run = incl - sum;
	for (uint32_t w = 0; w < wave; w++)
		run += waveTot[w];
#pragma unroll
	for (int j = 0; j < SCAN_ITEMS; j++)
	{
		const uint32_t x = v[j];
		v[j] = run;
		run += x;
	}
	if (vec)
	{
		*reinterpret_cast<uint4*>(data + b0) = make_uint4(v[0], v[1], v[2], v[3]);
		*reinterpret_cast<uint4*>(data + b0 + 4) = make_uint4(v[4], v[5], v[6], v[7]);
	}
	else
	{
#pragma unroll
		for (int j = 0; j < SCAN_ITEMS; j++)
			if (b0 + j < n) data[b0 + j] = v[j];
	}
	if (t == SCAN_THREADS - 1)
	{
		uint32_t bs = 0;
		unsigned long long w = 0;
		for (int k = 0; k < SCAN_THREADS / 64; k++)
		{
			bs += waveTot[k];
			w += waveSum[k];
		}
		sums[(size_t) blockIdx.y * nb + blockIdx.x] = bs;
		sums64[(size_t) blockIdx.y * nb + blockIdx.x] = w;
	}
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_finish(uint32_t* base, size_t stride, size_t n,
                                                              const uint32_t* sums, const unsigned long long* sums64,
                                                              uint32_t nb, uint32_t* totals,
                                                              unsigned long long* totals64)
{
	uint32_t* data = base + (size_t) blockIdx.y * stride;
	const uint32_t* s = sums + (size_t) blockIdx.y * nb;
	__shared__ uint32_t red[SCAN_THREADS / 64];
	const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
	uint32_t acc = 0;
	for (uint32_t k = t; k < blockIdx.x; k += SCAN_THREADS) acc += s[k];
	for (int off = 32; off > 0; off >>= 1)
		acc += __shfl_xor(acc, off);
	if (lane == 0)
		red[wave] = acc;
	__syncthreads();
	uint32_t add = 0;
	for (int w = 0; w < SCAN_THREADS / 64; w++)
		add += red[w];
	const size_t b0 = (size_t) blockIdx.x * SCAN_CHUNK;
	if (add)
		for (int j = 0; j < SCAN_ITEMS; j++)
		{
			const size_t i = b0 + (size_t) j * SCAN_THREADS + t;
			if (i < n) data[i] += add;
		}
	if (blockIdx.x == nb - 1)
	{
		if (t == 0)
			totals[blockIdx.y] = add + s[blockIdx.x];
		if (totals64)
		{
			/* the array's exact total (the last block reduces every block's) */
			__shared__ unsigned long long red64[SCAN_THREADS / 64];
			const unsigned long long* s64 = sums64 + (size_t) blockIdx.y * nb;
			unsigned long long acc64 = 0;
			for (uint32_t k = t; k < nb; k += SCAN_THREADS) acc64 += s64[k];
			for (int off = 32; off > 0; off >>= 1)
				acc64 += __shfl_xor(acc64, off);
			if (lane == 0)
				red64[wave] = acc64;
			__syncthreads();
			if (t == 0)
			{
				unsigned long long tot = 0;
				for (int w = 0; w < SCAN_THREADS / 64; w++)
					tot += red64[w];
				totals64[blockIdx.y] = tot;
			}
		}
	}
}

/* Exact (64-bit) sum of n u32 values added into *out (zeroed by the caller). */
__global__ __launch_bounds__(SCAN_THREADS) void k_sum64(const uint32_t* data, size_t n, unsigned long long* out)
{
	unsigned long long acc = 0;
	for (size_t i = (size_t) blockIdx.x * SCAN_THREADS + threadIdx.x; i < n; i += (size_t) gridDim.x * SCAN_THREADS)
		acc += data[i];
	for (int off = 32; off > 0; off >>= 1)
		acc += __shfl_xor(acc, off);
	if ((threadIdx.x & 63u) == 0 && acc)
		atomicAdd(out, acc);
}

void scanExclusiveMulti(uint32_t* base, size_t stride, int nArrays, size_t n, uint32_t* totals, ScanScratch& scratch,
                        hipStream_t s, unsigned long long* totals64)
{
	if (n == 0)
	{
		fillU32(totals, 0u, (size_t) nArrays, s);
		if (totals64)
			fillU32(reinterpret_cast<uint32_t*>(totals64), 0u, (size_t) nArrays * 2, s);
		return;
	}
	const size_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
	if (nb > 8192)
	{
		if (totals64)
		{
			fillU32(reinterpret_cast<uint32_t*>(totals64), 0u, (size_t) nArrays * 2, s);
			const size_t blocks = (n + SCAN_THREADS - 1) / SCAN_THREADS;
			const uint32_t g = (uint32_t) (blocks < 2048 ? blocks : 2048);
			for (int a = 0; a < nArrays; a++)
				hipLaunchKernelGGL(k_sum64, dim3(g), dim3(SCAN_THREADS), 0, s, base + (size_t) a * stride, n,
				                   totals64 + a);
		}
		for (int a = 0; a < nArrays; a++)
			scanExclusive(base + (size_t) a * stride, n, totals + a, scratch, s);
		return;
	}
	scanReserve(scratch, n, nArrays);
	const dim3 grid((uint32_t) nb, (uint32_t) nArrays);
	hipLaunchKernelGGL(k_scan_blocks, grid, dim3(SCAN_THREADS), 0, s, base, stride, n, scratch.buf, scratch.buf64,
	                   (uint32_t) nb);
	hipLaunchKernelGGL(k_scan_finish, grid, dim3(SCAN_THREADS), 0, s, base, stride, n, scratch.buf, scratch.buf64,
	                   (uint32_t) nb, totals, totals64);
}

/* Tile ranges of a list sorted by tile: [start, end) per tile; tiles that
 * received nothing keep start == end == 0 (set by the caller's fill). */
__global__ void k_tile_bounds(const uint32_t* keys, size_t n, const uint32_t* nDev, uint32_t* start, uint32_t* end)
{
	if (nDev && *nDev < n) n = *nDev;
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t k = keys[i];
	if (i == 0 || keys[i - 1] != k) start[k] = (uint32_t) i;
	if (i + 1 == n || keys[i + 1] != k) end[k] = (uint32_t) (i + 1);
}

void tileBounds(const uint32_t* keys, size_t n, uint32_t* start, uint32_t* end, hipStream_t s, const uint32_t* nDev)
{
	if (n) hipLaunchKernelGGL(k_tile_bounds, dim3((uint32_t) ((n + 255) / 256)), dim3(256), 0, s, keys, n, nDev, start, end);
}

/* ------------------------------------------------------------ radix sort */

constexpr int RS_THREADS = 256;
constexpr int RS_SLICES = 2;     /* slices of RS_THREADS items ranked per round */
constexpr int RS_CHUNK = 1024;   /* items per block: more blocks than CUs for the 1080p mesh's ~350 k pairs
                                   (measured: 4096 0.234 ms/frame, 2048 0.223, 1024 0.219) */
/* Items per block of a sort of more than RS_LONG_FROM items: each block
 * then writes longer runs per digit (RS_LONG_CHUNK / 256 keys on average,
 * whole cache lines rather than pieces of them), the digit counts to scan
 * shrink by the same factor, and there are still thousands of blocks. */
constexpr int RS_LONG_CHUNK = 32768;
constexpr size_t RS_LONG_FROM = (size_t) 4 << 20;

/* The block size a sort of n items uses (identical for its reservation). */
inline uint32_t rsChunk(size_t n)
{
	return n > RS_LONG_FROM ? (uint32_t) RS_LONG_CHUNK : (uint32_t) RS_CHUNK;
}

/* Up to this many blocks, each scatter block sums the per-block digit counts
 * itself (reading all of them from L2) instead of a separate scan over them:
 * two launches per pass fewer, at nBlocks KiB of L2 reads per block. */
constexpr uint32_t RS_FUSED_MAX_BLOCKS = 512;

/* Up to RS_FUSED_MAX_BLOCKS blocks, the digit counts are also summed per
 * group of RS_GROUP consecutive blocks (grp[group * 256 + digit], atomics),
 * so that a scatter block finds its offsets from the group rows and at most
 * RS_GROUP - 1 block rows.  The groups' sums are zero when a pass starts:
 * the last scatter block of each pass to read them clears them (and the
 * pass's counter of such blocks, grp[RS_GROUP_WORDS]). */
constexpr uint32_t RS_GROUP = 16;
constexpr uint32_t RS_GROUP_WORDS = RS_FUSED_MAX_BLOCKS / RS_GROUP * 256u;

/* Digit counts of one block of `chunk` keys, written at
 * hist[digit * digitStride + block * blockStride]; with grp, also added to
 * the block's group row. */
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t* keys, size_t n, const uint32_t* nDev,
                                                           uint32_t shift, uint32_t* hist, uint32_t digitStride,
                                                           uint32_t blockStride, uint32_t chunk, uint32_t* grp)
{
	if (nDev && *nDev < n) n = *nDev;
	__shared__ uint32_t h[256];
	const uint32_t t = threadIdx.x;
	h[t] = 0;
	__syncthreads();
	const size_t begin = (size_t) blockIdx.x * chunk;
	const size_t end = (begin + chunk < n) ? begin + chunk : n;
	for (size_t i = begin + t; i < end; i += RS_THREADS)
		atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
	__syncthreads();
	hist[(size_t) t * digitStride + (size_t) blockIdx.x * blockStride] = h[t];
	if (grp && h[t])
		atomicAdd(grp + (blockIdx.x / RS_GROUP) * 256u + t, h[t]);
}

/* Stable scatter: items keep their input order within a digit.  Each tile
 * of 256 items ranks its items per digit with wave ballots, then by wave.
 * FUSED: hist holds raw block-major counts (hist[block * 256 + digit]), grp
 * their sums per group of blocks, and the block finds its digits' offsets
 * itself; otherwise hist holds the exclusive scan of the digit-major
 * counts. */
template <bool FUSED>
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(const uint32_t* keysIn, const uint32_t* valsIn,
                                                              uint32_t* keysOut, uint32_t* valsOut, size_t n,
                                                              const uint32_t* nDev, uint32_t shift, const uint32_t* hist,
                                                              uint32_t nBlocks, uint32_t chunk, uint32_t* grp)
{
	if (nDev && *nDev < n) n = *nDev;
	/* blocks past the device count (a capacity-sized grid) have nothing to
	 * place; whole blocks leave together, before any barrier */
	if ((size_t) blockIdx.x * chunk >= n)
		return;
	/* hist rows of blocks past the count are zero: not read */
	const uint32_t usedBlocks = (uint32_t) ((n + chunk - 1) / chunk);
	__shared__ uint32_t running[256];
	__shared__ uint32_t sliceCnt[RS_SLICES * 4][256];
	__shared__ uint32_t tileTotal[256];
	__shared__ bool lastReader;
	uint32_t (*waveCnt)[256] = sliceCnt;   /* the offsets prologue's per-wave sums */
	const uint32_t t = threadIdx.x;
	const uint32_t lane = t & 63u, wave = t >> 6;
	if (FUSED)
	{
		/* Digit t: items of all blocks (total) and of the blocks before
		 * this one (before), from the group rows and then the block rows
		 * of this block's group before it.  Wave w reads every fourth row,
		 * a lane four digits at once; the four waves' sums meet in LDS. */
		__shared__ uint32_t befCnt[4][256];
		const uint32_t d4 = 4u * lane;
		const uint32_t myGroup = blockIdx.x / RS_GROUP;
		const uint32_t nGroups = (usedBlocks + RS_GROUP - 1) / RS_GROUP;
		uint32_t tot[4] = {0u, 0u, 0u, 0u}, bef[4] = {0u, 0u, 0u, 0u};
		for (uint32_t g = wave; g < nGroups; g += 4u)
		{
			const uint4 v = *reinterpret_cast<const uint4*>(grp + (size_t) g * 256u + d4);
			const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
			const uint32_t take = (g < myGroup) ? ~0u : 0u;
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				tot[k] += vv[k];
				bef[k] += vv[k] & take;
			}
		}
		for (uint32_t b = myGroup * RS_GROUP + wave; b < blockIdx.x; b += 4u)
		{
			const uint4 v = *reinterpret_cast<const uint4*>(hist + (size_t) b * 256u + d4);
			bef[0] += v.x;
			bef[1] += v.y;
			bef[2] += v.z;
			bef[3] += v.w;
		}
#pragma unroll
		for (int k = 0; k < 4; k++)
		{
			waveCnt[wave][d4 + k] = tot[k];
			befCnt[wave][d4 + k] = bef[k];
		}
		__syncthreads();
		const uint32_t total = waveCnt[0][t] + waveCnt[1][t] + waveCnt[2][t] + waveCnt[3][t];
		const uint32_t before = befCnt[0][t] + befCnt[1][t] + befCnt[2][t] + befCnt[3][t];
		/* exclusive scan of the totals over the 256 digits */
		uint32_t incl = total;
		for (int off = 1; off < 64; off <<= 1)
		{
			const uint32_t v = __shfl_up(incl, off);
			if ((int) lane >= off) incl += v;
		}
		if (lane == 63u)
			tileTotal[wave] = incl;
		__syncthreads();   /* also: every total read before waveCnt is reused below */
		uint32_t waveBase = 0;
		for (uint32_t w = 0; w < wave; w++)
			waveBase += tileTotal[w];
		running[t] = waveBase + incl - total + before;
		/* Every group row this block reads has been read (the sums above
		 * are done); the block that finishes reading last clears the rows
		 * and the counter for the next pass. */
		if (t == 0)
			lastReader = atomicAdd(grp + RS_GROUP_WORDS, 1u) == usedBlocks - 1u;
		__syncthreads();
	}
	else
		running[t] = hist[(size_t) t * nBlocks + blockIdx.x];
	const size_t begin = (size_t) blockIdx.x * chunk;
	const size_t end = (begin + chunk < n) ? begin + chunk : n;
	const uint64_t ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64u - lane));
	/* Rounds of RS_SLICES slices of RS_THREADS items; item order is slice,
	 * then wave, then lane, and so is the order of the per-(slice, wave)
	 * digit counts summed below. */
	for (size_t base = begin; base < end; base += (size_t) RS_SLICES * RS_THREADS)
	{
		for (int k = 0; k < RS_SLICES * 4; k++) sliceCnt[k][t] = 0;
		__syncthreads();
		uint32_t key[RS_SLICES], val[RS_SLICES], d[RS_SLICES], rank[RS_SLICES];
		bool valid[RS_SLICES];
#pragma unroll
		for (int sl = 0; sl < RS_SLICES; sl++)
		{
			const size_t i = base + (size_t) sl * RS_THREADS + t;
			valid[sl] = i < end;
			key[sl] = 0; val[sl] = 0; d[sl] = 0;
			if (valid[sl])
			{
				key[sl] = keysIn[i];
				val[sl] = valsIn[i];
				d[sl] = (key[sl] >> shift) & 255u;
			}
		}
#pragma unroll
		for (int sl = 0; sl < RS_SLICES; sl++)
		{
			uint64_t peers = __ballot(valid[sl]);
#pragma unroll
			for (int b = 0; b < 8; b++)
			{
				const bool bit = (d[sl] >> b) & 1u;
				const uint64_t m = __ballot(bit);
				peers &= bit ? m : ~m;
			}
			rank[sl] = (uint32_t) __popcll(peers & ltMask);
			if (valid[sl] && rank[sl] == 0)
				sliceCnt[sl * 4 + wave][d[sl]] = (uint32_t) __popcll(peers);
		}
		__syncthreads();
		{
			uint32_t sum = 0;
			for (int k = 0; k < RS_SLICES * 4; k++)
			{
				const uint32_t c = sliceCnt[k][t];
				sliceCnt[k][t] = sum;
				sum += c;
			}
			tileTotal[t] = sum;
		}
		__syncthreads();
#pragma unroll
		for (int sl = 0; sl < RS_SLICES; sl++)
			if (valid[sl])
			{
				const uint32_t dst = running[d[sl]] + sliceCnt[sl * 4 + wave][d[sl]] + rank[sl];
				keysOut[dst] = key[sl];
				valsOut[dst] = val[sl];
			}
		__syncthreads();
		running[t] += tileTotal[t];
	}
	if (FUSED && lastReader)
	{
		const uint32_t nGroups = (usedBlocks + RS_GROUP - 1) / RS_GROUP;
		for (uint32_t i = t; i < nGroups * 256u; i += RS_THREADS)
			grp[i] = 0u;
		if (t == 0)
			grp[RS_GROUP_WORDS] = 0u;
	}
}

/* Scratch of a multi-array scan of nArrays x n values, allocated now. */
void scanReserve(ScanScratch& sc, size_t n, int nArrays)
{
	if (n == 0)
		return;
	const size_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
	size_t need;
	if (nb > 8192)
	{
		need = 2;
		for (size_t m = n; m > 1; m = (m + SCAN_CHUNK - 1) / SCAN_CHUNK)
			need += (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
	}
	else
		need = nb * (size_t) nArrays + 2;
	if (need > sc.cap)
	{
		if (sc.buf) hipFree(sc.buf);
		sc.cap = need * 2;
		hipMalloc(&sc.buf, sc.cap * sizeof(uint32_t));
	}
	if (nb <= 8192 && need > sc.cap64)
	{
		if (sc.buf64) hipFree(sc.buf64);
		sc.cap64 = need * 2;
		hipMalloc(&sc.buf64, sc.cap64 * sizeof(unsigned long long));
	}
}

void radixSortReserve(SortScratch& sc, size_t n)
{
	if (n <= 1)
		return;
	const uint32_t chunk = rsChunk(n);
	const uint32_t nBlocks = (uint32_t) ((n + chunk - 1) / chunk);
	if (n > sc.cap)
	{
		if (sc.keys) hipFree(sc.keys);
		if (sc.vals) hipFree(sc.vals);
		sc.cap = n + n / 2;
		hipMalloc(&sc.keys, sc.cap * sizeof(uint32_t));
		hipMalloc(&sc.vals, sc.cap * sizeof(uint32_t));
	}
	const size_t histN = (size_t) 256 * nBlocks;
	if (histN + 1 > sc.histCap)
	{
		if (sc.hist) hipFree(sc.hist);
		sc.histCap = (histN + 1) * 2;
		hipMalloc(&sc.hist, sc.histCap * sizeof(uint32_t));
	}
	if (nBlocks > RS_FUSED_MAX_BLOCKS)
		scanReserve(sc.scan, histN, 1);
	else if (!sc.grp)
	{
		/* zeroed once (a copy, complete on return); every fused pass leaves
		 * it zero again */
		static const uint32_t zeros[RS_GROUP_WORDS + 1] = {};
		hipMalloc(&sc.grp, sizeof(zeros));
		hipMemcpy(sc.grp, zeros, sizeof(zeros), hipMemcpyHostToDevice);
	}
}

void radixSort(uint32_t* keys, uint32_t* vals, size_t n, int bits, SortScratch& sc,
               uint32_t** outKeys, uint32_t** outVals, hipStream_t s, const uint32_t* nDev)
{
	*outKeys = keys;
	*outVals = vals;
	if (n <= 1 || bits <= 0)
		return;
	const uint32_t chunk = rsChunk(n);
	const uint32_t nBlocks = (uint32_t) ((n + chunk - 1) / chunk);
	radixSortReserve(sc, n);
	const size_t histN = (size_t) 256 * nBlocks;
	uint32_t* kIn = keys;
	uint32_t* vIn = vals;
	uint32_t* kOut = sc.keys;
	uint32_t* vOut = sc.vals;
	const bool fused = nBlocks <= RS_FUSED_MAX_BLOCKS;
	for (int shift = 0; shift < bits; shift += 8)
	{
		if (fused)
		{
			hipLaunchKernelGGL(k_radix_hist, dim3(nBlocks), dim3(RS_THREADS), 0, s, kIn, n, nDev, (uint32_t) shift,
			                   sc.hist, 1u, 256u, chunk, sc.grp);
			hipLaunchKernelGGL(k_radix_scatter<true>, dim3(nBlocks), dim3(RS_THREADS), 0, s, kIn, vIn, kOut, vOut, n,
			                   nDev, (uint32_t) shift, sc.hist, nBlocks, chunk, sc.grp);
		}
		else
		{
			hipLaunchKernelGGL(k_radix_hist, dim3(nBlocks), dim3(RS_THREADS), 0, s, kIn, n, nDev, (uint32_t) shift,
			                   sc.hist, nBlocks, 1u, chunk, (uint32_t*) nullptr);
			scanExclusiveMulti(sc.hist, histN, 1, histN, sc.hist + histN, sc.scan, s);
			hipLaunchKernelGGL(k_radix_scatter<false>, dim3(nBlocks), dim3(RS_THREADS), 0, s, kIn, vIn, kOut, vOut, n,
			                   nDev, (uint32_t) shift, sc.hist, nBlocks, chunk, (uint32_t*) nullptr);
		}
		uint32_t* tk = kIn; kIn = kOut; kOut = tk;
		uint32_t* tv = vIn; vIn = vOut; vOut = tv;
	}
	*outKeys = kIn;
	*outVals = vIn;
}

/* ------------------------------------------------------------------ fills */

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	const size_t stride = (size_t) gridDim.x * blockDim.x;
	for (size_t k = i; k < n; k += stride) p[k] = v;
}

__global__ void k_fill_u8(uint8_t* p, uint8_t v, size_t n)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	const size_t stride = (size_t) gridDim.x * blockDim.x;
	for (size_t k = i; k < n; k += stride) p[k] = v;
}

__global__ void k_iota(uint32_t* p, size_t n)
{
	const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	const size_t stride = (size_t) gridDim.x * blockDim.x;
	for (size_t k = i; k < n; k += stride) p[k] = (uint32_t) k;
}

static dim3 fillGrid(size_t n)
{
	size_t b = (n + 255) / 256;
	if (b > 4096) b = 4096;
	if (b == 0) b = 1;
	return dim3((uint32_t) b);
}

void fillU32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
	if (n) hipLaunchKernelGGL(k_fill_u32, fillGrid(n), dim3(256), 0, s, p, v, n);
}

void fillU8(uint8_t* p, uint8_t v, size_t n, hipStream_t s)
{
	if (n) hipLaunchKernelGGL(k_fill_u8, fillGrid(n), dim3(256), 0, s, p, v, n);
}

/* Colour, depth and stencil fills of one framebuffer in one launch (a null
 * array is left alone). */
__global__ void k_fill_fb(uint32_t* color, uint32_t cv, uint32_t* depth, uint32_t dv, uint8_t* stencil, uint8_t sv, size_t n)
{
	const size_t i0 = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	const size_t stride = (size_t) gridDim.x * blockDim.x;
	for (size_t i = i0; i < n; i += stride)
	{
		if (color) color[i] = cv;
		if (depth) depth[i] = dv;
	}
	/* stencil bytes four at a time, then the tail */
	const size_t n4 = n / 4;
	const uint32_t sv4 = (uint32_t) sv * 0x01010101u;
	if (stencil)
	{
		uint32_t* s4 = reinterpret_cast<uint32_t*>(stencil);
		for (size_t i = i0; i < n4; i += stride) s4[i] = sv4;
		for (size_t i = n4 * 4 + i0; i < n; i += stride) stencil[i] = sv;
	}
}

void fillFramebuffer(uint32_t* color, uint32_t cv, uint32_t* depth, uint32_t dv, uint8_t* stencil, uint8_t sv, size_t n,
                     hipStream_t s)
{
	if (n) hipLaunchKernelGGL(k_fill_fb, fillGrid(n), dim3(256), 0, s, color, cv, depth, dv, stencil, sv, n);
}

/* Block b fills framebuffer d, the last descriptor with firstBlock <= b,
 * with the blocks from its firstBlock to the next one's. */
__global__ void k_fill_many(const FillDesc* desc, uint32_t nDesc, uint32_t blocks)
{
	uint32_t a = 0, z = nDesc;
	while (z - a > 1)
	{
		const uint32_t m = (a + z) >> 1;
		if (desc[m].firstBlock <= blockIdx.x) a = m;
		else z = m;
	}
	const FillDesc D = desc[a];
	const uint32_t nb = ((a + 1 < nDesc) ? desc[a + 1].firstBlock : blocks) - D.firstBlock;
	const size_t i0 = (size_t) (blockIdx.x - D.firstBlock) * blockDim.x + threadIdx.x;
	const size_t stride = (size_t) nb * blockDim.x;
	for (size_t i = i0; i < D.n; i += stride)
	{
		if (D.color) D.color[i] = D.colorValue;
		if (D.depth) D.depth[i] = D.depthValue;
		if (D.stencil) D.stencil[i] = (uint8_t) D.stencilValue;
	}
}

uint32_t fillBlocksFor(size_t n)
{
	const size_t b = (n + 1023) / 1024;
	return (uint32_t) (b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void fillFramebuffers(const FillDesc* desc, uint32_t nDesc, uint32_t blocks, hipStream_t s)
{
	if (nDesc && blocks)
		hipLaunchKernelGGL(k_fill_many, dim3(blocks), dim3(256), 0, s, desc, nDesc, blocks);
}

void iota(uint32_t* p, size_t n, hipStream_t s)
{
	if (n) hipLaunchKernelGGL(k_iota, fillGrid(n), dim3(256), 0, s, p, n);
}

__global__ void k_publish(const uint32_t* totals, uint32_t* hostCopy, uint32_t seq)
{
	publish(totals, hostCopy, seq);
}

void launchPublish(const uint32_t* totals, uint32_t* hostCopy, uint32_t seq, hipStream_t s)
{
	hipLaunchKernelGGL(k_publish, dim3(1), dim3(1), 0, s, totals, hostCopy, seq);
}


}  // namespace lib
}  // namespace srp
