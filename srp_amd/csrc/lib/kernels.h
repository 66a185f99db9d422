/* Library-internal kernels and their host launchers. */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srp/buffer.h"
#include "srp/framebuffer.h"
#include "srp/message_callback.h"
#include "srp/device/abi.h"
#include "srp/device/pipeline.hpp"

#define SRP_POOL 32       /* vertices one triangle's clip may hold: 3 originals + up to 29 generated */
#define SRP_MAX_POLY 16   /* vertices of a clipped polygon */
#define SRP_SMALL_BBOX 24 /* largest box side (P->maskBox) for exact tile masks at binning */

namespace srp {
namespace lib {

/* One draw of an assembly launch (device memory).  A launch covers one
 * draw or a batch of them; its blocks are split between the draws by
 * firstBlock, and the draws' inputs are laid out one after another in the
 * count arrays from primBase on. */
struct AsmDraw
{
	const SrpDrawParams* P;
	const uint8_t* indices;
	uint64_t start;       /* first stream index */
	uint64_t count;       /* stream indices of the draw (line loops wrap) */
	uint32_t prim;        /* SRPPrimitive */
	uint32_t nIn;         /* input primitives */
	uint32_t indexElem;   /* 0: non-indexed, else bytes per index */
	uint32_t primBase;    /* first slot of the draw in recOff/genOff/itemOff */
	uint32_t tileBase;    /* first tile of the draw's framebuffer in the tile keys */
	uint32_t firstBlock;  /* first block of the draw in the launch */
};

/* A triangle binned to every tile of its box (no exact masks): its tile
 * pairs are items item .. item + nx*ny - 1, tile (dx, dy) of the box at
 * item + dy*nx + dx with key keyBase + dy*stride + dx (k_expand_spans). */
struct SrpSpan
{
	uint32_t item;
	uint32_t rec;         /* triangle record, 0xFFFFFFFF: none (capacity exceeded) */
	uint32_t keyBase;     /* tile key of the box's first tile */
	uint32_t stride;      /* tiles per framebuffer row */
	uint32_t nx, ny;      /* box size in tiles */
};

/* Arguments of the assembly kernels (assemble.hip). */
/* Capacities an emitting pass was sized for before its counts were known,
 * and the number of triangles its clipping kernels handle (0: they were not
 * launched, so any clipped triangle makes the draw not fit). */
struct FitCaps
{
	unsigned long long records, generated, items, spans, clipped;
};

struct AsmArgs
{
	const AsmDraw* draws; /* device */
	uint32_t nDraws;
	uint32_t nBlocks;     /* blocks of 128 threads over all draws */
	/* COUNT pass: per-input counts; EMIT pass: their exclusive scans */
	uint32_t* recOff;
	uint32_t* genOff;
	uint32_t* itemOff;
	/* EMIT outputs */
	SrpTri* tris;
	SrpLine* lines;
	SrpPoint* points;
	uint32_t* pairKeys;       /* tile of each (tile, triangle) pair */
	uint32_t* pairVals;       /* pair index (sorted along with the keys) */
	uint32_t* pairRec;        /* triangle record | 0x80000000 when the mask is exact */
	unsigned long long* pairMask; /* coverage of the tile's 64 pixels */
	SrpFrag* frags;
	uint32_t* fragKeys;
	uint8_t* genVaryings;
	uint32_t* error;      /* bit 0: clip overflow; bit 1: a capacity below was exceeded;
	                         bit 2: a vertex index outside the prepared cache range */
	/* triangles (input slots) that need clipping, appended by the counting
	 * pass; the counter starts at zero and has room for every input */
	uint32_t* clipList;
	uint32_t* clipCount;
	/* spans: the counting pass counts them into *spanCount; the emitting
	 * pass appends them (from a zeroed *spanCount) for k_expand_spans */
	SrpSpan* spans;
	uint32_t* spanCount;
	uint32_t spanCap;
	/* capacities of the triangle outputs (records, tile pairs, generated
	 * vertices); what lies beyond them is dropped and flagged */
	uint32_t recCap, itemCap, genCap;
	/* Emitting pass of a single draw: its counter block (u32 words: [0..2]
	 * the counting pass's totals mod 2^32, [4] clipped triangles, [6] spans,
	 * [7] the outcome, [8..15] the three exact u64 totals).  Every block
	 * first checks the totals against fitCaps and does nothing unless they
	 * fit and there is something to emit; block 0 of the leading kernel
	 * records the outcome, totals[7] = 0 (fits), 1 (does not fit) or 2
	 * (nothing to emit), and totals[2] = the item count if it fits, else 0
	 * (the sort and tile ranges read it on the device), then with fitHost
	 * publishes the 16 words as launchPublish does.  A draw that does not
	 * fit thus leaves every buffer unchanged.  nullptr: no check. */
	uint32_t* fitTotals;
	FitCaps fitCaps;
	uint32_t* fitHost;
	uint32_t fitSeq;
	/* triangle counting pass: also sets zero[0..zeroCount) to 0 (the tile
	 * ranges the emitting half fills), saving a launch */
	uint32_t* zero;
	uint32_t zeroCount;
};

/* kind: 0 triangles, 1 lines, 2 points.  clipped: how many triangles the
 * counting pass listed for clipping, when the caller knows it (0 skips the
 * clipped-triangle kernel); with kClipUnknown its grid is sized from
 * clipHint (an expected count; any value is correct, the grid strides). */
constexpr uint32_t kClipUnknown = 0xFFFFFFFFu;
void launchAssemble(const AsmArgs& A, int kind, bool emit, hipStream_t s, uint32_t clipped = kClipUnknown,
                    uint32_t clipHint = 0);
/* The clipped-triangle kernel of a triangle launch alone (the same grid
 * rules), for a caller that runs it on a stream of its own. */
void launchClipped(const AsmArgs& A, bool emit, hipStream_t s, uint32_t clipped, uint32_t clipHint);
/* After the emitting pass of filled triangles: write the spans' pairs
 * (`blocks` blocks of 4 waves stride over the device-side span count). */
void launchExpandSpans(const AsmArgs& A, uint32_t blocks, hipStream_t s);

/* util.hip */
struct ScanScratch
{
	uint32_t* buf = nullptr;   /* block sums of every level */
	size_t cap = 0;
	unsigned long long* buf64 = nullptr;   /* exact block sums (multi-array scan) */
	size_t cap64 = 0;
};
/* Exclusive prefix sum of n u32 values in place; *total (device) = sum. */
void scanExclusive(uint32_t* data, size_t n, uint32_t* total, ScanScratch& scratch, hipStream_t s);
/* The same for nArrays arrays of n values `stride` apart; totals[a] = sums
 * (mod 2^32).  With totals64, totals64[a] receives the exact sums: the
 * scanned offsets are valid only while those fit 32 bits. */
void scanExclusiveMulti(uint32_t* base, size_t stride, int nArrays, size_t n, uint32_t* totals,
                        ScanScratch& scratch, hipStream_t s, unsigned long long* totals64 = nullptr);
/* Allocate now the scratch a multi-array scan of nArrays x n values (or a
 * radix sort of n items) needs, so that later calls allocate nothing. */
void scanReserve(ScanScratch& scratch, size_t n, int nArrays);
/* [start, end) of each tile's run in a tile-sorted key array.  With nDev,
 * n is a capacity and the count is min(n, *nDev), read on the device. */
void tileBounds(const uint32_t* keys, size_t n, uint32_t* start, uint32_t* end, hipStream_t s,
                const uint32_t* nDev = nullptr);

struct SortScratch
{
	uint32_t* keys = nullptr;
	uint32_t* vals = nullptr;
	uint32_t* hist = nullptr;
	uint32_t* grp = nullptr;   /* per-group digit sums and a counter (fused passes) */
	size_t cap = 0, histCap = 0;
	ScanScratch scan;
};
/* Stable LSD radix sort of (keys, vals) on the low `bits` bits of the keys.
 * The sorted arrays are returned through *outKeys / *outVals (either the
 * inputs or the scratch buffers).  With nDev, n is a capacity and the count
 * is min(n, *nDev), read on the device (no host synchronisation). */
void radixSort(uint32_t* keys, uint32_t* vals, size_t n, int bits, SortScratch& scratch,
               uint32_t** outKeys, uint32_t** outVals, hipStream_t s, const uint32_t* nDev = nullptr);
void radixSortReserve(SortScratch& scratch, size_t n);

/* 16 words of a draw's counter block copied to host-mapped memory
 * (hostCopy: its device address), then hostCopy[16] = seq once they are
 * visible to the host, which waits for that word. */
__device__ __forceinline__ void publish(const uint32_t* totals, uint32_t* hostCopy, uint32_t seq)
{
	for (int k = 0; k < 16; k++)
		hostCopy[k] = totals[k];
	__threadfence_system();
	__atomic_store_n(hostCopy + 16, seq, __ATOMIC_RELEASE);
}
void launchPublish(const uint32_t* totals, uint32_t* hostCopy, uint32_t seq, hipStream_t s);


void fillU32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
/* *dst = *src by a one-thread kernel (dst may be host-mapped memory) */
void copyU32(const uint32_t* src, uint32_t* dst, hipStream_t s);
void fillU8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
void iota(uint32_t* p, size_t n, hipStream_t s);
void fillFramebuffer(uint32_t* color, uint32_t cv, uint32_t* depth, uint32_t dv, uint8_t* stencil, uint8_t sv, size_t n,
                     hipStream_t s);
/* One framebuffer of a multi-framebuffer fill: its arrays (null: not
 * filled), values, pixels and first block in the launch. */
struct FillDesc
{
	uint32_t* color;
	uint32_t* depth;
	uint8_t* stencil;
	uint32_t colorValue, depthValue, stencilValue;
	uint32_t firstBlock;
	uint64_t n;
};
/* Blocks a framebuffer of n pixels gets in fillFramebuffers. */
uint32_t fillBlocksFor(size_t n);
/* Every framebuffer of desc (device, nDesc entries, blocks in all) in one
 * launch. */
void fillFramebuffers(const FillDesc* desc, uint32_t nDesc, uint32_t blocks, hipStream_t s);

/* runtime.cpp: the library stream, and the device arrays of a framebuffer
 * (pending deferred fills applied first; a synchronous framebuffer's host
 * arrays are uploaded first). */
hipStream_t libStream();
/* Report through the user's message callback (printf-style). */
void libMessage(SRPMessageType type, SRPMessageSeverity sev, const char* fn, const char* fmt, ...);
void framebufferDevice(const SRPFramebuffer* fb, uint32_t** color, float** depth, uint8_t** stencil);

}  // namespace lib
}  // namespace srp
