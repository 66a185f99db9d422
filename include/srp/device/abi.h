/* Device-side data layouts shared by libsrp_amd and HIP shader modules.
 *
 * The library runs the shader-independent stages (assembly, clipping,
 * setup, binning, sorting) and writes the records below; the kernels that
 * call user shaders (vertex stage, rasteriser, fragment lists) are
 * instantiated inside the shader module from srp/device/pipeline.hpp and
 * read the same records.  SRP_ABI_VERSION guards against mixing builds. */
#pragma once

#include <stddef.h>
#include <stdint.h>
#include "srp/context.h"

#define SRP_ABI_VERSION 16u
#define SRP_MAX_VARYINGS 16
#define SRP_MAX_WORDS 32          /* fast interpolation path: 4-byte words */
#define SRP_MAX_VARYING_BYTES 256 /* general path: bytes per fragment */
/* Fragment counters (SrpDrawParams.stats): SRP_STAT_SLOTS slots of
 * SRP_STAT_STRIDE u64 (one 128-byte line each), slot s holding covered at
 * [s * STRIDE] and shaded at [s * STRIDE + 1]; waves add to slot
 * (block % SLOTS) so that their atomics do not all meet on one address. */
/* The largest float not above the reference's EPSILON (1e-9, as a double;
 * math/utils.h:45 tests fabs(x) <= EPSILON on a float promoted to double).
 * For every float x, |x| <= 1e-9 holds exactly when |x| <= this, so the
 * test needs no double arithmetic. */
#define SRP_EPSILON_F 0x1.12e0bep-30f
/* The reference's edge test (raster/triangle.c:78-110) admits a pixel on
 * edge e when l > 0, or when |l| <= EPSILON and the edge is top-left.  For
 * a top-left edge that is l >= -SRP_EPSILON_F, i.e. l above the next float
 * below it, this; for any other edge l > 0.  So each edge is one compare
 * with a per-edge threshold (NaN fails both forms). */
#define SRP_EDGE_TL_THRESHOLD (-0x1.12e0c0p-30f)
#define SRP_STAT_SLOTS 64
#define SRP_STAT_STRIDE 16

/* interpolation kinds of a 4-byte word in the fast path */
enum SrpWordKind
{
	SRP_WORD_PERSPECTIVE = 0,  /* float, perspective-correct */
	SRP_WORD_AFFINE = 1,       /* float, screen-space linear */
	SRP_WORD_PROVOKING = 2,    /* flat float or integer bytes: provoking vertex */
	SRP_WORD_UNUSED = 3        /* padding beyond the described varyings */
};

typedef struct SrpVaryingDesc
{
	uint32_t nItems;
	uint32_t type;    /* SRPType */
	uint32_t mode;    /* SRPInterpolationMode */
	uint32_t offset;  /* byte offset of the first element */
} SrpVaryingDesc;

/* Render state captured when the draw is issued. */
typedef struct SrpState
{
	uint32_t provokingLast;
	uint32_t cullFace, frontFace, polygonMode;
	float pointSize;
	uint32_t scissorEnabled;
	uint64_t sx0, sx1, sy0, sy1;      /* scissor box, half-open */
	uint32_t stencilEnabled;
	SRPStencilFaceState stencil[2];   /* [0] front, [1] back */
	uint32_t depthTest, depthWrite, depthOp;
	uint32_t lateZ;                   /* fragment shader may overwrite depth */
} SrpState;

/* Triangle after clipping and setup (one per assembled primitive). */
typedef struct SrpTri
{
	float l0[3];           /* barycentrics at the bbox's first pixel centre */
	float dx[3], dy[3];    /* per-pixel and per-row increments */
	uint16_t x0, y0, x1, y1; /* bounding box [x0,x1) x [y0,y1) */
	uint32_t flags;        /* bits 0-2: top-left edge; bit 3: front facing */
	uint32_t id;           /* primitive ID */
	float invW[3];
	float zw[3];           /* ndc z * invW per vertex */
	uint32_t vref[3];      /* varying slots after the winding fix */
	uint32_t pad;
} SrpTri;

typedef struct SrpLine
{
	float ssx[2], ssy[2];
	float invW[2];
	float zw[2];
	uint32_t vref[2];
	uint32_t id;
	uint32_t pad;
} SrpLine;

typedef struct SrpPoint
{
	float ssx, ssy;
	float z, w;            /* ndc z and w (w is 1 after the divide) */
	uint32_t vref;
	uint32_t id;
	uint32_t pad[2];
} SrpPoint;

/* One line or point fragment, in primitive order. */
typedef struct SrpFrag
{
	int32_t px, py;        /* unwrapped pixel (a line may round to x == W) */
	uint32_t prim;         /* record index */
	float t;               /* line parameter (lines only) */
} SrpFrag;

typedef struct SrpTexDesc
{
	const uint32_t* texels;  /* RGBA8, row-major, R in the low byte */
	int32_t w, h, wM1, hM1;
	int32_t wrapX, wrapY;
} SrpTexDesc;

/* Everything a draw's kernels read, in one device block. */
typedef struct SrpDrawParams
{
	SrpState st;
	uint32_t width, height;
	uint32_t* color;
	float* depth;
	uint8_t* stencil;
	const uint8_t* uniform;

	/* vertex stage */
	const uint8_t* vertices;
	uint64_t vertexStride;
	uint64_t cacheBase;      /* vertex index of cache slot 0 */
	uint32_t cacheSize;
	float* clip;             /* 4 floats per cache slot */
	uint8_t* varyings;       /* varyings of the cache slots */
	uint8_t* genVaryings;    /* varyings of clip-generated vertices; slot
	                            cacheSize + k is genVaryings[k] */
	uint32_t varyingsSize;
	uint32_t describedSize;
	uint32_t nVaryings;
	SrpVaryingDesc var[SRP_MAX_VARYINGS];
	uint32_t wordPath;       /* 1: every element 4-byte and aligned */
	uint32_t nWords;
	uint8_t wordKind[SRP_MAX_WORDS];

	/* tiles */
	uint32_t tileShift, tilesX, tilesY;
	uint32_t maskBox;        /* triangles whose box is at most this many pixels
	                            on each side get exact 8x8 masks at binning */

	unsigned long long* stats;   /* fragment counters, SRP_STAT_SLOTS slots */
} SrpDrawParams;

/* One draw of a multi-draw launch: its parameters and its range of the
 * launch's work items (cache slots, or tiles in the global tile numbering)
 * and of its blocks. */
typedef struct SrpSeg
{
	const SrpDrawParams* P;
	uint32_t first;        /* first work item of the draw */
	uint32_t count;        /* work items of the draw */
	uint32_t firstBlock;   /* first block of the draw (vertex stage) */
	uint32_t pad;
} SrpSeg;

/* Launch arguments of the shader-module kernels. */
typedef struct SrpVsLaunch
{
	const SrpDrawParams* params;
	uint32_t count;
	/* multi-draw: nSeg > 0 draws, blocks split by SrpSeg.firstBlock */
	const SrpSeg* seg;
	uint32_t nSeg;
	uint32_t nBlocks;
	/* when set, block 0 first clears zero[0..zeroCount) (zeroCount <= 256):
	 * the draw's counters, when its parameters were not uploaded again */
	uint32_t* zero;
	uint32_t zeroCount;
	/* single draws: the framebuffer's pending fill rides along in fillBlocks
	 * blocks after the vertex blocks, setting fillCount pixels of each
	 * non-null array (depth as its bits, stencil bytes) to the value given */
	uint32_t* fillColor;
	uint32_t* fillDepth;
	uint8_t* fillStencil;
	uint32_t fillColorValue, fillDepthValue, fillStencilValue;
	uint32_t fillBlocks;
	uint64_t fillCount;
} SrpVsLaunch;

typedef struct SrpRasterLaunch
{
	const SrpDrawParams* params;
	const SrpTri* tris;
	const uint32_t* tileList;   /* pair indices, grouped by tile, in primitive order */
	const uint32_t* tileStart;  /* first list entry of each tile */
	const uint32_t* tileEnd;    /* one past its last entry (start == end: empty) */
	const uint32_t* pairRec;    /* triangle record of a pair | 0x80000000 if pairMask is exact */
	const unsigned long long* pairMask;  /* exact coverage of the tile's 8x8 pixels */
	uint32_t nTiles;
	uint32_t lateZ;
	uint32_t words;             /* fast interpolation path usable */
	uint32_t tile;              /* tile side: 8 or 16 */
	uint32_t nBlocks;           /* grid size (for the XCD-aware tile order) */
	uint32_t tileOrder;         /* single draws: 0 = one contiguous range of tiles per
	                               XCD; n >= 1 = runs of n tiles dealt to the XCDs in
	                               turn (srpTileOfBlock) */
	uint32_t batched;           /* batched three-phase kernel (tile 8, one wave per tile) */
	uint32_t bigPixel;          /* batched: triangles without exact masks are shaded
	                               pixel-parallel in runs instead of via fragment lists */
	uint32_t depthCull;         /* batched runs: skip a triangle for a whole tile when
	                               every fragment it could have there fails the depth
	                               test (no stencil, no late depth); the fragments so
	                               skipped are not counted as covered */
	/* multi-draw persistent mode (nSeg > 0): nTiles counts the tiles of all
	 * draws, seg[d] covers draw d's tiles [first, first + count), and the
	 * nBlocks waves take tiles from `queues` (8 counters, one per range of
	 * tiles when xcdQueues, else only queues[0]) until all are done. */
	const SrpSeg* seg;
	uint32_t nSeg;
	uint32_t xcdQueues;
	uint32_t* queues;
	/* Deferred shading (single draws of a shader with a wave entry and no
	 * depth output): the batched rasteriser does every fragment operation
	 * but shades nothing.  Tile t lists the pixels whose colour the draw
	 * writes, as (pixel index, triangle record of the last fragment that
	 * wrote it), in its 64 two-word slots of deferList from 64 t on, and
	 * their number in deferCounts[t] (written for every tile).  The library
	 * scans deferCounts in place (deferCounts[nTiles] = the total); the
	 * shader's resolve kernels then rebuild those fragments, compacted, as
	 * word arrays of deferStride entries in deferFrags (fragCoord x, y, z,
	 * w; front facing; primitive ID; pixel index; then the interpolated
	 * varyings, one array per 4-byte word) and shade them, 64 per wave. */
	uint32_t* deferList;
	uint32_t* deferCounts;
	uint32_t* deferFrags;
	uint32_t deferStride;
} SrpRasterLaunch;

typedef struct SrpFragLaunch
{
	const SrpDrawParams* params;
	const SrpFrag* frags;        /* in primitive order */
	const uint32_t* keys;        /* pixel index per fragment, sorted (stable) */
	const uint32_t* order;       /* fragment index for each sorted key */
	const SrpLine* lines;        /* kind 0 */
	const SrpPoint* points;      /* kind 1 */
	uint32_t count;              /* fragments, or a capacity when countDev is set */
	uint32_t kind;
	uint32_t lateZ;
	uint32_t words;
	const uint32_t* countDev;    /* if set: the fragment count is min(count, *countDev) */
} SrpFragLaunch;

typedef void (*SrpVsLaunchFn)(const SrpVsLaunch*, void* stream);
typedef void (*SrpRasterLaunchFn)(const SrpRasterLaunch*, void* stream);
typedef void (*SrpFragLaunchFn)(const SrpFragLaunch*, void* stream);

typedef struct SrpVsEntry
{
	uint32_t abiVersion;
	size_t uniformSize;
	SrpVsLaunchFn launch;
} SrpVsEntry;

typedef struct SrpFsEntry
{
	uint32_t abiVersion;
	size_t uniformSize;
	SrpRasterLaunchFn raster;
	SrpFragLaunchFn frags;
	SrpRasterLaunchFn resolve;   /* deferred shading of a raster launch's list (wave shaders), else null */
} SrpFsEntry;
