/* MI355X extensions to the srp C API.
 *
 * Nothing here is needed to run a program written for the reference
 * renderer: with the defaults every draw is synchronous and the host arrays
 * of SRPFramebuffer hold the result when the call returns.  These calls add
 * what a GPU renderer needs on top:
 *  - a deferred mode in which the device copy of a framebuffer is
 *    authoritative and draws only enqueue work (no per-draw PCIe copies);
 *  - device-side fills of depth and stencil (for deferred mode);
 *  - textures from memory;
 *  - fragment/primitive counters and a stream handle for timing;
 *  - registration of HIP device shaders (used by the SRP_HIP_*_SHADER
 *    macros of srp/srp_device.h; not called directly by users);
 *  - batches: many independent triangle draws (each with its own
 *    framebuffer, buffers, uniform and render state) rendered together,
 *    with one persistent rasteriser launch for all their tiles, optionally
 *    captured once into a HIP graph and replayed, and tuned by measurement. */
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include "srp/buffer.h"
#include "srp/context.h"
#include "srp/framebuffer.h"
#include "srp/shaders.h"
#include "srp/texture.h"
#include "srp/detail/qualifiers.h"

SRP_EXTERN_C_BEGIN

/** Deferred: draws and clears act on the device copy only; the host arrays
 *  are refreshed by srpFramebufferDownload and pushed by srpFramebufferUpload.
 *  Switching deferred mode on uploads the host arrays once. */
void srpFramebufferSetDeferred(const SRPFramebuffer* fb, bool deferred);
void srpFramebufferDownload(const SRPFramebuffer* fb);
void srpFramebufferUpload(const SRPFramebuffer* fb);
/** Write the colour buffer as an RGBA PNG, alpha 255 (what the reference's
 *  test harness does with its frames: tests/utils/save.c:4-33).  Deferred
 *  framebuffers are downloaded first.  0 on success, else -1 and a
 *  message. */
int srpFramebufferSaveImage(const SRPFramebuffer* fb, const char* path);
/** Fill the whole depth / stencil buffer (device side in deferred mode). */
void srpFramebufferFillDepth(const SRPFramebuffer* fb, float value);
void srpFramebufferFillStencil(const SRPFramebuffer* fb, uint8_t value);

/** Texture from RGB8 bytes (row-major, top row first), as if loaded from a file. */
SRPTexture* srpNewTextureFromMemory(const uint8_t* rgb, int width, int height,
                                    SRPTextureWrappingMode wrappingModeX,
                                    SRPTextureWrappingMode wrappingModeY);
/** Host-side sampling (what srpTextureGetFilteredColor does on the host). */
void srpTextureGetFilteredColorHost(const SRPTexture* texture, float u, float v, float out[4]);

/** Counters since the last reset: covered fragments (fragments that reach
 *  the per-fragment tests), shaded fragments (fragment shader invocations),
 *  assembled primitives, draws. */
void srpHipGetStats(uint64_t out[4]);
void srpHipResetStats(void);

/** The HIP stream (hipStream_t) all library work is queued on. */
void* srpHipGetStream(void);
/** Wait for all queued work. */
void srpHipSynchronize(void);

/** Optional timing with HIP events on the library stream (no host sync is
 *  added to draws): enable, then read accumulated milliseconds and counts
 *  per slot: 0 triangle rasteriser kernel (with the deferred shading that
 *  follows it), 1 whole draw (first to last GPU operation), 2 line/point
 *  fragment kernel, 3 vertex stage + assembly count pass. */
void srpHipProfileEnable(bool enable);
void srpHipProfileRead(double ms[4], uint64_t launches[4]);
/** Which slots record while profiling is on (bit s: slot s; all by
 *  default).  Each recorded interval costs the stream two event markers,
 *  and the device waits at each for the work before it, so timing only
 *  the kernel of interest disturbs the rest of the frame least. */
void srpHipProfileSlots(unsigned mask);

/** Raster tuning: tile side in pixels (8 or 16). */
void srpHipSetTileSize(int tile);
int srpHipGetTileSize(void);
/** Raster variant of single draws (none changes the result): batched = 1:
 *  the three-phase wave-per-tile kernel (tile 8), 0: one lane per pixel;
 *  pixelRuns = 1: in the batched kernel, triangles without exact binning
 *  masks are shaded pixel-parallel in runs, 0: through fragment lists;
 *  maskBox: exact 8x8 binning masks for triangle boxes up to this side
 *  (0..24).  A negative argument leaves that setting unchanged. */
void srpHipSetRasterOptions(int batched, int pixelRuns, int maskBox);
/** Tile-level depth culling (on by default): in the batched kernel, a
 *  triangle none of whose fragments in an 8x8 tile can pass the depth test
 *  (a conservative bound on its interpolated depth against the tile's
 *  stored depth range; LESS/LEQUAL/GREATER/GEQUAL/NEVER, stencil off, the
 *  shader not writing depth) is skipped for that tile.  The buffers and the
 *  shaded count are unchanged; the covered count then omits the skipped
 *  fragments (exact again with culling off). */
void srpHipSetTileDepthCull(int enable);
int srpHipGetTileDepthCull(void);
/** Order in which a single draw's tiles are spread over the GPU's 8 XCDs
 *  (each has its own L2).  `run` >= 1 deals runs of `run` consecutive
 *  tiles (row-major) to the XCDs in turn, so a frame whose busy tiles
 *  cluster in one region still loads every XCD (default 8); 0 gives each
 *  XCD one contiguous range of tiles.  Values above 64 are reported through the
 *  message callback and leave the setting unchanged.  The frame is the
 *  same for every setting. */
void srpHipSetTileOrder(int run);
int srpHipGetTileOrder(void);
/** Deferred shading (off by default) of fragment shaders registered with a
 *  wave entry (SRP_HIP_FRAGMENT_SHADER_WAVE) that do not write depth, in
 *  single triangle draws with the batched rasteriser: the rasteriser does
 *  every fragment operation but shades nothing, remembering for each pixel
 *  the fragment whose colour would be the last written; a second kernel
 *  then evaluates the shader once per such pixel, 64 pixels per wave, on
 *  the inputs it would have had.  The frame and the fragment counts are
 *  the same either way; off, the shader runs inside the rasteriser for
 *  every fragment that passes the depth test (on the 1080p terrain with
 *  the learned shader the faster of the two: 0.433 against 0.439 ms per
 *  frame). */
void srpHipSetDeferredShading(int enable);
int srpHipGetDeferredShading(void);

/** Shader registration (see srp/srp_device.h).  `entry` points to a
 *  struct SrpVsEntry / SrpFsEntry of srp/device/abi.h. */
void srpHipRegisterVertexShader(const void* hostFunction, const void* entry);
void srpHipRegisterFragmentShader(const void* hostFunction, const void* entry);

/* ------------------------------------------------------------------ batches */

/** One draw of a batch: the arguments of srpDrawIndexBuffer (ib NULL:
 *  srpDrawVertexBuffer) plus its render state and an optional clear of its
 *  framebuffer at the start of every frame (colour 0, depth clearDepth,
 *  stencil clearStencil).  Framebuffers must be in deferred mode; draws of
 *  a batch share one vertex and one fragment shader; primitives are
 *  SRP_PRIM_TRIANGLES / _STRIP / _FAN in SRP_POLYGON_MODE_FILL. */
typedef struct SRPDrawCommand
{
	const SRPIndexBuffer* ib;
	const SRPVertexBuffer* vb;
	const SRPFramebuffer* fb;
	const SRPShaderProgram* sp;
	SRPPrimitive primitive;
	size_t startIndex;
	size_t count;
	const SRPContext* context;   /**< NULL: the current context when the batch is made */
	int clear;
	float clearDepth;
	uint8_t clearStencil;
} SRPDrawCommand;

/** Implementation choices of a batch; none changes the rendered result. */
typedef struct SRPBatchTuning
{
	int wavesPerCU;   /**< persistent rasteriser waves per compute unit (1..32) */
	int maskBox;      /**< exact 8x8 binning masks for triangle boxes up to this side (0..24 px) */
	int xcdQueues;    /**< 1: one tile queue per XCD range (work stealing between them); 0: one queue */
	int pixelRuns;    /**< 1: triangles without exact masks are shaded pixel-parallel in runs; 0: via fragment lists */
	int depthCull;    /**< 1: tile-level depth culling (see srpHipSetTileDepthCull; covered counts become a lower bound) */
	int tileSize;     /**< raster tile side: 8 (one wave a tile, the batched kernel) or 16 (four waves a tile, a lane
	                       a pixel; no binning masks or pixel runs) */
} SRPBatchTuning;

typedef struct SRPBatch SRPBatch;

/** Validate and prepare the draws (device memory is sized by one counting
 *  pass over them; nothing is drawn).  NULL, with a message, on invalid
 *  input.  The commands and each draw's render state (its context, or the
 *  current one) are copied, and so are the uniforms' contents; vertex and
 *  index data are read by every frame.  Giving a vertex or index buffer
 *  new data (srp*BufferCopyData) makes the next frame prepare the batch
 *  again from the copied commands and state (re-reading the uniforms,
 *  keeping the tuning, capturing the graph again if there was one);
 *  freeing a buffer or framebuffer the batch uses makes its frames draw
 *  nothing, with a message.  Errors a frame meets on the device (a clipped
 *  polygon too large, a capacity exceeded) are reported through the message
 *  callback by the next batch call. */
SRPBatch* srpNewBatch(const SRPDrawCommand* cmds, size_t n);
void srpFreeBatch(SRPBatch* batch);
/** Render one frame of every draw in command order (a replay of the HIP
 *  graph once captured).  Queued on the library stream. */
void srpBatchRender(SRPBatch* batch);
/** Capture the frame into a HIP graph (nothing is drawn now); later frames
 *  replay it.  0 on success. */
int srpBatchCapture(SRPBatch* batch);
/** Change the tuning (drops a captured graph) / read it. */
void srpBatchSetTuning(SRPBatch* batch, const SRPBatchTuning* tuning);
void srpBatchGetTuning(const SRPBatch* batch, SRPBatchTuning* tuning);
/** Time `trials` frames of every candidate tuning and keep the fastest: 48
 *  with 8x8 tiles (a wave a tile: waves per CU 4, 8, 16, 28 x mask box 0,
 *  16, 24 x both queue modes x both run modes) and 8 with 16x16 tiles (four
 *  waves a tile: waves per CU x queue modes); returns its mean frame time in
 *  ms and writes it to *best (if not NULL). */
double srpBatchAutotune(SRPBatch* batch, int trials, SRPBatchTuning* best);
/** draws, assembled primitives, binned (tile, triangle) pairs, tiles of the last frame sizing */
void srpBatchInfo(const SRPBatch* batch, uint64_t out[4]);

SRP_EXTERN_C_END
