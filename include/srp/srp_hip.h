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
 *    macros of srp/srp_device.h; not called directly by users). */
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include "srp/framebuffer.h"
#include "srp/texture.h"
#include "srp/detail/qualifiers.h"

SRP_EXTERN_C_BEGIN

/** Deferred: draws and clears act on the device copy only; the host arrays
 *  are refreshed by srpFramebufferDownload and pushed by srpFramebufferUpload.
 *  Switching deferred mode on uploads the host arrays once. */
void srpFramebufferSetDeferred(const SRPFramebuffer* fb, bool deferred);
void srpFramebufferDownload(const SRPFramebuffer* fb);
void srpFramebufferUpload(const SRPFramebuffer* fb);
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
 *  per slot: 0 triangle rasteriser kernel, 1 whole draw (first to last GPU
 *  operation, including the wait for the host between the two passes),
 *  2 line/point fragment kernel, 3 vertex stage + assembly count pass. */
void srpHipProfileEnable(bool enable);
void srpHipProfileRead(double ms[4], uint64_t launches[4]);

/** Raster tuning: tile side in pixels (8 or 16). */
void srpHipSetTileSize(int tile);
int srpHipGetTileSize(void);

/** Shader registration (see srp/srp_device.h).  `entry` points to a
 *  struct SrpVsEntry / SrpFsEntry of srp/device/abi.h. */
void srpHipRegisterVertexShader(const void* hostFunction, const void* entry);
void srpHipRegisterFragmentShader(const void* hostFunction, const void* entry);

SRP_EXTERN_C_END
