/* Minimal PNG writer for framebuffer images (the reference's test harness
 * writes frames with stb_image_write as 8-bit RGBA PNGs, alpha 255:
 * tests/utils/save.c:4-33). */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Writes `width` x `height` RGBA8 pixels (row-major, top row first) to
 *  `path` as a non-interlaced 8-bit RGBA PNG.  0 on success, -1 on a bad
 *  argument or a failed write. */
int srpPngSaveRGBA(const char* path, int width, int height, const uint8_t* rgba);

/** A framebuffer's colour array (one u32 per pixel, R in the top byte as
 *  colorPack stores it) as an RGBA PNG with alpha 255, the channels the
 *  reference's harness writes.  0 on success, -1 on failure. */
int srpPngSaveColorBuffer(const char* path, int width, int height, const uint32_t* color);

#ifdef __cplusplus
}
#endif
