/* Minimal PNG reader used by srpNewTexture.
 *
 * The reference decodes images with the vendored stb_image (lib/stb_image.h
 * v2.30, core/texture.c:34) asking for 3 channels.  PNG decoding is
 * lossless, so any conforming decoder yields the same RGB bytes; this one
 * handles non-interlaced PNGs of every colour type at bit depth 8 (and
 * palette/grey at 1/2/4 bits, 16-bit samples keep their high byte, as stb
 * does when converting to 8 bits). */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Decodes `path` into a malloc'ed RGB8 buffer (row-major, 3 bytes/texel).
 *  Returns NULL on failure and points *err at a static message. */
uint8_t* srpPngLoadRGB(const char* path, int* width, int* height, const char** err);
/** The same from a PNG file's bytes in memory (not taken over). */
uint8_t* srpPngDecodeRGB(const uint8_t* data, size_t size, int* width, int* height, const char** err);

#ifdef __cplusplus
}
#endif
