/* Image files for srpNewTexture.
 *
 * The reference decodes textures with the vendored stb_image (lib/stb_image.h
 * v2.30, called from core/texture.c:34 with 3 channels requested).  Of its
 * formats, the lossless ones are read here with the same results: PNG
 * (png_decode.c), BMP (uncompressed 1/4/8-bit palette, 16/32-bit bit
 * fields, 24/32-bit; OS/2 and Windows V3-V5 headers), TGA (colour-mapped,
 * true-colour and grey, raw or run-length encoded, 8/15/16/24/32 bits) and
 * binary PNM (P5 grey, P6 RGB, 8-bit samples).  Every pixel comes out as
 * RGB8: grey is replicated, alpha is dropped, narrow channels are widened
 * by repeating their bits, as stb does.  JPEG, GIF, PSD, HDR, PIC and 16-bit
 * PNM are refused with a message (lossy or rarely used for textures; a JPEG
 * decoder would have to repeat stb's fixed-point IDCT, upsampling and
 * colour conversion bit for bit to give the reference's texels). */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/** Decodes `path` (any format above, recognised by its content) into a
 *  malloc'ed RGB8 buffer (row-major, top row first, 3 bytes per texel).
 *  Returns NULL on failure and points *err at a static message (stb's
 *  short reason where stb fails the same way: "can't fopen", "unknown image
 *  type", "bad BMP", ...). */
uint8_t* srpImageLoadRGB(const char* path, int* width, int* height, const char** err);

/** The same from bytes in memory (PNG excepted: use a file). */
uint8_t* srpImageDecodeRGB(const uint8_t* data, size_t size, int* width, int* height, const char** err);

#ifdef __cplusplus
}
#endif
