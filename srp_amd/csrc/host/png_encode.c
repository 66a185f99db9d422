/* PNG writer: signature, IHDR (8-bit RGBA), one IDAT holding the zlib
 * stream of the scanlines (filter type 0 on every row), IEND; each chunk
 * with its CRC-32 (zlib's). */
#include "png_encode.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static void put32(uint8_t* p, uint32_t v)
{
	p[0] = (uint8_t) (v >> 24);
	p[1] = (uint8_t) (v >> 16);
	p[2] = (uint8_t) (v >> 8);
	p[3] = (uint8_t) v;
}

static int writeChunk(FILE* f, const char* type, const uint8_t* data, uint32_t len)
{
	uint8_t head[8];
	put32(head, len);
	memcpy(head + 4, type, 4);
	uLong crc = crc32(0L, (const Bytef*) type, 4);
	if (len)
		crc = crc32(crc, (const Bytef*) data, len);
	uint8_t tail[4];
	put32(tail, (uint32_t) crc);
	return fwrite(head, 1, 8, f) == 8 && (len == 0 || fwrite(data, 1, len, f) == len) && fwrite(tail, 1, 4, f) == 4;
}

int srpPngSaveRGBA(const char* path, int width, int height, const uint8_t* rgba)
{
	if (!path || !rgba || width <= 0 || height <= 0 || width > (1 << 24) || height > (1 << 24))
		return -1;
	const size_t row = (size_t) width * 4 + 1;
	const size_t rawLen = row * (size_t) height;
	uint8_t* raw = (uint8_t*) malloc(rawLen);
	if (!raw)
		return -1;
	for (size_t y = 0; y < (size_t) height; y++)
	{
		raw[y * row] = 0;   /* filter: none */
		memcpy(raw + y * row + 1, rgba + y * (size_t) width * 4, (size_t) width * 4);
	}
	uLongf zLen = compressBound((uLong) rawLen);
	uint8_t* z = (uint8_t*) malloc(zLen);
	if (!z || compress2(z, &zLen, raw, (uLong) rawLen, 6) != Z_OK || zLen > 0x7FFFFFFFu)
	{
		free(raw);
		free(z);
		return -1;
	}
	free(raw);
	FILE* f = fopen(path, "wb");
	if (!f)
	{
		free(z);
		return -1;
	}
	static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
	uint8_t ihdr[13];
	put32(ihdr, (uint32_t) width);
	put32(ihdr + 4, (uint32_t) height);
	ihdr[8] = 8;    /* bit depth */
	ihdr[9] = 6;    /* RGBA */
	ihdr[10] = 0;   /* deflate */
	ihdr[11] = 0;   /* adaptive filtering */
	ihdr[12] = 0;   /* no interlace */
	const int ok = fwrite(sig, 1, 8, f) == 8 && writeChunk(f, "IHDR", ihdr, 13) &&
	               writeChunk(f, "IDAT", z, (uint32_t) zLen) && writeChunk(f, "IEND", NULL, 0);
	free(z);
	return (fclose(f) == 0 && ok) ? 0 : -1;
}

int srpPngSaveColorBuffer(const char* path, int width, int height, const uint32_t* color)
{
	if (!color || width <= 0 || height <= 0 || width > (1 << 24) || height > (1 << 24))
		return -1;
	const size_t n = (size_t) width * (size_t) height;
	uint8_t* px = (uint8_t*) malloc(n * 4);
	if (!px)
		return -1;
	for (size_t i = 0; i < n; i++)
	{
		const uint32_t c = color[i];
		px[4 * i + 0] = (uint8_t) (c >> 24);
		px[4 * i + 1] = (uint8_t) (c >> 16);
		px[4 * i + 2] = (uint8_t) (c >> 8);
		px[4 * i + 3] = 0xFF;
	}
	const int rc = srpPngSaveRGBA(path, width, height, px);
	free(px);
	return rc;
}
