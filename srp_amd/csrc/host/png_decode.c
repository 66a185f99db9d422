/* Minimal PNG reader: chunk walk, zlib inflate, scanline unfiltering and
 * conversion to RGB8.  See png_decode.h for the supported subset. */
#include "png_decode.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static uint32_t be32(const uint8_t* p)
{
	return ((uint32_t) p[0] << 24) | ((uint32_t) p[1] << 16) | ((uint32_t) p[2] << 8) | p[3];
}

static int paeth(int a, int b, int c)
{
	const int p = a + b - c;
	const int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
	if (pa <= pb && pa <= pc) return a;
	if (pb <= pc) return b;
	return c;
}

static uint8_t* fail(const char** err, const char* msg, void* a, void* b)
{
	free(a);
	free(b);
	if (err) *err = msg;
	return NULL;
}

uint8_t* srpPngLoadRGB(const char* path, int* width, int* height, const char** err)
{
	FILE* f = fopen(path, "rb");
	if (!f) return fail(err, "can't fopen", NULL, NULL);
	fseek(f, 0, SEEK_END);
	const long fsize = ftell(f);
	fseek(f, 0, SEEK_SET);
	if (fsize < 8) { fclose(f); return fail(err, "file too small", NULL, NULL); }
	uint8_t* file = (uint8_t*) malloc((size_t) fsize);
	if (!file || fread(file, 1, (size_t) fsize, f) != (size_t) fsize) { fclose(f); return fail(err, "read error", file, NULL); }
	fclose(f);
	uint8_t* out = srpPngDecodeRGB(file, (size_t) fsize, width, height, err);
	free(file);
	return out;
}

uint8_t* srpPngDecodeRGB(const uint8_t* file, size_t fsize, int* width, int* height, const char** err)
{
	static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
	if (fsize < 8 || memcmp(file, sig, 8) != 0) return fail(err, "not a PNG file", NULL, NULL);

	uint32_t w = 0, h = 0;
	int depth = 0, ctype = -1, interlace = 0;
	uint8_t palette[256 * 3];
	memset(palette, 0, sizeof(palette));
	uint8_t* idat = NULL;
	size_t idatLen = 0, idatCap = 0;

	size_t pos = 8;
	while (pos + 12 <= (size_t) fsize)
	{
		const uint32_t len = be32(file + pos);
		const uint8_t* type = file + pos + 4;
		const uint8_t* data = file + pos + 8;
		if (pos + 12 + (size_t) len > (size_t) fsize) return fail(err, "truncated chunk", NULL, idat);
		if (!memcmp(type, "IHDR", 4) && len >= 13)
		{
			w = be32(data); h = be32(data + 4);
			depth = data[8]; ctype = data[9]; interlace = data[12];
		}
		else if (!memcmp(type, "PLTE", 4))
			memcpy(palette, data, len > sizeof(palette) ? sizeof(palette) : len);
		else if (!memcmp(type, "IDAT", 4))
		{
			if (idatLen + len > idatCap)
			{
				idatCap = (idatLen + len) * 2;
				uint8_t* grown = (uint8_t*) realloc(idat, idatCap);
				if (!grown) return fail(err, "out of memory", NULL, idat);
				idat = grown;
			}
			memcpy(idat + idatLen, data, len);
			idatLen += len;
		}
		else if (!memcmp(type, "IEND", 4))
			break;
		pos += 12 + (size_t) len;
	}
	if (w == 0 || h == 0 || ctype < 0) return fail(err, "missing IHDR", NULL, idat);
	/* as stb_image (the reference's loader): sides above 2^24 are refused,
	 * so no size product below can overflow or truncate */
	if (w > (1u << 24) || h > (1u << 24)) return fail(err, "image too large", NULL, idat);
	if (interlace) return fail(err, "interlaced PNG not supported", NULL, idat);

	int channels;
	switch (ctype)
	{
		case 0: channels = 1; break;
		case 2: channels = 3; break;
		case 3: channels = 1; break;
		case 4: channels = 2; break;
		case 6: channels = 4; break;
		default: return fail(err, "bad colour type", NULL, idat);
	}
	if (!(depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4))))
		return fail(err, "unsupported bit depth", NULL, idat);

	const size_t bitsPerPixel = (size_t) channels * (size_t) depth;
	const size_t rowBytes = ((size_t) w * bitsPerPixel + 7) / 8;
	const size_t bpp = (bitsPerPixel + 7) / 8;  /* filter byte distance */
	/* w, h <= 2^24 and at most 64 bits per pixel: (rowBytes + 1) * h and
	 * w * h * 3 stay below 2^55; zlib's length type must hold the former,
	 * and the output's side lengths must fit an int */
	if ((rowBytes + 1) * (size_t) h > (size_t) (uLongf) -1 || (size_t) w * h * 3 > ((size_t) 1 << 48))
		return fail(err, "image too large", NULL, idat);
	uLongf rawLen = (uLongf) ((rowBytes + 1) * h);
	uint8_t* raw = (uint8_t*) malloc(rawLen);
	if (!raw) return fail(err, "out of memory", NULL, idat);
	if (uncompress(raw, &rawLen, idat, (uLong) idatLen) != Z_OK || rawLen != (rowBytes + 1) * h)
	{
		free(raw);
		return fail(err, "corrupt image data", NULL, idat);
	}
	free(idat);

	/* Undo the per-scanline filters in place. */
	for (size_t y = 0; y < h; y++)
	{
		uint8_t* row = raw + y * (rowBytes + 1);
		const int filter = row[0];
		uint8_t* cur = row + 1;
		const uint8_t* prev = (y > 0) ? raw + (y - 1) * (rowBytes + 1) + 1 : NULL;
		for (size_t i = 0; i < rowBytes; i++)
		{
			const int a = (i >= bpp) ? cur[i - bpp] : 0;
			const int b = prev ? prev[i] : 0;
			const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
			switch (filter)
			{
				case 0: break;
				case 1: cur[i] = (uint8_t) (cur[i] + a); break;
				case 2: cur[i] = (uint8_t) (cur[i] + b); break;
				case 3: cur[i] = (uint8_t) (cur[i] + ((a + b) >> 1)); break;
				case 4: cur[i] = (uint8_t) (cur[i] + paeth(a, b, c)); break;
				default: free(raw); if (err) *err = "bad filter type"; return NULL;
			}
		}
	}

	uint8_t* out = (uint8_t*) malloc((size_t) w * h * 3);
	if (!out) { free(raw); if (err) *err = "out of memory"; return NULL; }
	for (size_t y = 0; y < h; y++)
	{
		const uint8_t* row = raw + y * (rowBytes + 1) + 1;
		for (size_t x = 0; x < w; x++)
		{
			uint8_t rgb[3];
			if (depth < 8)
			{
				const size_t bit = x * (size_t) depth;
				const int v = (row[bit >> 3] >> (8 - depth - (int) (bit & 7))) & ((1 << depth) - 1);
				if (ctype == 3) { rgb[0] = palette[3 * v]; rgb[1] = palette[3 * v + 1]; rgb[2] = palette[3 * v + 2]; }
				else { const uint8_t g = (uint8_t) (v * (255 / ((1 << depth) - 1))); rgb[0] = rgb[1] = rgb[2] = g; }
			}
			else
			{
				const size_t step = (size_t) (depth / 8);
				const uint8_t* px = row + x * (size_t) channels * step;
				/* High byte of each sample (16-bit samples are big-endian). */
				#define SAMPLE(k) px[(k) * step]
				switch (ctype)
				{
					case 0: case 4: rgb[0] = rgb[1] = rgb[2] = SAMPLE(0); break;
					case 2: case 6: rgb[0] = SAMPLE(0); rgb[1] = SAMPLE(1); rgb[2] = SAMPLE(2); break;
					default: { const int v = px[0]; rgb[0] = palette[3 * v]; rgb[1] = palette[3 * v + 1]; rgb[2] = palette[3 * v + 2]; }
				}
				#undef SAMPLE
			}
			memcpy(out + (y * w + x) * 3, rgb, 3);
		}
	}
	free(raw);
	*width = (int) w;
	*height = (int) h;
	return out;
}
