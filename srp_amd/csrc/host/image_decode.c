/* BMP, TGA and binary PNM readers and the format dispatch of srpNewTexture
 * (see image_decode.h).  What each reader accepts, and the value of every
 * texel, follow the reference's loader (stb_image v2.30): the same header
 * checks, the same bit widening, the same quirks (a TGA colour map's first
 * index is skipped as bytes, reads past the end of the data give 0). */
#include "image_decode.h"
#include "png_decode.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Sides above 2^24 are refused (stb's STBI_MAX_DIMENSIONS). */
#define IMG_MAX_SIDE (1u << 24)

typedef struct
{
	const uint8_t* p;
	size_t size, pos;
} Reader;

static int rd8(Reader* r) { return (r->pos < r->size) ? r->p[r->pos++] : (r->pos++, 0); }
static int rd16le(Reader* r) { const int lo = rd8(r); return lo | (rd8(r) << 8); }
static uint32_t rd32le(Reader* r) { const uint32_t lo = (uint32_t) rd16le(r); return lo | ((uint32_t) rd16le(r) << 16); }
static void skip(Reader* r, long n) { if (n > 0) r->pos += (size_t) n; }

static uint8_t* fail(const char** err, const char* msg, void* a)
{
	free(a);
	if (err) *err = msg;
	return NULL;
}

/* w * h * 3 bytes, or NULL if that does not fit. */
static uint8_t* allocRGB(uint32_t w, uint32_t h)
{
	if (w == 0 || h == 0 || w > IMG_MAX_SIDE || h > IMG_MAX_SIDE)
		return NULL;
	return (uint8_t*) malloc((size_t) w * h * 3);
}

static void flipRows(uint8_t* img, uint32_t w, uint32_t h)
{
	const size_t row = (size_t) w * 3;
	uint8_t* tmp = (uint8_t*) malloc(row);
	if (!tmp) return;
	for (uint32_t y = 0; y < h / 2; y++)
	{
		uint8_t* a = img + (size_t) y * row;
		uint8_t* b = img + (size_t) (h - 1 - y) * row;
		memcpy(tmp, a, row);
		memcpy(a, b, row);
		memcpy(b, tmp, row);
	}
	free(tmp);
}

/* ------------------------------------------------------------------ BMP */

static int highBit(uint32_t v)
{
	int n = -1;
	while (v) { n++; v >>= 1; }
	return n;
}

static int popCount(uint32_t v)
{
	int n = 0;
	while (v) { n += (int) (v & 1u); v >>= 1; }
	return n;
}

/* A channel of `bits` bits (its mask's top bit moved to bit 7 by `shift`)
 * widened to 8 bits by repeating its bits from the top. */
static uint8_t widen(uint32_t v, int shift, int bits)
{
	v = (shift < 0) ? v << -shift : v >> shift;
	v &= 0xFFu;
	if (bits <= 0)
		return 0;
	v >>= 8 - bits;               /* the channel's value, `bits` wide */
	uint32_t out = 0;
	for (int have = 0; have < 8; have += bits)
		out = (out << bits) | v;  /* repeat until at least 8 bits */
	const int total = ((8 + bits - 1) / bits) * bits;
	return (uint8_t) (out >> (total - 8));
}

static uint8_t* decodeBmp(Reader* r, int* width, int* height, const char** err)
{
	if (rd8(r) != 'B' || rd8(r) != 'M') return fail(err, "not BMP", NULL);
	rd32le(r);   /* file size */
	rd16le(r);
	rd16le(r);
	const int32_t offset = (int32_t) rd32le(r);
	const uint32_t hsz = rd32le(r);
	int extra = 14;
	uint32_t mr = 0, mg = 0, mb = 0, ma = 0;
	if (offset < 0) return fail(err, "bad BMP", NULL);
	if (hsz != 12 && hsz != 40 && hsz != 56 && hsz != 108 && hsz != 124) return fail(err, "unknown BMP", NULL);
	int32_t w, h;
	if (hsz == 12) { w = rd16le(r); h = rd16le(r); }
	else { w = (int32_t) rd32le(r); h = (int32_t) rd32le(r); }
	if (rd16le(r) != 1) return fail(err, "bad BMP", NULL);
	const int bpp = rd16le(r);
	if (hsz != 12)
	{
		const uint32_t compress = rd32le(r);
		if (compress == 1 || compress == 2) return fail(err, "BMP RLE", NULL);
		if (compress >= 4) return fail(err, "BMP JPEG/PNG", NULL);
		if (compress == 3 && bpp != 16 && bpp != 32) return fail(err, "bad BMP", NULL);
		for (int k = 0; k < 5; k++) rd32le(r);   /* image size, resolutions, colours */
		const int defaults32 = (bpp == 32);
		if (hsz == 40 || hsz == 56)
		{
			if (hsz == 56) for (int k = 0; k < 4; k++) rd32le(r);
			if (bpp == 16 || bpp == 32)
			{
				if (compress == 0)
				{
					if (defaults32) { mr = 0xFFu << 16; mg = 0xFFu << 8; mb = 0xFFu; ma = 0xFFu << 24; }
					else { mr = 31u << 10; mg = 31u << 5; mb = 31u; ma = 0; }
				}
				else if (compress == 3)
				{
					mr = rd32le(r); mg = rd32le(r); mb = rd32le(r);
					extra += 12;
					if (mr == mg && mg == mb) return fail(err, "bad BMP", NULL);
				}
				else
					return fail(err, "bad BMP", NULL);
			}
		}
		else
		{
			mr = rd32le(r); mg = rd32le(r); mb = rd32le(r); ma = rd32le(r);
			if (compress != 3)
			{
				if (bpp == 32) { mr = 0xFFu << 16; mg = 0xFFu << 8; mb = 0xFFu; ma = 0xFFu << 24; }
				else if (bpp == 16) { mr = 31u << 10; mg = 31u << 5; mb = 31u; ma = 0; }
				else { mr = mg = mb = ma = 0; }
			}
			for (int k = 0; k < 13; k++) rd32le(r);   /* colour space and its parameters */
			if (hsz == 124) for (int k = 0; k < 4; k++) rd32le(r);
		}
	}
	const int flip = h > 0;   /* bottom-up rows */
	const uint32_t W = (uint32_t) w, H = (uint32_t) (h < 0 ? -(int64_t) h : h);
	if (W > IMG_MAX_SIDE || H > IMG_MAX_SIDE) return fail(err, "too large", NULL);
	int psize = 0;
	if (hsz == 12) { if (bpp < 24) psize = (offset - extra - 24) / 3; }
	else if (bpp < 16) psize = (offset - extra - (int) hsz) >> 2;
	if (psize == 0)
	{
		/* the pixels may start a little after the header (at most a
		 * palette's worth of padding), never inside it */
		const long read = (long) r->pos;
		if (offset < read || offset - read > 1024) return fail(err, "bad offset", NULL);
		skip(r, offset - read);
	}
	uint8_t* out = allocRGB(W, H);
	if (!out) return fail(err, W && H ? "too large" : "bad BMP", NULL);
	size_t z = 0;
	if (bpp < 16)
	{
		uint8_t pal[256][3];
		if (psize <= 0 || psize > 256) return fail(err, "invalid", out);
		for (int i = 0; i < psize; i++)
		{
			pal[i][2] = (uint8_t) rd8(r);
			pal[i][1] = (uint8_t) rd8(r);
			pal[i][0] = (uint8_t) rd8(r);
			if (hsz != 12) rd8(r);
		}
		skip(r, offset - extra - (long) hsz - (long) psize * (hsz == 12 ? 3 : 4));
		uint32_t rowBytes;
		if (bpp == 1) rowBytes = (W + 7) >> 3;
		else if (bpp == 4) rowBytes = (W + 1) >> 1;
		else if (bpp == 8) rowBytes = W;
		else return fail(err, "bad bpp", out);
		const int pad = (int) ((4u - (rowBytes & 3u)) & 3u);
		for (uint32_t y = 0; y < H; y++)
		{
			if (bpp == 1)
			{
				int bit = 7, v = rd8(r);
				for (uint32_t x = 0; x < W; x++)
				{
					const int c = (v >> bit) & 1;
					memcpy(out + z, pal[c], 3);
					z += 3;
					if (x + 1 == W) break;
					if (--bit < 0) { bit = 7; v = rd8(r); }
				}
			}
			else
			{
				for (uint32_t x = 0; x < W; x += 2)
				{
					int v = rd8(r), v2 = 0;
					if (bpp == 4) { v2 = v & 15; v >>= 4; }
					memcpy(out + z, pal[v < psize ? v : 0], 3);   /* an index past the palette reads entry 0 */
					z += 3;
					if (x + 1 == W) break;
					v = (bpp == 8) ? rd8(r) : v2;
					memcpy(out + z, pal[v < psize ? v : 0], 3);
					z += 3;
				}
			}
			skip(r, pad);
		}
	}
	else
	{
		skip(r, offset - extra - (long) hsz);
		const uint32_t rowBytes = (bpp == 24) ? 3 * W : (bpp == 16 ? 2 * W : 0);
		const int pad = (int) ((4u - (rowBytes & 3u)) & 3u);
		const int easy = (bpp == 24) ? 1 : ((bpp == 32 && mb == 0xFFu && mg == 0xFF00u && mr == 0xFF0000u &&
		                                     ma == 0xFF000000u) ? 2 : 0);
		int rs = 0, gs = 0, bs = 0, rc = 0, gc = 0, bc = 0;
		if (!easy)
		{
			if (!mr || !mg || !mb) return fail(err, "bad masks", out);
			rs = highBit(mr) - 7; rc = popCount(mr);
			gs = highBit(mg) - 7; gc = popCount(mg);
			bs = highBit(mb) - 7; bc = popCount(mb);
			if (rc > 8 || gc > 8 || bc > 8 || popCount(ma) > 8) return fail(err, "bad masks", out);
		}
		for (uint32_t y = 0; y < H; y++)
		{
			for (uint32_t x = 0; x < W; x++)
			{
				if (easy)
				{
					out[z + 2] = (uint8_t) rd8(r);
					out[z + 1] = (uint8_t) rd8(r);
					out[z + 0] = (uint8_t) rd8(r);
					if (easy == 2) rd8(r);   /* alpha */
				}
				else
				{
					const uint32_t v = (bpp == 16) ? (uint32_t) rd16le(r) : rd32le(r);
					out[z + 0] = widen(v & mr, rs, rc);
					out[z + 1] = widen(v & mg, gs, gc);
					out[z + 2] = widen(v & mb, bs, bc);
				}
				z += 3;
			}
			skip(r, pad);
		}
	}
	if (flip)
		flipRows(out, W, H);
	*width = (int) W;
	*height = (int) H;
	return out;
}

/* ------------------------------------------------------------------ TGA */

/* The header tests of a TGA file, which has no signature. */
static int isTga(const uint8_t* p, size_t n)
{
	if (n < 18) return 0;
	const int cmapType = p[1], type = p[2];
	if (cmapType > 1) return 0;
	if (cmapType == 1)
	{
		if (type != 1 && type != 9) return 0;
		const int e = p[7];
		if (e != 8 && e != 15 && e != 16 && e != 24 && e != 32) return 0;
	}
	else if (type != 2 && type != 3 && type != 10 && type != 11)
		return 0;
	const int w = p[12] | (p[13] << 8), h = p[14] | (p[15] << 8), bpp = p[16];
	if (w < 1 || h < 1) return 0;
	if (cmapType == 1 && bpp != 8 && bpp != 16) return 0;
	return bpp == 8 || bpp == 15 || bpp == 16 || bpp == 24 || bpp == 32;
}

/* Channels of a TGA pixel (1 grey, 2 grey + alpha, 3 RGB, 4 RGBA); *rgb16
 * for 15/16-bit colour. */
static int tgaChannels(int bits, int grey, int* rgb16)
{
	*rgb16 = 0;
	switch (bits)
	{
		case 8: return 1;
		case 15: case 16:
			if (bits == 16 && grey) return 2;
			*rgb16 = 1;
			return 3;
		case 24: case 32: return bits / 8;
		default: return 0;
	}
}

/* 15/16-bit colour: 5 bits a channel scaled to 0..255 (truncated). */
static void tgaRgb16(Reader* r, uint8_t* o)
{
	const int px = rd16le(r);
	o[0] = (uint8_t) ((((px >> 10) & 31) * 255) / 31);
	o[1] = (uint8_t) ((((px >> 5) & 31) * 255) / 31);
	o[2] = (uint8_t) (((px & 31) * 255) / 31);
}

static uint8_t* decodeTga(Reader* r, int* width, int* height, const char** err)
{
	const int idLen = rd8(r), indexed = rd8(r);
	int type = rd8(r);
	const int palStart = rd16le(r), palLen = rd16le(r), palBits = rd8(r);
	rd16le(r);
	rd16le(r);   /* origin */
	const uint32_t W = (uint32_t) rd16le(r), H = (uint32_t) rd16le(r);
	const int bpp = rd8(r), desc = rd8(r);
	const int rle = type >= 8;
	if (rle) type -= 8;
	const int bottomUp = !((desc >> 5) & 1);
	int rgb16;
	const int comp = indexed ? tgaChannels(palBits, 0, &rgb16) : tgaChannels(bpp, type == 3, &rgb16);
	if (!comp) return fail(err, "bad format", NULL);
	uint8_t* px = (uint8_t*) calloc((size_t) W * H, (size_t) comp);
	if (!px) return fail(err, "outofmem", NULL);
	skip(r, idLen);
	uint8_t* pal = NULL;
	if (!indexed && !rle && !rgb16)
	{
		for (uint32_t i = 0; i < H; i++)
		{
			const uint32_t row = bottomUp ? H - 1 - i : i;
			uint8_t* d = px + (size_t) row * W * comp;
			for (size_t k = 0; k < (size_t) W * comp; k++)
				d[k] = (uint8_t) rd8(r);
		}
	}
	else
	{
		if (indexed)
		{
			if (palLen == 0) return fail(err, "bad palette", px);
			skip(r, palStart);   /* the reference's loader skips the first index as bytes */
			pal = (uint8_t*) malloc((size_t) palLen * comp);
			if (!pal) return fail(err, "outofmem", px);
			for (int i = 0; i < palLen; i++)
			{
				if (rgb16) tgaRgb16(r, pal + (size_t) i * comp);
				else for (int j = 0; j < comp; j++) pal[(size_t) i * comp + j] = (uint8_t) rd8(r);
			}
		}
		uint8_t raw[4] = {0, 0, 0, 0};
		int count = 0, repeat = 0, fresh = 1;
		for (size_t i = 0; i < (size_t) W * H; i++)
		{
			if (rle)
			{
				if (count == 0)
				{
					const int cmd = rd8(r);
					count = 1 + (cmd & 127);
					repeat = cmd >> 7;
					fresh = 1;
				}
				else if (!repeat)
					fresh = 1;
			}
			else
				fresh = 1;
			if (fresh)
			{
				if (indexed)
				{
					int idx = (bpp == 8) ? rd8(r) : rd16le(r);
					if (idx >= palLen) idx = 0;
					memcpy(raw, pal + (size_t) idx * comp, (size_t) comp);
				}
				else if (rgb16)
					tgaRgb16(r, raw);
				else
					for (int j = 0; j < comp; j++) raw[j] = (uint8_t) rd8(r);
				fresh = 0;
			}
			memcpy(px + i * comp, raw, (size_t) comp);
			count--;
		}
		if (bottomUp)
		{
			const size_t row = (size_t) W * comp;
			for (uint32_t j = 0; 2 * j < H; j++)
			{
				uint8_t* a = px + (size_t) j * row;
				uint8_t* b = px + (size_t) (H - 1 - j) * row;
				for (size_t k = 0; k < row; k++) { const uint8_t t = a[k]; a[k] = b[k]; b[k] = t; }
			}
		}
	}
	free(pal);
	uint8_t* out = allocRGB(W, H);
	if (!out) return fail(err, "too large", px);
	for (size_t i = 0; i < (size_t) W * H; i++)
	{
		const uint8_t* s = px + i * comp;
		uint8_t* o = out + i * 3;
		if (comp <= 2) o[0] = o[1] = o[2] = s[0];                       /* grey (alpha dropped) */
		else if (rgb16) { o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; }      /* already RGB */
		else { o[0] = s[2]; o[1] = s[1]; o[2] = s[0]; }                 /* BGR(A) */
	}
	free(px);
	*width = (int) W;
	*height = (int) H;
	return out;
}

/* ------------------------------------------------------------------ PNM */

static int pnmSpace(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

/* Skips blanks and '#' comments, then reads a decimal number; *c holds the
 * character after it. */
static long pnmNumber(Reader* r, int* c)
{
	for (;;)
	{
		while (r->pos <= r->size && pnmSpace(*c)) *c = rd8(r);
		if (*c != '#' || r->pos > r->size) break;
		while (r->pos <= r->size && *c != '\n' && *c != '\r') *c = rd8(r);
	}
	long v = 0;
	while (*c >= '0' && *c <= '9')
	{
		v = v * 10 + (*c - '0');
		if (v > (1L << 30)) return -1;
		*c = rd8(r);
	}
	return v;
}

static uint8_t* decodePnm(Reader* r, int* width, int* height, const char** err)
{
	if (rd8(r) != 'P') return fail(err, "unknown image type", NULL);
	const int kind = rd8(r);
	if (kind != '5' && kind != '6') return fail(err, "unknown image type", NULL);
	const int comp = (kind == '6') ? 3 : 1;
	int c = rd8(r);
	const long w = pnmNumber(r, &c), h = pnmNumber(r, &c);
	if (w <= 0 || h <= 0) return fail(err, "invalid width", NULL);
	const long maxv = pnmNumber(r, &c);
	if (maxv > 65535) return fail(err, "max value > 65535", NULL);
	if (maxv > 255) return fail(err, "16-bit PNM not supported", NULL);
	if ((uint64_t) w > IMG_MAX_SIDE || (uint64_t) h > IMG_MAX_SIDE) return fail(err, "too large", NULL);
	/* the sample data follow the single whitespace after the header */
	const size_t n = (size_t) w * (size_t) h * (size_t) comp;
	if (r->pos + n > r->size) return fail(err, "bad PNM", NULL);
	uint8_t* out = allocRGB((uint32_t) w, (uint32_t) h);
	if (!out) return fail(err, "outofmem", NULL);
	const uint8_t* s = r->p + r->pos;
	for (size_t i = 0; i < (size_t) w * (size_t) h; i++)
	{
		if (comp == 3) memcpy(out + 3 * i, s + 3 * i, 3);
		else out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = s[i];
	}
	*width = (int) w;
	*height = (int) h;
	return out;
}

/* ------------------------------------------------------------- dispatch */

uint8_t* srpImageDecodeRGB(const uint8_t* data, size_t size, int* width, int* height, const char** err)
{
	static const uint8_t png[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
	Reader r = {data, size, 0};
	if (size >= 8 && memcmp(data, png, 8) == 0)
		return srpPngDecodeRGB(data, size, width, height, err);
	if (size >= 2 && data[0] == 'B' && data[1] == 'M')
		return decodeBmp(&r, width, height, err);
	if (size >= 2 && data[0] == 'P' && (data[1] == '5' || data[1] == '6'))
		return decodePnm(&r, width, height, err);
	if (size >= 3 && data[0] == 0xFF && data[1] == 0xD8 && data[2] == 0xFF)
		return fail(err, "JPEG not supported", NULL);
	if (size >= 6 && (memcmp(data, "GIF87a", 6) == 0 || memcmp(data, "GIF89a", 6) == 0))
		return fail(err, "GIF not supported", NULL);
	if (isTga(data, size))
		return decodeTga(&r, width, height, err);
	return fail(err, "unknown image type", NULL);
}

uint8_t* srpImageLoadRGB(const char* path, int* width, int* height, const char** err)
{
	FILE* f = fopen(path, "rb");
	if (!f) return fail(err, "can't fopen", NULL);
	if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return fail(err, "read error", NULL); }
	const long n = ftell(f);
	fseek(f, 0, SEEK_SET);
	if (n < 0) { fclose(f); return fail(err, "read error", NULL); }
	uint8_t* buf = (uint8_t*) malloc((size_t) n + 1);
	if (!buf || fread(buf, 1, (size_t) n, f) != (size_t) n) { fclose(f); return fail(err, "read error", buf); }
	fclose(f);
	uint8_t* out = srpImageDecodeRGB(buf, (size_t) n, width, height, err);
	free(buf);
	return out;
}
