/* Limits on what one draw may produce on the GPU path, checked on the host
 * from the counting pass's exact (64-bit) totals before any buffer is sized
 * or the emitting pass runs.  The emitting pass writes at 32-bit scanned
 * offsets, so a total that does not fit would wrap them; the draw is then
 * refused through the message callback instead, the reference's convention
 * for a draw it cannot perform (pipeline/draw.c:167-189: a message and an
 * early return, nothing drawn).
 *
 * Header-only and free of HIP so that a host test can exercise it with
 * synthetic totals well beyond 2^32. */
#pragma once

#include <stdint.h>
#include <stdio.h>

namespace srp {
namespace lib {

/* Record indices carry a flag in bit 31 of the tile pairs' record word;
 * list indices and tile ranges are 32-bit.  2^31 - 1 items of 28 bytes
 * (pair key, value, record, mask, sort copies) is already ~60 GB. */
constexpr uint64_t kMaxDrawRecords = 0x7FFFFFFFull;
constexpr uint64_t kMaxDrawItems = 0x7FFFFFFFull;
/* post-VS cache slots plus clip-generated vertices share one 32-bit slot space */
constexpr uint64_t kMaxDrawVertexSlots = 0xFFFFFFFEull;

/* Exact totals of a draw's counting pass. */
struct DrawTotals
{
	uint64_t records;     /* assembled primitives */
	uint64_t generated;   /* clip-generated vertices */
	uint64_t items;       /* tile pairs (filled triangles) or fragments (lines, points) */
	uint64_t cacheSlots;  /* post-VS cache slots of the draw */
};

enum DrawLimit
{
	DRAW_FITS = 0,
	DRAW_TOO_MANY_RECORDS = 1,
	DRAW_TOO_MANY_ITEMS = 2,
	DRAW_TOO_MANY_VERTICES = 3
};

inline DrawLimit checkDrawTotals(const DrawTotals& t)
{
	if (t.records > kMaxDrawRecords)
		return DRAW_TOO_MANY_RECORDS;
	if (t.items > kMaxDrawItems)
		return DRAW_TOO_MANY_ITEMS;
	if (t.cacheSlots > kMaxDrawVertexSlots || t.generated > kMaxDrawVertexSlots - t.cacheSlots)
		return DRAW_TOO_MANY_VERTICES;
	return DRAW_FITS;
}

/* Text of the message for a draw over a limit (into buf; returns buf). */
inline const char* drawLimitText(DrawLimit code, const DrawTotals& t, char* buf, size_t size)
{
	switch (code)
	{
		case DRAW_TOO_MANY_RECORDS:
			snprintf(buf, size, "Draw assembles %llu primitives; the GPU path draws at most %llu per call",
			         (unsigned long long) t.records, (unsigned long long) kMaxDrawRecords);
			break;
		case DRAW_TOO_MANY_ITEMS:
			snprintf(buf, size,
			         "Draw needs %llu tile entries or fragments; the GPU path handles at most %llu per call "
			         "(split the draw)",
			         (unsigned long long) t.items, (unsigned long long) kMaxDrawItems);
			break;
		case DRAW_TOO_MANY_VERTICES:
			snprintf(buf, size, "Draw uses %llu cached and %llu clip-generated vertices; the GPU path allows %llu",
			         (unsigned long long) t.cacheSlots, (unsigned long long) t.generated,
			         (unsigned long long) kMaxDrawVertexSlots);
			break;
		default:
			snprintf(buf, size, "Draw fits");
			break;
	}
	return buf;
}

}  // namespace lib
}  // namespace srp
