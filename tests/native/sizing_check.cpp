/* Host check of the per-draw limits (srp_amd/csrc/lib/sizing.h) that the
 * library applies to the counting pass's exact 64-bit totals before it
 * sizes buffers or runs the emitting pass: totals at, just past and far
 * beyond 2^31 and 2^32 (synthetic; no draw is run).  Prints one line per
 * case and "ok" when every answer is the expected one. */
#include <stdio.h>
#include <string.h>
#include "sizing.h"

using namespace srp::lib;

static int failures = 0;

static void expect(const char* what, DrawTotals t, DrawLimit want, const char* textPart)
{
	const DrawLimit got = checkDrawTotals(t);
	char buf[256];
	drawLimitText(got, t, buf, sizeof(buf));
	const bool ok = got == want && (!textPart || strstr(buf, textPart));
	printf("%-44s -> %d %s: %s\n", what, (int) got, ok ? "ok" : "WRONG", buf);
	if (!ok)
		failures++;
}

int main()
{
	const uint64_t two31 = 1ull << 31, two32 = 1ull << 32;
	expect("the 1080p mesh frame", {261120, 0, 354000, 131841}, DRAW_FITS, nullptr);
	expect("items at the limit", {1000, 0, kMaxDrawItems, 100}, DRAW_FITS, nullptr);
	expect("items one past the limit", {1000, 0, kMaxDrawItems + 1, 100}, DRAW_TOO_MANY_ITEMS, "2147483648 tile");
	expect("items 2^32 (wraps to 0 in 32 bits)", {1000, 0, two32, 100}, DRAW_TOO_MANY_ITEMS, "4294967296");
	expect("items 2^32 + 5 (wraps to 5)", {1000, 0, two32 + 5, 100}, DRAW_TOO_MANY_ITEMS, "4294967301");
	expect("items ~430k points of size 100", {430000, 0, 430000ull * 10000ull, 430000}, DRAW_TOO_MANY_ITEMS, nullptr);
	expect("items 2^40", {1000, 0, 1ull << 40, 100}, DRAW_TOO_MANY_ITEMS, nullptr);
	expect("records 2^31", {two31, 0, 100, 100}, DRAW_TOO_MANY_RECORDS, "2147483648 primitives");
	expect("records 2^31 - 1", {two31 - 1, 0, 100, 100}, DRAW_FITS, nullptr);
	expect("records and items both over", {two32, 0, two32, 100}, DRAW_TOO_MANY_RECORDS, nullptr);
	expect("generated vertices fill the slot space", {100, kMaxDrawVertexSlots - 100, 100, 100}, DRAW_FITS, nullptr);
	expect("generated vertices one past", {100, kMaxDrawVertexSlots - 99, 100, 100}, DRAW_TOO_MANY_VERTICES,
	       "clip-generated");
	expect("cache slots past the slot space", {100, 0, 100, two32}, DRAW_TOO_MANY_VERTICES, nullptr);
	expect("generated 2^33 (would wrap the sum)", {100, 1ull << 33, 100, 100}, DRAW_TOO_MANY_VERTICES, nullptr);
	printf(failures ? "FAILED %d\n" : "ok\n", failures);
	return failures ? 1 : 0;
}
