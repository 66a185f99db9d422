/* RGBA8 colour (reference include/srp/color.h:16-19). */
#pragma once

#include <stdint.h>

typedef struct SRPColor
{
	uint8_t r, g, b, a;
} SRPColor;
