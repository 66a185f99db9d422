/* Framebuffer.  Interface of reference include/srp/framebuffer.h:17-39.
 *
 * `color`, `depth` and `stencil` are host arrays the caller may read and
 * write between API calls, exactly as with the reference.  The device copy
 * lives behind them (see srp_hip.h for the deferred, device-resident mode). */
#pragma once

#include <stddef.h>
#include <stdint.h>
#include "srp/detail/qualifiers.h"

typedef struct SRPFramebuffer
{
	size_t width;
	size_t height;
	size_t size;       /**< width * height */
	uint32_t* color;   /**< R<<24 | G<<16 | B<<8 | A per pixel, row-major */
	float* depth;
	uint8_t* stencil;
} SRPFramebuffer;

SRP_EXTERN_C_BEGIN
SRPFramebuffer* srpNewFramebuffer(size_t width, size_t height);
void srpFreeFramebuffer(SRPFramebuffer* this_);
/** Colour := 0, depth := -1.0; stencil is left as it is. */
void srpFramebufferClear(const SRPFramebuffer* this_);
SRP_EXTERN_C_END
