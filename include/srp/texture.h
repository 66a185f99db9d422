/* Textures.  Interface of reference include/srp/texture.h:18-67.
 *
 * An SRPTexture* returned by this library is a device address of the
 * texture descriptor, so a uniform block that embeds it can be copied to the
 * GPU as raw bytes and dereferenced there by srpTextureGetFilteredColor. */
#pragma once

#include <stdint.h>
#include <stddef.h>
#include "srp/color.h"
#include "srp/detail/qualifiers.h"

typedef enum
{
	TW_REPEAT,
	TW_CLAMP_TO_EDGE
} SRPTextureWrappingMode;

typedef struct SRPTexture SRPTexture;

typedef enum SRPTextureParameter
{
	SRP_TEXTURE_WRAPPING_MODE_X,
	SRP_TEXTURE_WRAPPING_MODE_Y
} SRPTextureParameter;

SRP_EXTERN_C_BEGIN
/** Loads an image file (PNG) as RGB8, like the reference's 3-channel load. */
SRPTexture* srpNewTexture(
	const char* image, SRPTextureWrappingMode wrappingModeX, SRPTextureWrappingMode wrappingModeY);
void srpFreeTexture(SRPTexture* this_);
int srpTextureGet(SRPTexture* this_, SRPTextureParameter parameter);
void srpTextureSet(SRPTexture* this_, SRPTextureParameter parameter, int data);
#if !defined(__HIPCC__)
/** Nearest-texel lookup; out = RGBA in [0,1], alpha 1. */
void srpTextureGetFilteredColor(const SRPTexture* this_, float u, float v, float out[4]);
#endif
SRP_EXTERN_C_END

#if defined(__HIPCC__)
/* Shader modules: a host+device definition (device side reads the
 * descriptor directly, host side goes through the library). */
#  include "srp/detail/texture_sample.h"
#endif
