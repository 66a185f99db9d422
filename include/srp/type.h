/* Scalar type tags used for opaque data (index buffers, varyings).
 * Interface of reference include/srp/type.h:17-29 (same enumerators, same values). */
#pragma once

#include <stdint.h>
#include <stddef.h>

typedef enum SRPType
{
	SRP_FLOAT = 0,
	SRP_DOUBLE,
	SRP_INT8,
	SRP_INT16,
	SRP_INT32,
	SRP_INT64,
	SRP_UINT8,
	SRP_UINT16,
	SRP_UINT32,
	SRP_UINT64
} SRPType;
