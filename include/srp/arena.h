/* Opaque per-context scratch handle (reference include/srp/arena.h:14).
 * In this library the slot in SRPContext holds the device runtime state. */
#pragma once

typedef struct SRPArena SRPArena;
