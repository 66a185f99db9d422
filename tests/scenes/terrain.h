/* Terrain workload: geometry, camera poses and render state (terrain.c). */
#pragma once

#include "shaders.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TERRAIN_TRAIN_POSES 256
#define TERRAIN_TEST_POSES 64

void srpTerrainBuild(int nx, int nz, VtxTerrain** verts, size_t* nVerts, uint32_t** idx, size_t* nIdx,
                     TerrainBounds* bounds);
void srpTerrainPose(int pose, int width, int height, UniTerrain* u);
void srpTerrainState(void);
extern SRPVaryingInfo g_terrainVaryings[3];

#ifdef __cplusplus
}
#endif
