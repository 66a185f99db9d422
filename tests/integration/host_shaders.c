/* Host build of the INTEGRATION.md sample's shaders (for the CPU oracle):
 * the same sample_shaders.h with SHADER empty, exported under the names the
 * shader module registers on the GPU. */
#define SHADER static
#include "sample_shaders.h"

void sampleVS(SRPVertexShaderIn* in, SRPVertexShaderOut* out) { vs(in, out); }
void sampleFS(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out) { fs(in, out); }
