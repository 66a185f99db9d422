/* Host (function-pointer) instances of the scene shaders, for the CPU
 * builds: the oracle and the compiled reference. */
#include "shaders.h"

#define HOST_VS(name, fn, usize, part) void name(SRPVertexShaderIn* in, SRPVertexShaderOut* out) { fn(in, out); }
#define HOST_FS(name, fn, usize, part) void name(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out) { fn(in, out); }

#include "shader_list.h"
