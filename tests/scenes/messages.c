/* Error and warning paths of the API: a fixed sequence of misuses with a
 * message callback that records what the library reports (type, severity,
 * source function, text).  Runs on the oracle, the compiled reference and
 * the GPU library; their transcripts must agree. */
#include <stdio.h>
#include <string.h>

#include "shaders.h"
#include "scenes.h"

void sc_vsColorPass(SRPVertexShaderIn* in, SRPVertexShaderOut* out);
void sc_fsColor(SRPFragmentShaderIn* in, SRPFragmentShaderOut* out);

typedef struct Transcript { char* buf; size_t cap, len; } Transcript;

static void record(SRPMessageType type, SRPMessageSeverity sev, const char* fn, const char* msg, void* user)
{
	Transcript* t = (Transcript*) user;
	const int n = snprintf(t->buf + t->len, t->cap - t->len, "%d|%d|%s|%s\n", (int) type, (int) sev, fn, msg);
	if (n > 0 && (size_t) n < t->cap - t->len) t->len += (size_t) n;
}

int srpSceneMessages(const char* resDir, char* buf, size_t cap)
{
	Transcript t = {buf, cap, 0};
	buf[0] = 0;
	srpNewContext(&srpContext);
	SRPMessageCallback cb;
	cb.func = record;
	cb.userParameter = &t;
	srpSetMessageCallback(cb);

	VtxPosColor v[4];
	memset(v, 0, sizeof(v));
	for (int i = 0; i < 4; i++) { v[i].position = VEC3(-0.5f + 0.3f * i, (i & 1) ? 0.5f : -0.5f, 0.f); v[i].color = VEC3(1, 0, 0); }
	SRPVaryingInfo info[1] = {{3, SRP_FLOAT, SRP_INTERPOLATION_MODE_PERSPECTIVE}};
	SRPVertexShader vs; vs.shader = sc_vsColorPass; vs.nVaryings = 1; vs.varyingsInfo = info; vs.varyingsSize = sizeof(VarColor);
	SRPFragmentShader fs; fs.shader = sc_fsColor; fs.mayOverwriteDepth = false;
	SRPShaderProgram sp; sp.uniform = NULL; sp.vs = &vs; sp.fs = &fs;
	SRPFramebuffer* fb = srpNewFramebuffer(64, 64);
	memset(fb->stencil, 0, fb->size);
	srpFramebufferClear(fb);
	SRPVertexBuffer* vb = srpNewVertexBuffer();
	srpVertexBufferCopyData(vb, sizeof(VtxPosColor), sizeof(v), v);
	const uint16_t idx[5] = {0, 1, 2, 3, 0};
	SRPIndexBuffer* ib = srpNewIndexBuffer();
	srpIndexBufferCopyData(ib, SRP_UINT16, sizeof(idx), idx);

	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 0);         /* nothing */
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 2, 3);         /* OOB vertex buffer */
	srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_TRIANGLES, 3, 5);      /* OOB index buffer */
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 4);         /* excess vertex warning */
	srpDrawIndexBuffer(ib, vb, fb, &sp, SRP_PRIM_LINES, 0, 5);          /* odd line count warning */
	srpDrawVertexBuffer(vb, fb, &sp, (SRPPrimitive) 42, 0, 3);          /* unknown primitive */
	srpRasterCullFace(SRP_FACE_FRONT_AND_BACK);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_TRIANGLES, 0, 4);         /* culled: no warning either */
	srpRasterCullFace(SRP_FACE_NONE);
	srpRasterPointSize(0.0f);
	srpDrawVertexBuffer(vb, fb, &sp, SRP_PRIM_POINTS, 0, 4);            /* size 0: nothing */

	char path[4096];
	snprintf(path, sizeof(path), "%s/does_not_exist.png", resDir ? resDir : ".");
	SRPTexture* missing = srpNewTexture(path, TW_REPEAT, TW_REPEAT);
	if (missing != NULL) record(SRP_MESSAGE_ERROR, SRP_MESSAGE_SEVERITY_HIGH, "probe", "texture unexpectedly loaded", &t);
	snprintf(path, sizeof(path), "%s/stoneWall.png", resDir ? resDir : ".");
	SRPTexture* tex = srpNewTexture(path, TW_REPEAT, TW_CLAMP_TO_EDGE);
	char line[128];
	snprintf(line, sizeof(line), "get %d %d %d", srpTextureGet(tex, SRP_TEXTURE_WRAPPING_MODE_X),
	         srpTextureGet(tex, SRP_TEXTURE_WRAPPING_MODE_Y), srpTextureGet(tex, (SRPTextureParameter) 7));
	record(SRP_MESSAGE_WARNING, SRP_MESSAGE_SEVERITY_LOW, "probe", line, &t);
	srpTextureSet(tex, SRP_TEXTURE_WRAPPING_MODE_X, 5);
	srpTextureSet(tex, (SRPTextureParameter) 9, 0);
	srpTextureSet(tex, SRP_TEXTURE_WRAPPING_MODE_X, TW_CLAMP_TO_EDGE);
	float c[4];
	srpTextureGetFilteredColor(tex, 1.25f, -0.5f, c);
	snprintf(line, sizeof(line), "sample %.9g %.9g %.9g %.9g", c[0], c[1], c[2], c[3]);
	record(SRP_MESSAGE_WARNING, SRP_MESSAGE_SEVERITY_LOW, "probe", line, &t);
	srpFreeTexture(tex);

	srpFreeIndexBuffer(ib);
	srpFreeVertexBuffer(vb);
	srpFreeFramebuffer(fb);
	return (int) t.len;
}
