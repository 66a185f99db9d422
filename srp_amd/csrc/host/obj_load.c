/* Wavefront OBJ loader (see obj_load.h). */
#include "obj_load.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct Grow { void* p; size_t n, cap, elem; } Grow;

static void* growPush(Grow* g)
{
	if (g->n == g->cap)
	{
		g->cap = g->cap ? 2 * g->cap : 1024;
		g->p = realloc(g->p, g->cap * g->elem);
	}
	return (char*) g->p + (g->n++) * g->elem;
}

bool srpObjLoad(const char* path, SRPObjMesh* mesh)
{
	FILE* f = fopen(path, "r");
	if (!f) return false;
	Grow pos = {NULL, 0, 0, sizeof(vec3)}, uv = {NULL, 0, 0, sizeof(vec2)}, nrm = {NULL, 0, 0, sizeof(vec3)};
	Grow verts = {NULL, 0, 0, sizeof(SRPObjVertex)}, idx = {NULL, 0, 0, sizeof(uint32_t)};
	char line[256];
	while (fgets(line, sizeof(line), f))
	{
		if (line[0] == 'v' && line[1] == ' ')
		{
			vec3 p = {{0, 0, 0}};
			sscanf(line, "v %f %f %f", &p.x, &p.y, &p.z);
			*(vec3*) growPush(&pos) = p;
		}
		else if (line[0] == 'v' && line[1] == 't')
		{
			vec2 t = {{0, 0}};
			sscanf(line, "vt %f %f", &t.x, &t.y);
			*(vec2*) growPush(&uv) = t;
		}
		else if (line[0] == 'v' && line[1] == 'n')
		{
			vec3 n = {{0, 0, 0}};
			sscanf(line, "vn %f %f %f", &n.x, &n.y, &n.z);
			*(vec3*) growPush(&nrm) = n;
		}
		else if (line[0] == 'f')
		{
			int v[3], t[3], n[3];
			if (sscanf(line, "f %d/%d/%d %d/%d/%d %d/%d/%d",
			           &v[0], &t[0], &n[0], &v[1], &t[1], &n[1], &v[2], &t[2], &n[2]) != 9)
			{
				fprintf(stderr, "Unsupported face format: %s\n", line);
				continue;
			}
			/* all three corners are checked before any is pushed, so a bad
			 * face is skipped whole and later triangles stay aligned */
			int ok = 1;
			for (int k = 0; k < 3; k++)
				if (v[k] < 1 || (size_t) v[k] > pos.n || t[k] < 1 || (size_t) t[k] > uv.n || n[k] < 1 || (size_t) n[k] > nrm.n)
					ok = 0;
			if (!ok)
			{
				fprintf(stderr, "Face index out of range: %s\n", line);
				continue;
			}
			for (int k = 0; k < 3; k++)
			{
				SRPObjVertex* ov = (SRPObjVertex*) growPush(&verts);
				ov->position = ((vec3*) pos.p)[v[k] - 1];
				ov->uv = ((vec2*) uv.p)[t[k] - 1];
				ov->normal = ((vec3*) nrm.p)[n[k] - 1];
				*(uint32_t*) growPush(&idx) = (uint32_t) (verts.n - 1);
			}
		}
	}
	fclose(f);
	free(pos.p); free(uv.p); free(nrm.p);
	mesh->vertices = (SRPObjVertex*) verts.p;
	mesh->vertexCount = verts.n;
	mesh->indices = (uint32_t*) idx.p;
	mesh->indexCount = idx.n;
	return true;
}

void srpObjFree(SRPObjMesh* mesh)
{
	free(mesh->vertices);
	free(mesh->indices);
	mesh->vertices = NULL;
	mesh->indices = NULL;
}
