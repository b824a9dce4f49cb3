/*
 * slimtracin_gen.h - seeded synthetic scenes of the benchmark workloads,
 * returned as scene blobs (st_blob.h). The blob is malloc'd; release it with
 * st_free(). These have no counterpart in the reference, whose scenes are
 * hand-written in each example's setupScene.
 */
#ifndef SLIMTRACIN_GEN_H
#define SLIMTRACIN_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 4 quads + 4 boxes + 4 spheres + 4 tetrahedra, 3 point lights, Blinn. */
int st_gen_implicit(uint64_t seed, void **out, size_t *size);
/* Noise-displaced geodesic icosphere with 20*k*k triangles + glass sphere,
   mirror quad, floor, 4 point lights, 2 emissive quads. */
int st_gen_pbr(uint64_t seed, uint32_t k, void **out, size_t *size);
/* Noise-displaced sphere mesh with about `triangles` triangles + floor. */
int st_gen_mesh_scene(uint64_t seed, uint32_t triangles, void **out, size_t *size);
/* `instances` mesh primitives over `mesh_count` noise-displaced meshes of
   mesh_triangles[i] triangles each, a floor, 3 point lights, materials
   using a Perlin albedo map and a Perlin normal map of tex_size^2 texels
   (power of two, full mip chains to 4x4). */
int st_gen_instanced(uint64_t seed, uint32_t instances, const uint32_t *mesh_triangles, uint32_t mesh_count,
                     uint32_t tex_size, void **out, size_t *size);
void st_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
