/*
 * slimtracin_io.h - the reference's on-disk asset formats: .mesh (mesh with
 * its prebuilt BVH and triangle records) and .texture (mip chain of texel
 * quads), their readers, and converters equivalent to the reference's
 * obj2mesh and bmp2texture tools. Host-only (no device needed).
 *
 * Layouts are the Windows build's (u32 = 4 bytes), which is what the
 * reference's tools write and its .mesh files hold:
 *   .mesh     scene/io.h:92-175 (saveMeshToFile / loadMeshFromFile)
 *   .texture  scene/io.h:34-57 (loadTextureFromFile), bmp2texture.c:361-376
 * Readers return views into the caller's buffer (no copies); writers return
 * malloc'd images to release with st_free (slimtracin_gen.h).
 */
#ifndef SLIMTRACIN_IO_H
#define SLIMTRACIN_IO_H

#include <stddef.h>
#include <stdint.h>

#include "st_blob.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct StMeshFile {
    float aabb[6];                    /* min xyz, max xyz */
    uint32_t vertex_count, triangle_count, edge_count, uvs_count, normals_count, node_count, height;
    const float *positions;           /* vertex_count x 3 */
    const uint32_t *position_indices; /* triangle_count x 3 */
    const uint32_t *edges;            /* edge_count x (from, to), from < to */
    const float *uvs;                 /* uvs_count x 2, NULL when uvs_count is 0 */
    const uint32_t *uv_indices;       /* triangle_count x 3, NULL when uvs_count is 0 */
    const float *normals;             /* normals_count x 3, NULL when normals_count is 0 */
    const uint32_t *normal_indices;   /* triangle_count x 3, NULL when normals_count is 0 */
    const StTriangleRec *triangles;   /* triangle_count, in BVH leaf order */
    const StBVHNodeRec *nodes;        /* node_count */
    const uint32_t *leaf_ids;         /* triangle_count: original triangle -> leaf-order index */
} StMeshFile;

/* Parse a .mesh image (loadMeshFromFile, scene/io.h:92-143). ST_ERR_ARG with
   the reason in st_last_error when the sizes do not add up. */
int st_mesh_file_parse(const void *data, size_t size, StMeshFile *out);
/* Serialise (saveMeshToFile, scene/io.h:145-172). */
int st_mesh_file_write(const StMeshFile *mesh, void **out, size_t *size);
/* obj2mesh (src/obj2mesh.c:25-284): Wavefront OBJ text with faces of the
   forms "f a b c", "f a/t b/t c/t" or "f a/t/n b/t/n c/t/n" (the form of
   the first face holds for all), bound from the origin and the vertices,
   vertices recentred on the bound's centre, unique undirected edges in
   first-seen order, the reference SAH mesh BVH (leaves <= 4) and triangle
   records. invert_winding swaps each face's 2nd and 3rd vertices. Rejects
   the inputs the tool would crash on (attribute lines without the matching
   face form). */
int st_obj_to_mesh(const char *text, size_t len, int invert_winding, void **out, size_t *size);

typedef struct StTextureMipView {
    uint16_t width, height;
    const uint8_t *quads;             /* (height+1) x (width+1) TexelQuads, 12 B each */
} StTextureMipView;

typedef struct StTextureFile {
    uint16_t width, height;
    uint8_t mipmap, wrap, mip_count;
    StTextureMipView mips[16];
} StTextureFile;

/* Parse a .texture image (loadTextureFromFile, scene/io.h:34-57). */
int st_texture_file_parse(const void *data, size_t size, StTextureFile *out);
/* bmp2texture's loadTexture + file write (bmp2texture.c:226-376): texels are
   width x height rows of bytes_per_texel (3 or 4) bytes whose first three
   are r, g, b (a fourth is skipped); texel quads with clamped or wrapped
   borders; with mipmap, halvings while both sides exceed 4, each texel the
   float mean of its 2x2 block of the finer level (kept in float between
   levels), stored truncated to u8. */
int st_texture_from_texels(const uint8_t *texels, uint16_t width, uint16_t height, uint32_t bytes_per_texel,
                           int wrap, int mipmap, void **out, size_t *size);
/* bmp2texture (bmp2texture.c:36-100, 361-376): an uncompressed 24- or
   32-bit BMP image as the tool reads it (rows in file order, B and R
   swapped, no row padding skipped). */
int st_bmp_to_texture(const void *bmp, size_t size, int wrap, int mipmap, void **out, size_t *out_size);

#ifdef __cplusplus
}
#endif
#endif
