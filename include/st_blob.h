/*
 * st_blob.h - flat, little-endian interchange format for one scene and one
 * render view.
 *
 * A scene blob holds everything the reference keeps in its `Scene` struct
 * that the per-pixel tracing path reads (SlimTracin core/types.h:144-152
 * Primitive, :300-326 Light/Material, :272-292 Mesh, :56-74 Texture), in
 * fixed-width fields so that one byte string can be handed unchanged to the
 * HIP library, to the CPU oracle and to the harness around the reference
 * build. Camera and frame settings are not part of the scene: they travel per
 * render request as an StView (the reference keeps them in `Viewport`,
 * core/types.h:239-268, and `Camera`, :169-174).
 *
 * Layout:
 *   StBlobHeader
 *   StPrimRec     [prim_count]
 *   StMaterialRec [material_count]
 *   StLightRec    [light_count]
 *   per mesh:    StMeshRec, then its arrays (see StMeshRec)
 *   per texture: StTextureRec, then per mip: u16 width, u16 height,
 *                (width+1)*(height+1) texel quads of 12 bytes, padded to 4 B
 *   if (flags & ST_BLOB_HAS_SCENE_BVH): u32 node_count, u32 height,
 *                StBVHNodeRec nodes[node_count], u32 leaf_ids[prim_count]
 *                (a scene BVH built elsewhere, e.g. by the Morton LBVH,
 *                used instead of the SAH build over the primitives)
 */
#ifndef ST_BLOB_H
#define ST_BLOB_H

#include <stdint.h>

#define ST_BLOB_MAGIC 0x43535453u /* "STSC" */
#define ST_BLOB_VERSION 1u

/* Primitive types, same values as `enum PrimitiveType` (core/types.h:127-135). */
enum { ST_PRIM_NONE = 0, ST_PRIM_MESH = 1, ST_PRIM_BOX = 2, ST_PRIM_TETRAHEDRON = 3,
       ST_PRIM_QUAD = 4, ST_PRIM_SPHERE = 5, ST_PRIM_LIGHT = 6 /* picking only */ };

/* Render modes, same values as `enum RenderMode` (core/types.h:137-142). */
enum { ST_MODE_NORMALS = 0, ST_MODE_BEAUTY = 1, ST_MODE_DEPTH = 2, ST_MODE_UVS = 3 };

/* BRDFs, same values as `enum BRDF` (core/types.h:313-318). */
enum { ST_BRDF_LAMBERT = 0, ST_BRDF_PHONG = 1, ST_BRDF_BLINN = 2, ST_BRDF_COOK_TORRANCE = 3 };

/* Primitive flags (core/base.h:130-137). */
#define ST_IS_VISIBLE 1u
#define ST_IS_TRANSLATED 2u
#define ST_IS_ROTATED 4u
#define ST_IS_SCALED 8u
#define ST_IS_SCALED_NON_UNIFORMLY 16u
#define ST_IS_SHADOWING 32u
#define ST_IS_TRANSPARENT 64u
#define ST_ALL_FLAGS 63u

/* Material `is` / `use` bits (core/base.h:154-159). */
#define ST_REFLECTIVE 1u
#define ST_REFRACTIVE 2u
#define ST_EMISSIVE 4u
#define ST_ALBEDO_MAP 1u
#define ST_NORMAL_MAP 2u

typedef struct StBlobHeader {
    uint32_t magic, version;
    uint32_t prim_count, material_count, light_count, mesh_count, texture_count;
    float ambient[3];
    uint32_t flags;
    uint32_t reserved[3];
} StBlobHeader; /* 56 B */

#define ST_BLOB_HAS_SCENE_BVH 1u

typedef struct StPrimRec {
    float position[3], scale[3];
    float rotation[4];          /* quaternion: axis.x, axis.y, axis.z, amount */
    uint32_t type, mesh_id, material_id, flags;
} StPrimRec; /* 56 B */

typedef struct StMaterialRec {
    float albedo[3], reflectivity[3], emission[3], uv_repeat[2];
    float metallic, roughness, n1_over_n2, n2_over_n1;
    uint32_t brdf;
    uint8_t is, use, texture_count, pad;
    uint8_t texture_ids[16];
} StMaterialRec; /* 84 B */

typedef struct StLightRec {
    float position[3], color[3], intensity;
} StLightRec; /* 28 B */

/*
 * One mesh. Followed by:
 *   f32 positions[vertex_count*3], u32 position_indices[triangle_count*3],
 *   f32 normals[normals_count*3], u32 normal_indices[triangle_count*3] (if normals_count),
 *   f32 uvs[uvs_count*2],         u32 uv_indices[triangle_count*3]     (if uvs_count),
 *   and when `prebuilt` is 1 (a BVH read from a reference .mesh file):
 *   StBVHNodeRec nodes[node_count], StTriangleRec triangles[triangle_count],
 *   u32 leaf_ids[triangle_count].
 * `aabb` is the mesh bound (min xyz, max xyz) the scene BVH and the SSB use.
 */
typedef struct StMeshRec {
    uint32_t vertex_count, triangle_count, normals_count, uvs_count;
    uint32_t prebuilt, node_count, height, reserved;
    float aabb[6];
} StMeshRec; /* 56 B */

/* BVH node in the reference's 4-byte-u32 layout (core/types.h:204-208): 32 B. */
typedef struct StBVHNodeRec {
    float min[3], max[3];
    uint32_t first_child_id;
    uint16_t child_count, depth;
} StBVHNodeRec;

/* Precomputed triangle in the reference's layout (core/types.h:272-278): 128 B. */
typedef struct StTriangleRec {
    float world_to_tangent[9]; /* rows X, Y, Z */
    float position[3], normal[3];
    float vertex_normals[9];
    float uvs[6];
    float area_of_parallelogram, area_of_uv;
} StTriangleRec;

typedef struct StTextureRec {
    uint16_t width, height;
    uint8_t mip_count, wrap, mipmap, pad;
} StTextureRec; /* 8 B */

/*
 * Per-request view: camera and frame. The camera rotation is applied as one
 * `rotateXform3(xform, yaw, pitch, roll)` from identity (scene/xform.h:8-20),
 * where each amount is the rational parameter of core/base.h:404-411.
 * width/height are window sizes; the traced grid is (width+1) x (height+1)
 * (QCAA, core/base.h:344-358).
 */
typedef struct StView {
    float position[3];
    float yaw, pitch, roll;
    float focal_length;
    float cone_angle;   /* projection-plane cone angle (0 in the reference, whose value is never set) */
    uint16_t width, height;
    uint32_t render_mode;
    uint32_t depth;     /* trace depth: 1 = no bounce (core/init.h:533 default 2) */
    uint32_t reserved;
} StView; /* 48 B */

/* Per-pixel output of the tracer: the reference's FloatPixel (core/base.h:207). */
typedef struct StFloatPixel {
    float r, g, b, opacity;
    double depth;
} StFloatPixel; /* 24 B */

/* Primary-ray hit record for parity checks. */
typedef struct StHitRec {
    int32_t prim_id;        /* scene primitive index of the closest primary hit, -1 = miss */
    int32_t object_type;    /* ST_PRIM_* of that primitive */
    int32_t triangle_id;    /* triangle index (BVH leaf order) for mesh hits, -1 otherwise */
    float distance;         /* world-space hit distance t */
} StHitRec; /* 16 B */

/* Result of picking one window point (the first-press branch of
   manipulateSelection, viewport/manipulation.h:55-124). */
typedef struct StPick {
    int32_t object_type;    /* ST_PRIM_* of the primitive, ST_PRIM_LIGHT, or 0 when nothing is hit */
    uint32_t object_id;     /* primitive index (scene order) or light index */
    float position[3];      /* world-space hit point (the selection's transformation plane origin) */
    float object_distance;  /* view-space z of the hit point (selection->object_distance) */
} StPick; /* 24 B */

/* Ray counters: every ray that reaches an intersection routine. */
typedef struct StRayCounts {
    uint64_t primary, shadow, bounce;
} StRayCounts;

#endif
