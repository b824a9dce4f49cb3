/*
 * st_layout.h - device memory layout of a resident scene and of one render
 * request. Plain C structs: the host fills them, the kernels read them.
 *
 * Every record is a multiple of 16 bytes so the kernels fetch it with
 * 128-bit loads. BVH nodes keep the reference's 32-byte node (AABB, first
 * child, count, depth; SlimTracin core/types.h:204-208) because the
 * traversal always reads both children of a node together: one pair is one
 * 64-byte segment, and node arrays are placed so that every pair starts on a
 * 64-byte boundary. Triangles are split into the 64-byte record every
 * triangle test reads (world_to_tangent, position, normal) and an 80-byte
 * record read once per closest hit (vertex normals, uvs, areas).
 */
#ifndef ST_LAYOUT_H
#define ST_LAYOUT_H

#include <stdint.h>

/* Traversal stack entries per ray (a launch keeps each stack's newest entries
   in LDS and the rest in the lane's private memory) */
#define ST_SCENE_STACK_MAX 64   /* scene BVH */
#define ST_MESH_STACK_MAX 128   /* mesh BVH */

/* Device scratch of st_render_noalloc (bytes, one array per GPU). */
#define ST_NOALLOC_SCRATCH_BYTES (128u << 20)

/* Checkpoint spacing of the primary directions along a row (pixels). */
#define ST_DIR_MARK 8u
#define ST_DIR_MARKS(width1) (((width1) + ST_DIR_MARK - 1u) / ST_DIR_MARK)

typedef struct StDPrim {
    float rot[4];          /* axis xyz, amount */
    float pos[3];
    uint32_t flags_mat;    /* flags (8 bits) | material_id << 8 */
    float scale[3];
    uint32_t type;
    float inv_scale[3];    /* 1/scale, the reference's oneOverVec3(scale) */
    uint32_t mesh_id;
} StDPrim; /* 64 B */

typedef struct StDMat {
    float albedo[3], reflectivity[3], emission[3];
    float uv_repeat[2];
    float metallic, roughness, n1_over_n2, n2_over_n1;
    uint32_t brdf;
    uint32_t bits;         /* is | use << 8 | texture_count << 16 */
    uint32_t tex0, tex1;   /* texture_ids[0], texture_ids[1] */
    uint32_t pad;
} StDMat; /* 80 B */

typedef struct StDLight {
    float pos[3], intensity;
    float color[3], pad;
} StDLight; /* 32 B */

typedef struct StDMesh {
    uint32_t node_offset;  /* first node of this mesh in the mesh node array */
    uint32_t tri_offset;   /* first triangle in the triangle arrays */
    uint32_t triangle_count;
    uint32_t attrs;        /* 1: has normals, 2: has uvs */
    uint32_t height, pad[3];
} StDMesh; /* 32 B */

typedef struct StDTriHot {
    float w2t[9];
    float pos[3];
    float normal[3];
    float pad;
} StDTriHot; /* 64 B */

typedef struct StDTriCold {
    float vn[9];
    float uv[6];
    float area, uv_area;
    float pad[3];
} StDTriCold; /* 80 B */

typedef struct StDTex {
    uint32_t width, height, mip_count, mipmap;
    uint32_t mip_base, pad[3];
} StDTex; /* 32 B */

typedef struct StDMip {
    uint32_t width, height;
    uint64_t quad_offset;  /* in 12-byte texel quads from the scene's texel base */
} StDMip; /* 16 B */

/* Device pointers of one resident scene, passed to kernels by value. */
typedef struct StDScene {
    const StDPrim *prims;
    const StDMat *mats;
    const StDLight *lights;
    const uint32_t *scene_nodes;   /* 8 words per node */
    const uint32_t *scene_leaf_ids;
    const StDMesh *meshes;
    const uint32_t *mesh_nodes;    /* 8 words per node, all meshes */
    const StDTriHot *tri_hot;
    const StDTriCold *tri_cold;
    const StDTex *textures;
    const StDMip *mips;
    const uint8_t *texels;
    uint32_t prim_count, light_count, mesh_count, texture_count;
    float ambient[3];
    uint32_t mesh_stack_size;      /* (u8)(max mesh BVH height + 2), core/init.h:523-527 */
    uint32_t has_emissive_quads;
    uint32_t scene_stack_need;     /* entries the scene-BVH traversal stack can reach */
    const float *mesh_aabbs;       /* 6 floats per mesh: the bound the SSB pass uses */
    uint32_t *fault;               /* per launch (set in the launch's scene table): a traversal that runs
                                      past its step bound or stack ORs ST_FAULT_* in and stops */
    uint32_t step_cap;             /* per launch: 0, or a lower step bound (st_debug_step_cap) */
    uint32_t main_step_limit;      /* per launch: mesh traversals of the main stages give up after this
                                      many steps and their pixel is redone (ST_ABANDON_STEPS) */
    uint32_t coop_min_tris;        /* per launch: the redo stage splits traversals of meshes of at least
                                      this many triangles across a group of blocks (ST_COOP_MIN_TRIS) */
    uint32_t stack_mask_s, stack_mask_m; /* per launch: LDS slots of the scene / mesh traversal stack - 1
                                      (a power of two less one; all ones: the whole stack is in LDS) */
    uint32_t *spill;               /* per launch, with a window: the stacks' older entries, per wave of a
                                      stage kernel (64 lanes x (spill_s + spill_m) words) */
    uint32_t spill_s, spill_m;     /* entries per lane of the scene and mesh stacks in `spill` */
} StDScene;

/* Mesh traversals longer than this (node-pair steps) give up in the main
   stages: the lane would hold its wave for the whole chain of dependent
   fetches (a ray with an exactly zero object-space direction component
   makes the reference's slab test admit every box on one side of the mesh,
   hundreds of thousands of steps). Its pixel is redone from the start by
   the redo stage, where a group of blocks traces it together. */
#define ST_ABANDON_STEPS 2048u
/* Meshes from this size are traversed by the pixel's group of blocks in the
   redo stage; smaller ones sequentially (a few hundred steps at most). */
#define ST_COOP_MIN_TRIS 4096u

/* Fault bits of StDScene.fault. A traversal of a valid tree never sets them:
   it loads each node pair at most once (so takes at most as many steps as the
   tree has nodes) and pushes at most its height. */
#define ST_FAULT_MESH_STEPS 1u
#define ST_FAULT_SCENE_STEPS 2u
#define ST_FAULT_SCENE_STACK 4u

/* One render request on the device. */
typedef struct StDView {
    float cam_pos[3], cone_angle;
    float cam_rot_inv[4];          /* camera rotation_inverted quaternion */
    float right[3];                /* projection-plane step along a row */
    uint32_t width1, height1;      /* traced grid (window + 1 each way) */
    uint32_t mode, depth;
    uint32_t scene;                /* index into the scene table (batched launches) */
    uint32_t tile_base;            /* first tile of this request in a batched launch */
    uint32_t tiles_x;              /* tiles per row */
    uint32_t tile_shift;           /* tile width = 1 << tile_shift lanes, height = 64 >> tile_shift */
    const float *row_starts;       /* 4 floats per row: sequential start vectors */
    float *dir_marks;              /* 4 floats per ST_DIR_MARK pixels of a row: the row's sequential
                                      `current` vector at x = 0, 8, 16, ... (rows of ST_DIR_MARKS(width1)) */
    const int32_t *rects;          /* 4 ints per primitive: SSB min.x, min.y, max.x, max.y */
    void *pixels;                  /* StFloatPixel per pixel */
    void *hits;                    /* StHitRec per pixel, or NULL */
    unsigned long long *counters;  /* primary, shadow, bounce */
} StDView;

#endif
