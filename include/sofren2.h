/*
 * sofren2.h — public C API of the MI355X-native sofren2 renderer.
 *
 * This header is the drop-in boundary: it re-declares the public API of the
 * reference single-header renderer (sofren2.c:794-864) with identical names,
 * argument meaning and value-type layouts, so host code written against the
 * reference links against libsofren2_gpu.so unchanged.  Every function is an
 * exported extern "C" symbol (the reference marks them `static` because it is
 * compiled into one translation unit; a shared library has to export them).
 *
 * Value types that cross the ABI keep the reference layout byte for byte:
 *   sftvec 12 B, sftquat 16 B, sftmat 64 B (row-vector convention, translation
 *   in row 3), sft_instance 88 B, sft_arena 40 B.  The model/mesh/meshlet
 *   structs are defined here unconditionally (the reference defines them only
 *   under SFT_IMPL, sofren2.c:870-1327) so that host code which builds models in
 *   memory — the only way to make skinned models without the glTF loader — keeps
 *   working.  Sizes are checked by static asserts at the end of this file.
 *
 * Ownership and errors follow the reference (sofren2.c:238-247): fatal errors
 * print to stderr and exit(1); recoverable ones print and return a sentinel
 * (NULL, SFT_INVALID_HANDLE, or nothing).  HIP errors are fatal.
 */
#ifndef SFT_H
#define SFT_H

#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include <math.h>

#ifdef __cplusplus
extern "C" {
#define SFT_ALIGNAS(n) alignas(n)
#else
#include <stdalign.h>
#define SFT_ALIGNAS(n) alignas(n)
#endif

/* ---- scalar typedefs (sofren2.c:20-55) ---- */
#ifdef SFT_PREFIXED_TYPES
    #define i8  sfri8_t
    #define u8  sfru8_t
    #define i16 sfri16_t
    #define u16 sfru16_t
    #define i32 sfri32_t
    #define u32 sfru32_t
    #define i64 sfri64_t
    #define u64 sfru64_t
    #define f32 sfrf32_t
    #define f64 sfrf64_t
#endif
typedef int8_t   i8;
typedef uint8_t  u8;
typedef int16_t  i16;
typedef uint16_t u16;
typedef int32_t  i32;
typedef uint32_t u32;
typedef int64_t  i64;
typedef uint64_t u64;
typedef float    f32;
typedef double   f64;

/* ---- public enums (sofren2.c:58-72) ---- */
enum sft_debug_mode {
    SFT_DEBUG_NONE,
    SFT_DEBUG_WIREFRAME,
    SFT_DEBUG_TILES,
    SFT_DEBUG_DEPTH,
    SFT_DEBUG_LIGHT_COUNT,
    SFT_DEBUG_CHECKERBOARD,
    SFT_DEBUG_COUNT
};

enum sft_dmesh_type {
    SFT_DMESH_QUADS,
    SFT_DMESH_TRIANGLES
};

typedef u32 sft_tex_handle;
typedef u32 sft_mat_handle;

/* handle layouts (sofren2.c:101-117) */
#define SFT_INVALID_HANDLE 0xFFFFFFFFu
#define SFT_HANDLE_INDEX_MASK 0x0000FFFFu
#define SFT_HANDLE_GEN_SHIFT  16
#define SFT_HANDLE_GEN_MASK   0xFFFF0000u
#define SFT_HANDLE_INDEX(h) ((h) & SFT_HANDLE_INDEX_MASK)
#define SFT_MAT_TRANSIENT_BIT 0x80000000u
#define SFT_MAT_GEN_SHIFT     16
#define SFT_MAT_GEN_MASK      0x7FFF0000u
#define SFT_MAT_INDEX_MASK    0x0000FFFFu
#define SFT_MAT_IS_TRANSIENT(h) (0 != ((h) & SFT_MAT_TRANSIENT_BIT))
#define SFT_MAT_INDEX(h)        ((h) & SFT_MAT_INDEX_MASK)

/* misc macros (sofren2.c:123-150) */
#define SFT_PI (3.14159265358979323846)
#define SFT_EPSILON (1e-6f)
#define SFT_ARRLEN(_arr)  (sizeof(_arr) / sizeof((_arr)[0]))
#define SFT_CLAMP(_x, _min, _max) ((_x) < (_min) ? (_min) : ((_x) > (_max) ? (_max) : (_x)))
#define SFT_MIN(a, b) ((a) < (b) ? (a) : (b))
#define SFT_MAX(a, b) ((a) > (b) ? (a) : (b))
#define SFT_LERPF(_a, _b, _t) ((_a) + (_t) * ((_b) - (_a)))
#ifndef SFT_SHADOW_LIGHT_SIZE
    #define SFT_SHADOW_LIGHT_SIZE 0.1f
#endif
#define SFT_CONE_LIMIT (-0.173648f)

/* visibility-ID bit split; must total 31, bit 31 flags a clipped triangle
   (sofren2.c:155-166).  The split is fixed when the library is built. */
#ifndef SFT_INSTANCE_BITS
    #define SFT_INSTANCE_BITS 6
#endif
#ifndef SFT_MESHLET_BITS
    #define SFT_MESHLET_BITS 18
#endif
#define SFT_TRIANGLE_BITS 7
#if 31 != (SFT_INSTANCE_BITS + SFT_MESHLET_BITS + SFT_TRIANGLE_BITS)
    #error "SFT_INSTANCE_BITS + SFT_MESHLET_BITS + SFT_TRIANGLE_BITS must equal 31"
#endif
#define SFT_MAX_INSTANCES_PER_FRAME (1 << SFT_INSTANCE_BITS)
#define SFT_VERTS_PER_MESHLET 64
#define SFT_MESHLETS_PER_JOB 8
#define SFT_TILE_SHIFT 6
#define SFT_TILE_SIZE (1 << SFT_TILE_SHIFT)
#define SFT_TILE_PIXELS (SFT_TILE_SIZE * SFT_TILE_SIZE)

/* ---- public value types (sofren2.c:871-899, 1183-1215) ---- */
typedef struct sftvec  { f32 x, y, z; } sftvec;
typedef struct sftquat { f32 x, y, z, w; } sftquat;
typedef union  sftmat  { sftquat rows[4]; f32 m[4][4]; } sftmat;

#ifdef __cplusplus
#define SFT_V(a, b, c) (sftvec{ (a), (b), (c) })
#define SFT_MAT_IDENTITY (sftmat{ { {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1} } })
#else
#define SFT_V(a, b, c) (sftvec){ .x = (a), .y = (b), .z = (c) }
#define SFT_MAT_IDENTITY (sftmat){ .m = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}} }
#endif

typedef struct sft_ray_hit {
    u8 hit;
    f32 distance;
    sftvec pos;
    sftvec normal;
} sft_ray_hit;

typedef struct sft_arena {
    u8* base;
    u64 max_size;
    u64 committed_size;
    u64 offset;
    volatile i32 commit_lock;
} sft_arena;

/* ---- model data (layouts as sofren2.c:906-1181) ---- */
struct sft_aabb { sftvec min, max; };

struct sft_meshlet {
    SFT_ALIGNAS(32) f32 x[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 y[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 z[SFT_VERTS_PER_MESHLET];
};

struct sft_meshlet_attribs {
    SFT_ALIGNAS(32) f32 nx[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 ny[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 nz[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 u[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 v[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 lu[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) f32 lv[SFT_VERTS_PER_MESHLET];
};

struct sft_meshlet_skin {
    SFT_ALIGNAS(32) u16 j0[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u16 j1[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u16 j2[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u16 j3[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u8 w0[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u8 w1[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u8 w2[SFT_VERTS_PER_MESHLET];
    SFT_ALIGNAS(32) u8 w3[SFT_VERTS_PER_MESHLET];
};

/* 16 batches of 8 triangles; each batch is 8x v0, 8x v1, 8x v2, 8x pad. */
struct sft_meshlet_inds {
    SFT_ALIGNAS(32) u8 inds[512];
};

/* 8-wide BVH node, SoA child boxes.  child_indices: bit 31 set = leaf payload
   (meshlet<<7 | triangle), clear = child node index, 0xFFFFFFFF = unused. */
struct sft_bvh_node {
    SFT_ALIGNAS(32) f32 min_x[8];
    SFT_ALIGNAS(32) f32 min_y[8];
    SFT_ALIGNAS(32) f32 min_z[8];
    SFT_ALIGNAS(32) f32 max_x[8];
    SFT_ALIGNAS(32) f32 max_y[8];
    SFT_ALIGNAS(32) f32 max_z[8];
    u32 child_indices[8];
    u8 child_count;
    u8 is_leaf_group;
};

struct sft_shadow_proxy {
    sftvec posa, posb;
    f32 radius;
    u32 bone_ind;
};

struct sft_bone {
    sftmat inverse_bind_matrix;
    sftmat local_bind_matrix;
    sftvec bind_pos;
    i32 parent_ind;
    sftvec base_t;
    sftquat base_r;
    sftvec base_s;
};

struct sft_skeleton {
    struct sft_bone* bones;
    i32 bone_count;
};

struct sft_anim_sampler {
    f32* inputs;
    f32* outputs;
    i32 count;
    u8 is_cubic;
    u8 is_step;
};

enum sft_anim_path {
    SFT_ANIM_PATH_TRANSLATION,
    SFT_ANIM_PATH_ROTATION,
    SFT_ANIM_PATH_SCALE
};

struct sft_anim_channel {
    i32 sampler_ind;
    i32 node_ind;
    enum sft_anim_path path;
};

struct sft_animation {
    struct sft_anim_sampler* samplers;
    struct sft_anim_channel* channels;
    i32 sampler_count;
    i32 channel_count;
    f32 duration;
};

typedef struct sft_mesh {
    struct sft_meshlet* meshlets;
    struct sft_meshlet_attribs* meshlet_attribs;
    struct sft_meshlet_skin* meshlet_skins;   /* NULL for static meshes */
    u32 meshlet_count;
    u32 model_meshlet_offset;
    struct sft_meshlet_inds* inds;
    f32* bounds_center_x;
    f32* bounds_center_y;
    f32* bounds_center_z;
    f32* bounds_radius;
    f32* bounds_min_x;
    f32* bounds_max_x;
    f32* bounds_min_y;
    f32* bounds_max_y;
    f32* bounds_min_z;
    f32* bounds_max_z;
    f32* cone_axis_x;
    f32* cone_axis_y;
    f32* cone_axis_z;
    f32* cone_cutoff;
    i32* culling_bone_inds;
    sft_mat_handle material;
    u32 lightmap_page;                        /* 0xFFFFFFFF = no lightmap */
} sft_mesh;

typedef struct sft_model {
    struct sft_mesh* meshes;
    u32 mesh_count;
    u32 total_meshlets;
    u16* meshlet_mesh_inds;
    struct sft_bvh_node* blas_nodes;
    u32 blas_node_count, blas_root_ind;
    sft_tex_handle* lightmaps;
    u32 lightmap_count;
    struct sft_shadow_proxy* shadow_proxies;
    u32 shadow_proxy_count;
    struct sft_skeleton skeleton;
    struct sft_animation* animations;
    i32 anim_count;
} sft_model;

typedef struct sft_pose {
    const struct sft_model* model;
    sftmat* global_matrices;
    sftmat* posed_bone_matrices;
} sft_pose;

typedef struct sft_dmesh {
    enum sft_dmesh_type type;
    sft_mat_handle material;
    struct sft_meshlet* meshlets;
    struct sft_meshlet_attribs* attribs;
    struct sft_meshlet_inds* inds;
    f32* bounds_center_x;
    f32* bounds_center_y;
    f32* bounds_center_z;
    f32* bounds_radius;
    f32* bounds_min_x;
    f32* bounds_min_y;
    f32* bounds_min_z;
    f32* bounds_max_x;
    f32* bounds_max_y;
    f32* bounds_max_z;
    f32* cone_axis_x;
    f32* cone_axis_y;
    f32* cone_axis_z;
    f32* cone_cutoff;
    u32 max_meshlets;
    u32 curr_meshlet;
    u32 curr_vert;
    u32 max_verts_per_chunk;
} sft_dmesh;

typedef struct sft_dmesh_qb {
    f32 *x, *y, *z;
    f32 *nx, *ny, *nz;
    f32 *u, *v;
    u32 vert_offset;
    u32 allocated_count;
} sft_dmesh_qb;

typedef struct sft_instance {
    const struct sft_model* model;
    sftmat transform;
    u32 ss_meshlet_offset;
    sftmat* posed_bone_matrices;              /* NULL if no skeleton */
} sft_instance;

typedef struct sft_directional_light {
    sftvec dir;
    sftvec color;
} sft_directional_light;

typedef struct sft_point_light {
    sftvec pos;
    sftvec color;
    f32 radius;
    f32 inv_radius_sqr;
} sft_point_light;

/* ---- core functions (sofren2.c:808-813) ---- */
void sft_init(i32 width, i32 height, i32 requested_threads);  /* threads: accepted, unused by the GPU backend */
void sft_release(void);
void sft_present(void);
void sft_frame_begin(void);
void sft_frame_end(void);
void sft_resize(i32 width, i32 height);

/* ---- misc (sofren2.c:816-821) ---- */
void sft_set_camera(sftvec cam_pos, sftmat view, sftmat proj);
void sft_set_debug_mode(enum sft_debug_mode mode);
const char* sft_get_debug_name(enum sft_debug_mode mode);
void sft_set_clear(u8 enabled);
void sft_set_clear_color(u32 color);
void sft_set_color_buffer(u32* buf, i32 pitch_in_bytes);

/* ---- drawing (sofren2.c:824-838) ---- */
void sft_draw_instance(const sft_instance* instance);
void sft_draw_cube(sft_dmesh* dm);
void sft_draw_cube_inv(sft_dmesh* dm);
void sft_draw_sphere(sft_dmesh* dm, i32 rings, i32 slices);
void sft_draw_capsule(sft_dmesh* dm, sftvec start_pos, sftvec end_pos, f32 radius, i32 slices, i32 rings);
void sft_draw_cylinder(sft_dmesh* dm, sftvec start_pos, sftvec end_pos, f32 start_radius, f32 end_radius, i32 sides);
sft_dmesh* sft_dmesh_begin(enum sft_dmesh_type type, sft_mat_handle material, u32 max_prims);
void sft_dmesh_push_quad(sft_dmesh* dm, sftvec p0, sftvec p1, sftvec p2, sftvec p3, f32 u0, f32 v0, f32 u1, f32 v1, sftvec normal);
void sft_dmesh_push_quad_4n(sft_dmesh* dm, sftvec p0, sftvec p1, sftvec p2, sftvec p3, f32 u0, f32 v0, f32 u1, f32 v1, sftvec n0, sftvec n1, sftvec n2, sftvec n3);
void sft_dmesh_push_tri(sft_dmesh* dm, sftvec p0, sftvec p1, sftvec p2, f32 u0, f32 v0, f32 u1, f32 v1, f32 u2, f32 v2, sftvec normal);
u8   sft_dmesh_alloc_quads(sft_dmesh* dm, u32 count, sft_dmesh_qb* out_batch);
void sft_dmesh_end(sft_dmesh* dm, const sftmat* opt_transform);
/* skeletal poses (sofren2.c:8940-9107): NULL for a model without a skeleton;
   sft_pose_update samples animation anim_id at `time` into
   posed_bone_matrices, the array an sft_instance passes for skinning */
sft_pose* sft_pose_create(const sft_model* model, sft_arena* arena);
void sft_pose_update(sft_pose* pose, i32 anim_id, f32 time);
/* closest hit of a ray with an instance's model through its BLAS
   (sofren2.c:3733-3838; public in the reference without a prototype) */
sft_ray_hit sft_raycast_instance(const sft_instance* inst, sftvec origin, sftvec dir, f32 max_dist);

/* ---- lights (sofren2.c:841-846) ---- */
void sft_set_ambient_light(f32 r, f32 g, f32 b);
void sft_draw_directional_light(sftvec dir, f32 r, f32 g, f32 b, f32 intensity);
void sft_draw_point_light(sftvec pos, f32 r, f32 g, f32 b, f32 intensity, f32 radius, u32 light_id);
void sft_set_shadows_enabled(u8 enabled);
void sft_set_lightmap_brightness(f32 b);

/* ---- textures (sofren2.c:849-858); loading from image files needs stb_image
   in the reference and is not provided ---- */
sft_tex_handle sft_texture_create(i32 width, i32 height, const u32* pixels);
sft_tex_handle sft_texture_create_solid(u32 color);
i32 sft_texture_get_width(sft_tex_handle tex);
i32 sft_texture_get_height(sft_tex_handle tex);
u32* sft_texture_get_linear_pixels(sft_tex_handle tex);   /* caller frees */
void sft_texture_release(sft_tex_handle tex);

/* ---- materials (sofren2.c:861-863) ---- */
sft_mat_handle sft_material_create(sft_tex_handle tex, u32 base_color, f32 shininess);
sft_mat_handle sft_material_clone(sft_mat_handle base_mat);
void sft_material_release(sft_mat_handle mat);

/* ---- arenas (sofren2.c:866-869) ---- */
void sft_arena_init(sft_arena* arena, u64 reserve_size);
void* sft_arena_alloc(sft_arena* arena, u64 size);
void sft_arena_commit_if_needed(sft_arena* arena, u64 required_bytes);
void sft_arena_frame_begin(sft_arena* arena);

/* ---- models (sofren2.c:872-876); glTF import needs cgltf and is not provided ---- */
sft_model* sft_model_load_qsf(const char* path, sft_arena* arena);

/* ================================================================
   Scalar math (sofren2.c:745-782, definitions 1503-1920).  Header-only, as in
   the reference.  Operation order follows the reference so results are
   bit-identical when compiled with -ffp-contract=off.
   ================================================================ */
#ifndef SFT_INLINE
#if defined(__HIPCC__)
/* the GPU kernels evaluate the same scalar math (shadow-ray packet set-up) */
#define SFT_INLINE static inline __host__ __device__
#else
#define SFT_INLINE static inline
#endif
#endif

SFT_INLINE sftvec sft_v_add(sftvec a, sftvec b) { sftvec r; r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; return r; }
SFT_INLINE sftvec sft_v_sub(sftvec a, sftvec b) { sftvec r; r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z; return r; }
SFT_INLINE sftvec sft_v_mul(sftvec a, f32 b) { sftvec r; r.x = a.x * b; r.y = a.y * b; r.z = a.z * b; return r; }
SFT_INLINE sftvec sft_v_div(sftvec a, f32 b) { const f32 ib = 1.f / b; return sft_v_mul(a, ib); }
SFT_INLINE f32 sft_v_dot(sftvec a, sftvec b) { return a.x*b.x + a.y*b.y + a.z*b.z; }
SFT_INLINE f32 sft_v_length2(sftvec v) { return sft_v_dot(v, v); }
SFT_INLINE f32 sft_v_length(sftvec v) { return sqrtf(sft_v_dot(v, v)); }
SFT_INLINE sftvec sft_v_cross(sftvec a, sftvec b) {
    sftvec r;
    r.x = a.y*b.z - a.z*b.y;
    r.y = a.z*b.x - a.x*b.z;
    r.z = a.x*b.y - a.y*b.x;
    return r;
}
SFT_INLINE sftvec sft_v_norm(sftvec v) {
    const f32 l = sft_v_length(v);
    if (l > SFT_EPSILON) return sft_v_mul(v, 1.f / l);
    sftvec z = { 0.f, 0.f, 0.f };
    return z;
}
SFT_INLINE sftvec sft_v_normf(f32 x, f32 y, f32 z) { sftvec v = { x, y, z }; return sft_v_norm(v); }
SFT_INLINE sftvec sft_v_face_normal(sftvec a, sftvec b, sftvec c) {
    return sft_v_norm(sft_v_cross(sft_v_sub(b, a), sft_v_sub(c, a)));
}
SFT_INLINE sftvec sft_v_lerp(sftvec a, sftvec b, f32 t) { return sft_v_add(a, sft_v_mul(sft_v_sub(b, a), t)); }
/* cross product with the cardinal axis of the smallest component (sofren2.c:1565) */
SFT_INLINE sftvec sft_v_perpendicular(sftvec v) {
    f32 mn = fabsf(v.x);
    sftvec ax = { 1.f, 0.f, 0.f };
    if (fabsf(v.y) < mn) { mn = fabsf(v.y); ax.x = 0.f; ax.y = 1.f; ax.z = 0.f; }
    if (fabsf(v.z) < mn) { ax.x = 0.f; ax.y = 0.f; ax.z = 1.f; }
    sftvec r;
    r.x = v.y*ax.z - v.z*ax.y;
    r.y = v.z*ax.x - v.x*ax.z;
    r.z = v.x*ax.y - v.y*ax.x;
    return r;
}
SFT_INLINE sftvec sft_v_rotate_by_quat(sftvec v, sftquat q) {
    sftvec r;
    r.x = v.x*(q.x*q.x + q.w*q.w - q.y*q.y - q.z*q.z) + v.y*(2*q.x*q.y - 2*q.w*q.z) + v.z*(2*q.x*q.z + 2*q.w*q.y);
    r.y = v.x*(2*q.w*q.z + 2*q.x*q.y) + v.y*(q.w*q.w - q.x*q.x + q.y*q.y - q.z*q.z) + v.z*(-2*q.w*q.x + 2*q.y*q.z);
    r.z = v.x*(-2*q.w*q.y + 2*q.x*q.z) + v.y*(2*q.w*q.x + 2*q.y*q.z) + v.z*(q.w*q.w - q.x*q.x - q.y*q.y + q.z*q.z);
    return r;
}
SFT_INLINE sftmat sft_m_identity(void) {
    sftmat r; memset(&r, 0, sizeof(r));
    r.m[0][0] = 1.f; r.m[1][1] = 1.f; r.m[2][2] = 1.f; r.m[3][3] = 1.f;
    return r;
}
SFT_INLINE sftmat sft_m_rot_x(f32 a) { sftmat r = sft_m_identity(); const f32 c = cosf(a), s = sinf(a); r.m[1][1] = c; r.m[1][2] = s; r.m[2][1] = -s; r.m[2][2] = c; return r; }
SFT_INLINE sftmat sft_m_rot_y(f32 a) { sftmat r = sft_m_identity(); const f32 c = cosf(a), s = sinf(a); r.m[0][0] = c; r.m[0][2] = s; r.m[2][0] = -s; r.m[2][2] = c; return r; }
SFT_INLINE sftmat sft_m_rot_z(f32 a) { sftmat r = sft_m_identity(); const f32 c = cosf(a), s = sinf(a); r.m[0][0] = c; r.m[0][1] = s; r.m[1][0] = -s; r.m[1][1] = c; return r; }
SFT_INLINE sftmat sft_m_translate(f32 x, f32 y, f32 z) { sftmat r = sft_m_identity(); r.m[3][0] = x; r.m[3][1] = y; r.m[3][2] = z; return r; }
SFT_INLINE sftmat sft_m_scale(f32 x, f32 y, f32 z) { sftmat r = sft_m_identity(); r.m[0][0] = x; r.m[1][1] = y; r.m[2][2] = z; return r; }
/* D3D-style projection, clip z in [0, w] (sofren2.c:1646-1656) */
SFT_INLINE sftmat sft_m_proj(f32 fov_deg, f32 aspect, f32 n, f32 f) {
    const f32 fov = 1.f / tanf(fov_deg * 0.5f / 180.f * SFT_PI);
    sftmat r; memset(&r, 0, sizeof(r));
    r.m[0][0] = fov / aspect;
    r.m[1][1] = fov;
    r.m[2][2] = f / (f - n);
    r.m[3][2] = (-f * n) / (f - n);
    r.m[2][3] = 1.f;
    r.m[3][3] = 0.f;
    return r;
}
SFT_INLINE sftmat sft_m_ortho(f32 l, f32 rt, f32 b, f32 t, f32 n, f32 f) {
    sftmat r; memset(&r, 0, sizeof(r));
    r.m[0][0] = 2.f / (rt - l);
    r.m[1][1] = 2.f / (t - b);
    r.m[2][2] = 1.f / (f - n);
    r.m[3][0] = -(rt + l) / (rt - l);
    r.m[3][1] = -(t + b) / (t - b);
    r.m[3][2] = -n / (f - n);
    r.m[3][3] = 1.f;
    return r;
}
SFT_INLINE sftmat sft_m_mul(sftmat a, sftmat b) {
    sftmat r;
    for (i32 i = 0; i < 4; i += 1) {
        const f32 a0 = a.m[i][0], a1 = a.m[i][1], a2 = a.m[i][2], a3 = a.m[i][3];
        for (i32 j = 0; j < 4; j += 1) {
            r.m[i][j] = a0 * b.m[0][j] + a1 * b.m[1][j] + a2 * b.m[2][j] + a3 * b.m[3][j];
        }
    }
    return r;
}
SFT_INLINE sftvec sft_m_mul_vec(sftmat m, sftvec v) {
    sftvec r;
    r.x = v.x * m.m[0][0] + v.y * m.m[1][0] + v.z * m.m[2][0];
    r.y = v.x * m.m[0][1] + v.y * m.m[1][1] + v.z * m.m[2][1];
    r.z = v.x * m.m[0][2] + v.y * m.m[1][2] + v.z * m.m[2][2];
    return r;
}
SFT_INLINE sftvec sft_m_mul_quat(sftmat m, sftquat v) {
    sftvec r;
    r.x = v.x * m.m[0][0] + v.y * m.m[1][0] + v.z * m.m[2][0] + v.w * m.m[3][0];
    r.y = v.x * m.m[0][1] + v.y * m.m[1][1] + v.z * m.m[2][1] + v.w * m.m[3][1];
    r.z = v.x * m.m[0][2] + v.y * m.m[1][2] + v.z * m.m[2][2] + v.w * m.m[3][2];
    return r;
}
/* inverse of a rigid transform (sofren2.c:1681) */
SFT_INLINE sftmat sft_m_qinv(sftmat m) {
    sftmat r;
    for (i32 i = 0; i < 3; i += 1) {
        for (i32 j = 0; j < 3; j += 1) r.m[i][j] = m.m[j][i];
        r.m[i][3] = 0.f;
    }
    for (i32 j = 0; j < 3; j += 1) {
        r.m[3][j] = -(m.m[3][0] * r.m[0][j] + m.m[3][1] * r.m[1][j] + m.m[3][2] * r.m[2][j]);
    }
    r.m[3][3] = 1.f;
    return r;
}
SFT_INLINE sftmat sft_m_model_look_at(sftvec pos, sftvec target, sftvec up) {
    const sftvec fwd = sft_v_norm(sft_v_sub(target, pos));
    up = sft_v_norm(sft_v_sub(up, sft_v_mul(fwd, sft_v_dot(up, fwd))));
    const sftvec right = sft_v_cross(up, fwd);
    sftmat r;
    r.m[0][0] = right.x; r.m[0][1] = right.y; r.m[0][2] = right.z; r.m[0][3] = 0.f;
    r.m[1][0] = up.x;    r.m[1][1] = up.y;    r.m[1][2] = up.z;    r.m[1][3] = 0.f;
    r.m[2][0] = fwd.x;   r.m[2][1] = fwd.y;   r.m[2][2] = fwd.z;   r.m[2][3] = 0.f;
    r.m[3][0] = pos.x;   r.m[3][1] = pos.y;   r.m[3][2] = pos.z;   r.m[3][3] = 1.f;
    return r;
}
/* camera view matrix: transposed basis plus inverse translation (sofren2.c:1737-1770) */
SFT_INLINE sftmat sft_m_view_look_at(sftvec pos, sftvec target, sftvec up) {
    const sftvec fwd = sft_v_norm(sft_v_sub(target, pos));
    if (fabsf(fwd.x) < SFT_EPSILON && fabsf(fwd.z) < SFT_EPSILON) {
        up.x = 0.f; up.y = 0.f; up.z = -1.f;
    }
    up = sft_v_norm(sft_v_sub(up, sft_v_mul(fwd, sft_v_dot(up, fwd))));
    const sftvec right = sft_v_cross(up, fwd);
    sftmat r;
    r.m[0][0] = right.x; r.m[0][1] = up.x; r.m[0][2] = fwd.x; r.m[0][3] = 0.f;
    r.m[1][0] = right.y; r.m[1][1] = up.y; r.m[1][2] = fwd.y; r.m[1][3] = 0.f;
    r.m[2][0] = right.z; r.m[2][1] = up.z; r.m[2][2] = fwd.z; r.m[2][3] = 0.f;
    r.m[3][0] = -sft_v_dot(pos, right);
    r.m[3][1] = -sft_v_dot(pos, up);
    r.m[3][2] = -sft_v_dot(pos, fwd);
    r.m[3][3] = 1.f;
    return r;
}
SFT_INLINE sftmat sft_m_from_quat(sftquat q) {
    sftmat m = sft_m_identity();
    const f32 x = q.x, y = q.y, z = q.z, w = q.w;
    const f32 x2 = x + x, y2 = y + y, z2 = z + z;
    const f32 xx = x * x2, xy = x * y2, xz = x * z2;
    const f32 yy = y * y2, yz = y * z2, zz = z * z2;
    const f32 wx = w * x2, wy = w * y2, wz = w * z2;
    m.m[0][0] = 1.f - (yy + zz); m.m[0][1] = xy + wz;         m.m[0][2] = xz - wy;         m.m[0][3] = 0.f;
    m.m[1][0] = xy - wz;         m.m[1][1] = 1.f - (xx + zz); m.m[1][2] = yz + wx;         m.m[1][3] = 0.f;
    m.m[2][0] = xz + wy;         m.m[2][1] = yz - wx;         m.m[2][2] = 1.f - (xx + yy); m.m[2][3] = 0.f;
    m.m[3][3] = 1.f;
    return m;
}
SFT_INLINE sftquat sft_q_mul(sftquat a, sftquat b) {
    sftquat r;
    r.x =  a.x * b.w + a.y * b.z - a.z * b.y + a.w * b.x;
    r.y = -a.x * b.z + a.y * b.w + a.z * b.x + a.w * b.y;
    r.z =  a.x * b.y - a.y * b.x + a.z * b.w + a.w * b.z;
    r.w = -a.x * b.x - a.y * b.y - a.z * b.z + a.w * b.w;
    return r;
}
SFT_INLINE sftquat sft_q_norm(sftquat q) {
    const f32 len2 = q.x*q.x + q.y*q.y + q.z*q.z + q.w*q.w;
    if (0.f != len2) {
        const f32 il = 1.f / sqrtf(len2);
        q.x *= il; q.y *= il; q.z *= il; q.w *= il;
    }
    return q;
}
SFT_INLINE sftquat sft_q_invert(sftquat q) {
    sftquat r = q;
    const f32 len2 = q.x*q.x + q.y*q.y + q.z*q.z + q.w*q.w;
    if (0.f != len2) {
        const f32 il = 1.f / len2;
        r.x *= -il; r.y *= -il; r.z *= -il; r.w *= il;
    }
    return r;
}
SFT_INLINE sftquat sft_q_slerp(sftquat a, sftquat b, f32 t) {
    f32 d = a.x*b.x + a.y*b.y + a.z*b.z + a.w*b.w;
    if (d < 0.f) { b.x = -b.x; b.y = -b.y; b.z = -b.z; b.w = -b.w; d = -d; }
    if (d > 0.9995f) {
        sftquat r;
        r.x = SFT_LERPF(a.x, b.x, t); r.y = SFT_LERPF(a.y, b.y, t);
        r.z = SFT_LERPF(a.z, b.z, t); r.w = SFT_LERPF(a.w, b.w, t);
        return sft_q_norm(r);
    }
    const f32 th0 = acosf(d);
    const f32 th = th0 * t;
    const f32 s = sinf(th);
    const f32 is0 = 1.f / sinf(th0);
    const f32 s0 = cosf(th) - d * s * is0;
    const f32 s1 = s * is0;
    sftquat r;
    r.x = a.x * s0 + b.x * s1; r.y = a.y * s0 + b.y * s1;
    r.z = a.z * s0 + b.z * s1; r.w = a.w * s0 + b.w * s1;
    return sft_q_norm(r);
}

#ifdef __cplusplus
} /* extern "C" */
#endif

#if defined(__cplusplus)
static_assert(sizeof(sftvec) == 12, "sftvec layout");
static_assert(sizeof(sftmat) == 64, "sftmat layout");
static_assert(sizeof(sft_instance) == 88, "sft_instance layout");
static_assert(sizeof(sft_arena) == 40, "sft_arena layout");
static_assert(sizeof(struct sft_meshlet) == 768, "meshlet layout");
static_assert(sizeof(struct sft_meshlet_attribs) == 1792, "attribs layout");
static_assert(sizeof(struct sft_meshlet_skin) == 768, "skin layout");
static_assert(sizeof(struct sft_meshlet_inds) == 512, "inds layout");
static_assert(sizeof(struct sft_bvh_node) == 256, "bvh node layout");
static_assert(sizeof(struct sft_shadow_proxy) == 32, "proxy layout");
static_assert(sizeof(sft_point_light) == 32, "point light layout");
#else
_Static_assert(sizeof(sftvec) == 12, "sftvec layout");
_Static_assert(sizeof(sftmat) == 64, "sftmat layout");
_Static_assert(sizeof(sft_instance) == 88, "sft_instance layout");
_Static_assert(sizeof(sft_arena) == 40, "sft_arena layout");
_Static_assert(sizeof(struct sft_meshlet) == 768, "meshlet layout");
_Static_assert(sizeof(struct sft_meshlet_attribs) == 1792, "attribs layout");
_Static_assert(sizeof(struct sft_meshlet_skin) == 768, "skin layout");
_Static_assert(sizeof(struct sft_meshlet_inds) == 512, "inds layout");
_Static_assert(sizeof(struct sft_bvh_node) == 256, "bvh node layout");
_Static_assert(sizeof(struct sft_shadow_proxy) == 32, "proxy layout");
_Static_assert(sizeof(sft_point_light) == 32, "point light layout");
#endif

#endif /* SFT_H */
