/*
 * scenes.h — seeded synthetic workloads driven through the public sofren2 API.
 *
 * The same source is compiled against any implementation of the API (the HIP
 * backend in this package, the CPU reference build used as a checker), so
 * every backend renders byte-identical scene descriptions.  All randomness
 * comes from a PCG32 stream seeded per scene.
 *
 * Workloads:
 *   SFTX_SCENE_PLANE_SPHERES (0): 640x360, ground plane of 20x20 unit quads and
 *       16 dmesh spheres (16 rings x 32 slices), radius U[0.5,1.5], centres
 *       U[-8,8]^2 at y = radius, flat per-object colour U[0,1]^3, one point
 *       light at (0,10,0), ambient 0.35, no textures, no shadows.
 *   SFTX_SCENE_BOX_ROOM (1): 1920x1080 textured interior, ~260k triangles in
 *       4096 axis-aligned boxes (log-uniform sizes in [0.1,4]) on a 16^3
 *       lattice inside a room, 16 value-noise 256x256 textures with full mip
 *       chains, a baked lightmap page (per-vertex light U[0.2,1]), 8 point
 *       lights without shadows, a 120-frame camera path.  The scene is written
 *       to a QSF2 file and loaded through sft_model_load_qsf.
 *   SFTX_SCENE_CROWD_SHADOWS (2): the box room (with a triangle BLAS) plus 32
 *       skinned humanoids (6144 triangles, 10-joint chain, 4 weights per
 *       vertex), 10 capsule shadow proxies each, 16 shadow-casting point
 *       lights, shadows enabled.
 *   SFTX_SCENE_SERVED (3): variants of workload 2 drawn from the seed for the
 *       request server: 50k-1M triangles (a 9^3 to 25^3 box lattice), 0-64
 *       skinned humanoids (64 and the room need the 7/17 ID split), 0-16
 *       shadow-casting lights.
 *   SFTX_SCENE_MAP (4): the QSF2 map file named by $SFTX_QSF (a real,
 *       lightmapped scene), lit by its own game-data lights, camera turning
 *       at (0, 2, 0).
 *   SFTX_SCENE_FEATURES (5): 640x360 dmesh scene for the shading branches the
 *       others miss: non-power-of-two textures, unlit materials, directional
 *       lights, more than 16 point lights on a tile, 64 draws (frames
 *       500-999: six more, which must be dropped), no clearing (the caller's
 *       buffer shows through and blends under MSAA).
 *   SFTX_SCENE_INSTANCED (6): the crowd scene on a 6^3 lattice, 8 humanoids,
 *       12 shadow-casting lights, the room model drawn 40 more times under
 *       rigid transforms (a TLAS over many instances with a BLAS, humanoids
 *       without one drawn first).
 */
#ifndef SFTX_SCENES_H
#define SFTX_SCENES_H

#include "sofren2.h"

#ifdef __cplusplus
extern "C" {
#endif

enum sftx_scene_kind {
    SFTX_SCENE_PLANE_SPHERES = 0,
    SFTX_SCENE_BOX_ROOM = 1,
    SFTX_SCENE_CROWD_SHADOWS = 2,
    SFTX_SCENE_SERVED = 3,
    SFTX_SCENE_MAP = 4,
    SFTX_SCENE_FEATURES = 5,
    SFTX_SCENE_INSTANCED = 6,
    SFTX_SCENE_COUNT
};

/* Creates the scene's resources; call after sft_init.  Returns 0 on success. */
i32 sftx_scene_create(i32 kind, u32 seed, i32 width, i32 height);

/* Records one frame of the scene (buffer, frame_begin, camera, lights, draws)
   without presenting it. */
void sftx_scene_record(i32 frame, u32* color, i32 pitch_bytes);

/* Records one frame of the scene (camera, lights, draws) and presents it into
   the caller's ARGB8888 buffer: frame_begin, camera, lights, draws,
   sft_present, frame_end. */
void sftx_scene_render(i32 frame, u32* color, i32 pitch_bytes);

/* Several scenes at once (handles 0..63): sftx_scene_new builds one and
   makes it current; sftx_scene_draw records a frame of a scene's camera path
   (camera for a width x height view, lights, draws) into the frame or view
   being recorded, without sft_frame_begin; sftx_scene_use selects the scene
   the single-scene calls work on.  sftx_scene_create frees every scene and
   builds one. */
i32 sftx_scene_new(i32 kind, u32 seed, i32 width, i32 height);
i32 sftx_scene_use(i32 handle);
void sftx_scene_draw(i32 handle, i32 frame, i32 width, i32 height);
void sftx_scene_free(i32 handle);
/* the camera sftx_scene_draw would set for that frame and view size; frame
   numbers 1000 * (p + 1) + c pose the scene at frame p with camera frame c */
void sftx_scene_camera(i32 handle, i32 frame, i32 width, i32 height, sftvec* pos, sftmat* view, sftmat* proj);

/* skinned humanoids the scene draws (each one instance, beside the room) */
i32 sftx_scene_characters(i32 handle);
/* shadow-casting point lights of the crowd kinds (0 for the others) */
i32 sftx_scene_shadow_lights(i32 handle);

/* Frame statistics of the scene description (not of the renderer). */
u32 sftx_scene_triangle_count(void);
u32 sftx_scene_meshlet_count(void);

void sftx_scene_destroy(void);

/* Host-side query checks, word-comparable between builds of the API:
   sftx_pose_check samples a seeded 12-bone skeleton's two animations at the
   given times (sft_pose_create / sft_pose_update) and writes 2 * 12 * 16
   floats per time (global then posed matrices); sftx_raycast_check casts
   seeded rays at the current scene's room model with sft_raycast_instance
   and writes 8 floats per ray (hit, distance, position, normal). */
i32 sftx_pose_check(u32 seed, i32 n_times, const f32* times, f32* out);
i32 sftx_raycast_check(u32 seed, i32 n_rays, f32* out);

/* Offline builder check, byte-comparable between builds of the API: builds
   meshlets (kinds 0-3: seeded triangle soups), the BLAS and shadow proxies
   (kinds 4-5: the crowd humanoid under two skeletons) and writes every
   output byte into out[0, cap).  Returns the bytes the output needs. */
i64 sftx_builder_check(u32 seed, i32 kind, u8* out, u64 cap);

/* Edits the current scene's room model in place (box room, crowd and served
   kinds), as an application editing a loaded model would: moves meshlets up
   by dy (vertex positions and bounds).  Mode 1 moves every meshlet, mode 2
   only those strictly between each mesh's first and last meshlet.  Returns
   the meshlets moved, -1 without a room model. */
i32 sftx_scene_edit_room(i32 mode, f32 dy);
const sft_model* sftx_scene_room_model(void);

#ifdef __cplusplus
}
#endif

#endif
