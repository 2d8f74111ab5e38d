/*
 * sofren2_ext.h — additive entry points of libsofren2_gpu.so.
 *
 * None of these exist in the reference API (sofren2.c:794-864); they sit
 * beside the drop-in boundary.  They give tests and tools read access to the
 * intermediate results of the last sft_present (visibility buffer, screen-space
 * meshlet cache, clipped-triangle records) and to per-frame counters and pass
 * timings.  The same capture functions are implemented by the CPU reference
 * build used as a checker, so a test can compare two backends field by field.
 */
#ifndef SFT_EXT_H
#define SFT_EXT_H

#include "sofren2.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Size in bytes of one screen-space meshlet cache entry (sofren2.c:1464-1477):
   sx, sy, inv_w, wx, wy, wz, wnx, wny, wnz, each f32[64]. */
#define SFTX_SS_MESHLET_BYTES 2304
/* Size in bytes of one clipped-triangle record (sofren2.c:1271-1280). */
#define SFTX_CLIP_RECORD_BYTES 164

/* Capture flags for sftx_set_capture. */
#define SFTX_CAPTURE_VISIBILITY 1u   /* per-sample 32-bit IDs and 1/w depth */
#define SFTX_CAPTURE_SS_CACHE   2u   /* screen-space meshlet cache + cull result */
#define SFTX_CAPTURE_CLIPS      4u   /* clipped-triangle records */
#define SFTX_CAPTURE_SHADOWS    8u   /* per-proxy light visibility counts */

/* Enables capture of intermediates on the following sft_present calls. */
void sftx_set_capture(u32 flags);

/* Tile grid of the last present: columns, rows (64 px tiles) and samples per
   pixel.  The captured visibility covers the whole grid (cols*64 x rows*64),
   including pixels past the screen edge inside border tiles. */
void sftx_cap_grid(i32* cols, i32* rows, i32* samples);

/* Per-sample reference 32-bit visibility IDs (inst<<25 | meshlet<<7 | tri under
   the default split, inst<<24 under 7/17 (sftx_set_id_split), or
   clip_slot | 1<<31, 0xFFFFFFFF = empty) and depth (interpolated 1/w, 0 =
   empty), laid out [sample][cols*64 * rows*64] row-major.  Either pointer may
   be NULL. */
void sftx_cap_visibility(u32* vid, f32* depth);

/* Number of screen-space meshlet slots of the last present (sum of
   total_meshlets over drawn instances), their contents (count *
   SFTX_SS_MESHLET_BYTES) and the per-slot cull result (1 = visible). */
u32 sftx_cap_ss_count(void);
void sftx_cap_ss_cache(void* out);
void sftx_cap_ss_visible(u8* out);

/* Clipped-triangle records of the last present, in slot order. */
u32 sftx_cap_clip_count(void);
void sftx_cap_clip_records(void* out);
/* World shadow proxies of the last frame and, per proxy, the 64 light
   visibility counts (0..3 of the 3 occlusion rays reach light l;
   struct sft_world_shadow_proxy::light_vis, sofren2.c:1432-1437). */
u32 sftx_cap_proxy_count(void);
void sftx_cap_proxy_visibility(u8* out);    /* proxy_count * 64 bytes */
u32 sftx_cap_tlas_count(void);               /* nodes of the frame's TLAS (SFTX_CAPTURE_SHADOWS) */
void sftx_cap_tlas(void* out);               /* tlas_count sft_bvh_node records as the occlusion pass read them */

/* Per-frame counters and pass timings of the last present. */
typedef struct sftx_frame_stats {
    u32 meshlets_tested;      /* N_t */
    u32 meshlets_visible;     /* N_v */
    u32 meshlets_skinned;     /* N_vs */
    u32 batch_payloads;       /* batch payloads emitted */
    u32 clip_triangles;       /* K: clipped-triangle slots used */
    u32 tile_payloads;        /* P: (payload, tile) pairs */
    u32 tiles;                /* 64x64 tiles */
    u32 samples;              /* S: width*height*samples_per_pixel */
    u32 pixels;               /* X */
    u32 shaded_pixels;        /* pixels covered by geometry */
    u32 shading_groups;       /* Q: (4x2 block, triangle) groups shaded */
    u32 tex_bilinear;         /* T_b: bilinear albedo samples */
    u32 tex_nearest;          /* T_n: nearest albedo samples */
    u32 lightmapped_pixels;   /* L_px: pixels with a lightmap sample */
    u32 occlusion_rays;       /* R: shadow-proxy occlusion rays traced */
    u32 occlusion_nodes;      /* BVH child boxes tested against ray-packet frusta */
    u32 occlusion_tri_tests;  /* ray-triangle tests */
    u32 occlusion_invalid;    /* packets with a degenerate frustum (corner behind the light) */
    u32 occlusion_overflow;   /* packets whose traversal stack ran out (result incomplete) */
    u32 shadow_proxies;       /* world capsule proxies of the frame */
    f32 ms_upload;            /* host->device copies of the frame's commands */
    f32 ms_geometry;          /* cull + skin + project + classify */
    f32 ms_binning;           /* clip + bin */
    f32 ms_raster_shade;      /* tile raster + resolve */
    f32 ms_download;          /* device->host colour copy */
    f32 ms_raster;            /* tile raster into the visibility buffer */
    f32 ms_shade;             /* deferred shading */
    f32 ms_total;             /* whole sft_present on the GPU stream */
    f32 ms_shadows;           /* proxy occlusion + per-light capsule lists */
    u32 tex_npot;             /* albedo samples of non-power-of-two textures (edge clamp) */
    u32 unlit_shadings;       /* shadings of materials with negative shininess */
    u32 dirlit_shadings;      /* shadings that evaluated directional lights */
    u32 light_cut_tiles;      /* tiles whose light list stopped at 16 with more lights in reach */
    u32 bg_blend_pixels;      /* MSAA pixels without clearing that blend the caller's colour */
    u32 instances_dropped;    /* draws beyond the instances of the visibility-ID split (64 by default) */
    u32 capsule_tests;        /* capsule ray tests: per shaded pixel and light reaching it, the light's cached capsules */
} sftx_frame_stats;

void sftx_get_stats(sftx_frame_stats* out);

/* Output target of sft_present.  SFTX_OUTPUT_HOST (default) copies the frame
   into the buffer given to sft_set_color_buffer, as the reference does, and
   sft_present returns when it is there.  SFTX_OUTPUT_DEVICE leaves the ARGB
   frame in HBM for consumers on the GPU; nothing crosses PCIe, and (without
   captures) sft_present returns once the frame is launched, so the next frame
   is recorded while it renders (four frames in flight).  sftx_device_color,
   sftx_get_stats and sftx_synchronize wait for the frames in flight; the
   frame's inputs are copied before sft_present returns, as always. */
#define SFTX_OUTPUT_HOST   0
#define SFTX_OUTPUT_DEVICE 1
void sftx_set_output(i32 mode);
const u32* sftx_device_color(void);     /* device pointer, width*height ARGB of the last frame */

/* Multisampling of the following frames: 1 (off), 2 or 4 samples per pixel
   at the reference's rotated-grid offsets (sofren2.c:186-196); the reference
   fixes this at compile time with SFT_USE_AA / SFT_AA_SAMPLES.  Other values
   are rejected with a message and leave the setting unchanged. */
void sftx_set_msaa(i32 samples);

/* Visibility-ID split of the following frames: instance_bits bits of
   instance index and 24 - instance_bits of model meshlet index above the 7
   triangle bits, so 1 << instance_bits instances per frame; draws beyond
   that are dropped.  Replaces the reference's compile-time SFT_INSTANCE_BITS /
   SFT_MESHLET_BITS (sofren2.c:154-165; default 6 / 18): an application built
   with -DSFT_INSTANCE_BITS=7 -DSFT_MESHLET_BITS=17 calls
   sftx_set_id_split(7) after sft_init.  Set it between frames; views of a
   batch and server templates keep the split in effect when they were
   recorded.  Values outside 1..16 are rejected with a message. */
void sftx_set_id_split(i32 instance_bits);
i32 sftx_get_id_split(void);

/* ---- batched views ---------------------------------------------------------
   Records several views with the ordinary calls and renders them with one
   pass of the kernels over all of them (each view is exactly what sft_present
   would have produced for it alone):

     sftx_batch_begin();
     for each view:
         sftx_view_begin(width, height, samples);
         sft_set_camera(...); sft_draw_*(...); lights ...
         sftx_view_end(argb, pitch_bytes, vis64_or_null);
     sftx_batch_submit();          blocking; outputs are written when it returns
     sftx_batch_end();             releases the views' transient storage

   vis64, when given, receives per sample (1/w bits << 32) | visibility ID
   over the view's 64-px tile grid: [samples][tile_rows*64][tile_cols*64].
   Views (and the sft_init / sft_resize display, and requests) are 1 to
   16384 pixels a side (256 tiles: the binning packs tile coordinates in 8
   bits); sftx_view_end returns -1 and drops a view outside that.
   The display size set by sftx_view_begin stays in effect after the batch.

   Pipelined form: sftx_batch_submit_async() in place of sftx_batch_submit
   launches the batch and returns a ticket (> 0; 0 if nothing was submitted)
   without waiting, so the next batch can be recorded (batch_end, batch_begin,
   views ...) while it runs.  Three batches can be in flight; submitting a
   fourth first completes the oldest.  sftx_batch_wait(ticket) completes batches in
   submission order up to that one and returns 0 (-1 for an unknown ticket);
   until then the batch's colour and vis64 buffers belong to the renderer
   (they are written as soon as the batch is done, by a thread of the
   renderer, and must not be read or freed before the wait).  A batch's
   kernels start after those of the batch submitted before it.
   sft_present, sftx_batch_submit, sftx_synchronize, sftx_server_start and
   sft_release complete any batch still in flight first. */
void sftx_batch_begin(void);
void sftx_view_begin(i32 width, i32 height, i32 samples);
i32 sftx_view_end(u32* color, i32 pitch_bytes, u64* vis64);
void sftx_batch_submit(void);
i32 sftx_batch_submit_async(void);
i32 sftx_batch_wait(i32 ticket);
void sftx_batch_end(void);

/* ---- request server ---------------------------------------------------------
   Renders requests against resident scene templates with continuous
   batching.  sftx_server_template keeps the frame recorded since
   sft_frame_begin (instances, poses, lights, materials, proxies; dmesh
   geometry is copied) as a template; a request renders that frame with its
   own camera, size and MSAA exactly as sft_present would.  While the server
   runs it owns the renderer: sft_present and sftx_batch_submit are refused,
   and textures, materials and models must not be created or released.
   Requests are admitted in arrival order into the next batch while the
   resident scene state, the framebuffer state of the batches in flight and
   the request's own fit in the HBM budget; up to three batches are in
   flight, so a batch is prepared and launched, and a finished one's frames
   copied to host memory, while another executes. */
typedef struct sftx_request {
    i32 scene;                  /* template handle */
    i32 width, height, samples; /* samples: 1, 2 or 4 */
    sftvec camera_pos;
    sftmat view, proj;
    u32* color;                 /* host ARGB output, pitch width * 4, or null */
    u64* vis64;                 /* host 64-bit visibility (see sftx_view_end), or null */
} sftx_request;

typedef struct sftx_server_stats {
    u64 submitted, completed;
    u64 batches, views_in_batches, max_batch;
    u64 admission_waits;        /* times waiting requests did not fit the budget */
    u64 resident_bytes;         /* scene, BVH and texture state in HBM */
    u64 pending, in_flight;
    /* time of the server thread: preparing and launching batches (host
       tables, TLAS builds, uploads), completing them (waiting for them and
       delivering their outputs), and the batches' GPU spans (first pass to
       last, not counting waits for the batch before) */
    f64 host_submit_ms, host_complete_ms, gpu_ms;
    /* views of templates with capsule shadows rendered from the template's
       traced shadow records (no occlusion rays, no frame TLAS) */
    u64 shadow_cache_views;
} sftx_server_stats;

i32 sftx_server_template(void);
i32 sftx_server_start(u64 hbm_budget_bytes, i32 max_views_per_batch);
u64 sftx_server_submit(const sftx_request* r);         /* request id, 0 if rejected */
i32 sftx_server_poll(u64 id);                          /* 1 done, 0 pending, -1 unknown */
i32 sftx_server_wait(u64 id, f64* latency_ms);         /* blocks; 0 on success; forgets the id */
void sftx_server_get_stats(sftx_server_stats* out);
void sftx_server_stop(void);                           /* finishes every admitted request */
void sftx_server_release(void);                        /* stops and drops the templates */

/* Frame launches replayed from a HIP graph (default on): the kernel sequence
   of a frame is captured once and replayed while launch sizes and buffers are
   unchanged.  sftx_graph_captures counts the captures made so far. */
void sftx_set_graphs(i32 enabled);

/* Colour delivered to host memory (sft_present with host output, host-output
   batches and served requests) streamed band by band (default on): the
   shading kernels write each pixel into pinned staging as well as HBM and
   raise a flag per tile row once its pixels are there; the host copies the
   finished rows into the caller's buffer while later rows shade, instead of
   downloading the frame after it.  An overflow re-run, or a band that is
   never flagged, falls back to the download after the frame.  The results
   are the same either way.  Off: the frame is downloaded after it.
   sftx_stream_stats counts the batches delivered streamed and those that
   fell back. */
void sftx_set_stream_download(i32 enabled);
void sftx_stream_stats(u64* streamed, u64* fallbacks);

/* The frame's top-level BVH over the drawn instances (shadow packets only):
   built on the device, one wave per view, by default; off, it is built on
   the host and uploaded with the frame's tables.  The nodes are the same
   bytes either way (the reference's builder, sofren2.c:2895-3062,
   3619-3703). */
void sftx_set_gpu_tlas(i32 enabled);

/* Autotuner.  No counterpart in the reference (its tiles and job sizes are
   compile-time constants, sofren2.c:124-130); it exists for batches and the
   request server, whose views span many resolutions.  sftx_batch_tune renders
   the recorded batch like sftx_batch_submit (blocking; outputs written), after
   timing candidate settings on it: per resolution class (views below 512x512
   pixels, below 1024x1024, larger) the payloads per visibility work item
   (candidates 64-512, any of 1..1024 settable; the item's triangles, 8 per
   payload, are staged in LDS 128 payloads at a time into one visibility
   plane; with MSAA an item covers a pair of sample planes and stages 32
   payloads at a time) and the row-part split threshold, then the persistent visibility
   grid (half, one or two occupancy-sized grids) and the occlusion
   traversal stack.  The
   chosen settings apply to every later sft_present, batch and server batch;
   no setting changes an output bit.  The bin tile stays 64 px: the
   reference's edge accumulation starts at the 64-px tile's clamped bounding
   box (sofren2.c:5148-5499), so another tile size would change depth and
   visibility bits.  A packet that fills the full occlusion stack (1024
   entries per wave; no BVH the builders make can) is counted in
   sftx_frame_stats. */
#define SFTX_RES_CLASSES 3
typedef struct sftx_tuning {
    u32 raster_chunk[SFTX_RES_CLASSES];   /* payloads per visibility work item, by class */
    /* by class: estimated sweep work above which a triangle is swept as 2,
       4 or 8 row parts by different lane groups; each part keeps the 64-px
       tile's accumulation origin (rows reached by add-only steps from it),
       so the split never changes a bit; 0xFFFFFFFF: never split */
    u32 raster_split[SFTX_RES_CLASSES];
    u32 raster_groups;                    /* persistent visibility workgroups */
    u32 raster_groups_resident;           /* workgroups resident at once on this device */
    u32 tile_px;                          /* bin / raster tile side (fixed: the reference's 64) */
    /* occlusion traversal stack entries per wave: 256, 512 or 1024; a batch
       whose traversal fills a smaller stack runs again with 1024 */
    u32 occlusion_stack;
    f32 ms_class[SFTX_RES_CLASSES];       /* tuned visibility-pass time of each class's views (0: none) */
    f32 ms_batch;                         /* tuned visibility-pass time of the whole batch */
    f32 ms_occlusion;                     /* tuned occlusion-pass time of the batch (0: no occlusion) */
} sftx_tuning;
i32 sftx_batch_tune(i32 repeats, sftx_tuning* report);   /* 0, or -1 without a recorded batch */
void sftx_get_tuning(sftx_tuning* out);
void sftx_set_tuning(const sftx_tuning* in);
u64 sftx_graph_captures(void);

/* Waits for all GPU work of the renderer. */
void sftx_synchronize(void);

/* Rendering without allocation.  No counterpart in the reference; it is for
   callers that cannot allocate device memory: code under HIP graph capture,
   tight serving loops, hosts that forbid allocation after start-up.
   sftx_present_noalloc renders the frame recorded since sft_frame_begin
   exactly as sft_present with host output does (the colour into the buffer
   of sft_set_color_buffer, at its pitch), but calls no hipMalloc, hipFree or
   pinned-memory allocation: every per-frame intermediate is carved from one
   device array of sftx_noalloc_scratch_bytes() (64 MiB) that the library
   holds on the GPU for the life of the process, tables are staged in
   ordinary host memory and kernels are launched eagerly.  It returns 0, or
   a negative SFTX_NOALLOC_* code with a message and nothing rendered:
     UNREADY        before sft_init or the first sft_present (which sets up
                    the renderer), inside a batch, or while the server runs;
     IN_FLIGHT      device-output frames or batches still in flight
                    (call sftx_synchronize first);
     NOT_RESIDENT   the frame draws models or textures not yet resident, or
                    more dmesh rows than any earlier frame (present the scene
                    once with sft_present);
     TOO_LARGE      the frame's intermediates exceed the scratch array (a
                    1920x1080 box-room frame without MSAA fits).
   Captures and 64-bit visibility are not written on this path.  Every call
   shares the array and the renderer's streams: calls must not overlap (from
   several threads, or with other rendering calls); each call clears what
   must start at zero itself, whatever the previous call left. */
#define SFTX_NOALLOC_UNREADY       (-1)
#define SFTX_NOALLOC_IN_FLIGHT     (-2)
#define SFTX_NOALLOC_NOT_RESIDENT  (-3)
#define SFTX_NOALLOC_TOO_LARGE     (-4)
i32 sftx_present_noalloc(void);
u64 sftx_noalloc_scratch_bytes(void);

/* Test hook: the capacities every frame set starts from for clipped
   triangles and tile-list entries (defaults 65536 and 1048576).  The kernels
   flag a batch whose intermediates exceed them and the batch is run again
   with larger ones (they then stay larger); small values make tests take
   that path.  sftx_overflow_reruns counts such re-runs.  Waits for every
   batch in flight first; not while the request server runs. */
void sftx_set_initial_caps(u32 clip_cap, u32 tile_list_cap);
u64 sftx_overflow_reruns(void);

/* Resident models and edits.  A model drawn by sft_draw_instance (not a
   dmesh) is copied to HBM at its first draw and drawn from there; the
   reference instead reads the model's arrays at every present
   (sofren2.c:8139-8289).  Every present compares each resident model it
   draws with what was uploaded: the model header, every mesh header except
   material and lightmap page (which are read at every present), and the
   first eight vertex positions of each mesh's first and last meshlet.  A
   difference uploads the model again, so replacing a model's arrays or
   rewriting all of its vertices is seen.  Other in-place edits (some
   vertices, bounds, attributes, indices, skin weights or BLAS nodes behind
   unchanged pointers) must be declared: sftx_model_changed drops the
   resident copy (after the batches in flight complete) and the next draw
   uploads the model again.  The rows of a dropped copy stay allocated until
   sft_release.  Returns 0, SFTX_MODEL_UNKNOWN for a model that is not
   resident (nothing to do), or SFTX_MODEL_BUSY while the request server
   runs (its templates hold copies of their models, which no edit reaches).
   sftx_model_edits counts the resident copies dropped for edits, declared
   or detected. */
#define SFTX_MODEL_UNKNOWN (-1)
#define SFTX_MODEL_BUSY    (-2)
i32 sftx_model_changed(const sft_model* model);
u64 sftx_model_edits(void);

/* ---- offline asset builders -------------------------------------------
   The reference builds meshlets, BLAS and shadow proxies inside its glTF
   importer with static functions; these entry points expose the same
   builders for applications that construct models in memory (the QSF2
   loader reads all of it from the file).  Host code, run once per asset.
   Results equal the reference's builders bit for bit; the bounding spheres
   draw from the C library's rand() as the reference's do.

   sftx_mesh_build replaces sft__mesh_build (sofren2.c:2526-2813): meshlets of
   a triangle soup (positions and indices only; meshlet_attribs, skins,
   material and lightmap page are left to the caller).  Arena memory for one
   meshlet per triangle, as the reference allocates.

   sftx_mesh_build_attribs replaces the glTF importer's per-mesh meshlet
   build (sofren2.c:7420-7660 with sft__builder_finalize_attribs,
   6916-7045): normals and UVs are required, skin data (joints and weights,
   all eight arrays) optional, bone_count bounds the culling-bone search.
   Returns the meshlet count, 0 on invalid input (message on stderr).

   sftx_model_link fills the model-relative meshlet offsets and the
   meshlet -> mesh table (sofren2.c:7668-7681) after the meshes are built.

   sftx_blas_build replaces sft__blas_build_model (sofren2.c:3526-3617);
   sftx_shadow_proxies_build replaces sft__generate_shadow_proxies
   (sofren2.c:3511-3524: bone clusters for a skinned model, BVH subtrees
   otherwise; nothing without a BLAS or a positive budget). */
typedef struct sftx_raw_geometry {
    const f32 *vx, *vy, *vz;
    const f32 *nx, *ny, *nz;
    const f32 *u, *v;
    const u16 *j0, *j1, *j2, *j3;   /* optional skinning data */
    const u8 *w0, *w1, *w2, *w3;
} sftx_raw_geometry;
void sftx_mesh_build(sft_mesh* out_mesh, sft_arena* arena, const f32* vx, const f32* vy, const f32* vz,
                     const i32* inds, i32 tri_count);
u32 sftx_mesh_build_attribs(sft_mesh* out_mesh, sft_arena* arena, const sftx_raw_geometry* raw, u32 vertex_count,
                            const i32* inds, i32 tri_count, u32 bone_count);
void sftx_model_link(sft_model* model, sft_arena* arena);
void sftx_blas_build(sft_model* model, sft_arena* arena);
void sftx_shadow_proxies_build(sft_model* model, sft_arena* arena, i32 budget);

/* Copies `bytes` between two device buffers with a plain streaming kernel
   and returns its duration in ms; a calibration point for HBM counters. */
f32 sftx_copy_calibration(u64 bytes);

/* ---- frame description -------------------------------------------------
   Everything the per-frame render path reads, as recorded by the front-end
   since sft_frame_begin: the reference's frame state (sofren2.c:1335-1495)
   minus its scratch.  Pointers stay valid until the next sft_frame_end or
   resource call.  The CPU restatement of the render path consumes this. */
/* The texture's mip chain as the GPU built it in HBM: all levels in the 4x4-block
   layout (sofren2.c:9229-9289), `capacity` texels at `out`.  Returns the chain's
   texel count (with out == NULL just the count), -1 for an invalid handle or a
   short buffer. */
i64 sftx_cap_texture(sft_tex_handle tex, u32* out, u64 capacity);

typedef struct sftx_texture_desc {
    const u32* texels;          /* all mip levels, 4x4-block swizzled, contiguous; NULL when the
                                   chain is built on the GPU from level0 */
    const u32* level0;          /* the w x h linear pixels the texture was created from */
    u32 total;                  /* texels of the whole chain */
    i32 w, h, mip_levels, is_pot;
    u32 mip_offset[14];         /* in texels from texels[0] */
    i32 blocks_w[14];
    i32 stride_shift[14];
} sftx_texture_desc;

typedef struct sftx_proxy_desc {
    sftvec posa, posb;
    f32 radius;
} sftx_proxy_desc;

typedef struct sftx_frame_desc {
    i32 width, height, tile_cols, tile_rows;
    f32 half_width, half_height, vp_x, vp_y;
    sftvec camera_pos;
    sftmat view_proj;
    f32 planes[6][4];               /* normalised nx, ny, nz, d */
    sftvec ambient;
    const sft_directional_light* dir_lights;
    i32 n_dir;
    const sft_point_light* point_lights;
    i32 n_point;
    f32 lightmap_brightness;
    u8 shadows_enabled, should_clear;
    u32 clear_color;
    i32 debug_mode;
    const sft_instance* instances;  /* ss_meshlet_offset filled in */
    i32 n_instances;
    u32 total_ss;
    const sftx_proxy_desc* proxies;
    i32 n_proxies;
    const sftx_texture_desc* textures;
    i32 n_textures;
    const u32* mat_tex;
    const u32* mat_color;
    const f32* mat_shininess;
    i32 n_mat;
    const u32* tmat_tex;
    const u32* tmat_color;
    const f32* tmat_shininess;
    i32 n_tmat;
    u32 base_tex;                   /* albedo of the base material */
    u32* color_buffer;
    i32 color_pitch;                /* in u32 */
    i32 samples;                    /* MSAA samples per pixel: 1, 2 or 4 */
} sftx_frame_desc;

/* Fills `out` with the frame recorded so far.  Needs no GPU. */
void sftx_export_frame(sftx_frame_desc* out);

#ifdef __cplusplus
}
#endif

#endif
