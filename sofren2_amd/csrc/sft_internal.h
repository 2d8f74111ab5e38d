/*
 * sft_internal.h — state shared by the host front-end (frontend.cpp) and the
 * HIP backend (backend.hip).  Not part of the public API.
 *
 * HBM layout (all resident pools are indexed by a "meshlet slot"; a model's
 * meshlets occupy slots [model_base, model_base + total_meshlets) in the
 * model-relative order used by visibility IDs, sofren2.c:8457-8477):
 *   pos   f32[slot][3][64]   positions, SoA (sofren2.c:914-918)
 *   attr  f32[slot][7][64]   nx ny nz u v lu lv (sofren2.c:920-933)
 *   inds  u8 [slot][512]     16 batches x (v0[8] v1[8] v2[8] pad[8]) (949-953)
 *   cull  f32[slot][8]       centre xyz, radius, cone axis xyz, cutoff
 *   slot_mesh u32[slot]      mesh record (material + lightmap) of the slot
 *   slot_skin u32[slot]      skin record or 0xFFFFFFFF
 *   skin  u8 [skin][768]     j0..j3 u16[64], w0..w3 u8[64] (935-947)
 * Per frame:
 *   work  WorkItem[n]        drawn meshlets in emission order (instance draw
 *                            order, mesh, meshlet: the reference's job order)
 *   ss    f32[ss][9][64]     screen-space cache, reference layout (1464-1477)
 *   tile lists               u64 payload descriptors per 64x64 tile
 *   clip records             reference layout (1271-1280) in slot order
 */
#pragma once
#include <stdint.h>
#include <deque>
#include <memory>
#include <vector>
#include <unordered_map>
#include "../../include/sofren2.h"
#include "../../include/sofren2_ext.h"

namespace sft {

constexpr u32 kInvalid = 0xFFFFFFFFu;
constexpr int kMaxTileLights = 16;
/* tile rectangles of payloads and clipped triangles pack each 64-px tile
   coordinate into 8 bits: a view is at most 256 tiles (16384 px) on a side */
constexpr i32 kMaxViewSide = 256 * SFT_TILE_SIZE;
inline bool view_size_ok(i32 w, i32 h) { return w > 0 && h > 0 && w <= kMaxViewSide && h <= kMaxViewSide; }

/* GPU-side per-frame constants. */
struct FrameConst {
    i32 width, height, screen_w, screen_h;
    i32 tile_cols, tile_rows, n_tiles, samples;
    f32 hw, hh, vp_x, vp_y;
    f32 bias_x, bias_y, neg_hh, pad0;
    f32 vp[16];
    f32 planes[6][4];
    f32 cam[4];
    f32 ambient[4];
    u32 clear_color, should_clear, debug_mode, shadows_enabled;
    f32 lm_brightness, pad1, pad2, pad3;
    u32 n_work, n_inst, n_point, n_dir;
    u32 n_proxies, base_tex, rcp_bits, rsq_bits;
    u32 clip_cap, tile_list_cap, capture, n_tex;
    u32 n_mat, n_tmat, tlas_root, occ_lights;   /* occ_lights = min(point lights, 64) when shadows run */
    u32 occ_batches, tlas_nodes, pad8, pad9;    /* proxy packets of 8 per light */
};

/* One view of a batch on the GPU: its frame constants and where its rows
   start in the batch-wide tables (instances, work items, lights, transient
   materials, proxies, TLAS nodes, 64x64 tiles) and output planes. */
struct ViewGpu {
    FrameConst fc;
    u32 inst_base, work_base, n_work, pl_base;
    u32 dl_base, tmat_base, tile_base, proxy_base;
    u32 tlas_base, occ_base, occ_count, seg_base;
    u32 n_segs;         /* binning segments of the view, its clip segment last */
    u32 vis64;          /* 1: the view's 64-bit visibility planes are written */
    u32 raster_chunk;   /* payloads per visibility work item (the view's resolution class) */
    u32 band_base;      /* colour streamed to host memory: the view's first band (tile row) of
                           FramePtrs::band_count / band_flag; kInvalid: not streamed */
    u32 raster_split;   /* estimated sweep above which a payload's rows are split into 2-8 row-band jobs */
    u32 tlas_build;     /* 1: k_tlas builds the view's TLAS into its (reserved) nodes */
    u32 id_shift;       /* visibility IDs: instance << id_shift | meshlet << SFT_TRIANGLE_BITS | triangle,
                           id_shift = SFT_TRIANGLE_BITS + meshlet bits of the view's ID split */
    u32 lp_cache;       /* first light of the view's capsule records in FramePtrs::cached_lp
                           (uploaded from a served template's ShadowCache); kInvalid: traced */
    u64 vid_base;       /* first element of the view's [samples][grid] planes */
    u64 color_base;     /* first pixel of the view's width x height colour */
};

/* batch-wide counts and capacities */
struct BatchGpu {
    u32 n_views, n_work, n_tiles, max_samples;
    u32 clip_cap, tile_list_cap, rcp_bits, rsq_bits;
    u32 n_shade_tiles, n_shade_tiles_ms, occ_total, n_lights;
    u32 vis64, raster_chunk, raster_groups, n_segs;
    u32 n_shade_sh, n_shade_ms_sh;  /* of the shaded tiles above, those of views with capsule shadows (listed last) */
    u32 occ_stack;                  /* occlusion traversal stack entries per wave: 256, 512 or 1024 */
    u32 tlas_gpu;                   /* 1: some view's TLAS is built on the device (k_tlas) */
    u32 items2_base;                /* first entry of the MSAA views' visibility work items */
    u32 n_views_ms;                 /* views with MSAA (their work items are sample pairs) */
};

/* A binning segment: up to kSegItems consecutive batch slots of one view, or
   (count == kInvalid) the view's clip slots.  m_base locates its per-tile
   counts (then cursors) in the binning matrix, n_tiles of the view wide. */
constexpr u32 kSegItems = 4096;
struct BinSeg {
    u32 view, first, count, m_base;
};

struct InstGpu {
    f32 t[16];
    f32 inv[16];        /* rigid inverse of t (sft_m_qinv) for shadow rays */
    u32 model_base;     /* first meshlet slot of the model */
    u32 ss_offset;      /* ss_meshlet_offset, batch-wide */
    u32 bone_base;      /* first matrix in the frame's bone buffer, kInvalid if none */
    u32 blas_base;      /* first node in the resident BLAS pool; kInvalid when the
                           model has no BLAS or has a skeleton (shadow rays test
                           static geometry only, sofren2.c:4130) */
    u32 blas_root;      /* root node relative to blas_base */
    u32 blas_count;     /* nodes of the model's BLAS */
    u32 n_meshlets;     /* total_meshlets of the model */
    u32 tlas_item;      /* 1: the model has a BLAS, so the instance is a TLAS item */
    f32 root_lo[3], root_hi[3];   /* model-space box of the BLAS root's children */
};

/* world-space capsule proxy as the GPU reads it */
struct ProxyGpu {
    f32 a[3], b[3], radius, pad;
};

/* per (light, proxy) shading constants (sofren2.c:5960-5985, 991-998); 56 B,
   unpadded, so that a tile's 16 x 16 records stay small in LDS */
struct ProxyRec {
    f32 pa[3], abdab;
    f32 ab[3], wdw;
    f32 w[3], rb2;
    f32 pr2, vis_mult;
};
static_assert(sizeof(ProxyRec) == 56, "ProxyRec layout");
constexpr int kMaxLightProxies = 16;

/* A served template's shadow data: per point light, the capsule records the
   shading reads (k_proxy_cache's selection over the proxies' occlusion
   counts).  Proxies, lights, instances and BLAS are the template's, the
   request's camera takes no part, so every request of the template gets the
   same records: they are read back from the first batch that traced them
   and uploaded with later batches instead of tracing again. */
struct ShadowCache {
    bool filled = false;
    std::vector<ProxyRec> recs;     /* [n_point][kMaxLightProxies] */
    std::vector<u32> counts;        /* [n_point] */
};

struct WorkItem {
    u32 slot;           /* resident meshlet slot */
    u32 ss;             /* screen-space cache index (batch-wide) */
    u32 inst;           /* instance slot in its view (visibility ID field) */
    u32 gm;             /* model-relative meshlet index (visibility ID field) */
    i32 cull_bone;      /* bone used for the 8-meshlet group's cull volume, -1 none */
    u32 skin;           /* skin record, kInvalid if unskinned or no pose */
    u32 view;           /* view of the batch */
    u32 ginst;          /* instance row in the batch-wide table */
};

struct MeshRec {
    u32 material;       /* sft_mat_handle */
    u32 lightmap_tex;   /* sft_tex_handle of the lightmap page */
    u32 lightmapped;    /* mesh has a lightmap page (sofren2.c:6414) */
    u32 pad;
};

/* texture metadata mirrored to the GPU (sofren2.c:1395-1412) */
struct TexMeta {
    i32 w, h, mip_levels, is_pot;
    u32 mip_offset[14];     /* in texels, from the start of the device texel heap */
    i32 blocks_w[14];
    i32 stride_shift[14];
};

struct PointLightGpu { f32 pos[3], color[3], radius, inv_r2; };
struct DirLightGpu { f32 dir[3], color[3], pad[2]; };

/* world capsule proxy (sofren2.c:1432-1437) */
struct WorldProxy {
    sftvec posa, posb;
    f32 radius;
    u8 light_vis[64];
};

/* host texture pool entry */
struct HostTexture {
    u32 generation = 0;
    i32 next_free = -1;
    bool live = false;
    i32 w = 0, h = 0, mip_levels = 0;
    u8 is_pot = 0;
    std::vector<u32> level0;    /* the linear pixels it was created from; the swizzled
                                   mip chain is built on the GPU */
    u64 total = 0;              /* texels of the chain, every level padded to 4x4 */
    u32 mip_offset[14] = {0};   /* in texels from the chain's first texel */
    i32 blocks_w[14] = {0};
    i32 stride_shift[14] = {0};
    u64 dev_offset = 0;         /* texel offset in the device heap */
    bool uploaded = false;
};

/* Everything the render of one view reads, snapshot from the front-end at
   sft_present (a batch of one) or at sftx_view_end (sofren2.c:1335-1495 minus
   the reference's scratch).  Output targets are host memory or null. */
struct ViewRecord {
    i32 width = 0, height = 0, tile_cols = 0, tile_rows = 0, samples = 1;
    f32 half_width = 0, half_height = 0, vp_x = 0, vp_y = 0;
    u32* color_buffer = nullptr;
    i32 color_pitch = 0;            /* in u32 */
    u64* vis64 = nullptr;           /* [samples][tile_rows*64][tile_cols*64] */
    bool host_output = false;       /* copy colour to color_buffer whatever the output mode */
    u8 should_clear = 0;
    u32 clear_color = 0;
    i32 debug_mode = 0;
    sftvec camera_pos{0, 0, 0};
    sftmat view_proj{};
    f32 planes[6][4] = {};
    sftvec ambient{0, 0, 0};
    std::vector<sft_directional_light> dir_lights;
    std::vector<sft_point_light> point_lights;
    f32 lightmap_brightness = 1.5f;
    u8 shadows_enabled = 0;
    std::vector<WorldProxy> proxies;
    std::vector<sft_instance> instances;
    u32 dropped = 0;                /* draws beyond the instance cap (sofren2.c:8412-8417) */
    u32 instance_bits = SFT_INSTANCE_BITS;  /* ID split: 1 << instance_bits instances per frame */
    std::shared_ptr<ShadowCache> shadow_cache;  /* served templates with capsule shadows */
    u32 total_ss = 0;
    std::vector<const sft_model*> transient_models;
    std::vector<u32> tmat_tex, tmat_color;
    std::vector<f32> tmat_shininess;
    u32 capture_flags = 0;
    /* batched views keep their own copy of the posed bone matrices, taken when
       the view is ended (the reference reads them at present time) */
    std::vector<sftmat> bone_store;
    std::vector<u32> bones_kept;    /* per instance: matrices in bone_store (batched views) */
};

struct Context;
extern Context* g_ctx;

/* view_proj = view * proj and the six normalised frustum planes
   (sofren2.c:8345-8376) */
void camera_setup(const sftmat& view, const sftmat& proj, sftmat* view_proj, f32 planes[6][4]);
/* the frame recorded since sft_frame_begin as one view */
void snapshot_view(const Context& c, ViewRecord& v);

/* frame TLAS over the drawn instances (tlas.cpp); returns the root node */
/* a BVH build item: payload id, box, box centre (sofren2.c:961-965) */
struct BvhItem {
    u32 id;
    f32 lo[3], hi[3];
    f32 mid[3];
};
u32 build_bvh(BvhItem* items, u32 n, std::vector<sft_bvh_node>& nodes);
u32 build_tlas(const std::vector<sft_instance>& instances, std::vector<sft_bvh_node>& nodes);
u32 tlas_node_count(u32 items);     /* nodes k_tlas writes for `items` items (at most 64) */

struct Backend;
Backend* backend_create();
void backend_destroy(Backend* b);
/* renders a batch of views (one pass of kernels over all of them) and writes
   each view's outputs; blocks until they are in host memory */
void backend_render(Backend* b, Context* ctx, std::vector<ViewRecord>& views);
/* sft_present may return before its frame completes (device output) */
bool backend_async_present(const Backend* b);
/* frame sets of the backend.  Device-output frames keep up to kFrameSets in
   flight (small frames overlap on the GPU); batches, which fill it alone, up
   to kBatchesInFlight (sets 0-2: a batch's frame download and copy-out and the
   next batch's recording and tables overlap the GPU work of the others; a
   caller cycling three view sets replays each set's graph).  The request
   server keeps as many. */
constexpr int kFrameSets = 4;
constexpr int kBatchesInFlight = 3;
constexpr int kServerBatchesInFlight = kBatchesInFlight;
/* completes every batch of sftx_batch_submit_async still in flight */
void drain_async_batches(Context& c);
/* asynchronous halves for batches and the serving loop: up to
   kBatchesInFlight batches (sets 0-2) can be in flight; complete blocks until
   its batch is done and its host outputs are written.  An ordered submission
   starts its kernels after the previous ordered one's (its tables upload
   meanwhile): large batches then run back to back rather than sharing the
   GPU, and the one waited for ends first. */
void backend_submit(Backend* b, Context* ctx, std::vector<ViewRecord>& views, int set, bool ordered = false);
void backend_complete(Backend* b, Context* ctx, std::vector<ViewRecord>& views, int set);
bool backend_batch_done(Backend* b, int set);
u64 backend_resident_bytes(const Backend* b);
/* forgets a model's residency (its slots stay allocated) */
void backend_forget_model(Backend* b, const sft_model* m);
void backend_release_texture(Backend* b, u32 index);
/* resolution class of a view (sftx_tuning) */
int res_class(i32 width, i32 height);
void backend_get_tuning(const Backend* b, sftx_tuning* t);
void backend_set_tuning(Backend* b, const sftx_tuning* t);
void backend_tune(Backend* b, Context* ctx, std::vector<ViewRecord>& views, int repeats, sftx_tuning* report);

/* host estimate tables for _mm_rcp_ps / _mm_rsqrt_ps (rcp_probe.cpp) */
struct EstimateTables {
    int rcp_bits = 0;
    int rsq_bits = 0;
    std::vector<f32> rcp;       /* 2^rcp_bits entries: estimate for mantissa bucket in [1,2) */
    std::vector<f32> rsq;       /* 2^(rsq_bits+1) entries: [1,2) then [2,4) */
    u64 rcp_mismatches = 0;     /* emulation vs instruction on the probe's checks */
    u64 rsq_mismatches = 0;
};
void probe_estimate_tables(EstimateTables* out);
f32 emulate_rcp(const EstimateTables& t, f32 a);
f32 emulate_rsqrt(const EstimateTables& t, f32 a);

struct Context {
    i32 width = 0, height = 0;
    f32 half_width = 0, half_height = 0, vp_x = 0, vp_y = 0;
    i32 tile_cols = 0, tile_rows = 0;
    u32* color_buffer = nullptr;
    i32 color_pitch = 0;            /* in u32 */
    u8 should_clear = 0;
    u32 clear_color = 0xFF222222u;
    i32 samples = 1;                /* MSAA samples per pixel */
    enum sft_debug_mode debug_mode = SFT_DEBUG_NONE;

    sftvec camera_pos{0, 0, 0};
    sftmat view{}, proj{}, view_proj{};
    f32 planes[6][4] = {};

    sftvec ambient{0, 0, 0};
    std::vector<sft_directional_light> dir_lights;
    std::vector<sft_point_light> point_lights;
    f32 lightmap_brightness = 1.5f;

    u8 shadows_enabled = 0;
    std::vector<WorldProxy> proxies;

    i32 instance_count = 0;         /* counts dropped instances too, like the reference */
    u32 instance_bits = SFT_INSTANCE_BITS;  /* visibility-ID split (sftx_set_id_split) */
    std::vector<sft_instance> instances;
    u32 total_ss = 0;

    sft_arena scratch_arena{};      /* dmesh storage, reset each frame */
    std::vector<const sft_model*> transient_models;

    /* textures */
    std::vector<HostTexture> tex;
    i32 tex_free_head = -1;
    /* static materials */
    std::vector<u32> mat_gen;
    std::vector<i32> mat_next_free;
    std::vector<u32> mat_tex, mat_color;
    std::vector<f32> mat_shininess;
    i32 mat_free_head = -1;
    /* transient materials, reset in frame_end */
    std::vector<u32> tmat_tex, tmat_color;
    std::vector<f32> tmat_shininess;

    sft_meshlet_inds quad_inds{}, tri_inds{};
    sft_mat_handle base_mat = kInvalid;
    sft_tex_handle base_tex = kInvalid;

    u32 capture_flags = 0;
    sftx_frame_stats stats{};

    /* batched views being recorded (sftx_batch_begin .. sftx_batch_submit) */
    bool batch_open = false;
    bool view_rejected = false;     /* the view begun last has an unsupported size */
    bool server_running = false;    /* the request server owns the backend */
    std::vector<ViewRecord> batch_views;
    /* batches submitted with sftx_batch_submit_async, oldest first (at most
       one per frame set) */
    struct AsyncBatch {
        i32 ticket;
        int set;
        std::vector<ViewRecord> views;
    };
    std::deque<AsyncBatch> async_batches;
    i32 next_ticket = 1;

    EstimateTables est;
    std::vector<sftx_texture_desc> export_tex;
    std::vector<sftx_proxy_desc> export_proxies;
    Backend* backend = nullptr;
};

}  // namespace sft
