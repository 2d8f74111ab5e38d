/*
 * kernels.h — device pointer bundle of one frame and the launch wrappers of the
 * render-path kernels (kernels.hip).  Host orchestration is in backend.cpp.
 */
#pragma once
#include <hip/hip_runtime.h>
#include "sft_internal.h"

namespace sft {

/* counters written by the kernels (u32 each) */
enum Counter : int {
    CNT_VISIBLE = 0,        /* meshlets surviving the cull */
    CNT_SKINNED,            /* visible skinned meshlets */
    CNT_BATCH_PAYLOADS,     /* batch payloads with at least one triangle */
    CNT_BATCH_TILE_PAIRS,   /* (batch payload, tile) pairs */
    CNT_CLIP_TRIS,          /* clipped-triangle slots used */
    CNT_CLIP_TILE_PAIRS,    /* (clip payload, tile) pairs */
    CNT_TILE_PAIRS,         /* total entries of the tile lists */
    CNT_OVERFLOW,           /* bit 0: clip pool, bit 1: tile lists, bit 2: a reduced occlusion stack */
    CNT_SHADED,             /* pixels with geometry */
    CNT_QGROUPS,            /* (block, triangle) shading groups */
    CNT_TEX_BILINEAR,
    CNT_TEX_NEAREST,
    CNT_LIGHTMAPPED,
    CNT_OCC_RAYS,           /* shadow-proxy occlusion rays traced */
    CNT_OCC_NODES,          /* BVH child boxes tested against packet frusta */
    CNT_OCC_TRIS,           /* ray-triangle tests */
    CNT_OCC_INVALID,        /* packets whose frustum the reference leaves unset */
    CNT_OCC_OVERFLOW,       /* packets whose traversal stack overflowed */
    CNT_TEX_NPOT,           /* albedo samples of non-power-of-two textures (edge clamp) */
    CNT_UNLIT,              /* shadings of unlit materials (negative shininess) */
    CNT_DIR_LIT,            /* shadings that evaluated directional lights */
    CNT_LIGHT_CUT,          /* tiles whose light list stopped at 16 with more lights in reach */
    CNT_BG_BLEND,           /* MSAA pixels blending the caller's colour under empty samples */
    CNT_CAP_TESTS,          /* capsule ray tests: per shaded pixel and reaching light, its cached capsules */
    CNT_RASTER_ITEMS,       /* visibility work items (tile, chunk) of views without MSAA listed */
    CNT_RASTER_NEXT,        /* their queue head */
    CNT_RASTER_ITEMS2,      /* visibility work items (tile, sample pair, chunk) of MSAA views listed */
    CNT_RASTER_NEXT2,       /* their queue head */
    CNT_COUNT
};

/* Statistics are accumulated in CNT_STAT_SLOTS cache-line-separated copies per
   counter (after the CNT_COUNT narrow words) so that thousands of waves do not
   serialise on one address; k_stats folds them into the narrow words. */
constexpr int CNT_STAT_SLOTS = 64;
constexpr int CNT_STAT_STRIDE = 32;     /* u32 per slot: one 128-byte line */
/* largest payload count of one visibility work item (its triangles are staged in LDS) */
constexpr u32 kRasterMaxChunk = 128;
constexpr u32 kRasterMaxItem = 1024;     /* payloads per visibility work item: up to 8 LDS stagings */
constexpr u32 kOccStackEntries = 1024;   /* occlusion traversal stack per wave */
constexpr u32 kRasterJobCost = 2048;     /* default row-band split threshold (ViewGpu::raster_split) */
constexpr size_t COUNTER_WORDS = CNT_COUNT + (size_t)CNT_COUNT * CNT_STAT_SLOTS * CNT_STAT_STRIDE;

struct FramePtrs {
    /* resident */
    const f32* pos;         /* [slot][3][64] */
    const f32* attr;        /* [slot][7][64] */
    const u8* inds;         /* [slot][512] */
    const f32* cull;        /* [slot][8] */
    const u32* slot_mesh;   /* [slot] */
    const u8* skin;         /* [skin][768] */
    const MeshRec* mesh_recs;
    const u32* texels;
    const TexMeta* tex_meta;
    const f32* est_rcp;
    const f32* est_rsq;
    /* per batch inputs */
    const BatchGpu* batch;
    const ViewGpu* views;
    const u32* tile_view;   /* [n_tiles] view of each batch-wide tile */
    const u32* shade_tiles; /* tiles shaded without MSAA, then tiles with MSAA */
    const InstGpu* insts;
    const WorkItem* work;
    const f32* bones;       /* [n][16] */
    const PointLightGpu* point_lights;
    const DirLightGpu* dir_lights;
    const u32* mat_tex;     /* static material pool */
    const u32* mat_color;
    const f32* mat_shin;
    const u32* tmat_tex;    /* transient material pool */
    const u32* tmat_color;
    const f32* tmat_shin;
    /* per frame outputs / scratch */
    f32* ss;                /* [ss][9][64] */
    u8* ss_vis;             /* [ss] */
    u32* clip_count;        /* [n_work] */
    u32* clip_base;         /* [n_work] */
    u64* clip_mask;         /* [n_work][2] */
    uint2* batch_pl;        /* [n_work][16]: x = lane mask, y = tile rect */
    const BinSeg* segs;     /* [n_segs] binning segments */
    u32* bin_m;             /* per segment: [n_tiles of its view] counts, then cursors */
    u32* tile_count;        /* [n_tiles] */
    u32* tile_off;          /* [n_tiles] */
    u64* tile_list;         /* [tile_list_cap] */
    f32* clip_rec;          /* [clip_cap][41] */
    u32* clip_key;          /* [clip_cap] */
    u32* clip_rect;         /* [clip_cap] */
    u32* color;             /* [height][width] */
    u32* vid;               /* [samples][grid] reference visibility IDs of the frame */
    u32* tile_lights;       /* [n_tiles][17]: count, then light indices */
    f32* tile_z;            /* [n_tiles][samples][2]: 1/w range per sample plane (MSAA) */
    /* shadows */
    const sft_bvh_node* blas;   /* resident BLAS pool */
    const sft_bvh_node* tlas;   /* frame TLAS */
    const ProxyGpu* proxies;    /* [n_proxies] */
    u8* proxy_vis;              /* [n_proxies][64] rays of 3 reaching each light */
    u32* light_proxy_count;     /* [n_point] */
    ProxyRec* light_proxies;    /* [n_point][16] */
    const ProxyRec* cached_lp;  /* capsule records of views with lp_cache set, [lights][16] */
    const u32* cached_lpc;      /* their counts, [lights] */
    f32* cap_depth;         /* [samples][grid] 1/w when capturing, else null */
    u64* vis64;             /* [samples][grid] (1/w bits << 32) | ID when requested, else null */
    u32* counters;          /* [CNT_COUNT] */
    u32* scan_sums;         /* per-4096-item block sums of the scans */
    uint2* raster_items;    /* visibility work items: tile, chunk << 2 | sample */
    u32* tile_nch;          /* [n_tiles] payload chunks of each tile */
    u32* tile_done;         /* [n_tiles][max_samples] chunks finished (zeroed per frame) */
    u64* merge;             /* [n_tiles][max_samples][4096] merge planes of split tiles */
    /* colour streamed to host memory while the frame shades (views with a
       band_base): pixels at color_base offsets, and per band (a view's tile
       row) the shading blocks finished (zeroed per frame) and a host-visible
       flag set when the band's pixels are all in host memory (1, or 2 when
       the frame's intermediates overflowed); null when nothing streams */
    u32* color_host;
    u32* band_count;
    u32* band_flag;
};

/* enqueues the kernels of one batch of views, starting with the clear of the
   `zero_words` counter words at p.counters; `ev` (10 events) brackets the
   passes when non-null */
/* `side` (optional, eager launches only): a second stream and its events
   (fork, join, start, end) on which the shadow-proxy occlusion and capsule
   lists run beside geometry, binning and rasterisation, joined before the
   shading kernels */
struct SideStream {
    hipStream_t stream;
    hipEvent_t fork, join, start, end;
};
void launch_frame(const FramePtrs& p, const BatchGpu& batch, u64 zero_words, hipStream_t s, hipEvent_t* ev,
                  const SideStream* side = nullptr);
void launch_copy(const uint4* src, uint4* dst, u64 n16, hipStream_t s);
/* builds texture `m`'s swizzled mip chain in `heap` (offsets in m are heap
   texel offsets, padding already zeroed) from its linear level-0 pixels */
void launch_texture_build(const u32* level0, u32* heap, const TexMeta& m, hipStream_t s);
u32 raster_groups_per_device();

}  // namespace sft
