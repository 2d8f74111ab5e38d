/*
 * port.h — CPU restatement of sofren2's per-frame render path (test
 * infrastructure: the checker of the HIP path, never linked into it).
 *
 * Input is the frame description exported by the front-end
 * (include/sofren2_ext.h); output is the ARGB frame plus the intermediates the
 * tests compare: per-sample visibility IDs and depth over the tile grid, the
 * screen-space meshlet cache with its cull result, and the clipped-triangle
 * records.  Single-threaded and in the reference's emission order, so payload
 * order, depth-tie winners and clip-slot numbers are those of the reference
 * run with one worker thread.
 */
#ifndef SFT_PORT_H
#define SFT_PORT_H

#include "sofren2.h"
#include "sofren2_ext.h"

typedef struct sftx_port_out {
    u32* color;         /* width*height ARGB, pitch = width; pre-filled by caller when not clearing */
    u32* vid;           /* samples * (cols*64)*(rows*64) */
    f32* depth;
    u8* ss;             /* total_ss * 2304 */
    u8* ss_vis;         /* total_ss */
    u8* clips;          /* clip_cap * 164 */
    u32 clip_cap;
    u32 clip_count;     /* set by the port */
    i32 samples;        /* 1, 2 or 4 */
    u8* proxy_vis;      /* n_proxies * 64 light visibility counts (0..3), or null */
    i32 invalid_packets;    /* set by the port: occlusion packets with a degenerate frustum */
} sftx_port_out;

/* Renders the frame.  Returns 0, or -1 if the clip buffer was too small. */
int sftx_port_render(const sftx_frame_desc* f, sftx_port_out* out);

/* The swizzled mip chain of a w x h texture from its linear level-0 pixels,
   laid out as sft_texture_create lays it out (sofren2.c:9181-9310): levels
   halve (floor, min 1) down to 1x1, at most 14, each padded to a multiple of 4
   in both axes and stored in 4x4 blocks.  Writes the chain (padding zero) to
   out when non-null; returns its texel count. */
u32 sftx_port_build_texture(const u32* level0, i32 w, i32 h, u32* out);

/* sizes the caller needs for the output buffers */
u32 sftx_port_ss_count(const sftx_frame_desc* f);
i32 sftx_port_proxy_count(const sftx_frame_desc* f);
i32 sftx_port_samples(const sftx_frame_desc* f);
void sftx_port_grid(const sftx_frame_desc* f, i32* cols, i32* rows, i32* width, i32* height);

#endif
