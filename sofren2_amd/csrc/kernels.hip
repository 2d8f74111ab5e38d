/*
 * kernels.hip — the per-frame render path on gfx950.
 *
 * Pipeline of one batch of views (one sft_present is a batch of one), all
 * kernels on the batch's HIP stream, every view's rows back to back:
 *   k_zero        clears counters and chunk completion counts (a kernel, so
 *                 the batch can be replayed from a HIP graph).
 *   k_occlude     proxy-occlusion packets: one wave per (light, 8-proxy packet,
 *                 ray), frustum-culled TLAS/BLAS traversal, any-hit OR
 *                 (sofren2.c:8163-8280, 3840-4243).
 *   k_proxy_cache per light: the proxies with visibility left and their cached
 *                 capsule constants (sofren2.c:5936-5990).
 *   k_geom        one wave64 per drawn meshlet, lane = vertex: frustum + normal-
 *                 cone cull, 4-joint skinning, world transform, projection into
 *                 the screen-space cache; then lane = triangle: near-plane
 *                 classification, area cull, tile bounds and the 8-triangle
 *                 batch payloads (sofren2.c:4901-5145, 4775-4884).
 *   k_scan_*      exclusive scan of clipped-triangle counts in emission order, so
 *                 clip slots are numbered as the reference's single worker
 *                 numbers them (sofren2.c:4351-4361).
 *   k_clip        one wave per meshlet with straddling triangles: near-plane
 *                 clipping into slot-numbered records (sofren2.c:4245-4381,
 *                 4691-4772).
 *   k_bin_count, k_bin_scan, k_bin
 *                 per-tile payload lists without global tile-counter atomics.
 *   k_raster_plan splits long tile lists into chunks of visibility work.
 *   k_raster      persistent workgroups over (tile, chunk) items, with MSAA
 *                 (tile, chunk, sample pair) items: rasterise into LDS 64-bit
 *                 visibility planes, merge split
 *                 tiles in HBM, the last chunk publishes IDs, the tile's 1/w
 *                 range and its light list (sofren2.c:5148-5499, 6756-6813).
 *   k_tile_lights light lists of MSAA tiles over every sample plane.
 *   k_shade       32x8 pixels per workgroup, lane = pixel in the reference's
 *                 4x2 block order: interpolation, LOD, texturing, lighting,
 *                 capsule shadows (capsules culled per wave), ARGB pack
 *                 (sofren2.c:5809-6753); specialised for views with and
 *                 without capsule shadows.
 *   k_shade_ms    the same over sample planes: per block, triangles in first
 *                 (sample, lane) order; (pixel, triangle) items shaded 64 at
 *                 a time; weighted accumulate in each pixel's triangle order.
 *   k_stats       folds the spread statistics counters.
 *
 * Bit-exactness: every float operation of the reference is replayed in the
 * same order with explicit fmaf where the reference uses an FMA intrinsic, the
 * file is compiled with -ffp-contract=off, and the host's rcpps/rsqrtps
 * estimates come from tables probed at init.  SSE min/max (return the second
 * operand unless the first compares smaller/larger) and cvttps (out of range or
 * NaN gives INT_MIN) are emulated.  The raster replays the reference's
 * incremental edge/depth accumulation (row steps of B, 8-pixel chunk steps of
 * 8A from an 8-aligned start inside the 64x64 tile).  Depth ties are resolved
 * without ordering the payload lists: each fragment does an LDS 64-bit atomic
 * max of (depth bits << 32 | ~emission_key), and the emission key (meshlet in
 * job order, batch, clipped-before-batch, lane, clip output) orders fragments
 * exactly as the reference's single-threaded payload order does, so the
 * earliest-drawn fragment wins at equal depth as with its strict depth test.
 */
#include "kernels.h"
#include "capsule_cull.h"

namespace sft {

#define DEV __device__ __forceinline__

/* element `idx` of an array whose byte offset fits 32 bits: lets the
   compiler use a scalar base with a 32-bit vector offset */
template <class T>
DEV T ld32(const T* base, u32 idx) {
    return *(const T*)((const char*)base + (u32)(idx * (u32)sizeof(T)));
}

DEV u32 f2u(f32 f) { return __float_as_uint(f); }
DEV f32 u2f(u32 u) { return __uint_as_float(u); }
/* _mm256_min_ps(a, b) / _mm256_max_ps(a, b) per lane */
DEV f32 smin(f32 a, f32 b) { return a < b ? a : b; }
DEV f32 smax(f32 a, f32 b) { return a > b ? a : b; }
/* _mm256_cvttps_epi32 per lane */
DEV i32 cvtt(f32 x) { return (x >= -2147483648.0f && x < 2147483648.0f) ? (i32)x : (i32)0x80000000; }

struct V3 { f32 x, y, z; };
DEV f32 dot3(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }   /* sft_vv_dot */
DEV V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEV V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }

/* sft_vv_transform with a row-major 4x4 (sofren2.c:1993-2012) */
DEV V3 xform(V3 v, const f32* m) {
    V3 r;
    r.x = fmaf(v.x, m[0], fmaf(v.y, m[4], fmaf(v.z, m[8], m[12])));
    r.y = fmaf(v.x, m[1], fmaf(v.y, m[5], fmaf(v.z, m[9], m[13])));
    r.z = fmaf(v.x, m[2], fmaf(v.y, m[6], fmaf(v.z, m[10], m[14])));
    return r;
}
/* sft_vv_transform_dir (sofren2.c:2013-2030) */
DEV V3 xform_dir(V3 v, const f32* m) {
    V3 r;
    r.x = fmaf(v.x, m[0], fmaf(v.y, m[4], v.z * m[8]));
    r.y = fmaf(v.x, m[1], fmaf(v.y, m[5], v.z * m[9]));
    r.z = fmaf(v.x, m[2], fmaf(v.y, m[6], v.z * m[10]));
    return r;
}

/* host rcpps estimate: table value for the mantissa bucket, exponent rescaled */
DEV f32 rcp_est(const f32* tab, u32 bits, f32 a) {
    const u32 b = f2u(a);
    const u32 sign = b & 0x80000000u;
    const u32 e = (b >> 23) & 0xFFu;
    const u32 m = b & 0x7FFFFFu;
    if (0u == e) return u2f(sign | 0x7F800000u);
    if (255u == e) return m ? u2f(b | 0x00400000u) : u2f(sign);
    const u32 tb = f2u(tab[m >> (23u - bits)]);
    const i32 oe = (i32)((tb >> 23) & 0xFFu) + 127 - (i32)e;
    if (oe <= 0) return u2f(sign);
    return u2f(sign | ((u32)oe << 23) | (tb & 0x7FFFFFu));
}

/* host rsqrtps estimate: parity of the exponent selects [1,2) or [2,4) */
DEV f32 rsq_est(const f32* tab, u32 bits, f32 a) {
    const u32 b = f2u(a);
    const u32 e = (b >> 23) & 0xFFu;
    const u32 m = b & 0x7FFFFFu;
    if (0u == e) return u2f((b & 0x80000000u) | 0x7F800000u);
    if (b & 0x80000000u) return u2f(0xFFC00000u);
    if (255u == e) return m ? u2f(b | 0x00400000u) : 0.f;
    const i32 u = (i32)e - 127;
    const i32 p = u & 1;
    const u32 tb = f2u(tab[((u32)p << bits) | (m >> (23u - bits))]);
    const i32 oe = (i32)((tb >> 23) & 0xFFu) - (u - p) / 2;
    return u2f(((u32)oe << 23) | (tb & 0x7FFFFFu));
}

struct Est {
    const f32* rcp;
    const f32* rsq;
    u32 rcp_bits, rsq_bits;
    /* sft_vf_rcp_refined: estimate plus one Newton step (sofren2.c:632-639) */
    DEV f32 rcp_refined(f32 a) const {
        const f32 x = rcp_est(rcp, rcp_bits, a);
        return x * fmaf(-a, x, 2.0f);
    }
    /* sft_vv_norm_fast (sofren2.c:1973-1992) */
    DEV V3 norm_fast(V3 v) const {
        const f32 sq = dot3(v, v);
        const f32 ap = rsq_est(rsq, rsq_bits, sq);
        const f32 ap2 = ap * ap;
        const f32 r = (0.5f * ap) * fmaf(-sq, ap2, 3.0f);
        return V3{v.x * r, v.y * r, v.z * r};
    }
};

DEV Est make_est(const FramePtrs& P) {
    Est e;
    e.rcp = P.est_rcp;
    e.rsq = P.est_rsq;
    e.rcp_bits = P.batch->rcp_bits;
    e.rsq_bits = P.batch->rsq_bits;
    return e;
}

/* first clip slot of a view: clip slots are numbered batch-wide in work order */
DEV u32 view_clip_base(const FramePtrs& P, const ViewGpu& V) {
    return V.n_work ? P.clip_base[V.work_base] : 0u;
}

/* view owning batch-wide row `i` of a table whose per-view starts are given
   by `start(view)`, increasing (binary search over the views) */
template <class F>
DEV u32 find_view(const FramePtrs& P, u32 i, F start) {
    u32 lo = 0, hi = P.batch->n_views;
    while (hi - lo > 1u) {
        const u32 mid = (lo + hi) >> 1;
        if (start(P.views[mid]) <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* adds to a statistics counter through one of its spread slots */
DEV void stat_add(u32* counters, int cnt, u32 v, u32 slot) {
    if (v) atomicAdd(&counters[CNT_COUNT + ((size_t)cnt * CNT_STAT_SLOTS + (slot & (CNT_STAT_SLOTS - 1))) * CNT_STAT_STRIDE], v);
}

DEV u32 stat_slot() { return blockIdx.x * 4u + blockIdx.y * 13u + (threadIdx.x >> 6); }

DEV void wave_sync() { __builtin_amdgcn_wave_barrier(); }
DEV u32 lanes_below(u64 m, int lane) { return (u32)__popcll(m & ((1ull << lane) - 1ull)); }

/* rotated-grid MSAA sample offsets from the pixel centre (sofren2.c:186-196) */
__constant__ f32 c_msaa_ox[3][4] = { { 0.f }, { -0.25f, 0.25f }, { -0.125f, 0.375f, -0.375f, 0.125f } };
__constant__ f32 c_msaa_oy[3][4] = { { 0.f }, { -0.25f, 0.25f }, { -0.375f, -0.125f, 0.125f, 0.375f } };
DEV int msaa_row(int samples) { return 4 == samples ? 2 : (2 == samples ? 1 : 0); }

DEV i32 wave_sum(i32 v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* ================================================================== k_geom */
__global__ __launch_bounds__(256) void k_geom(FramePtrs P) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const u32 wi = blockIdx.x * 4u + (u32)wave;
    if (wi >= P.batch->n_work) return;
    const Est est = make_est(P);

    const WorkItem w = P.work[wi];
    const FrameConst& fc = P.views[w.view].fc;
    const InstGpu* inst = &P.insts[w.ginst];
    f32 T[16];
    for (int i = 0; i < 16; i += 1) T[i] = inst->t[i];

    /* ---- cull (sofren2.c:4942-4994) ---- */
    const f32* cd = P.cull + (u64)w.slot * 8u;
    V3 c{cd[0], cd[1], cd[2]};
    V3 ax{cd[4], cd[5], cd[6]};
    if (w.cull_bone >= 0) {
        const f32* B = P.bones + (u64)(inst->bone_base + (u32)w.cull_bone) * 16u;
        c = xform(c, B);
        ax = xform_dir(ax, B);
    }
    c = xform(c, T);
    ax = xform_dir(ax, T);
    const f32 scale = sqrtf(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
    const f32 radius = cd[3] * scale;
    f32 alen = sqrtf(dot3(ax, ax));
    alen = smax(alen, 1e-6f);
    ax = V3{ax.x / alen, ax.y / alen, ax.z / alen};
    bool inside = true;
    for (int p = 0; p < 6; p += 1) {
        const V3 n{fc.planes[p][0], fc.planes[p][1], fc.planes[p][2]};
        const f32 dist = dot3(c, n) + fc.planes[p][3];
        inside = inside && (dist + radius >= 0.f);
    }
    const V3 raw = sub3(c, V3{fc.cam[0], fc.cam[1], fc.cam[2]});
    const f32 d = sqrtf(dot3(raw, raw));
    const f32 dv = dot3(raw, ax);
    inside = inside && (dv <= fmaf(d, cd[7], radius));

    if (0 == lane) P.ss_vis[w.ss] = inside ? 1 : 0;
    if (!inside) {
        if (lane < 16) P.batch_pl[wi * 16u + lane] = make_uint2(0u, 0u);
        if (lane < 2) P.clip_mask[wi * 2u + lane] = 0ull;
        if (0 == lane) P.clip_count[wi] = 0u;
        return;
    }

    /* ---- vertex: skin, world, project (sofren2.c:5016-5123) ---- */
    const f32* pp = P.pos + (u64)w.slot * 192u;
    const f32* ap = P.attr + (u64)w.slot * 448u;
    const V3 v{pp[lane], pp[64 + lane], pp[128 + lane]};
    const V3 n{ap[lane], ap[64 + lane], ap[128 + lane]};
    V3 vw, nw;
    if (kInvalid != w.skin) {
        const u8* sk = P.skin + (u64)w.skin * 768u;
        const u16* j = (const u16*)sk;
        const f32 inv255 = 1.f / 255.f;
        const f32 w0 = (f32)sk[512 + lane] * inv255, w1 = (f32)sk[576 + lane] * inv255;
        const f32 w2 = (f32)sk[640 + lane] * inv255, w3 = (f32)sk[704 + lane] * inv255;
        const f32* bb = P.bones + (u64)inst->bone_base * 16u;
        const f32* M0 = bb + (u64)j[lane] * 16u;
        const f32* M1 = bb + (u64)j[64 + lane] * 16u;
        const f32* M2 = bb + (u64)j[128 + lane] * 16u;
        const f32* M3 = bb + (u64)j[192 + lane] * 16u;
        const V3 p0 = xform(v, M0), q0 = xform_dir(n, M0);
        const V3 p1 = xform(v, M1), q1 = xform_dir(n, M1);
        const V3 p2 = xform(v, M2), q2 = xform_dir(n, M2);
        const V3 p3 = xform(v, M3), q3 = xform_dir(n, M3);
        vw.x = fmaf(p0.x, w0, fmaf(p1.x, w1, fmaf(p2.x, w2, p3.x * w3)));
        vw.y = fmaf(p0.y, w0, fmaf(p1.y, w1, fmaf(p2.y, w2, p3.y * w3)));
        vw.z = fmaf(p0.z, w0, fmaf(p1.z, w1, fmaf(p2.z, w2, p3.z * w3)));
        nw.x = fmaf(q0.x, w0, fmaf(q1.x, w1, fmaf(q2.x, w2, q3.x * w3)));
        nw.y = fmaf(q0.y, w0, fmaf(q1.y, w1, fmaf(q2.y, w2, q3.y * w3)));
        nw.z = fmaf(q0.z, w0, fmaf(q1.z, w1, fmaf(q2.z, w2, q3.z * w3)));
        vw = xform(vw, T);
        nw = xform_dir(nw, T);
    }
    else {
        vw = xform(v, T);
        nw = xform_dir(n, T);
    }
    const f32* VP = fc.vp;
    const f32 cw = fmaf(vw.x, VP[3], fmaf(vw.y, VP[7], fmaf(vw.z, VP[11], VP[15])));
    const f32 tx = fmaf(vw.x, VP[0], fmaf(vw.y, VP[4], fmaf(vw.z, VP[8], VP[12])));
    const f32 ty = fmaf(vw.x, VP[1], fmaf(vw.y, VP[5], fmaf(vw.z, VP[9], VP[13])));
    const f32 tz = fmaf(vw.x, VP[2], fmaf(vw.y, VP[6], fmaf(vw.z, VP[10], VP[14])));
    const f32 inv_w = est.rcp_refined(cw);
    const f32 sx = fmaf(tx, inv_w * fc.hw, fc.bias_x);
    const f32 sy = fmaf(ty, inv_w * fc.neg_hh, fc.bias_y);

    f32* ss = P.ss + (u64)w.ss * 576u;
    ss[lane] = sx;
    ss[64 + lane] = sy;
    ss[128 + lane] = inv_w;
    ss[192 + lane] = vw.x;
    ss[256 + lane] = vw.y;
    ss[320 + lane] = vw.z;
    ss[384 + lane] = nw.x;
    ss[448 + lane] = nw.y;
    ss[512 + lane] = nw.z;

    /* ---- triangles: classify, area cull, tile bounds (sofren2.c:4775-4884) ---- */
    const u8* ib = P.inds + (u64)w.slot * 512u;
    const i32 max_tx = fc.tile_cols - 1, max_ty = fc.tile_rows - 1;
    i32 n_clip = 0, n_payloads = 0, n_pairs = 0;
    for (int half = 0; half < 2; half += 1) {
        const int t = half * 64 + lane;
        const int b = t >> 3, l = t & 7;
        const i32 i0 = ib[b * 32 + l], i1 = ib[b * 32 + 8 + l], i2 = ib[b * 32 + 16 + l];
        const f32 tz0 = __shfl(tz, i0), tz1 = __shfl(tz, i1), tz2 = __shfl(tz, i2);
        const f32 sx0 = __shfl(sx, i0), sx1 = __shfl(sx, i1), sx2 = __shfl(sx, i2);
        const f32 sy0 = __shfl(sy, i0), sy1 = __shfl(sy, i1), sy2 = __shfl(sy, i2);

        const bool valid = (i0 | i1 | i2) > 0;
        const bool in0 = tz0 >= 0.f, in1 = tz1 >= 0.f, in2 = tz2 >= 0.f;
        const bool out0 = tz0 < 0.f, out1 = tz1 < 0.f, out2 = tz2 < 0.f;
        const bool all_in = in0 && in1 && in2;
        const bool all_out = out0 && out1 && out2;
        const bool inter = valid && !(all_in || all_out);
        const u32 cm = (u32)in0 | ((u32)in1 << 1) | ((u32)in2 << 2);
        if (inter && 0u != cm) n_clip += (1 == __popc(cm)) ? 1 : 2;
        const u64 inter_bits = __ballot(inter);
        if (0 == lane) P.clip_mask[wi * 2u + half] = inter_bits;

        const f32 dx1 = sx1 - sx0, dy1 = sy1 - sy0, dx2 = sx2 - sx0, dy2 = sy2 - sy0;
        const bool area_ok = fmaf(dx1, dy2, -(dy1 * dx2)) > 0.1f;
        const i32 tx0 = cvtt(smin(sx0, smin(sx1, sx2))) >> 6;
        const i32 tx1 = cvtt(smax(sx0, smax(sx1, sx2))) >> 6;
        const i32 ty0 = cvtt(smin(sy0, smin(sy1, sy2))) >> 6;
        const i32 ty1 = cvtt(smax(sy0, smax(sy1, sy2))) >> 6;
        const bool in_grid = tx1 >= 0 && tx0 <= max_tx && ty1 >= 0 && ty0 <= max_ty;
        const bool ok = all_in && area_ok && in_grid;

        const u64 okb = __ballot(ok);
        const u32 mask8 = (u32)(okb >> (lane & ~7)) & 0xFFu;
        i32 bx0 = ok ? tx0 : 9999, bx1 = ok ? tx1 : -9999;
        i32 by0 = ok ? ty0 : 9999, by1 = ok ? ty1 : -9999;
        for (int o = 1; o < 8; o <<= 1) {
            bx0 = min(bx0, __shfl_xor(bx0, o));
            bx1 = max(bx1, __shfl_xor(bx1, o));
            by0 = min(by0, __shfl_xor(by0, o));
            by1 = max(by1, __shfl_xor(by1, o));
        }
        uint2 pl = make_uint2(0u, 0u);
        if (mask8) {
            bx0 = max(bx0, 0); bx1 = min(bx1, max_tx);
            by0 = max(by0, 0); by1 = min(by1, max_ty);
            pl = make_uint2(mask8, (u32)bx0 | ((u32)bx1 << 8) | ((u32)by0 << 16) | ((u32)by1 << 24));
            const i32 nt = (bx1 - bx0 + 1) * (by1 - by0 + 1);
            if (0 == l) {
                n_payloads += 1;
                n_pairs += nt;
            }
        }
        if (0 == l) P.batch_pl[wi * 16u + (u32)b] = pl;
    }
    n_clip = wave_sum(n_clip);
    n_payloads = wave_sum(n_payloads);
    n_pairs = wave_sum(n_pairs);
    if (0 == lane) {
        P.clip_count[wi] = (u32)n_clip;
        const u32 sl = stat_slot();
        stat_add(P.counters, CNT_VISIBLE, 1u, sl);
        if (kInvalid != w.skin) stat_add(P.counters, CNT_SKINNED, 1u, sl);
        stat_add(P.counters, CNT_BATCH_PAYLOADS, (u32)n_payloads, sl);
        stat_add(P.counters, CNT_BATCH_TILE_PAIRS, (u32)n_pairs, sl);
    }
}

/* ================================================================== scans
   Exclusive scan of n u32 in two passes: k_scan_sums reduces blocks of 16384,
   k_scan_apply adds the sums of the blocks before it and scans its own block
   (a single block needs no first pass).  The last block also writes the total
   and flags totals over `cap`. */
constexpr u32 kScanPer = 16;
constexpr u32 kScanItems = 1024u * kScanPer;

DEV u32 block_sum_1024(u32 v, u32* s_wave) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (0 == lane) s_wave[wave] = v;
    __syncthreads();
    u32 t = 0;
    for (int k = 0; k < 16; k += 1) t += s_wave[k];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(1024) void k_scan_sums(const u32* in, u32 n, u32* sums) {
    __shared__ u32 s_wave[16];
    const u32 base = blockIdx.x * kScanItems + threadIdx.x * kScanPer;
    u32 v = 0;
    for (u32 k = 0; k < kScanPer; k += 1) v += (base + k < n) ? in[base + k] : 0u;
    const u32 t = block_sum_1024(v, s_wave);
    if (0 == threadIdx.x) sums[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_scan_apply(const u32* in, u32* out, u32 n, const u32* sums, u32* total,
                                                      u32 cap, u32* overflow, u32 overflow_bit) {
    __shared__ u32 s_wave[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    /* carry: the sums of all earlier blocks */
    u32 c = 0;
    for (u32 b = (u32)tid; b < blockIdx.x; b += 1024u) c += sums[b];
    const u32 carry = block_sum_1024(c, s_wave);
    const u32 base = blockIdx.x * kScanItems + (u32)tid * kScanPer;
    u32 v[kScanPer], sum = 0;
    for (u32 k = 0; k < kScanPer; k += 1) {
        v[k] = (base + (u32)k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    u32 incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const u32 t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (63 == lane) s_wave[wave] = incl;
    __syncthreads();
    u32 before = 0;
    for (int k = 0; k < wave; k += 1) before += s_wave[k];
    u32 run = carry + before + (incl - sum);
    for (u32 k = 0; k < kScanPer; k += 1) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1u && 1023 == tid) {
        *total = run;
        if (run > cap) atomicOr(overflow, overflow_bit);
    }
}

/* ================================================================== k_clip */
struct ClipVert { f32 tx, ty, tz, w, u, v, lu, lv, px, py, pz, nx, ny, nz; };

DEV ClipVert lerp_cv(const ClipVert& a, const ClipVert& b, f32 t) {
    ClipVert r;
    r.tx = a.tx + t * (b.tx - a.tx);
    r.ty = a.ty + t * (b.ty - a.ty);
    r.tz = a.tz + t * (b.tz - a.tz);
    r.w = a.w + t * (b.w - a.w);
    r.u = a.u + t * (b.u - a.u);
    r.v = a.v + t * (b.v - a.v);
    r.lu = a.lu + t * (b.lu - a.lu);
    r.lv = a.lv + t * (b.lv - a.lv);
    r.px = a.px + t * (b.px - a.px);
    r.py = a.py + t * (b.py - a.py);
    r.pz = a.pz + t * (b.pz - a.pz);
    r.nx = a.nx + t * (b.nx - a.nx);
    r.ny = a.ny + t * (b.ny - a.ny);
    r.nz = a.nz + t * (b.nz - a.nz);
    return r;
}

/* clip-space vertex rebuilt from the cached world position; same FMA order as
   k_geom, so the bits equal the ones the classification used */
DEV ClipVert load_cv(const FrameConst& fc, const f32* ss, const f32* ap, int i) {
    ClipVert c;
    const f32 x = ss[192 + i], y = ss[256 + i], z = ss[320 + i];
    const f32* VP = fc.vp;
    c.w = fmaf(x, VP[3], fmaf(y, VP[7], fmaf(z, VP[11], VP[15])));
    c.tx = fmaf(x, VP[0], fmaf(y, VP[4], fmaf(z, VP[8], VP[12])));
    c.ty = fmaf(x, VP[1], fmaf(y, VP[5], fmaf(z, VP[9], VP[13])));
    c.tz = fmaf(x, VP[2], fmaf(y, VP[6], fmaf(z, VP[10], VP[14])));
    c.u = ap[192 + i];
    c.v = ap[256 + i];
    c.lu = ap[320 + i];
    c.lv = ap[384 + i];
    c.px = x;
    c.py = y;
    c.pz = z;
    c.nx = ss[384 + i];
    c.ny = ss[448 + i];
    c.nz = ss[512 + i];
    return c;
}

DEV ClipVert pick_cv(const ClipVert& a, const ClipVert& b, const ClipVert& c, int s) {
    const bool s0 = 0 == s, s1 = 1 == s;
    ClipVert r;
    r.tx = s0 ? a.tx : (s1 ? b.tx : c.tx);
    r.ty = s0 ? a.ty : (s1 ? b.ty : c.ty);
    r.tz = s0 ? a.tz : (s1 ? b.tz : c.tz);
    r.w = s0 ? a.w : (s1 ? b.w : c.w);
    r.u = s0 ? a.u : (s1 ? b.u : c.u);
    r.v = s0 ? a.v : (s1 ? b.v : c.v);
    r.lu = s0 ? a.lu : (s1 ? b.lu : c.lu);
    r.lv = s0 ? a.lv : (s1 ? b.lv : c.lv);
    r.px = s0 ? a.px : (s1 ? b.px : c.px);
    r.py = s0 ? a.py : (s1 ? b.py : c.py);
    r.pz = s0 ? a.pz : (s1 ? b.pz : c.pz);
    r.nx = s0 ? a.nx : (s1 ? b.nx : c.nx);
    r.ny = s0 ? a.ny : (s1 ? b.ny : c.ny);
    r.nz = s0 ? a.nz : (s1 ? b.nz : c.nz);
    return r;
}

/* one clipped triangle: its record (screen x, y, 1/w and the interpolated
   attributes per vertex, then instance and mesh), its sort key (work item,
   batch, lane, output c) and its tile rectangle; returns its tile pairs */
DEV i32 emit_clip(const FramePtrs& P, const FrameConst& fc, const WorkItem& w, u32 wi, int b, int l, int c, u32 slot,
                  u32 clip_cap, i32 max_tx, i32 max_ty, const ClipVert& q0, const ClipVert& q1, const ClipVert& q2) {
    if (slot >= clip_cap) {
        atomicOr(&P.counters[CNT_OVERFLOW], 1u);
        return 0;
    }
    f32* rec = P.clip_rec + (u64)slot * 41u;
    f32 sxv[3], syv[3];
    auto put = [&](int k, const ClipVert& q) {
        const f32 iw = 1.f / q.w;
        sxv[k] = (q.tx * iw + 1.f) * fc.hw + fc.vp_x;
        syv[k] = (-q.ty * iw + 1.f) * fc.hh + fc.vp_y;
        rec[0 + k] = sxv[k];
        rec[3 + k] = syv[k];
        rec[6 + k] = iw;
        rec[9 + k] = q.u;
        rec[12 + k] = q.v;
        rec[15 + k] = q.lu;
        rec[18 + k] = q.lv;
        rec[21 + k] = q.px;
        rec[24 + k] = q.py;
        rec[27 + k] = q.pz;
        rec[30 + k] = q.nx;
        rec[33 + k] = q.ny;
        rec[36 + k] = q.nz;
    };
    put(0, q0);
    put(1, q1);
    put(2, q2);
    ((u32*)rec)[39] = w.inst;
    ((u32*)rec)[40] = w.gm;
    P.clip_key[slot] = (wi << 9) | ((u32)b << 5) | (u32)(l * 2 + c);

    u32 rect = kInvalid;
    i32 pairs = 0;
    const f32 area = (sxv[1] - sxv[0]) * (syv[2] - syv[0]) - (syv[1] - syv[0]) * (sxv[2] - sxv[0]);
    if (!(area <= 0.1f)) {
        const i32 x0 = cvtt(floorf(smin(sxv[0], smin(sxv[1], sxv[2])))) >> 6;
        const i32 x1 = cvtt(ceilf(smax(sxv[0], smax(sxv[1], sxv[2])))) >> 6;
        const i32 y0 = cvtt(floorf(smin(syv[0], smin(syv[1], syv[2])))) >> 6;
        const i32 y1 = cvtt(ceilf(smax(syv[0], smax(syv[1], syv[2])))) >> 6;
        if (!(x1 < 0 || x0 > max_tx || y1 < 0 || y0 > max_ty)) {
            const i32 cx0 = max(x0, 0), cx1 = min(x1, max_tx);
            const i32 cy0 = max(y0, 0), cy1 = min(y1, max_ty);
            rect = (u32)cx0 | ((u32)cx1 << 8) | ((u32)cy0 << 16) | ((u32)cy1 << 24);
            pairs = (cx1 - cx0 + 1) * (cy1 - cy0 + 1);
        }
    }
    P.clip_rect[slot] = rect;
    return pairs;
}

__global__ __launch_bounds__(256) void k_clip(FramePtrs P) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const u32 wi = blockIdx.x * 4u + (u32)wave;
    if (wi >= P.batch->n_work) return;
    if (0u == P.clip_count[wi]) return;
    const WorkItem w = P.work[wi];
    const FrameConst& fc = P.views[w.view].fc;
    const u32 clip_cap = P.batch->clip_cap;
    const f32* ss = P.ss + (u64)w.ss * 576u;
    const f32* ap = P.attr + (u64)w.slot * 448u;
    const u8* ib = P.inds + (u64)w.slot * 512u;
    const i32 max_tx = fc.tile_cols - 1, max_ty = fc.tile_rows - 1;
    u32 base = P.clip_base[wi];
    i32 pairs = 0;
    for (int half = 0; half < 2; half += 1) {
        const int t = half * 64 + lane;
        const int b = t >> 3, l = t & 7;
        const bool inter = (P.clip_mask[wi * 2u + half] >> lane) & 1ull;
        ClipVert in[3];
        u32 cm = 0u;
        int nout = 0;
        if (inter) {
            const int idx[3] = { ib[b * 32 + l] & 63, ib[b * 32 + 8 + l] & 63, ib[b * 32 + 16 + l] & 63 };
            for (int k = 0; k < 3; k += 1) {
                in[k] = load_cv(fc, ss, ap, idx[k]);
                cm |= (u32)(in[k].tz >= 0.f) << k;
            }
            nout = (0u == cm) ? 0 : ((1 == __popc(cm)) ? 1 : 2);
        }
        /* exclusive prefix of outputs in triangle order */
        i32 incl = nout;
        for (int o = 1; o < 64; o <<= 1) {
            const i32 tv = __shfl_up(incl, o);
            if (lane >= o) incl += tv;
        }
        const u32 first = base + (u32)(incl - nout);
        base += (u32)__shfl(incl, 63);
        if (0 == nout) continue;

        /* the triangle rotated so that its lone vertex (the one in front for
           one output, the one behind for two) comes first: R1 follows it and
           R2 precedes it in the triangle's order.  Selects, not indexed
           arrays, keep every vertex in registers. */
        const int s = (1 == nout) ? (int)(cm >> 1) : (int)((~cm & 7u) >> 1);
        const ClipVert R0 = pick_cv(in[0], in[1], in[2], s);
        const ClipVert R1 = pick_cv(in[0], in[1], in[2], (2 == s) ? 0 : s + 1);
        const ClipVert R2 = pick_cv(in[0], in[1], in[2], (0 == s) ? 2 : s - 1);
        if (1 == nout) {
            const f32 da = R0.tz;
            const f32 t_ab = da / (da - R1.tz);
            const f32 t_ac = da / (da - R2.tz);
            pairs += emit_clip(P, fc, w, wi, b, l, 0, first, clip_cap, max_tx, max_ty, R0, lerp_cv(R0, R1, t_ab),
                               lerp_cv(R0, R2, t_ac));
        }
        else {
            const f32 da = R2.tz, db = R0.tz, dc = R1.tz;
            const f32 t_ab = da / (da - db);
            const f32 t_cb = dc / (dc - db);
            const ClipVert nab = lerp_cv(R2, R0, t_ab);
            const ClipVert ncb = lerp_cv(R1, R0, t_cb);
            pairs += emit_clip(P, fc, w, wi, b, l, 0, first, clip_cap, max_tx, max_ty, R2, nab, R1);
            pairs += emit_clip(P, fc, w, wi, b, l, 1, first + 1u, clip_cap, max_tx, max_ty, nab, ncb, R1);
        }
    }
    pairs = wave_sum(pairs);
    if (0 == lane) stat_add(P.counters, CNT_CLIP_TILE_PAIRS, (u32)pairs, stat_slot());
}

/* ================================================================== binning
   Payload descriptors: bits 0-31 index (work*16+batch, or clip slot), bits
   32-39 lane mask (0 marks a clipped triangle, as in the reference payload).
   Binning runs over segments of consecutive payload items of one view
   (BinSeg) without global atomics on the tile counters:
     k_bin_count  each segment counts its payloads per tile of its view in LDS
                  and writes the counts to its row of the binning matrix (the
                  rows of a view's segments are consecutive, n_tiles apart);
     k_bin_scan   one block per view sums the rows per tile, scans the totals
                  into tile offsets, reserves the view's range of the tile list
                  and turns every row into that segment's start per tile;
     k_bin        each segment writes its descriptors from those starts.
   Order inside a tile list is free (depth ties are settled by emission keys). */
constexpr u32 kBinThreads = 1024;
constexpr u32 kBinPer = kSegItems / kBinThreads;   /* batch slots per thread */
constexpr u32 kBinLdsTiles = 8192;                  /* larger views count in HBM */

/* the clip slots of a view: the clip scan runs in work order */
DEV void view_clip_range(const FramePtrs& P, const ViewGpu& V, u32& first, u32& end) {
    const BatchGpu& B = *P.batch;
    const u32 total = min(P.counters[CNT_CLIP_TRIS], B.clip_cap);
    const u32 we = V.work_base + V.n_work;
    first = (V.n_work > 0u) ? min(P.clip_base[V.work_base], total) : 0u;
    end = (we < B.n_work) ? min(P.clip_base[we], total) : total;
    if (end < first) end = first;
}

/* calls f(tile index inside the view, descriptor) for every tile of every
   payload of segment `sg`, items spread over the block's threads */
template <class F>
DEV void for_seg_payloads(const FramePtrs& P, const BinSeg& sg, const ViewGpu& V, F&& f) {
    const i32 cols = V.fc.tile_cols;
    auto tiles = [&](u32 rect, u64 desc) {
        const i32 x0 = rect & 0xFF, x1 = (rect >> 8) & 0xFF, y0 = (rect >> 16) & 0xFF, y1 = rect >> 24;
        for (i32 ty = y0; ty <= y1; ty += 1) {
            for (i32 tx = x0; tx <= x1; tx += 1) f((u32)(ty * cols + tx), desc);
        }
    };
    if (kInvalid != sg.count) {
        uint2 pl[kBinPer];
        for (u32 k = 0; k < kBinPer; k += 1) {
            const u32 j = threadIdx.x + k * kBinThreads;
            pl[k] = (j < sg.count) ? P.batch_pl[sg.first + j] : make_uint2(0u, 0u);
        }
        for (u32 k = 0; k < kBinPer; k += 1) {
            if (0u == pl[k].x) continue;
            tiles(pl[k].y, ((u64)pl[k].x << 32) | (u64)(sg.first + threadIdx.x + k * kBinThreads));
        }
        return;
    }
    u32 c0, c1;
    view_clip_range(P, V, c0, c1);
    for (u32 slot = c0 + threadIdx.x; slot < c1; slot += kBinThreads) {
        const u32 rect = P.clip_rect[slot];
        if (kInvalid != rect) tiles(rect, (u64)slot);
    }
}

__global__ __launch_bounds__(kBinThreads) void k_bin_count(FramePtrs P) {
    __shared__ u32 s_t[kBinLdsTiles];
    const BinSeg sg = P.segs[blockIdx.x];
    const ViewGpu& V = P.views[sg.view];
    const u32 nt = (u32)V.fc.n_tiles;
    u32* row = P.bin_m + sg.m_base;
    if (nt > kBinLdsTiles) {
        /* the row itself holds the counts; only this block touches it */
        for (u32 t = threadIdx.x; t < nt; t += kBinThreads) row[t] = 0u;
        __syncthreads();
        for_seg_payloads(P, sg, V, [&](u32 t, u64) { atomicAdd(&row[t], 1u); });
        return;
    }
    for (u32 t = threadIdx.x; t < nt; t += kBinThreads) s_t[t] = 0u;
    __syncthreads();
    for_seg_payloads(P, sg, V, [&](u32 t, u64) { atomicAdd(&s_t[t], 1u); });
    __syncthreads();
    for (u32 t = threadIdx.x; t < nt; t += kBinThreads) row[t] = s_t[t];
}

__global__ __launch_bounds__(kBinThreads) void k_bin_scan(FramePtrs P) {
    __shared__ u32 s_wave[16];
    __shared__ u32 s_base;
    const ViewGpu& V = P.views[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 nt = (u32)V.fc.n_tiles;
    const u32 per = (nt + kBinThreads - 1u) / kBinThreads;
    const u32 t0 = min((u32)tid * per, nt), t1 = min(t0 + per, nt);
    const u32 ns = V.n_segs;
    u32* rows = P.bin_m + P.segs[V.seg_base].m_base;
    /* totals of this thread's tiles, eight rows in flight */
    u32 sum = 0;
    for (u32 t = t0; t < t1; t += 1) {
        u32 c = 0;
        for (u32 k0 = 0; k0 < ns; k0 += 8u) {
            u32 v[8];
            for (u32 j = 0; j < 8u; j += 1) v[j] = (k0 + j < ns) ? rows[(k0 + j) * nt + t] : 0u;
            for (u32 j = 0; j < 8u; j += 1) c += v[j];
        }
        P.tile_count[V.tile_base + t] = c;
        sum += c;
    }
    u32 incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const u32 v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (63 == lane) s_wave[wave] = incl;
    __syncthreads();
    u32 before = 0, total = 0;
    for (int k = 0; k < 16; k += 1) {
        before += (k < wave) ? s_wave[k] : 0u;
        total += s_wave[k];
    }
    if (0 == tid) {
        /* the view's range of the batch's tile list */
        const u32 base = atomicAdd(&P.counters[CNT_TILE_PAIRS], total);
        if (base + total > P.batch->tile_list_cap || base + total < base) atomicOr(&P.counters[CNT_OVERFLOW], 2u);
        s_base = base;
    }
    __syncthreads();
    u32 run = s_base + before + (incl - sum);
    for (u32 t = t0; t < t1; t += 1) {
        P.tile_off[V.tile_base + t] = run;
        for (u32 k0 = 0; k0 < ns; k0 += 8u) {
            u32 v[8];
            for (u32 j = 0; j < 8u; j += 1) v[j] = (k0 + j < ns) ? rows[(k0 + j) * nt + t] : 0u;
            for (u32 j = 0; j < 8u; j += 1) {
                if (k0 + j < ns) rows[(k0 + j) * nt + t] = run;
                run += v[j];
            }
        }
    }
}

__global__ __launch_bounds__(kBinThreads) void k_bin(FramePtrs P) {
    __shared__ u32 s_t[kBinLdsTiles];
    const BinSeg sg = P.segs[blockIdx.x];
    const ViewGpu& V = P.views[sg.view];
    const u32 nt = (u32)V.fc.n_tiles;
    u32* row = P.bin_m + sg.m_base;
    const u32 cap = P.batch->tile_list_cap;
    if (nt > kBinLdsTiles) {
        for_seg_payloads(P, sg, V, [&](u32 t, u64 desc) {
            const u32 pos = atomicAdd(&row[t], 1u);
            if (pos < cap) P.tile_list[pos] = desc;
        });
        return;
    }
    for (u32 t = threadIdx.x; t < nt; t += kBinThreads) s_t[t] = row[t];
    __syncthreads();
    for_seg_payloads(P, sg, V, [&](u32 t, u64 desc) {
        const u32 pos = atomicAdd(&s_t[t], 1u);
        if (pos < cap) P.tile_list[pos] = desc;
    });
}

/* ================================================================== k_tile */
struct Tri {
    f32 sx[3], sy[3], iw[3];
    f32 u[3], v[3], lu[3], lv[3];
    f32 px[3], py[3], pz[3];
    f32 nx[3], ny[3], nz[3];
    u32 inst, slot;
};

/* rasterises one triangle of the tile into kPl sample planes, lane j = x offset
   inside an 8-pixel chunk, 8 lanes at group `g` of the wave
   (sofren2.c:5213-5497).  With MSAA (`aa`) plane k's sample sits at
   (ox[k], oy[k]) from the pixel centre: its edge and depth values gain
   A*ox + B*oy, added to the plane-free values at each test, trivial accept
   needs a margin of 0.375 (|A| + |B|), and a chunk whose 8 pixel centres are
   all beyond that margin of an edge is skipped.
   Several groups, in this wave and others, can share a triangle: this group
   sweeps the chunks of rows min_y + phase, min_y + phase + stride, ...
   (stride a power of two) and steps the row values through every row in
   between with additions only, so each row's edge and depth values are the
   same sequence of additions as in the reference. */
template <bool aa, int kPl>
DEV void raster_tri(u64* vis, const Est& est, i32 tmx, i32 tmy, int j, int g, u32 key,
                    f32 sx0, f32 sy0, f32 iw0, f32 sx1, f32 sy1, f32 iw1, f32 sx2, f32 sy2, f32 iw2,
                    const f32* ox, const f32* oy, int stride, int phase) {
    const f32 A0 = sy0 - sy1, B0 = sx1 - sx0;
    const f32 A1 = sy1 - sy2, B1 = sx2 - sx1;
    const f32 A2 = sy2 - sy0, B2 = sx0 - sx2;
    const f32 m01 = sx0 * sy1, m10 = sy0 * sx1;
    const f32 C0 = (fmaf(sx0, sy1, -m10) - fmaf(sy0, sx1, -m01)) * 0.5f;
    const f32 m12 = sx1 * sy2, m21 = sy1 * sx2;
    const f32 C1 = (fmaf(sx1, sy2, -m21) - fmaf(sy1, sx2, -m12)) * 0.5f;
    const f32 m20 = sx2 * sy0, m02 = sy2 * sx0;
    const f32 C2 = (fmaf(sx2, sy0, -m02) - fmaf(sy2, sx0, -m20)) * 0.5f;
    const f32 dx1 = sx1 - sx0, dx2 = sx2 - sx0;
    const f32 dy1 = sy1 - sy0, dy2 = sy2 - sy0;
    const f32 det = dx1 * dy2 - dx2 * dy1;
    if (!(det > 1e-6f)) return;
    const f32 inv_det = est.rcp_refined(det);
    const f32 dzdx = ((iw1 - iw0) * dy2 - (iw2 - iw0) * dy1) * inv_det;
    const f32 dzdy = ((iw2 - iw0) * dx1 - (iw1 - iw0) * dx2) * inv_det;
    const f32 c0 = fmaf(B0, 0.5f, fmaf(A0, 0.5f, C0));
    const f32 c1 = fmaf(B1, 0.5f, fmaf(A1, 0.5f, C1));
    const f32 c2 = fmaf(B2, 0.5f, fmaf(A2, 0.5f, C2));
    const f32 zb = fmaf(0.5f - sy0, dzdy, fmaf(0.5f - sx0, dzdx, iw0));

    const i32 min_x = cvtt(smax(smin(sx0, smin(sx1, sx2)), (f32)tmx));
    const i32 max_x = cvtt(smin(smax(sx0, smax(sx1, sx2)), (f32)(tmx + 63)));
    const i32 min_y = cvtt(smax(smin(sy0, smin(sy1, sy2)), (f32)tmy));
    const i32 max_y = cvtt(smin(smax(sy0, smax(sy1, sy2)), (f32)(tmy + 63)));
    if (min_x > max_x || min_y > max_y) return;
    const i32 amin = min_x & ~7;
    const i32 amax = (max_x & ~7) + 8;
    const f32 famin = (f32)amin, fminy = (f32)min_y;
    const f32 se0 = c0 + A0 * famin + B0 * fminy;
    const f32 se1 = c1 + A1 * famin + B1 * fminy;
    const f32 se2 = c2 + A2 * famin + B2 * fminy;
    const f32 sz = zb + dzdx * famin + dzdy * fminy;
    const f32 box_w = (f32)(amax - 1 - amin);
    const f32 box_h = (f32)(max_y - min_y);
    const f32 me0 = se0 + (A0 < 0.f ? A0 * box_w : 0.f) + (B0 < 0.f ? B0 * box_h : 0.f);
    const f32 me1 = se1 + (A1 < 0.f ? A1 * box_w : 0.f) + (B1 < 0.f ? B1 * box_h : 0.f);
    const f32 me2 = se2 + (A2 < 0.f ? A2 * box_w : 0.f) + (B2 < 0.f ? B2 * box_h : 0.f);
    const f32 pad0 = aa ? 0.375f * (fabsf(A0) + fabsf(B0)) : 0.f;
    const f32 pad1 = aa ? 0.375f * (fabsf(A1) + fabsf(B1)) : 0.f;
    const f32 pad2 = aa ? 0.375f * (fabsf(A2) + fabsf(B2)) : 0.f;
    const bool trivial = me0 >= pad0 && me1 >= pad1 && me2 >= pad2;
    /* the sample offsets of each plane (kPl planes, SFT_TILE_PIXELS apart) */
    f32 de0[kPl], de1[kPl], de2[kPl], dzs[kPl];
    for (int k = 0; k < kPl; k += 1) {
        de0[k] = A0 * ox[k] + B0 * oy[k];
        de1[k] = A1 * ox[k] + B1 * oy[k];
        de2[k] = A2 * ox[k] + B2 * oy[k];
        dzs[k] = dzdx * ox[k] + dzdy * oy[k];
    }

    const u32 nkey = ~key;
    const f32 fj = (f32)j;
    const f32 dz_chunk = dzdx * 8.f;
    f32 z_row = sz + dzdx * fj;
    const i32 first = min_y + phase;
    if (first > max_y) return;
    const int row_step = 64 * stride;
    int row = (first - tmy) * 64 + (amin - tmx) + j;
    if (trivial) {
        /* only the depth row values are needed */
        for (i32 y = min_y; y < first; y += 1) z_row = z_row + dzdy;
        for (i32 y = first;;) {
            f32 z = z_row;
            for (i32 x = amin, p = row; x < amax; x += 8, p += 8) {
                for (int k = 0; k < kPl; k += 1) {
                    const f32 zs = aa ? z + dzs[k] : z;
                    atomicMax((unsigned long long*)&vis[p + k * SFT_TILE_PIXELS],
                              (zs > 0.f) ? (((u64)f2u(zs) << 32) | nkey) : 0ull);
                }
                z = z + dz_chunk;
            }
            const i32 next = y + stride;
            if (next > max_y) break;
            for (; y < next; y += 1) z_row = z_row + dzdy;
            row += row_step;
        }
    }
    else {
        const f32 A0c = A0 * 8.f, A1c = A1 * 8.f, A2c = A2 * 8.f;
        f32 e0r = se0 + A0 * fj, e1r = se1 + A1 * fj, e2r = se2 + A2 * fj;
        for (i32 y = min_y; y < first; y += 1) {
            e0r = e0r + B0;
            e1r = e1r + B1;
            e2r = e2r + B2;
            z_row = z_row + dzdy;
        }
        for (i32 y = first;;) {
            f32 e0 = e0r, e1 = e1r, e2 = e2r, z = z_row;
            for (i32 x = amin, p = row; x < amax; x += 8, p += 8) {
                if constexpr (!aa) {
                    /* a miss adds 0 to the max: no branch in the chunk loop */
                    const bool hit = e0 >= 0.f && e1 >= 0.f && e2 >= 0.f && z > 0.f;
                    atomicMax((unsigned long long*)&vis[p], hit ? (((u64)f2u(z) << 32) | nkey) : 0ull);
                }
                else {
                    /* the triangle's 8 lanes run the same rows and chunks */
                    const bool near = e0 >= -pad0 && e1 >= -pad1 && e2 >= -pad2;
                    const u32 any_near = (u32)(__ballot(near) >> (g * 8)) & 0xFFu;
                    if (any_near) {
                        for (int k = 0; k < kPl; k += 1) {
                            const f32 zs = z + dzs[k];
                            if (e0 + de0[k] >= 0.f && e1 + de1[k] >= 0.f && e2 + de2[k] >= 0.f && zs > 0.f) {
                                atomicMax((unsigned long long*)&vis[p + k * SFT_TILE_PIXELS], ((u64)f2u(zs) << 32) | nkey);
                            }
                        }
                    }
                }
                e0 = e0 + A0c;
                e1 = e1 + A1c;
                e2 = e2 + A2c;
                z = z + dz_chunk;
            }
            const i32 next = y + stride;
            if (next > max_y) break;
            for (; y < next; y += 1) {
                e0r = e0r + B0;
                e1r = e1r + B1;
                e2r = e2r + B2;
                z_row = z_row + dzdy;
            }
            row += row_step;
        }
    }
}

/* resolves a winning emission key to the reference's 32-bit visibility ID */
/* emission key -> the view's reference visibility ID; `clip_base` is the
   view's first batch-wide clip slot */
DEV u32 key_to_vid(const FramePtrs& P, u32 key, u32 n_clip, u32 clip_base, u32 id_shift) {
    const u32 sub = key & 31u;
    if (sub >= 16u) {
        const WorkItem w = P.work[key >> 9];
        const u32 tri = ((key >> 5) & 15u) * 8u + (sub - 16u);
        return (w.inst << id_shift) | (w.gm << SFT_TRIANGLE_BITS) | tri;
    }
    /* clip keys are increasing in slot order */
    u32 lo = 0u, hi = n_clip;
    while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (P.clip_key[mid] < key) lo = mid + 1u;
        else hi = mid;
    }
    return (lo - clip_base) | 0x80000000u;
}

/* fetches triangle `vid` (a reference visibility ID of view V) */
DEV void load_tri(const FramePtrs& P, const ViewGpu& V, u32 vid, Tri& t) {
    if (vid & 0x80000000u) {
        const f32* r = P.clip_rec + (u64)((vid & 0x7FFFFFFFu) + view_clip_base(P, V)) * 41u;
        for (int k = 0; k < 3; k += 1) {
            t.sx[k] = r[0 + k]; t.sy[k] = r[3 + k]; t.iw[k] = r[6 + k];
            t.u[k] = r[9 + k]; t.v[k] = r[12 + k]; t.lu[k] = r[15 + k]; t.lv[k] = r[18 + k];
            t.px[k] = r[21 + k]; t.py[k] = r[24 + k]; t.pz[k] = r[27 + k];
            t.nx[k] = r[30 + k]; t.ny[k] = r[33 + k]; t.nz[k] = r[36 + k];
        }
        t.inst = V.inst_base + ((const u32*)r)[39];
        t.slot = P.insts[t.inst].model_base + ((const u32*)r)[40];
        /* keeps these loads in this branch: merged with the meshlet loads
           below they would need a full 64-bit address each */
        asm volatile("" ::: "memory");
        return;
    }
    t.inst = V.inst_base + (vid >> V.id_shift);
    const u32 gm = (vid >> SFT_TRIANGLE_BITS) & ((1u << (V.id_shift - SFT_TRIANGLE_BITS)) - 1u);
    const u32 tri = vid & ((1u << SFT_TRIANGLE_BITS) - 1u);
    const InstGpu* inst = &P.insts[t.inst];
    t.slot = inst->model_base + gm;
    const u32 ib = t.slot * 512u + (tri >> 3) * 32u + (tri & 7u);
    const u32 ss = (inst->ss_offset + gm) * 576u;
    const u32 ap = t.slot * 448u;
    for (int k = 0; k < 3; k += 1) {
        const u32 i = ld32(P.inds, ib + (u32)k * 8u) & 63u;
        /* one address per vertex; the planes are constant offsets from it */
        const f32* sv = P.ss + (ss + i);
        const f32* av = P.attr + (ap + i);
        t.sx[k] = sv[0]; t.sy[k] = sv[64]; t.iw[k] = sv[128];
        t.px[k] = sv[192]; t.py[k] = sv[256]; t.pz[k] = sv[320];
        t.nx[k] = sv[384]; t.ny[k] = sv[448]; t.nz[k] = sv[512];
        t.u[k] = av[192]; t.v[k] = av[256]; t.lu[k] = av[320]; t.lv[k] = av[384];
    }
}

/* the fields of Tri that barycentrics and UVs need: screen positions, 1/w, uv */
DEV void load_tri_uv(const FramePtrs& P, const ViewGpu& V, u32 vid, Tri& t) {
    if (vid & 0x80000000u) {
        const f32* r = P.clip_rec + (u64)((vid & 0x7FFFFFFFu) + view_clip_base(P, V)) * 41u;
        for (int k = 0; k < 3; k += 1) {
            t.sx[k] = r[0 + k]; t.sy[k] = r[3 + k]; t.iw[k] = r[6 + k];
            t.u[k] = r[9 + k]; t.v[k] = r[12 + k];
        }
        asm volatile("" ::: "memory");
        return;
    }
    const u32 gm = (vid >> SFT_TRIANGLE_BITS) & ((1u << (V.id_shift - SFT_TRIANGLE_BITS)) - 1u);
    const u32 tri = vid & ((1u << SFT_TRIANGLE_BITS) - 1u);
    const InstGpu* inst = &P.insts[V.inst_base + (vid >> V.id_shift)];
    const u32 slot = inst->model_base + gm;
    const u32 ib = slot * 512u + (tri >> 3) * 32u + (tri & 7u);
    const u32 ss = (inst->ss_offset + gm) * 576u;
    const u32 ap = slot * 448u;
    for (int k = 0; k < 3; k += 1) {
        const u32 i = ld32(P.inds, ib + (u32)k * 8u) & 63u;
        const f32* sv = P.ss + (ss + i);
        const f32* av = P.attr + (ap + i);
        t.sx[k] = sv[0]; t.sy[k] = sv[64]; t.iw[k] = sv[128];
        t.u[k] = av[192]; t.v[k] = av[256];
    }
}

/* (pixel, triangle) items of one wave in the MSAA resolve: 64 pixels x 4 samples */
constexpr u32 kMsItems = 256;

/* barycentric setup at a pixel centre (sofren2.c:6273-6312, non-MSAA) */
struct Bary { f32 pw0, pw1, pw2; };
/* with MSAA (`aa`) the screen barycentrics are clamped to the triangle and
   renormalised so pixel centres outside it never extrapolate (6296-6305) */
DEV Bary bary_at(const Est& est, const Tri& t, f32 px, f32 py, bool aa = false) {
    const f32 dx1 = t.sx[1] - t.sx[0], dx2 = t.sx[2] - t.sx[0];
    const f32 dy1 = t.sy[1] - t.sy[0], dy2 = t.sy[2] - t.sy[0];
    const f32 inv_det = est.rcp_refined(dx1 * dy2 - dx2 * dy1);
    const f32 ex = px - t.sx[0], ey = py - t.sy[0];
    f32 bu = (ex * dy2 - ey * dx2) * inv_det;
    f32 bv = (ey * dx1 - ex * dy1) * inv_det;
    f32 bw = (1.f - bu) - bv;
    if (aa) {
        bu = smax(0.f, bu);
        bv = smax(0.f, bv);
        bw = smax(0.f, bw);
        const f32 inv_sum = est.rcp_refined(bu + (bv + bw));
        bu = bu * inv_sum;
        bv = bv * inv_sum;
        bw = bw * inv_sum;
    }
    const f32 inv_z = fmaf(bw, t.iw[0], fmaf(bu, t.iw[1], bv * t.iw[2]));
    const f32 z = est.rcp_refined(inv_z);
    Bary b;
    b.pw0 = (bw * t.iw[0]) * z;
    b.pw1 = (bu * t.iw[1]) * z;
    b.pw2 = (bv * t.iw[2]) * z;
    return b;
}

DEV f32 interp(const Bary& b, const f32* a) {
    return fmaf(b.pw0, a[0], fmaf(b.pw1, a[1], b.pw2 * a[2]));
}

DEV f32 unpack_ch(u32 c, int shift) { return (f32)((c >> shift) & 0xFFu) * (1.f / 255.f); }

/* texel address inside the 4x4-block layout (sofren2.c:5606-5630) */
DEV u32 texel_addr(const TexMeta& m, int lod, i32 x, i32 y) {
    const u32 ux = (u32)x, uy = (u32)y;
    if (m.is_pot) {
        return (((uy >> 2) << m.stride_shift[lod]) + (ux >> 2)) * 16u + ((uy & 3u) << 2) + (ux & 3u);
    }
    return (((uy >> 2) * (u32)m.blocks_w[lod]) + (ux >> 2)) * 16u + ((uy & 3u) << 2) + (ux & 3u);
}

/* bilinear sample of the 4 nearest texels (sofren2.c:5543-5660) */
DEV void sample_bilinear(const FramePtrs& P, u32 h, int lod, f32 cu, f32 cv, f32& r, f32& g, f32& b) {
    if (SFT_INVALID_HANDLE == h) { r = 1.f; g = 1.f; b = 1.f; return; }
    const TexMeta& m = P.tex_meta[h & SFT_HANDLE_INDEX_MASK];
    const u32 px = m.mip_offset[lod];
    const i32 cw = max(m.w >> lod, 1), ch = max(m.h >> lod, 1);
    const f32 upx = cu * (f32)cw - 0.5f;
    const f32 vpx = cv * (f32)ch - 0.5f;
    const f32 uf = floorf(upx), vf = floorf(vpx);
    const f32 ufr = upx - uf, vfr = vpx - vf;
    const i32 u0 = cvtt(uf), v0 = cvtt(vf);
    i32 u0w, v0w, u1, v1;
    if (m.is_pot) {
        u1 = (u0 + 1) & (cw - 1); v1 = (v0 + 1) & (ch - 1);
        u0w = u0 & (cw - 1); v0w = v0 & (ch - 1);
    }
    else {
        /* the reference clamps only the upper edge; a negative coordinate
           would index before the level, so it is held at 0 here */
        u1 = max(min(u0 + 1, cw - 1), 0); v1 = max(min(v0 + 1, ch - 1), 0);
        u0w = max(min(u0, cw - 1), 0); v0w = max(min(v0, ch - 1), 0);
    }
    const u32 c00 = ld32(P.texels, px + texel_addr(m, lod, u0w, v0w));
    const u32 c10 = ld32(P.texels, px + texel_addr(m, lod, u1, v0w));
    const u32 c01 = ld32(P.texels, px + texel_addr(m, lod, u0w, v1));
    const u32 c11 = ld32(P.texels, px + texel_addr(m, lod, u1, v1));
    f32 o[3];
    const int sh[3] = {16, 8, 0};
    for (int k = 0; k < 3; k += 1) {
        const f32 f00 = unpack_ch(c00, sh[k]), f10 = unpack_ch(c10, sh[k]);
        const f32 f01 = unpack_ch(c01, sh[k]), f11 = unpack_ch(c11, sh[k]);
        const f32 top = f00 + (f10 - f00) * ufr;
        const f32 bot = f01 + (f11 - f01) * ufr;
        o[k] = top + (bot - top) * vfr;
    }
    r = o[0]; g = o[1]; b = o[2];
}

/* nearest sample (sofren2.c:5662-5729) */
DEV void sample_nearest(const FramePtrs& P, u32 h, int lod, f32 cu, f32 cv, f32& r, f32& g, f32& b) {
    if (SFT_INVALID_HANDLE == h) { r = 1.f; g = 1.f; b = 1.f; return; }
    const TexMeta& m = P.tex_meta[h & SFT_HANDLE_INDEX_MASK];
    const u32 px = m.mip_offset[lod];
    const i32 cw = max(m.w >> lod, 1), ch = max(m.h >> lod, 1);
    const i32 u0 = cvtt(floorf(cu * (f32)cw));
    const i32 v0 = cvtt(floorf(cv * (f32)ch));
    i32 uw, vw;
    if (m.is_pot) { uw = u0 & (cw - 1); vw = v0 & (ch - 1); }
    else { uw = max(min(u0, cw - 1), 0); vw = max(min(v0, ch - 1), 0); }
    const u32 c = ld32(P.texels, px + texel_addr(m, lod, uw, vw));
    r = unpack_ch(c, 16); g = unpack_ch(c, 8); b = unpack_ch(c, 0);
}

/* mip level from the first covered lane's UV derivatives (sofren2.c:5501-5541);
   the dither term cannot change the rounded level, so it is omitted */
DEV int calc_lod(const FramePtrs& P, u32 h, f32 dux, f32 duy, f32 dvx, f32 dvy) {
    if (SFT_INVALID_HANDLE == h) return 0;
    const TexMeta& m = P.tex_meta[h & SFT_HANDLE_INDEX_MASK];
    const f32 ux = dux * (f32)m.w, vx = dvx * (f32)m.h;
    const f32 uy = duy * (f32)m.w, vy = dvy * (f32)m.h;
    const f32 a = fabsf(ux), bb = fabsf(vx), c = fabsf(uy), d = fabsf(vy);
    const f32 m1 = a > bb ? a : bb, m2 = c > d ? c : d;
    const f32 rho = m1 > m2 ? m1 : m2;
    i32 lod = 0;
    if (rho > 1.f) {
        const i32 ir = cvtt(rho);
        lod = 31 - __clz((u32)ir);
    }
    return min(max(lod, 0), m.mip_levels - 1);
}

/* a tile's light list and, for views with capsule shadows (kShadows: the
   view's shadows_enabled && n_proxies > 0), each light's capsule records,
   staged in LDS by the shading workgroup */
template <bool kShadows>
struct TileShading {
    int count;
    u32 np[kMaxTileLights];                             /* capsules per light (0 without shadows) */
    PointLightGpu lights[kMaxTileLights];
    ProxyRec prox[kShadows ? kMaxTileLights : 1][kMaxLightProxies];
};

/* fills `ts` for tile `tile` of view VW with the workgroup's threads */
template <bool kShadows>
DEV void stage_tile_shading(const FramePtrs& P, const ViewGpu& VW, u32 tile, TileShading<kShadows>& ts, u32* s_li) {
    const int tid = threadIdx.x, nthreads = blockDim.x;
    const bool shadows = kShadows;
    const u32* tl = P.tile_lights + (u64)tile * (kMaxTileLights + 1);
    const int count = (int)tl[0];
    if (tid < count) {
        const u32 li = tl[1 + tid];
        s_li[tid] = li;
        ts.lights[tid] = P.point_lights[VW.pl_base + li];
        ts.np[tid] = shadows ? P.light_proxy_count[VW.pl_base + li] : 0u;
    }
    if (0 == tid) ts.count = count;
    __syncthreads();
    if (shadows) {
        for (int e = tid; e < count * kMaxLightProxies; e += nthreads) {
            const int i = e / kMaxLightProxies, k = e % kMaxLightProxies;
            if ((u32)k < ts.np[i]) ts.prox[i][k] = P.light_proxies[(u64)(VW.pl_base + s_li[i]) * kMaxLightProxies + k];
        }
    }
    __syncthreads();
}


/* per tile light i (16 bits each, 4 lights per word): capsules near the span */
struct CapsuleMasks {
    u64 q[4];
};

DEV u32 near_capsules(const CapsuleMasks& cm, int i) {
    const u64 q = (i < 4) ? cm.q[0] : (i < 8) ? cm.q[1] : (i < 12) ? cm.q[2] : cm.q[3];
    return (u32)(q >> ((i & 3) * 16)) & 0xFFFFu;
}

/* box of the positions of lanes with `inside` set.  Every lane of the wave
   must execute this call: the butterfly reads all 64 lanes. */
DEV WaveSpan wave_span(const V3& p, bool inside) {
    const f32 inf = __builtin_inff();
    f32 lx = inside ? p.x : inf, ly = inside ? p.y : inf, lz = inside ? p.z : inf;
    f32 hx = inside ? p.x : -inf, hy = inside ? p.y : -inf, hz = inside ? p.z : -inf;
    for (int o = 1; o < 64; o <<= 1) {
        lx = fminf(lx, __shfl_xor(lx, o)); ly = fminf(ly, __shfl_xor(ly, o)); lz = fminf(lz, __shfl_xor(lz, o));
        hx = fmaxf(hx, __shfl_xor(hx, o)); hy = fmaxf(hy, __shfl_xor(hy, o)); hz = fmaxf(hz, __shfl_xor(hz, o));
    }
    return span_of_box(lx, ly, lz, hx, hy, hz);
}

DEV bool ray_near(const ProxyRec& c, const V3& lraw, f32 d2, f32 inv_d2) {
    return ray_near_xyz(c, lraw.x, lraw.y, lraw.z, d2, inv_d2);
}

/* lane (light 4r + lane/16, capsule lane%16) tests one pair per round; every
   lane of the wave must execute this call */
DEV CapsuleMasks capsule_masks(const TileShading<true>& ts, const WaveSpan& s, int lane) {
    CapsuleMasks cm{{0ull, 0ull, 0ull, 0ull}};
    const int k = lane & 15;
    for (int r = 0; r < 4 && r * 4 < ts.count; r += 1) {
        const int i = r * 4 + (lane >> 4);
        const bool near = i < ts.count && (u32)k < ts.np[i] && capsule_near(ts.lights[i], ts.prox[i][k], s);
        cm.q[r] = __ballot(near);
    }
    return cm;
}

/* shades one pixel of triangle t (sofren2.c:6190-6620, non-MSAA); dux..dvy are
   the block's first-lane UV derivatives.  `group` holds the lanes of this
   lane's 4x2 block (bits 0-7, block starting at lane `block_lane`) covered by
   the same triangle: the capsule-shadow loop couples them.  Every lane of a
   wave that reaches the lighting loop runs the same light and proxy
   iterations, so the group's lanes meet at each ballot.  With shadows, `cm`
   holds the capsules near the span of every calling lane's position.
   kShadows is the view's shadows_enabled && n_proxies > 0 (tiles are launched
   in runs of one kind). */
template <bool kShadows>
DEV V3 shade_pixel(const FramePtrs& P, const ViewGpu& VW, const Est& est, const Tri& t,
                   const Bary& b, f32 dux, f32 duy, f32 dvx, f32 dvy, const TileShading<kShadows>& ts, u32& fetch,
                   u32 group, int block_lane, const CapsuleMasks& cm) {
    const FrameConst& fc = VW.fc;
    const f32 cu = interp(b, t.u), cv = interp(b, t.v);
    const f32 lu = interp(b, t.lu), lv = interp(b, t.lv);
    const V3 pos{interp(b, t.px), interp(b, t.py), interp(b, t.pz)};
    const V3 nrm{interp(b, t.nx), interp(b, t.ny), interp(b, t.nz)};
    const V3 N = est.norm_fast(nrm);
    const V3 V = est.norm_fast(sub3(V3{fc.cam[0], fc.cam[1], fc.cam[2]}, pos));

    /* material (sofren2.c:6349-6382) */
    const MeshRec mr = P.mesh_recs[P.slot_mesh[t.slot]];
    u32 base_color = 0xFFFFFFFFu;
    u32 tex = SFT_INVALID_HANDLE;
    f32 shin = 4.f;
    if (SFT_INVALID_HANDLE != mr.material) {
        const u32 mi = mr.material & SFT_MAT_INDEX_MASK;
        if (mr.material & SFT_MAT_TRANSIENT_BIT) {
            if (mi < fc.n_tmat) {
                const u32 k = VW.tmat_base + mi;
                base_color = P.tmat_color[k]; tex = P.tmat_tex[k]; shin = P.tmat_shin[k];
            }
        }
        else if (mi < fc.n_mat) {
            base_color = P.mat_color[mi]; tex = P.mat_tex[mi]; shin = P.mat_shin[mi];
        }
    }
    if (SFT_INVALID_HANDLE != tex && (tex & SFT_HANDLE_INDEX_MASK) >= fc.n_tex) tex = SFT_INVALID_HANDLE;
    if (SFT_DEBUG_CHECKERBOARD == fc.debug_mode) {
        tex = fc.base_tex;
        base_color = 0xFFFFFFFFu;
    }
    const f32 base_r = (f32)((base_color >> 16) & 0xFFu) * (1.f / 255.f);
    const f32 base_g = (f32)((base_color >> 8) & 0xFFu) * (1.f / 255.f);
    const f32 base_b = (f32)(base_color & 0xFFu) * (1.f / 255.f);

    const int lod = calc_lod(P, tex, dux, duy, dvx, dvy);
    f32 tr, tg, tb;
    if (lod > 1) sample_nearest(P, tex, lod, cu, cv, tr, tg, tb);
    else sample_bilinear(P, tex, lod, cu, cv, tr, tg, tb);
    if (SFT_INVALID_HANDLE != tex) {
        fetch |= (lod > 1) ? 2u : 1u;
        if (!P.tex_meta[tex & SFT_HANDLE_INDEX_MASK].is_pot) fetch |= 8u;
    }
    const f32 alb_r = tr * base_r, alb_g = tg * base_g, alb_b = tb * base_b;

    f32 fr, fg, fb;
    if (shin < 0.f) {
        fetch |= 16u;
        fr = alb_r * -shin;
        fg = alb_g * -shin;
        fb = alb_b * -shin;
    }
    else {
        f32 dr, dg, db;
        f32 sr = 0.f, sg = 0.f, sb = 0.f;
        f32 lm_shadow = 1.f;
        const f32 spec_scale = ((shin + 8.f) / 8.f) * 0.04f;
        const bool lightmapped = mr.lightmapped != 0u;
        if (lightmapped) {
            sample_bilinear(P, mr.lightmap_tex, 0, lu, lv, dr, dg, db);
            fetch |= 4u;
            dr = dr * fc.lm_brightness;
            dg = dg * fc.lm_brightness;
            db = db * fc.lm_brightness;
        }
        else {
            dr = fc.ambient[0]; dg = fc.ambient[1]; db = fc.ambient[2];
            if (fc.n_dir > 0u) fetch |= 32u;
            for (u32 l = 0; l < fc.n_dir; l += 1) {
                const DirLightGpu& L = P.dir_lights[VW.dl_base + l];
                const V3 Ld{L.dir[0], L.dir[1], L.dir[2]};
                const f32 ndl = smax(dot3(N, Ld), 0.f);
                const bool lit = ndl > 0.f;
                dr = fmaf(ndl, L.color[0], dr);
                dg = fmaf(ndl, L.color[1], dg);
                db = fmaf(ndl, L.color[2], db);
                if (lit) {
                    const V3 H = est.norm_fast(add3(Ld, V));
                    const f32 ndh = smax(dot3(N, H), 0.f);
                    const f32 den = (shin - shin * ndh) + ndh;
                    const f32 spec = ndh * est.rcp_refined(den);
                    const f32 term = (spec * spec_scale) * ndl;
                    sr = fmaf(term, L.color[0], sr);
                    sg = fmaf(term, L.color[1], sg);
                    sb = fmaf(term, L.color[2], sb);
                }
            }
        }
        const bool shadows = kShadows;
        /* on a lightmapped surface point lights act only through capsule
           shadows (sofren2.c:6574-6579): without proxies every light leaves
           lm_shadow at exactly 1, so the loop is skipped */
        const int n_eval = (lightmapped && !shadows) ? 0 : ts.count;
        for (int i = 0; i < n_eval; i += 1) {
            const PointLightGpu& L = ts.lights[i];
            const V3 lraw = sub3(V3{L.pos[0], L.pos[1], L.pos[2]}, pos);
            const f32 d2 = dot3(lraw, lraw);
            f32 att = smax(0.f, fmaf(-d2, L.inv_r2, 1.f));
            att = att * att;
            const bool active = att > 0.f;
            /* the reference skips a light no lane of the block reaches; a lane
               with zero attenuation gains nothing from it either, so the skip
               is taken only when the whole wave is out of reach */
            if (0ull == __ballot(active)) continue;
            const u32 np = ts.np[i];
            /* capsule ray tests of the reference: the light's cached capsules
               against this pixel's segment to the light (counted from bit 8) */
            if (kShadows && active) fetch += np << 8;
            const u32 near_caps = (kShadows && np > 0u) ? near_capsules(cm, i) : 0u;
            /* no capsule near: a lightmapped lane keeps lm_shadow exactly (its
               shadow intensity is 1); only a lane with a NaN attenuation would
               turn it into NaN */
            /* the light's direction: every lane needs it for diffuse and
               specular or for a capsule; lightmapped lanes only for the latter */
            V3 Ld{0.f, 0.f, 0.f};
            f32 ndl = 0.f;
            bool lit = false;
            if (!kShadows || !lightmapped) {
                Ld = est.norm_fast(lraw);
                ndl = smax(dot3(N, Ld), 0.f);
                lit = (ndl > 0.f) && active;
            }
            /* capsules with an active ray on some lane of the wave (first pass,
               the same test as below) */
            u32 hit_caps = 0u;
            f32 inv_d2 = 0.f;
            if (kShadows && near_caps != 0u) {
                inv_d2 = est.rcp_refined(d2);
                const bool gate = lightmapped || lit;
                /* two capsules per step, their loads issued together */
                for (u32 rest = near_caps; rest != 0u;) {
                    const u32 k0 = (u32)__ffs(rest) - 1u;
                    rest &= rest - 1u;
                    const u32 k1 = (0u != rest) ? (u32)__ffs(rest) - 1u : k0;
                    rest &= rest - 1u;
                    const bool n0 = ray_near(ts.prox[i][k0], lraw, d2, inv_d2) & gate;
                    const bool n1 = ray_near(ts.prox[i][k1], lraw, d2, inv_d2) & gate;
                    if (0ull != __ballot(n0)) hit_caps |= 1u << k0;
                    if (0ull != __ballot(n1)) hit_caps |= 1u << k1;
                }
            }
            /* no capsule reached: a lightmapped lane keeps lm_shadow exactly (its
               shadow intensity is 1); only a lane with a NaN attenuation would
               turn it into NaN */
            if (kShadows && 0u == hit_caps && 0ull == __ballot(!lightmapped || !(att == att))) continue;

            /* capsule soft shadows (6509-6568): closest approach of the pixel-
               to-light segment and each cached capsule, smooth-max blended.
               Capsules outside hit_caps have no active ray on any lane of the
               wave, so the reference's per-group skip would pass over them. */
            f32 max_shadow = 0.f;
            if (kShadows && hit_caps != 0u) {
                if (lightmapped) Ld = est.norm_fast(lraw);
                const f32 lrl = sqrtf(d2);
                for (u32 rest = hit_caps; rest != 0u; rest &= rest - 1u) {
                    const u32 k = (u32)__ffs(rest) - 1u;
                    const ProxyRec& c = ts.prox[i][k];
                    const bool ray_active = ray_near(c, lraw, d2, inv_d2) && (lightmapped || lit);
                    const u64 am = __ballot(ray_active);
                    /* a capsule no lane of the group reaches is skipped for the group */
                    if (0u == ((u32)(am >> block_lane) & group)) continue;
                    const V3 pa{c.pa[0], c.pa[1], c.pa[2]}, ab{c.ab[0], c.ab[1], c.ab[2]};
                    const V3 ao = sub3(pos, pa);
                    const f32 abl = dot3(ab, Ld), abao = dot3(ab, ao), lao = dot3(Ld, ao);
                    const f32 den = fmaf(-abl, abl, c.abdab);
                    const f32 sden = smax(den, SFT_EPSILON);
                    f32 t2 = fmaf(-abl, lao, abao) * est.rcp_refined(sden);
                    t2 = smax(0.f, smin(t2, 1.f));
                    const f32 t1 = fmaf(t2, abl, -lao);
                    const bool caster = t1 > SFT_EPSILON && t1 < lrl;
                    const V3 pr{pos.x + Ld.x * t1, pos.y + Ld.y * t1, pos.z + Ld.z * t1};
                    const V3 pc{pa.x + ab.x * t2, pa.y + ab.y * t2, pa.z + ab.z * t2};
                    const V3 dd = sub3(pr, pc);
                    const f32 dist2 = dot3(dd, dd);
                    f32 sv = smax(0.f, 1.f - dist2 / c.pr2);
                    if (!ray_active || !caster) sv = 0.f;
                    sv = sv * c.vis_mult;
                    const f32 diff = max_shadow - sv;
                    const f32 h = smax(0.f, 0.2f - fabsf(diff));
                    const f32 blend = (h * h) * (0.25f / 0.2f);
                    max_shadow = smax(max_shadow, sv) + blend;
                }
            }
            const f32 si = smax(0.f, 1.f - max_shadow);
            if (lightmapped) {
                const f32 drop = (1.f - si) * att;
                lm_shadow = lm_shadow * (1.f - drop);
            }
            else {
                const f32 di = ndl * att;
                const f32 mdi = lit ? di * si : 0.f;
                dr = fmaf(mdi, L.color[0], dr);
                dg = fmaf(mdi, L.color[1], dg);
                db = fmaf(mdi, L.color[2], db);
                if (lit) {
                    const V3 H = est.norm_fast(add3(Ld, V));
                    const f32 ndh = smax(dot3(N, H), 0.f);
                    const f32 den = (shin - shin * ndh) + ndh;
                    const f32 spec = ndh * est.rcp_refined(den);
                    const f32 term = ((spec * spec_scale) * di) * si;
                    sr = fmaf(term, L.color[0], sr);
                    sg = fmaf(term, L.color[1], sg);
                    sb = fmaf(term, L.color[2], sb);
                }
            }
        }
        if (lightmapped) {
            const f32 floor_ = 0.4f;
            const f32 mod = fmaf(lm_shadow, 1.f - floor_, floor_);
            dr = dr * mod; dg = dg * mod; db = db * mod;
            fr = alb_r * dr; fg = alb_g * dg; fb = alb_b * db;
        }
        else {
            fr = fmaf(alb_r, dr, sr);
            fg = fmaf(alb_g, dg, sg);
            fb = fmaf(alb_b, db, sb);
        }
    }
    return V3{smax(0.f, smin(fr, 1.f)), smax(0.f, smin(fg, 1.f)), smax(0.f, smin(fb, 1.f))};
}

DEV u32 pack_argb(f32 r, f32 g, f32 b) {
    return 0xFF000000u | ((u32)cvtt(r) << 16) | ((u32)cvtt(g) << 8) | (u32)cvtt(b);
}

/* tile light culling for one light against the tile's 1/w range and screen
   rectangle (sofren2.c:5862-5933); scalar float math in C order */
DEV bool light_in_tile(const FrameConst& fc, const PointLightGpu& L, f32 tmin, f32 tmax, i32 tile_x, i32 tile_y) {
    const f32* vp = fc.vp;
    const f32 px = L.pos[0], py = L.pos[1], pz = L.pos[2];
    const f32 radius = L.radius;
    const f32 w = px * vp[3] + py * vp[7] + pz * vp[11] + vp[15];
    const f32 lmin = w - radius;
    const f32 lmax = w + radius;
    if (lmax <= 1e-6f) return false;
    const f32 safe_min = (lmin > 1e-6f) ? lmin : 1e-6f;
    const f32 inv_min = 1.f / lmax;
    const f32 inv_max = 1.f / safe_min;
    if (inv_max < tmin || inv_min > tmax) return false;
    if (w > radius) {
        const f32 tx = px * vp[0] + py * vp[4] + pz * vp[8] + vp[12];
        const f32 ty = px * vp[1] + py * vp[5] + pz * vp[9] + vp[13];
        const f32 sx = (tx / w + 1.f) * fc.hw + fc.vp_x;
        const f32 sy = (-ty / w + 1.f) * fc.hh + fc.vp_y;
        const f32 psx = sqrtf(vp[0] * vp[0] + vp[4] * vp[4] + vp[8] * vp[8]);
        const f32 psy = sqrtf(vp[1] * vp[1] + vp[5] * vp[5] + vp[9] * vp[9]);
        const f32 safe_w = ((w - radius) > 1e-6f) ? (w - radius) : 1e-6f;
        const f32 rx = ((radius * psx) + fabsf(tx) * (radius / w)) / safe_w * fc.hw;
        const f32 ry = ((radius * psy) + fabsf(ty) * (radius / w)) / safe_w * fc.hh;
        const f32 l_min_x = sx - rx, l_max_x = sx + rx;
        const f32 l_min_y = sy - ry, l_max_y = sy + ry;
        const f32 t_min_x = (f32)tile_x, t_max_x = (f32)(tile_x + 64);
        const f32 t_min_y = (f32)tile_y, t_max_y = (f32)(tile_y + 64);
        if (l_max_x < t_min_x || l_min_x > t_max_x || l_max_y < t_min_y || l_min_y > t_max_y) return false;
    }
    return true;
}

/* first 16 point lights touching the tile, in index order, by one wave
   (sofren2.c:5862-5934) */
/* `first` (optional): light `lane` of the view, loaded by the caller ahead of
   time, for the first 64 lights */
DEV void build_tile_lights(const FramePtrs& P, const ViewGpu& V, u32 tile, f32 zmin, f32 zmax, int lane,
                           const PointLightGpu* first = nullptr) {
    const FrameConst& fc = V.fc;
    const u32 local = tile - V.tile_base;
    const i32 tmx = (i32)(local % (u32)fc.tile_cols) * 64, tmy = (i32)(local / (u32)fc.tile_cols) * 64;
    u32* tl = P.tile_lights + (u64)tile * (kMaxTileLights + 1);
    u32 n = 0;
    bool cut = false;
    if (zmin <= zmax) {
        for (u32 base = 0; base < fc.n_point && n < (u32)kMaxTileLights; base += 64u) {
            const u32 l = base + (u32)lane;
            const bool pass = l < fc.n_point &&
                light_in_tile(fc, (first && 0u == base) ? *first : P.point_lights[V.pl_base + l], zmin, zmax, tmx, tmy);
            const u64 m = __ballot(pass);
            const u32 rank = n + lanes_below(m, lane);
            if (pass && rank < (u32)kMaxTileLights) tl[1 + rank] = l;
            cut = cut || n + (u32)__popcll(m) > (u32)kMaxTileLights;
            n = min(n + (u32)__popcll(m), (u32)kMaxTileLights);
        }
    }
    if (0 == lane) {
        tl[0] = n;
        if (cut) stat_add(P.counters, CNT_LIGHT_CUT, 1u, stat_slot());
    }
}

/* Visibility pass.  A work item is (tile, chunk of at most kRasterMaxItem
   payloads of the tile's list, staged in LDS kRasterMaxChunk at a time into
   the same plane), with MSAA (tile, sample pair, chunk), staged 32 payloads
   at a time into the pair's two planes; k_raster_plan lists the items,
   and a persistent k_raster takes them from a queue, so a tile covered by
   thousands of payloads is spread over many workgroups.  Each item rasterises
   into an LDS plane; a tile with one chunk publishes from LDS, the chunks of a
   split tile merge with 64-bit atomic max into a zeroed HBM plane and the
   last one to finish publishes from there.  Publishing writes per-sample
   reference IDs and the plane's 1/w range; without MSAA it also builds the
   tile's light list (sofren2.c:6756-6813, 5820-5934). */
constexpr int kRasterThreads = 512;
constexpr int kRasterWaves = kRasterThreads / 64;

DEV u32 tile_payloads(const FramePtrs& P, u32 tile, u32& off) {
    const u32 cap = P.batch->tile_list_cap;
    off = P.tile_off[tile];
    u32 cnt = P.tile_count[tile];
    if (off >= cap) cnt = 0;
    else if (off + cnt > cap) cnt = cap - off;
    return cnt;
}

/* one workgroup per tile: chunk count, work items, zeroed merge planes.
   Without MSAA an item is (tile, chunk); with MSAA (tile, chunk, first
   sample of a pair), in the second list, whose items rasterise two sample
   planes at once. */
__global__ __launch_bounds__(256) void k_raster_plan(FramePtrs P) {
    __shared__ u32 s_at;
    const u32 tile = blockIdx.x;
    const ViewGpu& V = P.views[P.tile_view[tile]];
    const u32 S = (u32)V.fc.samples, K = min(max(V.raster_chunk, 1u), kRasterMaxItem), maxS = P.batch->max_samples;
    const bool pairs = S > 1u;
    const u32 per = pairs ? S / 2u : 1u;
    u32 off;
    const u32 cnt = tile_payloads(P, tile, off);
    const u32 nch = max(1u, (cnt + K - 1u) / K);
    if (0 == threadIdx.x) {
        s_at = atomicAdd(&P.counters[pairs ? CNT_RASTER_ITEMS2 : CNT_RASTER_ITEMS], nch * per);
        P.tile_nch[tile] = nch;
    }
    __syncthreads();
    uint2* items = P.raster_items + (pairs ? P.batch->items2_base : 0u);
    for (u32 i = threadIdx.x; i < nch * per; i += 256u) {
        items[s_at + i] = make_uint2(tile, ((i / per) << 2) | (pairs ? 2u * (i % per) : 0u));
    }
    if (nch > 1u) {
        u64* m = P.merge + (u64)tile * maxS * SFT_TILE_PIXELS;
        for (u32 i = threadIdx.x; i < S * (u32)SFT_TILE_PIXELS; i += 256u) m[i] = 0ull;
    }
}

/* per-pixel publish of one plane of a tile; `get(i)` is the packed
   (1/w bits << 32 | ~key) of pixel i */
template <class Get>
DEV void publish_plane(const FramePtrs& P, const ViewGpu& V, u32 tile, int sample, i32 tmx, i32 tmy, Get get,
                       f32 (*s_red)[kRasterWaves]) {
    const FrameConst& fc = V.fc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u32 n_clip = min(P.counters[CNT_CLIP_TRIS], P.batch->clip_cap);
    const u32 clip_base = view_clip_base(P, V);
    const i32 gw = fc.tile_cols * 64;
    const u64 plane = V.vid_base + (u64)sample * (u64)gw * (u64)(fc.tile_rows * 64);
    f32 zmin = 1e9f, zmax = -1e9f;
    /* the light records the list needs, loaded while the pixels are published */
    PointLightGpu light{};
    if (0 == wave && fc.samples < 2 && (u32)lane < fc.n_point) light = P.point_lights[V.pl_base + (u32)lane];
    /* all of a thread's pixels first, then their work-item loads together,
       so the thread waits on one round of loads rather than one per pixel */
    constexpr int kPer = SFT_TILE_PIXELS / kRasterThreads;
    u64 pk[kPer];
    uint2 ig[kPer];
    for (int k = 0; k < kPer; k += 1) pk[k] = get(tid + k * kRasterThreads);
    for (int k = 0; k < kPer; k += 1) {
        const u32 key = ~(u32)pk[k];
        ig[k] = make_uint2(0u, 0u);
        if (0ull != pk[k] && (key & 31u) >= 16u) {
            const WorkItem& w = P.work[key >> 9];
            ig[k] = make_uint2(w.inst, w.gm);
        }
    }
    for (int k = 0; k < kPer; k += 1) {
        const int i = tid + k * kRasterThreads;
        const f32 z = u2f((u32)(pk[k] >> 32));
        if (z > 0.0001f) {
            zmin = fminf(zmin, z);
            zmax = fmaxf(zmax, z);
        }
        const u64 gi = plane + (u64)(tmy + (i >> 6)) * gw + (tmx + (i & 63));
        const u32 key = ~(u32)pk[k];
        u32 vid = kInvalid;
        if (0ull != pk[k]) {
            vid = ((key & 31u) >= 16u)
                ? (ig[k].x << V.id_shift) | (ig[k].y << SFT_TRIANGLE_BITS) |
                  (((key >> 5) & 15u) * 8u + ((key & 31u) - 16u))
                : key_to_vid(P, key, n_clip, clip_base, V.id_shift);
        }
        P.vid[gi] = vid;
        if (P.cap_depth) P.cap_depth[gi] = z;
        if (P.vis64 && V.vis64) P.vis64[gi] = ((u64)(u32)(pk[k] >> 32) << 32) | vid;
    }
    for (int o = 32; o > 0; o >>= 1) {
        zmin = fminf(zmin, __shfl_xor(zmin, o));
        zmax = fmaxf(zmax, __shfl_xor(zmax, o));
    }
    if (0 == lane) { s_red[0][wave] = zmin; s_red[1][wave] = zmax; }
    __syncthreads();
    if (0 == wave) {
        zmin = s_red[0][0];
        zmax = s_red[1][0];
        for (int k = 1; k < kRasterWaves; k += 1) {
            zmin = fminf(zmin, s_red[0][k]);
            zmax = fmaxf(zmax, s_red[1][k]);
        }
        if (fc.samples > 1) {
            if (0 == lane) {
                P.tile_z[((u64)tile * P.batch->max_samples + (u32)sample) * 2u] = zmin;
                P.tile_z[((u64)tile * P.batch->max_samples + (u32)sample) * 2u + 1u] = zmax;
            }
        }
        else {
            build_tile_lights(P, V, tile, zmin, zmax, lane, &light);
        }
    }
}

/* screen positions, 1/w and emission key of one triangle of a payload;
   key kInvalid marks a lane-mask hole (keys never reach it: sub <= 23) */
struct RasterTri {
    f32 v[9];   /* sx0 sy0 iw0 sx1 sy1 iw1 sx2 sy2 iw2 */
    u32 key;
};

/* triangle g of the payload at tile-list position `at` */
DEV RasterTri fetch_raster_tri(const FramePtrs& P, u32 at, u32 g) {
    RasterTri tr;
    tr.key = kInvalid;
    const u64 desc = P.tile_list[at];
    const u32 mask = (u32)(desc >> 32) & 0xFFu;
    const u32 idx = (u32)desc;
    if (0u == mask) {
        if (0u != g) return tr;
        const f32* r = P.clip_rec + (u64)idx * 41u;
        tr.v[0] = r[0]; tr.v[3] = r[1]; tr.v[6] = r[2];
        tr.v[1] = r[3]; tr.v[4] = r[4]; tr.v[7] = r[5];
        tr.v[2] = r[6]; tr.v[5] = r[7]; tr.v[8] = r[8];
        tr.key = P.clip_key[idx];
        /* keeps these loads in this branch (merged with the meshlet loads
           they would need a 64-bit address each) */
        asm volatile("" ::: "memory");
        return tr;
    }
    if (0u == ((mask >> g) & 1u)) return tr;
    const u32 wi = idx >> 4, b = idx & 15u;
    const WorkItem wk = P.work[wi];
    const u32 ib = wk.slot * 512u + b * 32u + g;
    const u32 ss = wk.ss * 576u;
    for (u32 k = 0; k < 3u; k += 1) {
        const f32* vp = P.ss + (ss + (ld32(P.inds, ib + k * 8u) & 63u));
        tr.v[k * 3u] = vp[0];
        tr.v[k * 3u + 1u] = vp[64];
        tr.v[k * 3u + 2u] = vp[128];
    }
    tr.key = (wi << 9) | (b << 5) | (16u + g);
    return tr;
}

/* bounding box of a staged triangle clamped to the tile, in rows and 8-px
   chunks; false when it misses the tile (sofren2.c:5236-5246) */
DEV bool tri_extent(const RasterTri& tr, i32 tmx, i32 tmy, i32& rows, i32& chunks) {
    if (kInvalid == tr.key) return false;
    const i32 mnx = cvtt(smax(smin(tr.v[0], smin(tr.v[3], tr.v[6])), (f32)tmx));
    const i32 mxx = cvtt(smin(smax(tr.v[0], smax(tr.v[3], tr.v[6])), (f32)(tmx + 63)));
    const i32 mny = cvtt(smax(smin(tr.v[1], smin(tr.v[4], tr.v[7])), (f32)tmy));
    const i32 mxy = cvtt(smin(smax(tr.v[1], smax(tr.v[4], tr.v[7])), (f32)(tmy + 63)));
    if (mnx > mxx || mny > mxy) return false;
    rows = mxy - mny + 1;
    chunks = (((mxx & ~7) + 8) - (mnx & ~7)) >> 3;
    return true;
}

/* Work of one item's rasterisation: a unit is a staged triangle that
   reaches the tile, or one of K row parts of a triangle whose estimated
   sweep exceeds the split threshold (part k of K takes every K-th of its
   rows).  Units are listed by estimated cost per unit, the most expensive
   class first, and a wave takes 8 consecutive units at a time from an LDS
   counter, one per group of 8 lanes: a wave's groups then sweep for about
   as long as each other.  The visibility plane's 64-bit atomic max makes
   the result independent of which group sweeps which unit. */
/* kPl sample planes per item: 1 for the views without MSAA, 2 (a sample
   pair, whose planes share the staging, the unit list and every edge and
   depth step; each sample adds its own offsets to them) for MSAA views,
   which stage 32 payloads at a time to leave LDS for the second plane.
   Four waves per SIMD (at most 128 VGPRs): with two 512-thread groups per CU
   (about 78 KB of LDS each) any more registers would leave one group per CU */
template <int kPl>
__global__ __launch_bounds__(kRasterThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_raster(FramePtrs P) {
    constexpr u32 kStage = (1 == kPl) ? kRasterMaxChunk : kRasterMaxChunk / 4u;   /* payloads staged at a time */
    __shared__ u64 s_vis[kPl * SFT_TILE_PIXELS];
    /* the work item's triangles, fetched by the whole group at once so the
       waves rasterise from LDS instead of each waiting on its own chain of
       descriptor, work-item, index and vertex loads */
    __shared__ RasterTri s_tri[kStage * 8];
    /* units: staged triangle | part << 10 | log2(parts) << 13; a sub-chunk
       whose split units exceed the list sweeps each triangle as one unit */
    constexpr u32 kUnitCap = kStage * 16u;
    __shared__ u16 s_units[kUnitCap];
    __shared__ u32 s_hist[32], s_cur[16], s_split;
    /* each staged triangle's rows | chunks << 8 in the tile, 0 if it misses it */
    __shared__ u16 s_ext[kStage * 8];
    __shared__ f32 s_red[2][kRasterWaves];
    __shared__ u32 s_item, s_last, s_njobs, s_next;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Est est = make_est(P);
    const u32 n_items = P.counters[(1 == kPl) ? CNT_RASTER_ITEMS : CNT_RASTER_ITEMS2];
    const uint2* items = P.raster_items + ((1 == kPl) ? 0u : P.batch->items2_base);
    const u32 maxS = P.batch->max_samples;
    for (;;) {
        if (0 == tid) s_item = atomicAdd(&P.counters[(1 == kPl) ? CNT_RASTER_NEXT : CNT_RASTER_NEXT2], 1u);
        for (int i = tid; i < kPl * SFT_TILE_PIXELS; i += kRasterThreads) s_vis[i] = 0ull;
        __syncthreads();
        const u32 it = s_item;
        if (it >= n_items) break;
        const uint2 w = items[it];
        const u32 tile = w.x, chunk = w.y >> 2;
        const int sample = (int)(w.y & 3u);
        const ViewGpu& V = P.views[P.tile_view[tile]];
        const FrameConst& fc = V.fc;
        f32 ox[kPl], oy[kPl];
        for (int k = 0; k < kPl; k += 1) {
            ox[k] = c_msaa_ox[msaa_row(fc.samples)][sample + k];
            oy[k] = c_msaa_oy[msaa_row(fc.samples)][sample + k];
        }
        const u32 local = tile - V.tile_base;
        const i32 tmx = (i32)(local % (u32)fc.tile_cols) * 64, tmy = (i32)(local / (u32)fc.tile_cols) * 64;
        const u32 K = min(max(V.raster_chunk, 1u), kRasterMaxItem);

        u32 off;
        const u32 cnt = tile_payloads(P, tile, off);
        const u32 tile_chunks = P.tile_nch[tile];
        const u32 e0 = chunk * K, e1 = min(cnt, e0 + K);
        /* the item's payloads staged kStage at a time into the same planes */
        for (u32 c0 = e0;; c0 += kStage) {
        const u32 c1 = min(e1, c0 + kStage);
        const u32 n_tri = (c1 > c0) ? (c1 - c0) * 8u : 0u;
        for (u32 q = (u32)tid; q < n_tri; q += kRasterThreads) {
            const RasterTri tr = fetch_raster_tri(P, off + c0 + (q >> 3), q & 7u);
            s_tri[q] = tr;
            i32 rows = 0, chunks = 0;
            s_ext[q] = tri_extent(tr, tmx, tmy, rows, chunks) ? (u16)(rows | (chunks << 8)) : (u16)0;
        }
        if (0 == tid) s_next = 0u;
        __syncthreads();
        /* units: each staged triangle that reaches the tile, as 1, 2, 4 or 8
           row parts (part k of K takes every K-th row) when its estimated
           sweep exceeds the split threshold, listed by cost class (the
           largest first) so that the 8 units of a wave's job cost about the
           same; more units than the list holds: one part per triangle */
        if (tid < 32) s_hist[tid] = 0u;
        __syncthreads();
        u32 ucls[2] = {0u, 0u}, ucls1[2] = {0u, 0u}, upl[2] = {0u, 0u};
        for (int h = 0; h < 2; h += 1) {
            const u32 q = (u32)tid + (u32)h * kRasterThreads;
            if (q < n_tri && 0u != s_ext[q]) {
                const u32 e = s_ext[q], rows = e & 0xFFu, chunks = e >> 8;
                /* a sample pair's unit sweeps two planes: twice the cost */
                const u32 cost = (rows * 4u + rows * chunks * 20u) * (u32)kPl;
                u32 pl = 0u;
                while (pl < 3u && (cost >> pl) > V.raster_split) pl += 1u;
                upl[h] = pl;
                ucls[h] = min(15u, 31u - (u32)__clz((int)max(cost >> pl, 1u)));
                ucls1[h] = min(15u, 31u - (u32)__clz((int)max(cost, 1u)));
                atomicAdd(&s_hist[ucls[h]], 1u << pl);
                atomicAdd(&s_hist[16u + ucls1[h]], 1u);
            }
        }
        __syncthreads();
        if (0 == tid) {
            u32 t0 = 0u, t1 = 0u;
            for (int c = 0; c < 16; c += 1) { t0 += s_hist[c]; t1 += s_hist[16 + c]; }
            const bool split = t0 <= kUnitCap;
            u32 at = 0u;
            for (int c = 15; c >= 0; c -= 1) {
                s_cur[c] = at;
                at += split ? s_hist[c] : s_hist[16 + c];
            }
            s_njobs = at;
            s_split = split ? 1u : 0u;
            s_next = 0u;
        }
        __syncthreads();
        {
            const bool split = 0u != s_split;
            for (int h = 0; h < 2; h += 1) {
                const u32 q = (u32)tid + (u32)h * kRasterThreads;
                if (q < n_tri && 0u != s_ext[q]) {
                    const u32 pl = split ? upl[h] : 0u, n = 1u << pl;
                    const u32 at = atomicAdd(&s_cur[split ? ucls[h] : ucls1[h]], n);
                    for (u32 k = 0; k < n; k += 1) s_units[at + k] = (u16)(q | (k << 10) | (pl << 13));
                }
            }
        }
        __syncthreads();
        /* a wave per job of 8 consecutive units, 8 lanes per group, lane =
           x inside the 8-px chunk; group g sweeps unit 8 * job + g */
        const int g = lane >> 3, j = lane & 7;
        const u32 n_units = s_njobs, n_jobs = (n_units + 7u) >> 3;
        for (;;) {
            u32 jn = 0u;
            if (0 == lane) jn = atomicAdd(&s_next, 1u);
            jn = __shfl(jn, 0);
            if (jn >= n_jobs) break;
            const u32 un = jn * 8u + (u32)g;
            if (un >= n_units) continue;
            const u32 unit = s_units[un];
            const u32 tq = unit & 1023u, part = (unit >> 10) & 7u, plog = (unit >> 13) & 3u;
            const RasterTri& tr = s_tri[tq];
            const u32 key = tr.key;
            const f32 sx0 = tr.v[0], sy0 = tr.v[1], iw0 = tr.v[2];
            const f32 sx1 = tr.v[3], sy1 = tr.v[4], iw1 = tr.v[5];
            const f32 sx2 = tr.v[6], sy2 = tr.v[7], iw2 = tr.v[8];
            const int stride = 1 << plog, phase = (int)part;
            /* the items of MSAA views are the sample pairs (kPl 2) */
            raster_tri<(kPl > 1), kPl>(s_vis, est, tmx, tmy, j, g, key, sx0, sy0, iw0, sx1, sy1, iw1, sx2, sy2, iw2, ox, oy,
                                       stride, phase);
        }
        __syncthreads();
        if (c1 >= e1) break;
        }
        __syncthreads();

        if (1u == tile_chunks) {
            for (int k = 0; k < kPl; k += 1) {
                if (k > 0) __syncthreads();   /* the previous plane's 1/w range is read from s_red */
                const u64* vk = s_vis + k * SFT_TILE_PIXELS;
                publish_plane(P, V, tile, sample + k, tmx, tmy, [&](int i) { return vk[i]; }, s_red);
            }
        }
        else {
            /* the sample planes of a split tile merge in HBM; the item pair's
               completions count on its first sample */
            u64* m = P.merge + ((u64)tile * maxS + (u32)sample) * SFT_TILE_PIXELS;
            for (int i = tid; i < kPl * SFT_TILE_PIXELS; i += kRasterThreads) {
                const u64 v = s_vis[i];
                if (v) atomicMax((unsigned long long*)&m[i], (unsigned long long)v);
            }
            __threadfence();
            __syncthreads();
            if (0 == tid) {
                const u32 done = atomicAdd(&P.tile_done[(u64)tile * maxS + (u32)sample], 1u);
                s_last = (done + 1u == tile_chunks) ? 1u : 0u;
            }
            __syncthreads();
            if (s_last) {
                __threadfence();
                for (int k = 0; k < kPl; k += 1) {
                    if (k > 0) __syncthreads();
                    const u64* mk = m + k * SFT_TILE_PIXELS;
                    publish_plane(P, V, tile, sample + k, tmx, tmy,
                                  [&](int i) { return __hip_atomic_load(&mk[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); },
                                  s_red);
                }
            }
        }
        __syncthreads();
    }
}

/* MSAA: the tile's 1/w range over all sample planes, then its light list */
__global__ __launch_bounds__(64) void k_tile_lights(FramePtrs P) {
    const u32 tile = blockIdx.x;
    const ViewGpu& V = P.views[P.tile_view[tile]];
    const FrameConst& fc = V.fc;
    if (fc.samples < 2) return;
    const int lane = threadIdx.x;
    f32 zmin = 1e9f, zmax = -1e9f;
    for (int s = 0; s < fc.samples; s += 1) {
        zmin = fminf(zmin, P.tile_z[((u64)tile * P.batch->max_samples + (u32)s) * 2u]);
        zmax = fmaxf(zmax, P.tile_z[((u64)tile * P.batch->max_samples + (u32)s) * 2u + 1u]);
    }
    build_tile_lights(P, V, tile, zmin, zmax, lane);
}

/* the wave's shading counters from each lane's `fetch` word of shade_pixel
   (bits: 1 bilinear, 2 nearest, 4 lightmap, 8 non-power-of-two albedo, 16
   unlit, 32 directional lights; from bit 8 the capsule ray tests); `act`
   holds the lanes that shaded */
DEV void count_fetches(const FramePtrs& P, u32 fetch, u64 act, int lane) {
    static constexpr int kCnt[6] = { CNT_TEX_BILINEAR, CNT_TEX_NEAREST, CNT_LIGHTMAPPED, CNT_TEX_NPOT, CNT_UNLIT,
                                     CNT_DIR_LIT };
    u32 n[6];
    for (int k = 0; k < 6; k += 1) n[k] = (u32)__popcll(__ballot((fetch >> k) & 1u));
    const u32 caps = (u32)wave_sum((i32)(fetch >> 8));
    if (lane == __ffsll((long long)act) - 1) {
        const u32 sl = stat_slot();
        for (int k = 0; k < 6; k += 1) stat_add(P.counters, kCnt[k], n[k], sl);
        stat_add(P.counters, CNT_CAP_TESTS, caps, sl);
    }
}

/* Colour streamed to host memory: a system-scope store, written through to
   host memory (no cache line of it stays dirty in an L2, so no cache
   write-back has to precede the band's flag). */
DEV void host_store(u32* p, u32 v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }

/* Views whose colour streams to host memory: every thread of a shading
   block (256 per block, 16 blocks per tile) counts itself done after its
   host store, the lanes of a wave that end at the same point together once
   their wave's stores have completed (the wait covers all of the wave's
   stores, those of lanes that ended earlier included), in the block's LDS
   counter `lanes` (zeroed before the block's first barrier).  The block's
   last lanes count the block done for its band (the view's tile row), and
   the band's last block raises the band's host flag: every store of the
   band completed before its block's count, so the host, reading the flag,
   finds the band's pixels in place and copies them out while later bands
   shade. */
DEV void band_lanes_done(const FramePtrs& P, const ViewGpu& VW, u32 local_tile, u32* lanes) {
    __builtin_amdgcn_s_waitcnt(0);
    const u64 act = __ballot(true);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)act) - 1) {
        const u32 n = (u32)__popcll(act);
        if (atomicAdd(lanes, n) + n == 256u) {
            const u32 band = VW.band_base + local_tile / (u32)VW.fc.tile_cols;
            const u32 need = (u32)VW.fc.tile_cols * 16u;
            if (atomicAdd(&P.band_count[band], 1u) + 1u == need) {
                /* the overflow bits are final: the passes that set them ran before */
                const u32 ovf = __hip_atomic_load(&P.counters[CNT_OVERFLOW], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                host_store(&P.band_flag[band], (0u == ovf) ? 1u : 2u);
            }
        }
    }
}

/* per-pixel deferred shading; a block covers 32x8 pixels of one tile, a wave a
   32x2 strip of eight 4x2 blocks, lane = pixel in the reference's block lane
   order (x offsets 0..3, rows 0..1) */
template <bool kShadows>
__global__ __launch_bounds__(256) void k_shade(FramePtrs P) {
    __shared__ TileShading<kShadows> s_ts;
    __shared__ u32 s_li[kMaxTileLights];
    __shared__ u32 s_band_lanes;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (0 == tid) s_band_lanes = 0u;
    /* 16 blocks of 32x8 pixels per tile; tiles of views with shadows last */
    const u32 run = kShadows ? P.batch->n_shade_tiles - P.batch->n_shade_sh : 0u;
    const u32 tile = P.shade_tiles[run + (blockIdx.x >> 4)];
    const ViewGpu& VW = P.views[P.tile_view[tile]];
    const FrameConst& fc = VW.fc;
    const u32 local = tile - VW.tile_base, sub = blockIdx.x & 15u;
    const i32 x0 = (i32)(local % (u32)fc.tile_cols) * 64 + (i32)(sub & 1u) * 32;
    const i32 y0 = (i32)(local / (u32)fc.tile_cols) * 64 + (i32)(sub >> 1) * 8;
    const Est est = make_est(P);
    stage_tile_shading(P, VW, tile, s_ts, s_li);
    const TileShading<kShadows>& tl = s_ts;

    const int blk = lane >> 3, l = lane & 7, ix = l & 3, iy = l >> 2;
    const i32 bx = x0 + blk * 4, by = y0 + wave * 2;
    const i32 gx = bx + ix, gy = by + iy;
    const i32 gw = fc.tile_cols * 64;
    const bool in_bounds = gx < fc.screen_w && gy < fc.screen_h;
    const u32 vid = P.vid[VW.vid_base + (u64)gy * gw + gx];
    const bool valid = in_bounds && kInvalid != vid;
    const u32 mine = valid ? vid : kInvalid;

    /* lanes of the block holding the same triangle; the first is the LOD lane */
    u32 group = 0;
    for (int k = 0; k < 8; k += 1) {
        if (__shfl(mine, (lane & ~7) | k) == mine) group |= 1u << k;
    }
    const int first = __ffs(group) - 1;
    const f32 fbx = (f32)bx + 0.5f, fby = (f32)by + 0.5f;
    Tri t;
    Bary b{0.f, 0.f, 0.f};
    f32 cu = 0.f, cv = 0.f;
    if (valid) {
        load_tri(P, VW, vid, t);
        b = bary_at(est, t, fbx + (f32)ix, fby + (f32)iy);
        cu = interp(b, t.u);
        cv = interp(b, t.v);
    }
    /* UV derivatives at the LOD lane: ddx pairs lanes l^1, ddy rows l^4
       (sofren2.c:272-273).  A neighbour covered by the same triangle already
       holds the value; otherwise it is evaluated here under this triangle. */
    const int src0 = (lane & ~7) | first;
    const int src1 = (lane & ~7) | (first ^ 1);
    const int src4 = (lane & ~7) | (first ^ 4);
    const f32 cu0 = __shfl(cu, src0), cv0 = __shfl(cv, src0);
    f32 cu1 = __shfl(cu, src1), cv1 = __shfl(cv, src1);
    f32 cu4 = __shfl(cu, src4), cv4 = __shfl(cv, src4);
    const u32 vid1 = __shfl(mine, src1), vid4 = __shfl(mine, src4);
    /* capsules near this wave's shading positions (all lanes still run here) */
    CapsuleMasks cm{{0ull, 0ull, 0ull, 0ull}};
    if constexpr (kShadows) {
        const V3 p = valid ? V3{interp(b, t.px), interp(b, t.py), interp(b, t.pz)} : V3{0.f, 0.f, 0.f};
        cm = capsule_masks(tl, wave_span(p, valid), lane);
    }
    const bool streamed = kInvalid != VW.band_base;
    if (!in_bounds) {
        if (streamed) band_lanes_done(P, VW, local, &s_band_lanes);
        return;
    }
    const u64 px = VW.color_base + (u64)gy * fc.width + gx;
    u32* out = &P.color[px];
    if (!valid) {
        if (fc.should_clear) *out = fc.clear_color;
        if (streamed) {
            host_store(&P.color_host[px], *out);
            band_lanes_done(P, VW, local, &s_band_lanes);
        }
        return;
    }
    if (vid1 != vid) {
        const int f1 = first ^ 1;
        const Bary q = bary_at(est, t, fbx + (f32)(f1 & 3), fby + (f32)(f1 >> 2));
        cu1 = interp(q, t.u);
        cv1 = interp(q, t.v);
    }
    if (vid4 != vid) {
        const int f4 = first ^ 4;
        const Bary q = bary_at(est, t, fbx + (f32)(f4 & 3), fby + (f32)(f4 >> 2));
        cu4 = interp(q, t.u);
        cv4 = interp(q, t.v);
    }
    u32 fetch = 0;
    const V3 col = shade_pixel<kShadows>(P, VW, est, t, b, cu0 - cu1, cu0 - cu4, cv0 - cv1, cv0 - cv4, tl, fetch, group,
                                         lane & ~7, cm);
    *out = pack_argb(col.x * 255.f, col.y * 255.f, col.z * 255.f);
    if (streamed) host_store(&P.color_host[px], *out);

    i32 shaded = 1, groups = (first == l) ? 1 : 0;
    /* counters: one atomic per wave (inactive lanes contribute nothing) */
    const u64 act = __ballot(true);
    shaded = __popcll(act);
    groups = __popcll(__ballot(groups != 0));
    if (lane == __ffsll((long long)act) - 1) {
        const u32 sl = stat_slot();
        stat_add(P.counters, CNT_SHADED, (u32)shaded, sl);
        stat_add(P.counters, CNT_QGROUPS, (u32)groups, sl);
    }
    count_fetches(P, fetch, act, lane);
    if (streamed) band_lanes_done(P, VW, local, &s_band_lanes);
}

/* MSAA resolve (sofren2.c:6048-6135, 6677-6743).  Lanes are laid out as in
   k_shade.  A block's triangles are taken in the order of their first
   (sample, lane) occurrence, as the reference iterates them; every lane of
   the block evaluates the triangle at its pixel centre (clamped barycentrics)
   so the LOD lane's neighbours are at hand, the lanes covering it shade once
   and add the colour weighted by their share of samples.  Each pixel starts
   from the background colour times its empty-sample share.  The shading
   itself runs over a compacted list of (pixel, triangle) items, pooled over
   the workgroup's four waves, so that lanes whose pixel is done do not idle
   through another pixel's extra triangles.  Four waves per SIMD: the LDS
   below (about 39 KB with capsule records) admits four workgroups per CU,
   and the register budget is held to match. */
template <bool kShadows>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_shade_ms(FramePtrs P) {
    __shared__ TileShading<kShadows> s_ts;
    __shared__ u32 s_li[kMaxTileLights];
    /* per wave: up to 64 pixels x 4 samples (pixel, triangle) items */
    __shared__ u32 s_item_vid[4][kMsItems];
    __shared__ u32 s_item_meta[4][kMsItems];     /* pixel lane | group start << 8 | group size << 16 | samples << 24 */
    __shared__ f32 s_item_der[4][4][kMsItems];   /* UV derivatives at the group's first item, then each item's colour */
    __shared__ u32 s_item_count[4];
    __shared__ u32 s_band_lanes;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (0 == tid) s_band_lanes = 0u;
    /* 16 blocks of 32x8 pixels per tile; tiles of views with shadows last */
    const u32 run = P.batch->n_shade_tiles + (kShadows ? P.batch->n_shade_tiles_ms - P.batch->n_shade_ms_sh : 0u);
    const u32 tile = P.shade_tiles[run + (blockIdx.x >> 4)];
    const ViewGpu& VW = P.views[P.tile_view[tile]];
    const FrameConst& fc = VW.fc;
    const u32 local = tile - VW.tile_base, sub = blockIdx.x & 15u;
    const i32 x0 = (i32)(local % (u32)fc.tile_cols) * 64 + (i32)(sub & 1u) * 32;
    const i32 y0 = (i32)(local / (u32)fc.tile_cols) * 64 + (i32)(sub >> 1) * 8;
    const Est est = make_est(P);
    stage_tile_shading(P, VW, tile, s_ts, s_li);
    const TileShading<kShadows>& tl = s_ts;

    const int S = fc.samples;
    const f32 inv_s = 1.f / (f32)S;
    const int blk = lane >> 3, l = lane & 7, ix = l & 3, iy = l >> 2;
    const int base = lane & ~7;
    const i32 bx = x0 + blk * 4, by = y0 + wave * 2;
    const i32 gx = bx + ix, gy = by + iy;
    const i32 gw = fc.tile_cols * 64;
    const u64 plane = (u64)gw * (u64)(fc.tile_rows * 64);
    const bool in_bounds = gx < fc.screen_w && gy < fc.screen_h;
    const u32* vid_planes = P.vid + VW.vid_base;
    u32 sv[4] = {kInvalid, kInvalid, kInvalid, kInvalid};
    u32 mine = 0;
    i32 empty = 0;
    for (int s = 0; s < 4; s += 1) {
        if (s < S) {
            const u32 v = vid_planes[(u64)s * plane + (u64)gy * gw + gx];
            sv[s] = in_bounds ? v : kInvalid;
            if (in_bounds && kInvalid != v) mine |= 1u << (s * 8 + l);
            if (in_bounds && kInvalid == v) empty += 1;
        }
    }
    /* block-wide (sample, lane) occupancy, bit s * 8 + lane */
    u32 remaining = mine;
    remaining |= __shfl_xor(remaining, 1);
    remaining |= __shfl_xor(remaining, 2);
    remaining |= __shfl_xor(remaining, 4);
    u32* out = &P.color[VW.color_base + (u64)gy * fc.width + gx];
    /* a block without triangles stays in the wave (idle) so that every lane
       reaches the wave-wide capsule culling below */
    const bool idle = 0u == remaining;
    if (idle && in_bounds && fc.should_clear) *out = fc.clear_color;
    const f32 ew = (f32)empty * inv_s;
    f32 acc_r = 0.f, acc_g = 0.f, acc_b = 0.f;
    if (fc.should_clear) {
        const u32 cc = fc.clear_color;
        acc_r = unpack_ch(cc, 16) * ew;
        acc_g = unpack_ch(cc, 8) * ew;
        acc_b = unpack_ch(cc, 0) * ew;
    }
    else {
        const u32 any_ew = (u32)(__ballot(ew > 0.f) >> base) & 0xFFu;
        if (any_ew && !idle) {
            const u32 c = in_bounds ? *out : 0u;
            acc_r = unpack_ch(c, 16) * ew;
            acc_g = unpack_ch(c, 8) * ew;
            acc_b = unpack_ch(c, 0) * ew;
        }
    }
    const bool write = !idle && (fc.should_clear ? in_bounds : (0u != mine));
    const f32 fbx = (f32)bx + 0.5f, fby = (f32)by + 0.5f;
    i32 shaded = 0, groups = 0;
    /* six 5-bit counts of shade_pixel's fetch bits (a lane shades at most
       16 items: 4 waves x 256 items over 64 lanes) */
    u32 fetches = 0, cap_tests = 0;
    u32* it_vid = s_item_vid[wave];
    u32* it_meta = s_item_meta[wave];

    /* 1. the (block, triangle) groups: one triangle per block and round, the
       wave running rounds while any block has triangles left.  Each lane
       covered by the round's triangle gets an item (its pixel, the triangle,
       its sample count); a group's items are consecutive, groups follow in
       round then block order, so a pixel's items are in its triangle order.
       The group's first item keeps the LOD lane's UV derivatives. */
    u32 n_items = 0, my_items = 0, my_idx = 0;
    while (0ull != __ballot(0u != remaining)) {
        u32 hits = 0, target = 0;
        f32 dux = 0.f, duy = 0.f, dvx = 0.f, dvy = 0.f;
        int first = 0;
        u32 group = 0;
        if (0u != remaining) {
            /* next triangle: the first remaining (sample, lane) of the block */
            const int pos = __ffs(remaining) - 1;
            const int s0 = pos >> 3, l0 = pos & 7;
            const u32 own = (0 == s0) ? sv[0] : (1 == s0) ? sv[1] : (2 == s0) ? sv[2] : sv[3];
            target = __shfl(own, base | l0);
            u32 match = 0;
            for (int s = 0; s < 4; s += 1) {
                if (s < S && sv[s] == target) {
                    hits += 1;
                    match |= 1u << (s * 8 + l);
                }
            }
            match |= __shfl_xor(match, 1);
            match |= __shfl_xor(match, 2);
            match |= __shfl_xor(match, 4);
            remaining &= ~match;
            group = (u32)(__ballot(hits > 0u) >> base) & 0xFFu;
            Tri t;
            load_tri_uv(P, VW, target, t);
            const Bary b = bary_at(est, t, fbx + (f32)ix, fby + (f32)iy, true);
            const f32 cu = interp(b, t.u), cv = interp(b, t.v);
            first = __ffs(group) - 1;
            const f32 cu0 = __shfl(cu, base | first), cv0 = __shfl(cv, base | first);
            const f32 cu1 = __shfl(cu, base | (first ^ 1)), cv1 = __shfl(cv, base | (first ^ 1));
            const f32 cu4 = __shfl(cu, base | (first ^ 4)), cv4 = __shfl(cv, base | (first ^ 4));
            dux = cu0 - cu1; duy = cu0 - cu4; dvx = cv0 - cv1; dvy = cv0 - cv4;
        }
        const u64 hm = __ballot(hits > 0u);
        if (hits > 0u) {
            const u32 idx = n_items + lanes_below(hm, lane);
            const u32 gs = n_items + lanes_below(hm, base);
            it_vid[idx] = target;
            it_meta[idx] = (u32)lane | (gs << 8) | ((u32)__popc(group) << 16) | (hits << 24);
            if (l == first) {
                s_item_der[wave][0][gs] = dux; s_item_der[wave][1][gs] = duy;
                s_item_der[wave][2][gs] = dvx; s_item_der[wave][3][gs] = dvy;
            }
            my_idx |= idx << (8u * my_items);
            my_items += 1;
        }
        n_items += (u32)__popcll(hm);
    }
    if (0 == lane) s_item_count[wave] = n_items;
    __syncthreads();

    /* 2. the workgroup's items (wave 0's, then wave 1's, ...) shaded 64 at a
       time, chunk k by wave k mod 4; a chunk ends at a group boundary so that
       a group's lanes meet at the capsule ballots (lanes gs - c0 ...).  An
       item's colour replaces its UV derivatives in LDS: its group's
       derivatives are read by the chunk's lanes before any of them writes. */
    const u32 off1 = s_item_count[0], off2 = off1 + s_item_count[1], off3 = off2 + s_item_count[2];
    const u32 n_all = off3 + s_item_count[3];
    /* wave of workgroup item g, and that wave's first item */
    auto item_wave = [&](u32 g) { return (int)(g >= off1) + (int)(g >= off2) + (int)(g >= off3); };
    auto wave_first = [&](int w) { return (w == 0) ? 0u : (w == 1) ? off1 : (w == 2) ? off2 : off3; };
    u32 chunk = 0;
    for (u32 c0 = 0; c0 < n_all; chunk += 1) {
        u32 end = min(n_all, c0 + 64u);
        if (end < n_all) {
            const u32 g = end - 1u;
            const int w = item_wave(g);
            const u32 m = s_item_meta[w][g - wave_first(w)];
            const u32 gs = wave_first(w) + ((m >> 8) & 0xFFu), gn = (m >> 16) & 0xFFu;
            if (gs + gn > end) end = gs;
        }
        if ((int)(chunk & 3u) == wave) {
            const u32 item = c0 + (u32)lane;
            const bool has = item < end;
            const int iw = item_wave(item);
            const u32 li = item - wave_first(iw);
            u32 meta = 0;
            Tri t;
            Bary b{0.f, 0.f, 0.f};
            if (has) {
                meta = s_item_meta[iw][li];
                const int pl = (int)(meta & 63u), pix = pl & 7;
                load_tri(P, VW, s_item_vid[iw][li], t);
                b = bary_at(est, t, (f32)(x0 + (pl >> 3) * 4) + 0.5f + (f32)(pix & 3),
                            (f32)(y0 + iw * 2) + 0.5f + (f32)(pix >> 2), true);
            }
            /* capsules near this chunk's shading positions (every lane runs here) */
            CapsuleMasks cm{{0ull, 0ull, 0ull, 0ull}};
            if constexpr (kShadows) {
                const V3 p = has ? V3{interp(b, t.px), interp(b, t.py), interp(b, t.pz)} : V3{0.f, 0.f, 0.f};
                cm = capsule_masks(tl, wave_span(p, has), lane);
            }
            if (has) {
                const u32 gs = (meta >> 8) & 0xFFu, gn = (meta >> 16) & 0xFFu;
                u32 fetch = 0;
                const V3 col = shade_pixel<kShadows>(P, VW, est, t, b, s_item_der[iw][0][gs], s_item_der[iw][1][gs],
                                                     s_item_der[iw][2][gs], s_item_der[iw][3][gs], tl, fetch,
                                                     (1u << gn) - 1u, (int)(wave_first(iw) + gs - c0), cm);
                s_item_der[iw][0][li] = col.x;
                s_item_der[iw][1][li] = col.y;
                s_item_der[iw][2][li] = col.z;
                shaded += 1;
                groups += (li == gs) ? 1 : 0;
                for (int k = 0; k < 6; k += 1) fetches += ((fetch >> k) & 1u) << (5 * k);
                if constexpr (kShadows) cap_tests += fetch >> 8;
            }
        }
        c0 = end;
    }
    __syncthreads();

    /* 3. each pixel adds its triangles' colours in order, weighted by their
       share of samples */
    for (u32 j = 0; j < my_items; j += 1) {
        const u32 idx = (my_idx >> (8u * j)) & 0xFFu;
        const f32 w = (f32)(it_meta[idx] >> 24) * inv_s;
        acc_r = fmaf(s_item_der[wave][0][idx], w, acc_r);
        acc_g = fmaf(s_item_der[wave][1][idx], w, acc_g);
        acc_b = fmaf(s_item_der[wave][2][idx], w, acc_b);
    }
    if (write) *out = pack_argb(smin(acc_r, 1.f) * 255.f, smin(acc_g, 1.f) * 255.f, smin(acc_b, 1.f) * 255.f);
    const bool streamed = kInvalid != VW.band_base;
    if (streamed && in_bounds) {
        /* the host copy's address from `out` plus a uniform offset */
        const uintptr_t delta = (uintptr_t)P.color_host - (uintptr_t)P.color;
        host_store((u32*)((uintptr_t)out + delta), *out);
    }

    const u64 act = __ballot(true);
    shaded = wave_sum(shaded);
    groups = wave_sum(groups);
    const i32 blended = wave_sum((write && !fc.should_clear && ew > 0.f) ? 1 : 0);
    u32 fsum[6];
    for (int k = 0; k < 6; k += 1) fsum[k] = (u32)wave_sum((i32)((fetches >> (5 * k)) & 31u));
    const u32 caps = kShadows ? (u32)wave_sum((i32)cap_tests) : 0u;
    if (lane == __ffsll((long long)act) - 1) {
        const u32 sl = stat_slot();
        stat_add(P.counters, CNT_SHADED, (u32)shaded, sl);
        stat_add(P.counters, CNT_QGROUPS, (u32)groups, sl);
        stat_add(P.counters, CNT_BG_BLEND, (u32)blended, sl);
        static constexpr int kCnt[6] = { CNT_TEX_BILINEAR, CNT_TEX_NEAREST, CNT_LIGHTMAPPED, CNT_TEX_NPOT, CNT_UNLIT,
                                         CNT_DIR_LIT };
        for (int k = 0; k < 6; k += 1) stat_add(P.counters, kCnt[k], fsum[k], sl);
        stat_add(P.counters, CNT_CAP_TESTS, caps, sl);
    }
    if (streamed) band_lanes_done(P, VW, local, &s_band_lanes);
}

/* ================================================================ shadows
   Proxy occlusion (sofren2.c:8163-8280): for each light l < 64 and packet of 8
   proxies, 3 rays per proxy (from posa, posb and the midpoint) to 3 points on
   a disk of radius SFT_SHADOW_LIGHT_SIZE facing the packet; each ray that hits
   static geometry takes one from the proxy's visibility count for l.  A packet
   is traced by one wave: the packet frustum (4 side planes around the origins
   as seen from the light, a near cap at the light, a far cap at its radius,
   sofren2.c:3970-4084) culls 8-wide BVH nodes, 8 nodes x 8 children are box-
   tested per step across the 64 lanes, and leaf triangles are tested 8 at a
   time against the 8 rays (lane = triangle x ray).  A ray's result is the OR
   over every triangle the frustum reaches, so the visiting order (the
   reference's depth-first stack) does not change it. */
constexpr int kOccWaves = 3;

struct Frustum {
    f32 n[6][3], an[6][3], d[6];
};

template <int kStack>
struct OccLds {
    u32 stack[kStack];
    u32 tris[72];
    u32 insts[64];
};


/* origin of ray `ray` of a proxy: posa, posb or the midpoint (8222-8245) */
DEV sftvec ray_origin(const ProxyGpu& p, int ray) {
    if (0 == ray) return SFT_V(p.a[0], p.a[1], p.a[2]);
    if (1 == ray) return SFT_V(p.b[0], p.b[1], p.b[2]);
    return SFT_V((p.a[0] + p.b[0]) * 0.5f, (p.a[1] + p.b[1]) * 0.5f, (p.a[2] + p.b[2]) * 0.5f);
}

/* sft__build_packet_frustum (3970-4084); returns false where the reference
   leaves the side planes unset (a corner of the origin box behind the light) */
DEV bool packet_frustum(const ProxyGpu* px, u32 n, int ray, sftvec lp, f32 lrad, Frustum& fr) {
    f32 mnx = 1e9f, mxx = -1e9f, mny = 1e9f, mxy = -1e9f, mnz = 1e9f, mxz = -1e9f;
    sftvec c = SFT_V(0.f, 0.f, 0.f);
    for (u32 k = 0; k < n; k += 1) {
        const sftvec o = ray_origin(px[k], ray);
        if (o.x < mnx) mnx = o.x;
        if (o.x > mxx) mxx = o.x;
        if (o.y < mny) mny = o.y;
        if (o.y > mxy) mxy = o.y;
        if (o.z < mnz) mnz = o.z;
        if (o.z > mxz) mxz = o.z;
        c.x += o.x; c.y += o.y; c.z += o.z;
    }
    c.x /= (f32)n; c.y /= (f32)n; c.z /= (f32)n;
    const sftvec fwd = sft_v_norm(sft_v_sub(c, lp));
    const sftvec up = (fabsf(fwd.y) > 0.99f) ? SFT_V(1, 0, 0) : SFT_V(0, 1, 0);
    const sftmat lv = sft_m_view_look_at(lp, c, up);
    f32 mnxz = 1e9f, mxxz = -1e9f, mnyz = 1e9f, mxyz = -1e9f;
    bool valid = true;
    for (int i = 0; i < 8; i += 1) {
        const sftquat q = { (i & 1) ? mxx : mnx, (i & 2) ? mxy : mny, (i & 4) ? mxz : mnz, 1.f };
        const sftvec cl = sft_m_mul_quat(lv, q);
        if (cl.z < SFT_EPSILON) { valid = false; break; }
        const f32 xz = cl.x / cl.z, yz = cl.y / cl.z;
        if (xz < mnxz) mnxz = xz;
        if (xz > mxxz) mxxz = xz;
        if (yz < mnyz) mnyz = yz;
        if (yz > mxyz) mxyz = yz;
    }
    const sftvec lr = SFT_V(lv.m[0][0], lv.m[1][0], lv.m[2][0]);
    const sftvec lu = SFT_V(lv.m[0][1], lv.m[1][1], lv.m[2][1]);
    const sftvec lf = SFT_V(lv.m[0][2], lv.m[1][2], lv.m[2][2]);
    for (int p = 0; p < 4; p += 1) {
        fr.n[p][0] = fr.n[p][1] = fr.n[p][2] = 0.f;
        fr.an[p][0] = fr.an[p][1] = fr.an[p][2] = 0.f;
        fr.d[p] = 0.f;
    }
    if (valid) {
        const sftvec nl[4] = { sft_v_norm(SFT_V(1.f, 0.f, -mnxz)), sft_v_norm(SFT_V(-1.f, 0.f, mxxz)),
                               sft_v_norm(SFT_V(0.f, 1.f, -mnyz)), sft_v_norm(SFT_V(0.f, -1.f, mxyz)) };
        for (int p = 0; p < 4; p += 1) {
            const sftvec nw = sft_v_add(sft_v_mul(lr, nl[p].x), sft_v_add(sft_v_mul(lu, nl[p].y), sft_v_mul(lf, nl[p].z)));
            fr.n[p][0] = nw.x; fr.n[p][1] = nw.y; fr.n[p][2] = nw.z;
            fr.d[p] = -sft_v_dot(nw, lp) + SFT_SHADOW_LIGHT_SIZE;
            fr.an[p][0] = fabsf(nw.x); fr.an[p][1] = fabsf(nw.y); fr.an[p][2] = fabsf(nw.z);
        }
    }
    fr.n[4][0] = lf.x; fr.n[4][1] = lf.y; fr.n[4][2] = lf.z;
    fr.d[4] = -sft_v_dot(lf, lp) + SFT_SHADOW_LIGHT_SIZE;
    fr.an[4][0] = fabsf(lf.x); fr.an[4][1] = fabsf(lf.y); fr.an[4][2] = fabsf(lf.z);
    fr.n[5][0] = -lf.x; fr.n[5][1] = -lf.y; fr.n[5][2] = -lf.z;
    fr.d[5] = -sft_v_dot(SFT_V(-lf.x, -lf.y, -lf.z), lp) + lrad;
    fr.an[5][0] = fabsf(lf.x); fr.an[5][1] = fabsf(lf.y); fr.an[5][2] = fabsf(lf.z);
    return valid;
}

/* world frustum into a model's space through its transform M (4146-4169) */
DEV void model_frustum(const Frustum& fr, const f32* M, Frustum& mf) {
    for (int p = 0; p < 6; p += 1) {
        const f32 nx = fr.n[p][0], ny = fr.n[p][1], nz = fr.n[p][2], d = fr.d[p];
        const f32 nmx = nx * M[0] + ny * M[1] + nz * M[2];
        const f32 nmy = nx * M[4] + ny * M[5] + nz * M[6];
        const f32 nmz = nx * M[8] + ny * M[9] + nz * M[10];
        const f32 dm = (nx * M[12] + ny * M[13] + nz * M[14]) + d;
        const f32 il = 1.f / sqrtf(nmx * nmx + nmy * nmy + nmz * nmz);
        mf.n[p][0] = nmx * il; mf.n[p][1] = nmy * il; mf.n[p][2] = nmz * il;
        mf.d[p] = dm * il;
        mf.an[p][0] = fabsf(nmx * il); mf.an[p][1] = fabsf(nmy * il); mf.an[p][2] = fabsf(nmz * il);
    }
}

/* child box c of a node against the 6 planes, centre/extent form (3906-3968) */
DEV bool box_in_frustum(const sft_bvh_node* nd, int c, const Frustum& f) {
    const f32 mnx = nd->min_x[c], mny = nd->min_y[c], mnz = nd->min_z[c];
    const f32 mxx = nd->max_x[c], mxy = nd->max_y[c], mxz = nd->max_z[c];
    const f32 cx = (mxx + mnx) * 0.5f, cy = (mxy + mny) * 0.5f, cz = (mxz + mnz) * 0.5f;
    const f32 ex = (mxx - mnx) * 0.5f, ey = (mxy - mny) * 0.5f, ez = (mxz - mnz) * 0.5f;
    bool in = true;
    for (int p = 0; p < 6; p += 1) {
        const f32 r = fmaf(f.an[p][0], ex, fmaf(f.an[p][1], ey, f.an[p][2] * ez));
        const f32 dc = fmaf(f.n[p][0], cx, fmaf(f.n[p][1], cy, f.n[p][2] * cz));
        in = in && ((dc + f.d[p]) + r >= 0.f);
    }
    return in;
}

/* two-sided Moller-Trumbore, one ray (3840-3904) */
DEV bool ray_hits_tri(const Est& est, V3 v0, V3 v1, V3 v2, V3 o, V3 d, f32 tmax) {
    const V3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    const V3 h{d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x};
    const f32 a = dot3(e1, h);
    if (!(fabsf(a) > SFT_EPSILON)) return false;
    const f32 f = est.rcp_refined(a);
    const V3 s = sub3(o, v0);
    const f32 u = f * dot3(s, h);
    if (!(u >= 0.f && u <= 1.f)) return false;
    const V3 q{s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x};
    const f32 v = f * dot3(d, q);
    if (!(v >= 0.f && u + v <= 1.f)) return false;
    const f32 t = f * dot3(e2, q);
    return t > SFT_EPSILON && t < tmax;
}

/* Pops up to 8 nodes, box-tests their 8 children (lane = node x child) and
   pushes the internal hits; leaf hits are returned in `leaf`/`child`.  The pop
   count is limited so that pushes never pass the stack's end; a stack with no
   room left sets `overflow`. */
template <int kStack>
DEV void bvh_step(OccLds<kStack>& W, u32& top, const sft_bvh_node* nodes, u32 n_nodes, const Frustum& f, int lane,
                  bool& leaf, u32& child, u32& tested, bool& overflow) {
    u32 k = min(8u, top);
    k = min(k, ((u32)kStack - top) / 7u);
    leaf = false;
    child = kInvalid;
    if (0u == k) {
        overflow = true;
        top = 0;
        return;
    }
    const u32 slot = (u32)lane >> 3, c = (u32)lane & 7u;
    bool push = false;
    if (slot < k) {
        const sft_bvh_node* nd = nodes + W.stack[top - 1u - slot];
        child = nd->child_indices[c];
        if (kInvalid != child && box_in_frustum(nd, (int)c, f)) {
            if (nd->is_leaf_group) leaf = true;
            else push = child < n_nodes;
        }
    }
    tested += k * 8u;
    wave_sync();
    top -= k;
    const u64 pm = __ballot(push);
    if (push) W.stack[top + lanes_below(pm, lane)] = child;
    top += (u32)__popcll(pm);
    wave_sync();
}

/* traces one packet (8 rays, `active` lanes) against the frame's static
   geometry; returns the rays that hit (4086-4243) */
template <int kStack>
DEV u32 trace_packet(const FramePtrs& P, const ViewGpu& VW, const Est& est, OccLds<kStack>& W, const Frustum& fr,
                     sftvec o_r, sftvec l_r, u32 active, int lane, u32& nodes_tested, u32& tris_tested, bool& overflow) {
    const FrameConst& fc = VW.fc;
    u32 sat = 0;
    /* TLAS: collect the instance leaves the frustum reaches */
    u32 top = (fc.tlas_root < fc.tlas_nodes) ? 1u : 0u, ninst = 0;
    if (0 == lane) W.stack[0] = fc.tlas_root;
    wave_sync();
    /* every step pops at least one node and a tree holds each node once, so a
       traversal longer than the node count means a malformed (cyclic) BVH */
    u32 steps = 0;
    while (top > 0u) {
        if (++steps > fc.tlas_nodes + 1u) {
            overflow = true;
            break;
        }
        bool leaf;
        u32 child;
        bvh_step(W, top, P.tlas + VW.tlas_base, fc.tlas_nodes, fr, lane, leaf, child, nodes_tested, overflow);
        const u64 lm = __ballot(leaf);
        if (leaf) W.insts[min(ninst + lanes_below(lm, lane), 63u)] = child & 0x7FFFFFFFu;
        ninst = min(ninst + (u32)__popcll(lm), 64u);
        wave_sync();
    }
    const int r = lane & 7, tslot = lane >> 3;
    for (u32 ii = 0; ii < ninst && sat != active; ii += 1) {
        const u32 inst_id = W.insts[ii];
        if (inst_id >= fc.n_inst) continue;
        const InstGpu& in = P.insts[VW.inst_base + inst_id];
        if (kInvalid == in.blas_base || in.blas_root >= in.blas_count) continue;
        Frustum mf;
        model_frustum(fr, in.t, mf);
        /* this lane's ray (r) in model space (4171-4178) */
        const V3 ow{o_r.x, o_r.y, o_r.z};
        const V3 ew{o_r.x + l_r.x, o_r.y + l_r.y, o_r.z + l_r.z};
        const V3 lo = xform(ow, in.inv);
        const V3 le = xform(ew, in.inv);
        const V3 L = sub3(le, lo);
        const f32 tmax = sqrtf(dot3(L, L));
        const V3 ld = est.norm_fast(L);
        const sft_bvh_node* nodes = P.blas + in.blas_base;
        const u32 slot_base = in.model_base;
        top = 1;
        u32 ntri = 0;
        if (0 == lane) W.stack[0] = in.blas_root;
        wave_sync();
        steps = 0;
        while ((top > 0u || ntri > 0u) && sat != active) {
            if (top > 0u && ++steps > in.blas_count + 1u) {
                overflow = true;
                top = 0;
            }
            if (top > 0u) {
                bool leaf;
                u32 child;
                bvh_step(W, top, nodes, in.blas_count, mf, lane, leaf, child, nodes_tested, overflow);
                const u64 lm = __ballot(leaf);
                if (leaf) W.tris[ntri + lanes_below(lm, lane)] = child & 0x7FFFFFFFu;
                ntri += (u32)__popcll(lm);
                wave_sync();
            }
            while (ntri >= 8u || (0u == top && ntri > 0u)) {
                const u32 n = min(ntri, 8u);
                bool hit = false;
                const u32 pl = ((u32)tslot < n) ? W.tris[ntri - 1u - (u32)tslot] : 0u;
                const u32 m = pl >> 7, tri = pl & 0x7Fu;
                if ((u32)tslot < n && (((active & ~sat) >> r) & 1u) && m < in.n_meshlets) {
                    const u32 slot = slot_base + m;
                    const u8* ib = P.inds + (u64)slot * 512u + (tri >> 3) * 32u + (tri & 7u);
                    const f32* pp = P.pos + (u64)slot * 192u;
                    const u32 i0 = ib[0] & 63u, i1 = ib[8] & 63u, i2 = ib[16] & 63u;
                    const V3 v0{pp[i0], pp[64 + i0], pp[128 + i0]};
                    const V3 v1{pp[i1], pp[64 + i1], pp[128 + i1]};
                    const V3 v2{pp[i2], pp[64 + i2], pp[128 + i2]};
                    hit = ray_hits_tri(est, v0, v1, v2, lo, ld, tmax);
                }
                tris_tested += n * (u32)__popc(active & ~sat);
                const u64 hm = __ballot(hit);
                u32 h8 = 0;
                for (int t = 0; t < 8; t += 1) h8 |= (u32)(hm >> (8 * t)) & 0xFFu;
                sat |= h8;
                ntri -= n;
                wave_sync();
                if (sat == active) break;
            }
        }
    }
    return sat;
}

/* one workgroup per (light, packet of 8 proxies); wave w traces ray w.
   kStack: traversal stack entries per wave (the autotuner's choice); a
   packet that finds a stack smaller than kOccStackEntries full flags the
   batch to run again with the full stack, so the results never depend on
   the choice. */
template <int kStack>
__global__ __launch_bounds__(64 * kOccWaves) void k_occlude(FramePtrs P) {
    __shared__ OccLds<kStack> s_w[kOccWaves];
    __shared__ u32 s_hit[kOccWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const ViewGpu& VW = P.views[find_view(P, blockIdx.x, [](const ViewGpu& v) { return v.occ_base; })];
    const FrameConst& fc = VW.fc;
    const u32 g = blockIdx.x - VW.occ_base;
    const u32 l = g / fc.occ_batches, batch = g % fc.occ_batches;
    const u32 pbase = batch * 8u;
    const u32 n = min(8u, fc.n_proxies - pbase);
    const ProxyGpu* px = P.proxies + VW.proxy_base + pbase;
    const Est est = make_est(P);
    const PointLightGpu& L = P.point_lights[VW.pl_base + l];
    const sftvec lp = SFT_V(L.pos[0], L.pos[1], L.pos[2]);

    /* light-disk targets facing the packet's centre (8178-8210) */
    sftvec bc = SFT_V(0.f, 0.f, 0.f);
    for (u32 k = 0; k < n; k += 1) {
        bc.x += (px[k].a[0] + px[k].b[0]) * 0.5f;
        bc.y += (px[k].a[1] + px[k].b[1]) * 0.5f;
        bc.z += (px[k].a[2] + px[k].b[2]) * 0.5f;
    }
    bc = sft_v_mul(bc, 1.f / (f32)n);
    const sftvec ldir = sft_v_norm(sft_v_sub(lp, bc));
    const sftvec lright = sft_v_norm(sft_v_perpendicular(ldir));
    const sftvec lup = sft_v_cross(lright, ldir);
    const f32 rr = SFT_SHADOW_LIGHT_SIZE;
    sftvec target;
    if (0 == wave) target = sft_v_add(lp, sft_v_mul(lup, rr));
    else if (1 == wave) target = sft_v_add(lp, sft_v_add(sft_v_mul(lright, rr * 0.866025f), sft_v_mul(lup, -rr * 0.5f)));
    else target = sft_v_add(lp, sft_v_add(sft_v_mul(lright, -rr * 0.866025f), sft_v_mul(lup, -rr * 0.5f)));

    Frustum fr;
    const bool valid = packet_frustum(px, n, wave, lp, L.radius, fr);
    const u32 r = (u32)lane & 7u;
    sftvec o_r = SFT_V(0.f, 0.f, 0.f), l_r = SFT_V(0.f, 0.f, 0.f);
    if (r < n) {
        o_r = ray_origin(px[r], wave);
        l_r = SFT_V(target.x - o_r.x, target.y - o_r.y, target.z - o_r.z);
    }
    const u32 active = (1u << n) - 1u;
    u32 nodes_tested = 0, tris_tested = 0;
    bool overflow = false;
    const u32 sat = trace_packet(P, VW, est, s_w[wave], fr, o_r, l_r, active, lane, nodes_tested, tris_tested, overflow);
    if (0 == lane) {
        s_hit[wave] = sat;
        const u32 sl = stat_slot();
        stat_add(P.counters, CNT_OCC_RAYS, n, sl);
        stat_add(P.counters, CNT_OCC_NODES, nodes_tested, sl);
        stat_add(P.counters, CNT_OCC_TRIS, tris_tested, sl);
        stat_add(P.counters, CNT_OCC_INVALID, valid ? 0u : 1u, sl);
        stat_add(P.counters, CNT_OCC_OVERFLOW, overflow ? 1u : 0u, sl);
        if (overflow && kStack < (int)kOccStackEntries) atomicOr(&P.counters[CNT_OVERFLOW], 4u);
    }
    __syncthreads();
    if (0 == wave && r < n && lane < 8) {
        const u32 hits = ((s_hit[0] >> r) & 1u) + ((s_hit[1] >> r) & 1u) + ((s_hit[2] >> r) & 1u);
        P.proxy_vis[(u64)(VW.proxy_base + pbase + r) * 64u + l] = (u8)(3u - hits);
    }
}

/* per point light: the first 16 proxies, in index order, with visibility left
   for the light and a bounding sphere reaching the light's sphere, and their
   shading constants (sofren2.c:5936-5990).  The selection depends on the
   light only, so it is done once per light instead of once per tile. */
__global__ __launch_bounds__(64) void k_proxy_cache(FramePtrs P) {
    const int lane = threadIdx.x;
    const ViewGpu& VW = P.views[find_view(P, blockIdx.x, [](const ViewGpu& v) { return v.pl_base; })];
    const FrameConst& fc = VW.fc;
    if (!(fc.shadows_enabled && fc.n_proxies > 0u)) return;
    const u32 l = blockIdx.x - VW.pl_base;
    if (kInvalid != VW.lp_cache) {
        /* a served template's records, traced by an earlier batch */
        const u32 n = P.cached_lpc[VW.lp_cache + l];
        if ((u32)lane < n) {
            P.light_proxies[(u64)(VW.pl_base + l) * kMaxLightProxies + (u32)lane] =
                P.cached_lp[(u64)(VW.lp_cache + l) * kMaxLightProxies + (u32)lane];
        }
        if (0 == lane) P.light_proxy_count[VW.pl_base + l] = n;
        return;
    }
    const PointLightGpu& L = P.point_lights[VW.pl_base + l];
    const sftvec lpos = SFT_V(L.pos[0], L.pos[1], L.pos[2]);
    u32 count = 0;
    for (u32 base = 0; base < fc.n_proxies && count < (u32)kMaxLightProxies; base += 64u) {
        const u32 p = base + (u32)lane;
        bool ok = false;
        ProxyRec rec;
        if (p < fc.n_proxies) {
            const ProxyGpu& px = P.proxies[VW.proxy_base + p];
            /* the visibility count exists for the first 64 lights only */
            const u32 vis = (l < 64u) ? (u32)P.proxy_vis[(u64)(VW.proxy_base + p) * 64u + l] : 3u;
            if (0u != vis) {
                const sftvec a = SFT_V(px.a[0], px.a[1], px.a[2]), b = SFT_V(px.b[0], px.b[1], px.b[2]);
                const sftvec cc = SFT_V((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f);
                const sftvec ab = sft_v_sub(b, a);
                const f32 len = sft_v_length(ab);
                const f32 br = px.radius + (len * 0.5f);
                const f32 cr = br + L.radius;
                const sftvec w = sft_v_sub(cc, lpos);
                const f32 d2 = sft_v_length2(w);
                if (d2 <= cr * cr) {
                    ok = true;
                    rec.pa[0] = a.x; rec.pa[1] = a.y; rec.pa[2] = a.z;
                    rec.ab[0] = ab.x; rec.ab[1] = ab.y; rec.ab[2] = ab.z;
                    rec.w[0] = w.x; rec.w[1] = w.y; rec.w[2] = w.z;
                    rec.abdab = sft_v_dot(ab, ab);
                    rec.wdw = d2;
                    rec.rb2 = br * br;
                    rec.pr2 = px.radius * px.radius;
                    rec.vis_mult = (f32)vis * 0.3333333f;
                }
            }
        }
        const u64 m = __ballot(ok);
        const u32 rank = count + lanes_below(m, lane);
        if (ok && rank < (u32)kMaxLightProxies) P.light_proxies[(u64)(VW.pl_base + l) * kMaxLightProxies + rank] = rec;
        count = min(count + (u32)__popcll(m), (u32)kMaxLightProxies);
    }
    if (0 == lane) P.light_proxy_count[VW.pl_base + l] = count;
}

/* folds the spread statistics slots into the narrow counter words */
__global__ __launch_bounds__(64) void k_stats(u32* counters) {
    const int lane = threadIdx.x, c = blockIdx.x;
    u32 v = counters[CNT_COUNT + ((size_t)c * CNT_STAT_SLOTS + lane) * CNT_STAT_STRIDE];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (0 == lane && v) counters[c] += v;
}

/* streaming copy used to calibrate HBM byte counters */
/* clears the frame's counters, tile counters and chunk completion counts; a
   kernel rather than a memset so that the whole frame can live in a HIP graph */
__global__ __launch_bounds__(256) void k_zero(u32* __restrict__ p, u64 n) {
    for (u64 i = (u64)blockIdx.x * 256u + threadIdx.x; i < n; i += (u64)gridDim.x * 256u) p[i] = 0u;
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, u64 n) {
    for (u64 i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (u64)gridDim.x * 256ull) dst[i] = src[i];
}

/* ================================================================== textures
   The mip chain of sft_texture_create built in HBM (sofren2.c:9229-9289):
   level 0 is the caller's pixels in the 4x4-block layout, every next level a
   [1 2 2 1] x [1 2 2 1] tent over the previous one with edge clamp and
   truncating integer division per channel.  One thread per destination
   texel; the previous level is read back from its swizzled place. */
DEV u32 swz(i32 x, i32 y, i32 bw) {
    return (u32)(((y >> 2) * bw + (x >> 2)) * 16 + ((y & 3) << 2) + (x & 3));
}

__global__ __launch_bounds__(256) void k_tex_level0(const u32* __restrict__ src, u32* __restrict__ heap, u32 off,
                                                    i32 w, i32 h, i32 bw) {
    const u32 i = blockIdx.x * 256u + threadIdx.x;
    if (i >= (u32)w * (u32)h) return;
    const i32 x = (i32)(i % (u32)w), y = (i32)(i / (u32)w);
    heap[off + swz(x, y, bw)] = src[i];
}

__global__ __launch_bounds__(256) void k_tex_mip(u32* heap, u32 src_off, i32 cw, i32 ch, i32 sbw, u32 dst_off,
                                                 i32 nw, i32 nh, i32 dbw) {
    const u32 i = blockIdx.x * 256u + threadIdx.x;
    if (i >= (u32)nw * (u32)nh) return;
    const i32 x = (i32)(i % (u32)nw), y = (i32)(i / (u32)nw);
    u32 r = 0, g = 0, b = 0, a = 0, wsum = 0;
    for (i32 ky = -1; ky <= 2; ky += 1) {
        const i32 sy = min(max(y * 2 + ky, 0), ch - 1);
        const u32 wy = (-1 == ky || 2 == ky) ? 1u : 2u;
        for (i32 kx = -1; kx <= 2; kx += 1) {
            const i32 sx = min(max(x * 2 + kx, 0), cw - 1);
            const u32 wt = ((-1 == kx || 2 == kx) ? 1u : 2u) * wy;
            const u32 px = heap[src_off + swz(sx, sy, sbw)];
            r += ((px >> 16) & 255u) * wt;
            g += ((px >> 8) & 255u) * wt;
            b += (px & 255u) * wt;
            a += ((px >> 24) & 255u) * wt;
            wsum += wt;
        }
    }
    heap[dst_off + swz(x, y, dbw)] = ((a / wsum) << 24) | ((r / wsum) << 16) | ((g / wsum) << 8) | (b / wsum);
}

/* ================================================================== host */
void launch_copy(const uint4* src, uint4* dst, u64 n16, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, s, src, dst, n16);
}

void launch_texture_build(const u32* level0, u32* heap, const TexMeta& m, hipStream_t s) {
    i32 cw = m.w, ch = m.h;
    const u32 n0 = (u32)cw * (u32)ch;
    hipLaunchKernelGGL(k_tex_level0, dim3((n0 + 255u) / 256u), dim3(256), 0, s, level0, heap, m.mip_offset[0], cw, ch,
                       m.blocks_w[0]);
    for (i32 l = 1; l < m.mip_levels; l += 1) {
        const i32 nw = max(cw / 2, 1), nh = max(ch / 2, 1);
        const u32 n = (u32)nw * (u32)nh;
        hipLaunchKernelGGL(k_tex_mip, dim3((n + 255u) / 256u), dim3(256), 0, s, heap, m.mip_offset[l - 1], cw, ch,
                           m.blocks_w[l - 1], m.mip_offset[l], nw, nh, m.blocks_w[l]);
        cw = nw;
        ch = nh;
    }
}

/* resident workgroups of the persistent visibility kernel on the device */
u32 raster_groups_per_device() {
    int cus = 0, per_cu = 0, per_cu2 = 0;
    if (hipSuccess != hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0)) cus = 1;
    if (hipSuccess != hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_raster<1>, kRasterThreads, 0)) per_cu = 1;
    if (hipSuccess != hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, k_raster<2>, kRasterThreads, 0)) per_cu2 = 1;
    per_cu = std::min(per_cu, per_cu2);
    return (u32)std::max(cus, 1) * (u32)std::max(per_cu, 1);
}

/* ================================================================ TLAS
   The frame's top-level BVH of a view that traces shadow packets, built on
   the device (sofren2.c:3619-3703) into the view's nodes (reserved by the
   host, which knows their count: it depends on the item count only).  Items
   are the draw slots whose model has a BLAS, numbered over the first
   `built` slots as the reference numbers them (a slot without a BLAS drawn
   before one with a BLAS leaves a zeroed item).  The lanes compute the
   items' world boxes (the root box's 8 corners through the transform, the
   reference's operation order); lane 0 then groups them as the reference's
   builder does (sofren2.c:2895-3062): three levels of halving, each range
   sorted by the Hoare quicksort around the middle element along its
   centroids' widest axis, ranges of one item or level 3 becoming groups.
   The partitions of a range touch only that range, so sorting them in any
   order gives the reference's permutation.  With at most 64 items (the
   instance cap) every group holds at most 8: the root, then one leaf per
   group in group order.  Each node is written by the 64 lanes, a 32-bit
   word each (256 bytes, padding zero). */
constexpr int kTlasMaxItems = 64;

struct TlasItem {
    u32 id;
    f32 lo[3], hi[3], mid[3];
};

struct TlasLds {
    TlasItem it[kTlasMaxItems];
    u32 gstart[8], gcount[8], ng;
    u32 rs[2][8], rc[2][8], fin[2][8];   /* ranges of the current and next halving level */
    u16 stk[2 * kTlasMaxItems];
};

/* sorts it[0, n) by mid[axis] (sofren2.c:2895-2924), explicit stack */
DEV void tlas_sort(TlasItem* it, u16* stk, int n0, int axis) {
    int sp = 0;
    stk[sp++] = 0;
    stk[sp++] = (u16)n0;
    while (sp > 0) {
        int n = stk[--sp];
        int off = stk[--sp];
        while (n > 1) {
            TlasItem* a = it + off;
            const f32 pivot = a[n / 2].mid[axis];
            int i = 0, j = n - 1;
            while (i <= j) {
                while (a[i].mid[axis] < pivot) i += 1;
                while (a[j].mid[axis] > pivot) j -= 1;
                if (i <= j) {
                    u32* pi = (u32*)&a[i];
                    u32* pj = (u32*)&a[j];
                    for (int w = 0; w < (int)(sizeof(TlasItem) / 4); w += 1) {
                        const u32 x = pi[w];
                        pi[w] = pj[w];
                        pj[w] = x;
                    }
                    i += 1;
                    j -= 1;
                }
            }
            if (j > 0) {
                stk[sp++] = (u16)off;
                stk[sp++] = (u16)(j + 1);
            }
            if (i >= n) break;
            off += i;
            n -= i;
        }
    }
}

DEV void tlas_grow(f32* b, const f32* lo, const f32* hi) {
    for (int k = 0; k < 3; k += 1) {
        if (lo[k] < b[k]) b[k] = lo[k];
        if (hi[k] > b[3 + k]) b[3 + k] = hi[k];
    }
}

/* lane = 32-bit word of an sft_bvh_node whose children c < count have boxes
   lo(c)/hi(c) and indices idx(c) */
template <class Lo, class Hi, class Idx>
DEV void tlas_write_node(sft_bvh_node* nd, int lane, u32 count, u32 leaf, Lo lo, Hi hi, Idx idx) {
    u32 w = 0u;
    if (lane < 48) {
        const int arr = lane >> 3, c = lane & 7;
        f32 v = (arr < 3) ? 1e30f : -1e30f;
        if ((u32)c < count) v = (arr < 3) ? lo(c)[arr] : hi(c)[arr - 3];
        w = f2u(v);
    }
    else if (lane < 56) {
        const int c = lane - 48;
        w = ((u32)c < count) ? idx(c) : kInvalid;
    }
    else if (lane == 56) {
        w = count | (leaf << 8);
    }
    ((u32*)nd)[lane] = w;
}

__global__ __launch_bounds__(64) void k_tlas(FramePtrs P) {
    __shared__ TlasLds S;
    const int lane = threadIdx.x;
    const ViewGpu& VW = P.views[blockIdx.x];
    const FrameConst& fc = VW.fc;
    if (0u == VW.tlas_build || 0u == fc.tlas_nodes) return;
    const u32 n_inst = min(fc.n_inst, (u32)kTlasMaxItems);
    /* items by draw slot */
    bool has = false;
    if ((u32)lane < n_inst) {
        const InstGpu& in = P.insts[VW.inst_base + (u32)lane];
        has = 0u != in.tlas_item;
        TlasItem t;
        t.id = 0u;
        for (int k = 0; k < 3; k += 1) t.lo[k] = t.hi[k] = t.mid[k] = 0.f;
        if (has) {
            for (int k = 0; k < 3; k += 1) {
                t.lo[k] = 1e30f;
                t.hi[k] = -1e30f;
            }
            const f32* T = in.t;
            for (int corner = 0; corner < 8; corner += 1) {
                const f32 x = (corner & 1) ? in.root_hi[0] : in.root_lo[0];
                const f32 y = (corner & 2) ? in.root_hi[1] : in.root_lo[1];
                const f32 z = (corner & 4) ? in.root_hi[2] : in.root_lo[2];
                for (int k = 0; k < 3; k += 1) {
                    const f32 wk = x * T[k] + y * T[4 + k] + z * T[8 + k] + T[12 + k];
                    if (wk < t.lo[k]) t.lo[k] = wk;
                    if (wk > t.hi[k]) t.hi[k] = wk;
                }
            }
            t.id = (u32)lane;
            for (int k = 0; k < 3; k += 1) t.mid[k] = (t.lo[k] + t.hi[k]) * 0.5f;
        }
        u32* dst = (u32*)&S.it[lane];
        const u32* src = (const u32*)&t;
        for (int w = 0; w < (int)(sizeof(TlasItem) / 4); w += 1) dst[w] = src[w];
    }
    const u32 built = (u32)__popcll(__ballot(has));
    __syncthreads();
    /* grouping (lane 0) */
    if (0 == lane && built > 8u) {
        u32 nr = 1;
        S.rs[0][0] = 0u;
        S.rc[0][0] = built;
        S.fin[0][0] = 0u;
        for (u32 level = 0; level < 3u; level += 1) {
            const u32 cur = level & 1u, nxt = cur ^ 1u;
            u32* rs = S.rs[cur];
            u32* rc = S.rc[cur];
            u32* fin = S.fin[cur];
            u32* ns = S.rs[nxt];
            u32* nc = S.rc[nxt];
            u32* nf = S.fin[nxt];
            u32 nn = 0;
            for (u32 r = 0; r < nr; r += 1) {
                const u32 st = rs[r], ct = rc[r];
                if (fin[r] || 1u == ct) {
                    ns[nn] = st; nc[nn] = ct; nf[nn] = 1u; nn += 1;
                    continue;
                }
                f32 lo[3] = { 1e30f, 1e30f, 1e30f }, hi[3] = { -1e30f, -1e30f, -1e30f };
                for (u32 i = st; i < st + ct; i += 1) {
                    for (int k = 0; k < 3; k += 1) {
                        if (S.it[i].mid[k] < lo[k]) lo[k] = S.it[i].mid[k];
                        if (S.it[i].mid[k] > hi[k]) hi[k] = S.it[i].mid[k];
                    }
                }
                const f32 ext0 = hi[0] - lo[0], ext1 = hi[1] - lo[1], ext2 = hi[2] - lo[2];
                int axis = 0;
                if (ext1 > ext0 && ext1 > ext2) axis = 1;
                else if (ext2 > ext0 && ext2 > ext1) axis = 2;
                tlas_sort(S.it + st, S.stk, (int)ct, axis);
                const u32 half = ct / 2u;
                ns[nn] = st; nc[nn] = half; nf[nn] = 0u; nn += 1;
                ns[nn] = st + half; nc[nn] = ct - half; nf[nn] = 0u; nn += 1;
            }
            nr = nn;
        }
        /* three levels: the ranges are in buffer 1 */
        for (u32 r = 0; r < nr; r += 1) {
            S.gstart[r] = S.rs[1][r];
            S.gcount[r] = S.rc[1][r];
        }
        S.ng = nr;
    }
    __syncthreads();
    sft_bvh_node* nodes = const_cast<sft_bvh_node*>(P.tlas) + VW.tlas_base;
    if (built <= 8u) {
        tlas_write_node(nodes, lane, built, 1u, [&](int c) { return S.it[c].lo; }, [&](int c) { return S.it[c].hi; },
                        [&](int c) { return 0x80000000u | S.it[c].id; });
        return;
    }
    const u32 ng = S.ng;
    /* leaf g: its group's items; the root: the leaves' boxes */
    for (u32 g = 0; g < ng; g += 1) {
        const u32 s0 = S.gstart[g];
        tlas_write_node(nodes + 1u + g, lane, S.gcount[g], 1u, [&](int c) { return S.it[s0 + c].lo; },
                        [&](int c) { return S.it[s0 + c].hi; }, [&](int c) { return 0x80000000u | S.it[s0 + c].id; });
    }
    f32 gb[6] = { 1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f };
    if ((u32)lane < ng) {
        for (u32 i = 0; i < S.gcount[lane]; i += 1) tlas_grow(gb, S.it[S.gstart[lane] + i].lo, S.it[S.gstart[lane] + i].hi);
    }
    /* the root's words come from the lanes holding each leaf's box */
    u32 w = 0u;
    const int arr = lane >> 3, c = lane & 7;
    /* array `arr` of leaf c's box, from lane c */
    f32 v = 0.f;
    for (int k = 0; k < 6; k += 1) {
        const f32 g = __shfl(gb[k], c);
        v = (arr == k) ? g : v;
    }
    if (lane < 48) w = f2u(((u32)c < ng) ? v : ((arr < 3) ? 1e30f : -1e30f));
    else if (lane < 56) w = ((u32)(lane - 48) < ng) ? 1u + (u32)(lane - 48) : kInvalid;
    else if (lane == 56) w = ng;
    ((u32*)nodes)[lane] = w;
}

/* the occlusion pass with the batch's traversal stack size, after the
   device-built TLAS */
static void launch_occlude(const FramePtrs& p, const BatchGpu& B, hipStream_t s) {
    if (B.tlas_gpu) hipLaunchKernelGGL(k_tlas, dim3(B.n_views), dim3(64), 0, s, p);
    const dim3 grid(B.occ_total), block(64 * kOccWaves);
    if (B.occ_stack <= 256u) hipLaunchKernelGGL(k_occlude<256>, grid, block, 0, s, p);
    else if (B.occ_stack <= 512u) hipLaunchKernelGGL(k_occlude<512>, grid, block, 0, s, p);
    else hipLaunchKernelGGL(k_occlude<(int)kOccStackEntries>, grid, block, 0, s, p);
}

void launch_frame(const FramePtrs& p, const BatchGpu& B, u64 zero_words, hipStream_t s, hipEvent_t* ev,
                  const SideStream* side) {
    auto mark = [&](int k) { if (ev) hipEventRecord(ev[k], s); };
    hipLaunchKernelGGL(k_zero, dim3((u32)std::min<u64>((zero_words + 255u) / 256u, 1024u)), dim3(256), 0, s, p.counters,
                       zero_words);
    auto scan = [&](const u32* in, u32* out, u32 n, u32* total, u32 cap, u32 bit) {
        const u32 blocks = (n + kScanItems - 1u) / kScanItems;
        if (0u == blocks) return;
        if (blocks > 1u) hipLaunchKernelGGL(k_scan_sums, dim3(blocks), dim3(1024), 0, s, in, n, p.scan_sums);
        hipLaunchKernelGGL(k_scan_apply, dim3(blocks), dim3(1024), 0, s, in, out, n, (const u32*)p.scan_sums, total, cap,
                           p.counters + CNT_OVERFLOW, bit);
    };
    const u32 nw = B.n_work;
    const u32 geom_blocks = (nw + 3u) / 4u;
    mark(1);
    /* shadow-proxy visibility and the per-light capsule lists: on the side
       stream when there are rays to trace (their long, narrow traversal then
       overlaps the passes below), else in order */
    const bool fork = side && B.occ_total > 0u;
    if (fork) {
        hipEventRecord(side->fork, s);
        hipStreamWaitEvent(side->stream, side->fork, 0);
        hipEventRecord(side->start, side->stream);
        launch_occlude(p, B, side->stream);
        if (B.n_lights > 0u) hipLaunchKernelGGL(k_proxy_cache, dim3(B.n_lights), dim3(64), 0, side->stream, p);
        hipEventRecord(side->end, side->stream);
        hipEventRecord(side->join, side->stream);
    }
    else {
        if (B.occ_total > 0u) launch_occlude(p, B, s);
        if (B.n_lights > 0u) hipLaunchKernelGGL(k_proxy_cache, dim3(B.n_lights), dim3(64), 0, s, p);
    }
    mark(7);
    if (nw > 0) {
        hipLaunchKernelGGL(k_geom, dim3(geom_blocks), dim3(256), 0, s, p);
        scan(p.clip_count, p.clip_base, nw, p.counters + CNT_CLIP_TRIS, B.clip_cap, 1u);
    }
    mark(2);
    if (nw > 0) hipLaunchKernelGGL(k_clip, dim3(geom_blocks), dim3(256), 0, s, p);
    if (B.n_segs) hipLaunchKernelGGL(k_bin_count, dim3(B.n_segs), dim3(kBinThreads), 0, s, p);
    if (B.n_views) hipLaunchKernelGGL(k_bin_scan, dim3(B.n_views), dim3(kBinThreads), 0, s, p);
    if (B.n_segs) hipLaunchKernelGGL(k_bin, dim3(B.n_segs), dim3(kBinThreads), 0, s, p);
    mark(3);
    hipLaunchKernelGGL(k_raster_plan, dim3(B.n_tiles), dim3(256), 0, s, p);
    if (B.n_views_ms < B.n_views) hipLaunchKernelGGL(k_raster<1>, dim3(B.raster_groups), dim3(kRasterThreads), 0, s, p);
    if (B.n_views_ms > 0u) hipLaunchKernelGGL(k_raster<2>, dim3(B.raster_groups), dim3(kRasterThreads), 0, s, p);
    if (B.max_samples > 1u) hipLaunchKernelGGL(k_tile_lights, dim3(B.n_tiles), dim3(64), 0, s, p);
    mark(6);
    if (fork) hipStreamWaitEvent(s, side->join, 0);
    const u32 plain = B.n_shade_tiles - B.n_shade_sh, plain_ms = B.n_shade_tiles_ms - B.n_shade_ms_sh;
    if (plain) hipLaunchKernelGGL(k_shade<false>, dim3(plain * 16u), dim3(256), 0, s, p);
    if (B.n_shade_sh) hipLaunchKernelGGL(k_shade<true>, dim3(B.n_shade_sh * 16u), dim3(256), 0, s, p);
    if (plain_ms) hipLaunchKernelGGL(k_shade_ms<false>, dim3(plain_ms * 16u), dim3(256), 0, s, p);
    if (B.n_shade_ms_sh) hipLaunchKernelGGL(k_shade_ms<true>, dim3(B.n_shade_ms_sh * 16u), dim3(256), 0, s, p);
    mark(4);
    hipLaunchKernelGGL(k_stats, dim3(CNT_COUNT), dim3(64), 0, s, p.counters);
}

}  // namespace sft
