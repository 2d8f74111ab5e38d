/*
 * capsule_cull.h — the capsule culling of the shadowed shading pass, shared
 * by the kernels and a host build (a CPU property test checks that the wave
 * test never drops a capsule that some lane's exact test would keep).
 */
#pragma once
#include <math.h>
#include "sft_internal.h"

#ifdef __HIPCC__
#define SFT_HD __host__ __device__ __forceinline__
#else
#define SFT_HD inline
#endif

namespace sft {

/* capsule culling for a wave.  The shading positions of the wave's lanes span
   a box of centre M and half-diagonal R, and every segment from such a
   position to a light lies within R of the segment M-light.  A lane's ray is
   active for a capsule only if the capsule centre projects inside the lane's
   pixel-to-light segment at a distance below the bounding radius
   (sofren2.c:6521-6530); a capsule whose centre keeps that distance from
   segment M-light plus R plus a margin (4e-3 of the lengths involved and
   1e-5 of the coordinates' size, far above the float rounding of the lane's
   test) has no active ray on any lane, so the reference's per-group skip
   would pass over it on every group of the wave. */
struct WaveSpan {
    f32 m[3];       /* box centre */
    f32 r;          /* half-diagonal, widened for the rounding of the centre */
    f32 mag;        /* |m| in the 1-norm */
};

/* the span of a box [l, h] of shading positions */
SFT_HD WaveSpan span_of_box(f32 lx, f32 ly, f32 lz, f32 hx, f32 hy, f32 hz) {
    WaveSpan s;
    s.m[0] = 0.5f * (lx + hx); s.m[1] = 0.5f * (ly + hy); s.m[2] = 0.5f * (lz + hz);
    const f32 ex = hx - lx, ey = hy - ly, ez = hz - lz;
    s.mag = fabsf(s.m[0]) + fabsf(s.m[1]) + fabsf(s.m[2]);
    s.r = 0.5f * sqrtf(ex * ex + ey * ey + ez * ez) * 1.0001f + 1e-6f * s.mag;
    return s;
}

/* whether capsule c (cached for light L) may be near some lane's ray */
SFT_HD bool capsule_near(const PointLightGpu& L, const ProxyRec& c, const WaveSpan& s) {
    const f32 dx = s.m[0] - L.pos[0], dy = s.m[1] - L.pos[1], dz = s.m[2] - L.pos[2];
    const f32 dd = dx * dx + dy * dy + dz * dz;
    const f32 ud = c.w[0] * dx + c.w[1] * dy + c.w[2] * dz;
    const f32 t = (dd > 0.f) ? fminf(fmaxf(ud / dd, 0.f), 1.f) : 0.f;
    const f32 ex = c.w[0] - t * dx, ey = c.w[1] - t * dy, ez = c.w[2] - t * dz;
    const f32 dist = sqrtf(ex * ex + ey * ey + ez * ez);
    const f32 absl = fabsf(L.pos[0]) + fabsf(L.pos[1]) + fabsf(L.pos[2]);
    const f32 thr = sqrtf(c.rb2) + s.r + 4e-3f * (sqrtf(c.wdw) + sqrtf(dd) + s.r) + 1e-5f * (s.mag + absl) + 1e-4f;
    /* a NaN anywhere keeps the capsule */
    return !(dist >= thr);
}

/* the cheap half of the reference's capsule test (sofren2.c:6521-6530): the
   capsule's bounding centre projects strictly inside the pixel-to-light
   segment (lx, ly, lz) = light - position, of squared length d2 and
   approximate reciprocal inv_d2, closer to it than the bounding radius */
SFT_HD bool ray_near_xyz(const ProxyRec& c, f32 lx, f32 ly, f32 lz, f32 d2, f32 inv_d2) {
    const f32 wdv = 0.f - (c.w[0] * lx + (c.w[1] * ly + c.w[2] * lz));
    const f32 tf = wdv * inv_d2;
    const f32 cd2 = fmaf(-wdv, tf, c.wdw);
    return (wdv > 0.f) & (wdv < d2) & (cd2 < c.rb2);
}

}  // namespace sft
