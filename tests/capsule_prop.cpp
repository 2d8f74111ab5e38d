// Property check of the shadowed shading pass's capsule culling, built on the
// host from the kernels' own header (capsule_cull.h):
//
//   for every shading position p of a wave, light L and capsule c:
//   ray_near(c, L - p) implies capsule_near(L, c, span of the wave's positions)
//
// i.e. the wave test never drops a capsule whose cheap ray test some lane
// passes (sofren2.c:6521-6530); such a capsule would change that lane's
// shadow term.  Capsules are placed near the thresholds on purpose: their
// bounding sphere grazes one lane's pixel-to-light segment from just inside
// or just outside, or sits just inside or outside the wave test's own reach
// (bounding radius plus span radius from the span centre's segment), with
// the projection anywhere along the segment or near either end.
//
// The reciprocal is the one the kernels compute (estimate plus one Newton
// step); the estimate is this host's rcpss (the renderer probes its tables
// from the host it runs on), and alternatively any value within 2^-20 of
// 1/d2, which bounds the refined estimate of any x86 rcpss (its error is at
// most 1.5 * 2^-12 before the step).
//
// usage: capsule_prop <cases> <seed> [tight]
//   prints: cases rays ray_near_true violations culled
//   "tight" checks a planted mistake instead (the threshold without the
//   span radius and the margins), which must show violations.
#include <immintrin.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "capsule_cull.h"

using namespace sft;

static uint64_t g_state;
static uint32_t next_u32() {
    g_state = g_state * 6364136223846793005ULL + 1442695040888963407ULL;
    const uint64_t x = g_state;
    return (uint32_t)((x ^ (x >> 22)) >> (22 + (x >> 61)));
}
static double uni() { return (next_u32() >> 8) * (1.0 / 16777216.0); }
static double uni(double lo, double hi) { return lo + (hi - lo) * uni(); }
static double pow10u(double lo, double hi) { return __builtin_pow(10.0, uni(lo, hi)); }

static f32 rcp_refined_host(f32 a) {
    const f32 x = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(a)));
    return x * fmaf(-a, x, 2.0f);
}

static bool capsule_near_tight(const PointLightGpu& L, const ProxyRec& c, const WaveSpan& s) {
    const f32 dx = s.m[0] - L.pos[0], dy = s.m[1] - L.pos[1], dz = s.m[2] - L.pos[2];
    const f32 dd = dx * dx + dy * dy + dz * dz;
    const f32 ud = c.w[0] * dx + c.w[1] * dy + c.w[2] * dz;
    const f32 t = (dd > 0.f) ? fminf(fmaxf(ud / dd, 0.f), 1.f) : 0.f;
    const f32 ex = c.w[0] - t * dx, ey = c.w[1] - t * dy, ez = c.w[2] - t * dz;
    return !(sqrtf(ex * ex + ey * ey + ez * ez) >= sqrtf(c.rb2));
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const long cases = atol(argv[1]);
    g_state = strtoull(argv[2], nullptr, 10) * 0x9E3779B97F4A7C15ULL + 1;
    const bool tight = argc > 3 && 0 == strcmp(argv[3], "tight");
    long rays = 0, near_true = 0, violations = 0, culled = 0;
    for (long n = 0; n < cases; n += 1) {
        const double S = pow10u(-1.0, 4.0);
        double C[3], H[3];
        const double hs = S * pow10u(-5.0, 0.0);
        for (int k = 0; k < 3; k += 1) {
            C[k] = uni(-S, S);
            H[k] = hs * uni();
        }
        const int np = 1 + (int)(next_u32() % 8u);
        f32 P[8][3];
        f32 lo[3] = { 1e30f, 1e30f, 1e30f }, hi[3] = { -1e30f, -1e30f, -1e30f };
        for (int j = 0; j < np; j += 1) {
            for (int k = 0; k < 3; k += 1) {
                P[j][k] = (f32)(C[k] + H[k] * uni(-1.0, 1.0));
                lo[k] = fminf(lo[k], P[j][k]);
                hi[k] = fmaxf(hi[k], P[j][k]);
            }
        }
        const WaveSpan sp = span_of_box(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
        // light at a random direction and distance from the span
        double dir[3], dn = 0.0;
        for (int k = 0; k < 3; k += 1) {
            dir[k] = uni(-1.0, 1.0);
            dn += dir[k] * dir[k];
        }
        dn = __builtin_sqrt(dn) + 1e-12;
        const double dist = S * pow10u(-3.0, 1.0);
        PointLightGpu L{};
        for (int k = 0; k < 3; k += 1) L.pos[k] = (f32)(C[k] + dir[k] / dn * dist);
        // a capsule grazing the segment to the light from one lane's position
        // (half the cases) or from the span's centre at the span radius plus
        // the bounding radius (the wave test's own threshold)
        const int js = (int)(next_u32() % (uint32_t)np);
        const bool at_span = next_u32() & 1u;
        double seg[3], sl = 0.0, from[3];
        for (int k = 0; k < 3; k += 1) {
            from[k] = at_span ? (double)sp.m[k] : (double)P[js][k];
            seg[k] = (double)L.pos[k] - from[k];
            sl += seg[k] * seg[k];
        }
        sl = __builtin_sqrt(sl);
        if (!(sl > 0.0)) continue;
        const int ukind = (int)(next_u32() % 3u);
        const double eps_u = pow10u(-7.0, -0.5);
        const double u = (0 == ukind) ? uni(0.0, 1.0) : (1 == ukind ? eps_u : 1.0 - eps_u);
        double perp[3], pd = 0.0, pn = 0.0;
        for (int k = 0; k < 3; k += 1) perp[k] = uni(-1.0, 1.0);
        for (int k = 0; k < 3; k += 1) pd += perp[k] * seg[k] / sl;
        for (int k = 0; k < 3; k += 1) {
            perp[k] -= pd * seg[k] / sl;
            pn += perp[k] * perp[k];
        }
        pn = __builtin_sqrt(pn);
        if (!(pn > 1e-9)) continue;
        const double br = sl * pow10u(-4.0, 0.0);
        const double rel = pow10u(-7.0, -0.5);
        const double reach = at_span ? br + (double)sp.r : br;
        const double delta = reach * ((next_u32() & 1u) ? 1.0 - rel : 1.0 + rel);
        const f32 brf = (f32)br;
        f32 cc[3];
        for (int k = 0; k < 3; k += 1) cc[k] = (f32)(from[k] + u * seg[k] + delta * perp[k] / pn);
        ProxyRec rec{};
        const sftvec w = sft_v_sub(SFT_V(cc[0], cc[1], cc[2]), SFT_V(L.pos[0], L.pos[1], L.pos[2]));
        rec.w[0] = w.x; rec.w[1] = w.y; rec.w[2] = w.z;
        rec.wdw = sft_v_length2(w);
        rec.rb2 = brf * brf;
        const bool keep = tight ? capsule_near_tight(L, rec, sp) : capsule_near(L, rec, sp);
        culled += keep ? 0 : 1;
        for (int j = 0; j < np; j += 1) {
            const f32 lx = L.pos[0] - P[j][0], ly = L.pos[1] - P[j][1], lz = L.pos[2] - P[j][2];
            const f32 d2 = lx * lx + (ly * ly + lz * lz);
            for (int variant = 0; variant < 2; variant += 1) {
                f32 inv;
                if (0 == variant) {
                    inv = rcp_refined_host(d2);
                }
                else {
                    inv = (f32)((1.0 / (double)d2) * (1.0 + uni(-1.0, 1.0) * (1.0 / 1048576.0)));
                }
                rays += 1;
                const bool rn = ray_near_xyz(rec, lx, ly, lz, d2, inv);
                near_true += rn ? 1 : 0;
                if (rn && !keep) {
                    violations += 1;
                    if (!tight && violations <= 5) {
                        fprintf(stderr, "violation: case %ld lane %d S %g u %g br %g rel %g\n", n, j, S, u, br, rel);
                    }
                }
            }
        }
    }
    printf("%ld %ld %ld %ld %ld\n", cases, rays, near_true, violations, culled);
    return 0;
}
