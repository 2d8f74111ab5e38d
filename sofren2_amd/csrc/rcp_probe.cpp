/*
 * rcp_probe.cpp — tables that let the GPU reproduce this host's reciprocal and
 * reciprocal-square-root estimates bit for bit.
 *
 * The reference computes screen positions, depth slopes, barycentric
 * normalisation and specular terms with _mm256_rcp_ps followed by one Newton
 * step (sofren2.c:632-639) and normalises vectors with _mm256_rsqrt_ps
 * (sofren2.c:1973-1992).  Both instructions return implementation-defined
 * estimates (about 12 bits), so a bit-exact visibility buffer needs the
 * estimate of the CPU the reference runs on.  At sft_init the estimates of the
 * host CPU are read from the instruction itself:
 *   - for every mantissa of an input in [1,2) (2^23 inputs) the output of
 *     VRCPPS is recorded, and the smallest k for which the output depends
 *     only on the top k mantissa bits is found (the table then has 2^k
 *     entries);
 *   - the same for VRSQRTPS over [1,4), where the exponent parity matters;
 *   - the emulation (table value with the exponent rescaled, zero/denormal
 *     inputs to infinity, results below the normal range flushed to zero) is
 *     then checked against the instruction over every exponent and a set of
 *     mantissas; the number of disagreements is kept for the caller.
 */
#include <immintrin.h>
#include <string.h>
#include "sft_internal.h"

namespace sft {

static inline u32 f2u(f32 f) { u32 u; memcpy(&u, &f, 4); return u; }
static inline f32 u2f(u32 u) { f32 f; memcpy(&f, &u, 4); return f; }

__attribute__((target("avx")))
static void rcp8(const f32* in, f32* out) {
    _mm256_storeu_ps(out, _mm256_rcp_ps(_mm256_loadu_ps(in)));
}

__attribute__((target("avx")))
static void rsqrt8(const f32* in, f32* out) {
    _mm256_storeu_ps(out, _mm256_rsqrt_ps(_mm256_loadu_ps(in)));
}

static void rcp_n(const f32* in, f32* out, u64 n) {
    u64 i = 0;
    for (; i + 8 <= n; i += 8) rcp8(in + i, out + i);
    if (i < n) {
        f32 a[8] = {1, 1, 1, 1, 1, 1, 1, 1}, b[8];
        memcpy(a, in + i, (n - i) * 4);
        rcp8(a, b);
        memcpy(out + i, b, (n - i) * 4);
    }
}

static void rsqrt_n(const f32* in, f32* out, u64 n) {
    u64 i = 0;
    for (; i + 8 <= n; i += 8) rsqrt8(in + i, out + i);
    if (i < n) {
        f32 a[8] = {1, 1, 1, 1, 1, 1, 1, 1}, b[8];
        memcpy(a, in + i, (n - i) * 4);
        rsqrt8(a, b);
        memcpy(out + i, b, (n - i) * 4);
    }
}

f32 emulate_rcp(const EstimateTables& t, f32 a) {
    const u32 b = f2u(a);
    const u32 sign = b & 0x80000000u;
    const u32 e = (b >> 23) & 0xFFu;
    const u32 m = b & 0x7FFFFFu;
    if (0 == e) return u2f(sign | 0x7F800000u);
    if (255 == e) return m ? u2f(b | 0x00400000u) : u2f(sign);
    const u32 tb = f2u(t.rcp[m >> (23 - t.rcp_bits)]);
    const i32 oe = (i32)((tb >> 23) & 0xFFu) + 127 - (i32)e;
    if (oe <= 0) return u2f(sign);
    return u2f(sign | ((u32)oe << 23) | (tb & 0x7FFFFFu));
}

f32 emulate_rsqrt(const EstimateTables& t, f32 a) {
    const u32 b = f2u(a);
    const u32 e = (b >> 23) & 0xFFu;
    const u32 m = b & 0x7FFFFFu;
    if (0 == e) return u2f((b & 0x80000000u) | 0x7F800000u);
    if (b & 0x80000000u) return u2f(0xFFC00000u);
    if (255 == e) return m ? u2f(b | 0x00400000u) : 0.f;
    const i32 u = (i32)e - 127;
    const i32 p = u & 1;
    const u32 tb = f2u(t.rsq[((u32)p << t.rsq_bits) | (m >> (23 - t.rsq_bits))]);
    const i32 oe = (i32)((tb >> 23) & 0xFFu) - (u - p) / 2;
    return u2f(((u32)oe << 23) | (tb & 0x7FFFFFu));
}

/* smallest k such that out[m] depends only on the top k bits of m */
static int dependent_bits(const std::vector<u32>& out, u32 base, u32 count) {
    for (int k = 1; k <= 23; k += 1) {
        const u32 mask = ~((1u << (23 - k)) - 1u);
        bool ok = true;
        for (u32 m = 0; m < count && ok; m += 1) {
            ok = out[base + m] == out[base + (m & mask)];
        }
        if (ok) return k;
    }
    return 23;
}

void probe_estimate_tables(EstimateTables* t) {
    if (!__builtin_cpu_supports("avx")) {
        fprintf(stderr, "fatal sofren2 error: host CPU lacks AVX; the reference's estimate instructions cannot be probed\n");
        exit(1);
    }
    const u32 n = 1u << 23;
    std::vector<f32> in(2 * (u64)n);
    std::vector<u32> out(2 * (u64)n);

    /* reciprocal over [1,2) */
    for (u32 m = 0; m < n; m += 1) in[m] = u2f(0x3F800000u | m);
    rcp_n(in.data(), (f32*)out.data(), n);
    t->rcp_bits = dependent_bits(out, 0, n);
    t->rcp.resize((u64)1 << t->rcp_bits);
    for (u32 i = 0; i < (1u << t->rcp_bits); i += 1) t->rcp[i] = u2f(out[(u64)i << (23 - t->rcp_bits)]);

    /* reciprocal square root over [1,2) and [2,4) */
    for (u32 m = 0; m < n; m += 1) {
        in[m] = u2f(0x3F800000u | m);
        in[n + m] = u2f(0x40000000u | m);
    }
    rsqrt_n(in.data(), (f32*)out.data(), 2 * (u64)n);
    const int k0 = dependent_bits(out, 0, n), k1 = dependent_bits(out, n, n);
    t->rsq_bits = k0 > k1 ? k0 : k1;
    t->rsq.resize((u64)2 << t->rsq_bits);
    for (u32 i = 0; i < (1u << t->rsq_bits); i += 1) {
        t->rsq[i] = u2f(out[(u64)i << (23 - t->rsq_bits)]);
        t->rsq[(1u << t->rsq_bits) + i] = u2f(out[n + ((u64)i << (23 - t->rsq_bits))]);
    }

    /* check the emulation over all exponents, both signs, bucket edges and
       a stride of mantissas */
    std::vector<f32> probe;
    for (u32 e = 0; e < 256; e += 1) {
        for (u32 s = 0; s < 2; s += 1) {
            for (u32 m = 0; m < n; m += 4099) probe.push_back(u2f((s << 31) | (e << 23) | m));
            for (u32 i = 0; i < 64; i += 1) {
                const u32 bucket = (i * 977u) & ((1u << t->rcp_bits) - 1u);
                const u32 lo = bucket << (23 - t->rcp_bits);
                const u32 hi = lo | ((1u << (23 - t->rcp_bits)) - 1u);
                probe.push_back(u2f((s << 31) | (e << 23) | lo));
                probe.push_back(u2f((s << 31) | (e << 23) | hi));
            }
        }
    }
    std::vector<f32> got(probe.size());
    rcp_n(probe.data(), got.data(), probe.size());
    t->rcp_mismatches = 0;
    for (u64 i = 0; i < probe.size(); i += 1) {
        const f32 em = emulate_rcp(*t, probe[i]);
        const bool nan_both = (em != em) && (got[i] != got[i]);
        if (!nan_both && f2u(em) != f2u(got[i])) t->rcp_mismatches += 1;
    }
    rsqrt_n(probe.data(), got.data(), probe.size());
    t->rsq_mismatches = 0;
    for (u64 i = 0; i < probe.size(); i += 1) {
        const f32 em = emulate_rsqrt(*t, probe[i]);
        const bool nan_both = (em != em) && (got[i] != got[i]);
        if (!nan_both && f2u(em) != f2u(got[i])) t->rsq_mismatches += 1;
    }
}

}  // namespace sft
