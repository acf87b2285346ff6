// pv_device.hpp -- device math and per-geometry pair-interaction policies for gfx950.
//
// Each policy describes one geometry in one arithmetic:
//   Src   the per-source record staged in LDS (read by every lane of a wave at once,
//         i.e. an LDS broadcast), with all source-only factors folded in;
//   Tgt   the per-target record held in registers;
//   Acc   the per-target accumulators; SymAcc those of the symmetric kernels (the mixed
//         policies fold into a plain FP64 sum there, BlockSum);
//   pair()          one source's contribution to one target (j != i guaranteed);
//   pair_checked()  the same, for the few source tiles that contain the target itself;
//   fold()          called after each LDS tile (mixed: FP32 tile sum -> FP64 TwoSum).
//   kBatched        FP64 policies split pair() into prep() (the pair's denominators) and
//                   apply() (given their inverses), so that the kernel inverts all the
//                   denominators of a group of pairs with one rcp64 (TreeInverse).
//   kRingUnroll     unroll of the symmetric ring's step loop (measured per policy: 2 for the
//                   FP64 and disk policies, 1 for the FP32-heavy infinite and periodic ones).
//
// The pair formulas restate src/compute.cpp:117-126 (infinite), 223-249 (periodic) and
// 306-329 (disk).  The periodic image sum uses the exact Fourier form of the reference's
// truncated row sum (see PerF64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pv {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kInvTwoPi = 0.5 / kPi;  // src/compute.cpp:97 (inverseTwoPi)
constexpr int kMaxTerms = 16;            // capacity of the Fourier weight tables
constexpr int kTermsF64 = 12;            // Hamiltonian Fourier terms, rows |m| >= 1
constexpr int kImageTermsF64 = 4;        // velocity Fourier terms, rows |m| >= 2, FP64
constexpr int kImageTermsF32 = 2;        // velocity Fourier terms, rows |m| >= 2, mixed

// Geometry constants, filled on the host (pv_api.cpp) from pv_config.
struct GeoParams {
    int geometry;        // 0 infinite, 1 periodic, 2 disk
    int layers;          // periodicImageLayers
    double eps2;         // coreRadius^2 (infinite)
    double L, invL;      // box length and its reciprocal (periodic)
    double wave;         // 2 pi / L                (src/compute.cpp:213)
    double hwave;        // pi / L: half angles X/2 of the periodic kernels
    double scale;        // 0.5 / L                 (src/compute.cpp:214)
    double R, invR, R2;  // disk radius, 1/R, R^2
    double R2lo;         // R^2 - R2 exactly (fma(R, R, -R2)): R2 + R2lo is R*R
    double G4[kMaxTerms + 1];   // sum_{m=2..L} exp(-2 pi n m), n = 1..kMaxTerms (velocity)
    double H4[kMaxTerms + 1];   // 4 * sum_{m=1..L} exp(-2 pi n m) / n (Hamiltonian series)
    double CL;                  // 2 * sum_{m=1..L} log1p(exp(-4 pi m)) (Hamiltonian)
    double W1;                  // 2 when the rows m = +-1 exist (L >= 1), else 0
    double C1, C2, S4;          // cosh 2 pi, 2 cosh 2 pi, 4 sinh^2 2 pi
    float G4f[kMaxTerms + 1];
    float W1f, C1f, C2f, S4f;
};

// ---------------------------------------------------------------- device math helpers

// FP64 reciprocal: the hardware v_rcp_f64 seed r (relative error |e| <= 2^-24.4 on gfx950,
// measured over 40 binades) refined by one third-order step of the series
// 1/a = r / (1 - e) = r (1 + e + e^2 + ...), e = 1 - a r:  r' = r + r (e + e^2).  The
// truncation is e^3 <= 2^-73; e is formed by one FMA (relative error 2^-53), so r' is within
// about 1 ulp, as two Newton steps are, with one FMA fewer.  Infinite or zero arguments give
// NaN, as the Newton form does (e = -inf * 0).
__device__ __forceinline__ double rcp64(double a) {
    const double r = __builtin_amdgcn_rcp(a);
    const double e = __builtin_fma(-a, r, 1.0);
    return __builtin_fma(r, __builtin_fma(e, e, e), r);
}

__device__ __forceinline__ float rcp32(float a) { return __builtin_amdgcn_rcpf(a); }

// Reciprocals of K positive values with one rcp64: a product tree (an odd element at a level
// is passed up unpaired), one refined reciprocal of the root, then back-multiplication down the
// tree; about 2 log2(K) ulps.
// Returns false (inverses then unusable) when the root leaves [2^-960, 2^960], where a partial
// product could have lost bits; callers then fall back to one rcp64 per value.
template <int K> struct TreeInverse {
    __device__ static __forceinline__ bool run(const double *d, double *inv) {
        static_assert(K >= 2, "K must be at least 2");
        constexpr int H = K / 2, R = K - 2 * H;   // pairs, and an odd element passed up as is
        double lo[H + R], lo_inv[H + R];
#pragma unroll
        for (int k = 0; k < H; ++k)
            lo[k] = d[2 * k] * d[2 * k + 1];
        if constexpr (R)
            lo[H] = d[K - 1];
        const bool ok = TreeInverse<H + R>::run(lo, lo_inv);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            inv[2 * k] = lo_inv[k] * d[2 * k + 1];
            inv[2 * k + 1] = lo_inv[k] * d[2 * k];
        }
        if constexpr (R)
            inv[K - 1] = lo_inv[H];
        return ok;
    }
};
template <> struct TreeInverse<1> {
    __device__ static __forceinline__ bool run(const double *d, double *inv) {
        inv[0] = rcp64(d[0]);
        return d[0] >= 0x1p-960 && d[0] <= 0x1p960;
    }
};

// True when the product tree left some lane of the wave out of range (a wave-uniform
// condition: the ballot of the failing lanes is a scalar mask).
__device__ __forceinline__ bool tree_fallback(bool ok) {
    return __builtin_amdgcn_ballot_w64(!ok) != 0;
}

// Error-free addition (Knuth's TwoSum): s + e == a + b exactly.  Contraction is off here: a
// caller's product fused into these additions (a + x*y as one fma) would make s exact while e
// still measures the rounded product, and the pair would no longer be error-free.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

// Compensated accumulator: value = hi + lo.
struct Comp {
    double hi, lo;
    __device__ __forceinline__ void zero() { hi = lo = 0.0; }
    __device__ __forceinline__ void add(double x) {
        double s, e;
        two_sum(hi, x, s, e);
        hi = s;
        lo += e;
    }
    __device__ __forceinline__ double value() const { return hi + lo; }
};

// Plain FP64 sum.  A symmetric block or ensemble item folds at most 8 FP32 tile sums per
// target (a 1024-vortex source tile in 128-record chunks; 4 in an ensemble item); 24-bit
// significands add in FP64 without rounding while their exponents span under 26 binades, and
// otherwise round at 2^-53 of the largest, far below the FP32 sums' own 2^-24.  The compensated
// (TwoSum) sum of these per-block results runs in the reduction over the tile rows.
struct BlockSum {
    double hi;
    __device__ __forceinline__ void zero() { hi = 0.0; }
    __device__ __forceinline__ void add(double x) { hi += x; }
    __device__ __forceinline__ double value() const { return hi; }
};

// Per-target accumulators of the mixed policies: FP32 sums over the current tile, folded into
// S (Comp in the kernels that fold many tiles per target, BlockSum in the symmetric ones).
template <class S> struct MixedAcc {
    float u32, v32;
    S u, v;
};

// Exact IEEE remainder(d, L) for the common case (the quotient fits comfortably in a double):
// d - n L with n = d/L rounded to nearest; fma makes the subtraction exact.  A quotient that
// lands on the other side of a half-integer only moves the result by +-L, to which every
// periodic formula below is invariant.
__device__ __forceinline__ double remainder_fast(double d, double L, double invL) {
    const double n = __builtin_rint(d * invL);
    return __builtin_fma(-n, L, d);
}

// Taylor polynomials (coefficients 1/k!) on |h| <= pi/2 (+ a rounding margin), for the
// periodic Hamiltonian: truncation below 3e-17 relative.
__device__ __forceinline__ double sinh_f64(double h, double h2) {
    double p = 1.0 / 51090942171709440000.0;        // 1/21!
    p = __builtin_fma(p, h2, 1.0 / 121645100408832000.0);  // 1/19!
    p = __builtin_fma(p, h2, 1.0 / 355687428096000.0);     // 1/17!
    p = __builtin_fma(p, h2, 1.0 / 1307674368000.0);       // 1/15!
    p = __builtin_fma(p, h2, 1.0 / 6227020800.0);          // 1/13!
    p = __builtin_fma(p, h2, 1.0 / 39916800.0);            // 1/11!
    p = __builtin_fma(p, h2, 1.0 / 362880.0);              // 1/9!
    p = __builtin_fma(p, h2, 1.0 / 5040.0);                // 1/7!
    p = __builtin_fma(p, h2, 1.0 / 120.0);                 // 1/5!
    p = __builtin_fma(p, h2, 1.0 / 6.0);                   // 1/3!
    p = __builtin_fma(p, h2, 1.0);
    return h * p;
}
__device__ __forceinline__ double sin_f64(double h, double h2) {
    double p = 1.0 / 51090942171709440000.0;
    p = __builtin_fma(p, h2, -1.0 / 121645100408832000.0);
    p = __builtin_fma(p, h2, 1.0 / 355687428096000.0);
    p = __builtin_fma(p, h2, -1.0 / 1307674368000.0);
    p = __builtin_fma(p, h2, 1.0 / 6227020800.0);
    p = __builtin_fma(p, h2, -1.0 / 39916800.0);
    p = __builtin_fma(p, h2, 1.0 / 362880.0);
    p = __builtin_fma(p, h2, -1.0 / 5040.0);
    p = __builtin_fma(p, h2, 1.0 / 120.0);
    p = __builtin_fma(p, h2, -1.0 / 6.0);
    p = __builtin_fma(p, h2, 1.0);
    return h * p;
}
// Periodic-kernel trigonometry on |h| <= pi/2, z = h^2 <= Z = (pi/2)^2.  The polynomials are
// Chebyshev economizations of the Taylor series: the function interpolated at 40 Chebyshev
// nodes of its z-range in 60-digit arithmetic, truncated to the degree below, expanded in
// powers of z and rounded to the working precision (the hex literals are those values).
// The truncation costs less than the coefficient rounding, so they are as accurate as the
// longer Taylor forms and take 2 (cos) and 1 (sinh/sin) fewer FMAs.

// 2 cos(h) = C(z), degree 9 on 0 <= z <= Z: truncation 1.3e-20, 2.9e-17 absolute in all.
__device__ __forceinline__ double cos2_f64(double h2) {
    double p = -0x1.5ca3e79f9306cp-52;                    // z^9
    p = __builtin_fma(p, h2, 0x1.ae432ee9be395p-44);
    p = __builtin_fma(p, h2, -0x1.93969ca697656p-36);
    p = __builtin_fma(p, h2, 0x1.1eed8dcefb537p-28);
    p = __builtin_fma(p, h2, -0x1.27e4fb74f552cp-21);
    p = __builtin_fma(p, h2, 0x1.a01a01a013307p-15);
    p = __builtin_fma(p, h2, -0x1.6c16c16c16b71p-9);
    p = __builtin_fma(p, h2, 0x1.5555555555554p-4);
    p = __builtin_fma(p, h2, -0x1p+0);
    return __builtin_fma(p, h2, 0x1p+1);                  // z^0
}

// 2 sinh(h) and 2 sin(h) together: with P(z) = 2 sinh(sqrt z)/sqrt z (2 sin(sqrt -z)/sqrt -z
// for z < 0), 2 sinh(h) = h P(z) and 2 sin(h) = h P(-z).  P has degree 9 on -Z <= z <= Z
// (truncation 3.2e-19, 4.2e-17 relative in all); splitting it into even and odd powers,
// P(+-z) = A(z^2) +- z B(z^2), evaluates both signs in one pass.  On |h| <= pi/2 the sin
// branch A - z B stays above 4/pi: no cancellation.
__device__ __forceinline__ void sinh_sin2_f64(double h, double h2, double &sh, double &sn) {
    const double w = h2 * h2;
    double a = 0x1.953774b9570acp-48;                     // z^8
    a = __builtin_fma(a, w, 0x1.612460fff9d65p-32);
    a = __builtin_fma(a, w, 0x1.71de3a556e707p-18);
    a = __builtin_fma(a, w, 0x1.1111111111110p-6);
    a = __builtin_fma(a, w, 0x1p+1);                      // z^0
    double b = 0x1.2f4fd2496d099p-56;                     // z^9
    b = __builtin_fma(b, w, 0x1.ae7f3e4dfe3d8p-40);
    b = __builtin_fma(b, w, 0x1.ae64567f5514ap-25);
    b = __builtin_fma(b, w, 0x1.a01a01a01a01ap-12);
    b = __builtin_fma(b, w, 0x1.5555555555555p-2);        // z^1
    const double zb = h2 * b;
    sh = h * (a + zb);
    sn = h * (a - zb);
}
// FP32 forms, fitted the same way and rounded to FP32: 2 cos(h) of degree 5 (5.4e-8
// absolute) and P of degree 5 (2.2e-8 relative).
__device__ __forceinline__ void sinh_sin2_f32(float h, float h2, float &sh, float &sn) {
    const float w = h2 * h2;
    float a = 0x1.7210ap-18f;                             // z^4
    a = __builtin_fmaf(a, w, 0x1.11110ap-6f);
    a = __builtin_fmaf(a, w, 0x1p+1f);                    // z^0
    float b = 0x1.ae882cp-25f;                            // z^5
    b = __builtin_fmaf(b, w, 0x1.a019fep-12f);
    b = __builtin_fmaf(b, w, 0x1.555556p-2f);             // z^1
    const float zb = h2 * b;
    sh = h * (a + zb);
    sn = h * (a - zb);
}
__device__ __forceinline__ float cos2_f32(float h2) {
    float p = -0x1.17b8a8p-21f;                           // z^5
    p = __builtin_fmaf(p, h2, 0x1.9f67f8p-15f);
    p = __builtin_fmaf(p, h2, -0x1.6c1336p-9f);
    p = __builtin_fmaf(p, h2, 0x1.555544p-4f);
    p = __builtin_fmaf(p, h2, -0x1p+0f);
    return __builtin_fmaf(p, h2, 0x1p+1f);                // z^0
}

// ---------------------------------------------------------------- infinite plane

// src/compute.cpp:117-126: u -= G/(2 pi) dy / (r^2 + eps^2), v += G/(2 pi) dx / (...).
struct InfF64 {
    struct Src {
        double x, y, g, pad;
    };
    struct Tgt {
        double x, y;
    };
    struct Acc {
        double u, v;
    };
    using SymAcc = Acc;
    static constexpr int kFoldEvery = 0;
    static constexpr bool kBatched = true;
    static constexpr bool kNearTiles = false;
    static constexpr bool kSymNear = false;
    static constexpr int kRingUnroll = 2;
    __device__ static Src src(const GeoParams &, double x, double y, double g) {
        return {x, y, kInvTwoPi * g, 0.0};
    }
    __device__ static Tgt tgt(const GeoParams &, double x, double y) { return {x, y}; }
    __device__ static void zero(Acc &a) { a.u = a.v = 0.0; }
    __device__ static void pair(const GeoParams &p, const Tgt &t, const Src &s, Acc &a) {
        const double dx = t.x - s.x, dy = t.y - s.y;
        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, p.eps2));
        const double c = s.g * rcp64(r2);
        a.u = __builtin_fma(-c, dy, a.u);
        a.v = __builtin_fma(c, dx, a.v);
    }
    __device__ static void pair_checked(const GeoParams &p, const Tgt &t, const Src &s, Acc &a,
                                        bool self) {
        const double dx = t.x - s.x, dy = t.y - s.y;
        double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, p.eps2));
        r2 = self ? 1.0 : r2;
        const double c = self ? 0.0 : s.g * rcp64(r2);
        a.u = __builtin_fma(-c, dy, a.u);
        a.v = __builtin_fma(c, dx, a.v);
    }
    __device__ static void fold(Acc &) {}
    __device__ static void result(const Acc &a, double &u, double &v) {
        u = a.u;
        v = a.v;
    }
    // Batched form: prep() returns the pair's denominators, apply() consumes their inverses.
    static constexpr int kDenoms = 1;
    struct Tmp {
        double dx, dy;
    };
    __device__ static void prep(const GeoParams &p, const Tgt &t, const Src &s, Tmp &m,
                                double *d) {
        m.dx = t.x - s.x;
        m.dy = t.y - s.y;
        d[0] = __builtin_fma(m.dx, m.dx, __builtin_fma(m.dy, m.dy, p.eps2));
    }
    __device__ static void apply(const Src &s, const Tmp &m, const double *inv, Acc &a) {
        const double c = s.g * inv[0];
        a.u = __builtin_fma(-c, m.dy, a.u);
        a.v = __builtin_fma(c, m.dx, a.v);
    }
    // Symmetric evaluation (each unordered pair once): the term on j from i reuses the
    // difference and the inverse with i's circulation and the opposite sign of (dx, dy).
    static constexpr int kSymDenoms = 1;
    using SymTgt = Tgt;
    using SymTmp = Tmp;
    using SymScalar = double;
    using SymR = Acc;
    __device__ static double sym_g(const GeoParams &, double g) { return kInvTwoPi * g; }
    __device__ static SymTgt sym_tgt(const GeoParams &p, double x, double y) { return tgt(p, x, y); }
    __device__ static void sym_prep(const GeoParams &p, const SymTgt &t, const Src &s, SymTmp &m,
                                    double *d) {
        prep(p, t, s, m, d);
    }
    __device__ static void sym_apply(const Src &s, double gi, const SymTmp &m, const double *inv,
                                     Acc &a, Acc &r) {
        apply(s, m, inv, a);
        const double c = gi * inv[0];
        r.u = __builtin_fma(c, m.dy, r.u);
        r.v = __builtin_fma(-c, m.dx, r.v);
    }
};

// Mixed: FP64 difference, FP32 1/r^2 and products, FP32 tile sums folded into FP64.
struct InfMixed {
    struct Src {
        double x, y;
        float g, pad0;
        float pad1, pad2;
    };
    struct Tgt {
        double x, y;
    };
    using Acc = MixedAcc<Comp>;
    using SymAcc = MixedAcc<BlockSum>;
    static constexpr int kFoldEvery = 1;
    static constexpr bool kBatched = false;
    static constexpr bool kNearTiles = false;
    static constexpr bool kSymNear = false;
    static constexpr int kRingUnroll = 1;
    static constexpr bool kSymTwoPhase = false;
    __device__ static Src src(const GeoParams &, double x, double y, double g) {
        return {x, y, (float)(kInvTwoPi * g), 0.f, 0.f, 0.f};
    }
    __device__ static Tgt tgt(const GeoParams &, double x, double y) { return {x, y}; }
    template <class A> __device__ static void zero(A &a) {
        a.u32 = a.v32 = 0.f;
        a.u.zero();
        a.v.zero();
    }
    template <class A> __device__ static void pair(const GeoParams &p, const Tgt &t, const Src &s, A &a) {
        const float dx = (float)(t.x - s.x), dy = (float)(t.y - s.y);
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, (float)p.eps2));
        const float c = s.g * rcp32(r2);
        a.u32 = __builtin_fmaf(-c, dy, a.u32);
        a.v32 = __builtin_fmaf(c, dx, a.v32);
    }
    template <class A> __device__ static void pair_checked(const GeoParams &p, const Tgt &t, const Src &s, A &a,
                                        bool self) {
        const float dx = (float)(t.x - s.x), dy = (float)(t.y - s.y);
        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, (float)p.eps2));
        r2 = self ? 1.f : r2;
        const float c = self ? 0.f : s.g * rcp32(r2);
        a.u32 = __builtin_fmaf(-c, dy, a.u32);
        a.v32 = __builtin_fmaf(c, dx, a.v32);
    }
    template <class A> __device__ static void fold(A &a) {
        a.u.add((double)a.u32);
        a.v.add((double)a.v32);
        a.u32 = a.v32 = 0.f;
    }
    template <class A> __device__ static void result(const A &a, double &u, double &v) {
        u = a.u.value();
        v = a.v.value();
    }
    // Symmetric evaluation: one FP64 difference, one FP32 reciprocal per unordered pair; the
    // reaction on the source is an FP32 sum (over at most 64 x 4 targets) like the target
    // side's per-tile FP32 sums, and is added to the FP64 partials after that.
    static constexpr int kSymDenoms = 0;
    using SymTgt = Tgt;
    struct SymTmp {};
    using SymScalar = float;
    struct SymR {
        float u, v;
    };
    __device__ static float sym_g(const GeoParams &, double g) { return (float)(kInvTwoPi * g); }
    __device__ static SymTgt sym_tgt(const GeoParams &p, double x, double y) { return tgt(p, x, y); }
    // The pair in two parts, the FP64 difference (sym_diff) and the FP32 rest (sym_rest), so
    // that a policy's ring step can issue every target's difference before any FP32 work
    // (sym_step, kSymTwoPhase).  kDirect only selects a form in DiskMixed.
    __device__ static void sym_diff(const GeoParams &, const SymTgt &t, const Src &s, float &dx,
                                    float &dy) {
        dx = (float)(t.x - s.x);
        dy = (float)(t.y - s.y);
    }
    template <bool kDirect, class A>
    __device__ static void sym_rest(const GeoParams &p, const SymTgt &, float gi, const Src &s,
                                    float dx, float dy, A &a, SymR &r) {
        const float inv = rcp32(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, (float)p.eps2)));
        const float c = s.g * inv;
        a.u32 = __builtin_fmaf(-c, dy, a.u32);
        a.v32 = __builtin_fmaf(c, dx, a.v32);
        const float ci = gi * inv;
        r.u = __builtin_fmaf(ci, dy, r.u);
        r.v = __builtin_fmaf(-ci, dx, r.v);
    }
};

// ---------------------------------------------------------------- disk

// src/compute.cpp:306-329: direct term (j != i) plus the circle-theorem image of every
// source (j == i included).  t = x_i/R, s = x_j/R, a = 1 - t.s, b = t_y s_x - t_x s_y,
// u -= c(-a s_y + b s_x), v += c(-a s_x - b s_y), c = -G/(2 pi R (a^2+b^2)).
// FP64 disk (src/compute.cpp:286-336).  The circle-theorem image of source j (strength G_j at
// x_j, |x_j| < R) is a vortex of strength -G_j at the image point x_j* = R^2 x_j / |x_j|^2:
// with t = x_i/R, s = x_j/R, a = 1 - t.s and b = t x s, the reference's image velocity
// -G/(2 pi R (a^2 + b^2)) (-a s_y + b s_x, -a s_x - b s_y) equals the Biot-Savart velocity of
// -G at x_j* (1 - t conj(s) = (a - i b) and x_i - x_j* = R (t - s/|s|^2)).  So each pair is two
// plane-vortex terms with the image point computed once per staged source.  A source within
// 2^-100 R of the centre has its image point near overflow; a tile holding one is computed
// with the reference's a, b form instead (the tile flag is uniform, so no per-pair branch).
// Offset of the image point x* = x R^2/|x|^2 from x: delta = x (R^2 - |x|^2)/|x|^2.  R^2 - |x|^2
// is formed from the rounded squares with their exact rounding errors (fma) and exact sums
// (two_sum), so it keeps a few ulps of relative accuracy however close x is to the wall, and so
// does delta.  The image difference x_i - x_j* = d - delta_j (d = x_i - x_j) is then accurate to
// a few ulps as well: for x_i inside the disk, |d| + |delta_j| <= 5 |x_i - x_j*| (|delta_j| <=
// 2 dist(x_j*, disk) <= 2 |x_i - x_j*|, and |d| <= |x_i - x_j*| + |delta_j|), so the subtraction
// cannot cancel.  Forming x_j* itself instead rounds it by an ulp of R, a relative error of
// ulp(R) / |x_i - x_j*| in the image term, about 1e-12 for pairs 1e-4 from the wall.
// R^2 - |x|^2 to a few ulps, however close x is to the wall.  Contraction is off: R2 - x*x
// fused into one fma would no longer pair with the fma that measures x*x's rounding.
__device__ __forceinline__ double wall_gap(const GeoParams &p, double x, double y) {
#pragma clang fp contract(off)
    const double px = x * x, ex = __builtin_fma(x, x, -px);
    const double py = y * y, ey = __builtin_fma(y, y, -py);
    double s1, e1, s2, e2;
    two_sum(p.R2, -px, s1, e1);
    two_sum(s1, -py, s2, e2);
    return s2 + (((e1 + e2) + p.R2lo) - (ex + ey));
}

__device__ __forceinline__ void image_offset(const GeoParams &p, double x, double y, double r2,
                                             double &ox, double &oy) {
    const double q = wall_gap(p, x, y) / r2;
    ox = x * q;
    oy = y * q;
}

struct DiskF64 {
    // g = G/(2 pi); (ox, oy) = delta, the image point's offset x* - x (image_offset; the image
    // has strength -g); w = 1/|x|^2.  A source within 2^-100 R of the centre has no finite
    // image: (ox, oy, w) = 0, and every tile that holds one is evaluated in the reference's
    // form instead.
    struct Src {
        double x, y, g, ox, oy, w;
    };
    struct Tgt {
        double x, y, tx, ty;
    };
    struct Acc {
        double u, v;
    };
    using SymAcc = Acc;
    static constexpr int kFoldEvery = 0;
    static constexpr bool kBatched = true;
    static constexpr bool kNearTiles = true;
    static constexpr bool kSymNear = false;
    static constexpr int kRingUnroll = 2;
    __device__ static bool is_near(const GeoParams &p, double x, double y) {
        return __builtin_fma(x, x, y * y) < 0x1p-200 * p.R2;
    }
    __device__ static Src src(const GeoParams &p, double x, double y, double g) {
        const double r2 = __builtin_fma(x, x, y * y);
        const double gg = kInvTwoPi * g;
        if (r2 < 0x1p-200 * p.R2)
            return {x, y, gg, 0.0, 0.0, 0.0};
        double ox, oy;
        image_offset(p, x, y, r2, ox, oy);
        return {x, y, gg, ox, oy, 1.0 / r2};
    }
    __device__ static Tgt tgt(const GeoParams &p, double x, double y) {
        return {x, y, x * p.invR, y * p.invR};
    }
    __device__ static void zero(Acc &a) { a.u = a.v = 0.0; }
    __device__ static void point(double dx, double dy, double g, Acc &a) {
        const double c = g * rcp64(__builtin_fma(dx, dx, dy * dy));
        a.u = __builtin_fma(-c, dy, a.u);
        a.v = __builtin_fma(c, dx, a.v);
    }
    __device__ static void pair(const GeoParams &, const Tgt &t, const Src &s, Acc &a) {
        const double dx = t.x - s.x, dy = t.y - s.y;
        point(dx, dy, s.g, a);
        point(dx - s.ox, dy - s.oy, -s.g, a);
    }
    __device__ static void pair_checked(const GeoParams &, const Tgt &t, const Src &s, Acc &a,
                                        bool self) {
        const double dx = t.x - s.x, dy = t.y - s.y;
        double r2 = __builtin_fma(dx, dx, dy * dy);
        r2 = self ? 1.0 : r2;
        const double c = self ? 0.0 : s.g * rcp64(r2);
        a.u = __builtin_fma(-c, dy, a.u);
        a.v = __builtin_fma(c, dx, a.v);
        point(dx - s.ox, dy - s.oy, -s.g, a);     // a vortex's own image too (d = 0)
    }
    // The reference's form (src/compute.cpp:318-329), for tiles holding a near-centre source.
    __device__ static void pair_ref_checked(const GeoParams &p, const Tgt &t, const Src &s,
                                            Acc &a, bool self) {
        const double dx = t.x - s.x, dy = t.y - s.y;
        double r2 = __builtin_fma(dx, dx, dy * dy);
        r2 = self ? 1.0 : r2;
        const double c = self ? 0.0 : s.g * rcp64(r2);
        a.u = __builtin_fma(-c, dy, a.u);
        a.v = __builtin_fma(c, dx, a.v);
        const double sx = s.x * p.invR, sy = s.y * p.invR, gr = s.g * p.invR;
        const double a_ = 1.0 - __builtin_fma(t.tx, sx, t.ty * sy);
        const double b_ = __builtin_fma(t.ty, sx, -(t.tx * sy));
        const double k = gr * rcp64(__builtin_fma(a_, a_, b_ * b_));
        const double pu = __builtin_fma(a_, sy, -(b_ * sx));
        const double pv = __builtin_fma(a_, sx, b_ * sy);
        a.u = __builtin_fma(-k, pu, a.u);
        a.v = __builtin_fma(k, pv, a.v);
    }
    __device__ static void fold(Acc &) {}
    __device__ static void result(const Acc &a, double &u, double &v) {
        u = a.u;
        v = a.v;
    }
    static constexpr int kDenoms = 2;
    struct Tmp {
        double dx, dy, ex, ey;
    };
    __device__ static void prep(const GeoParams &, const Tgt &t, const Src &s, Tmp &m,
                                double *d) {
        m.dx = t.x - s.x;
        m.dy = t.y - s.y;
        d[0] = __builtin_fma(m.dx, m.dx, m.dy * m.dy);
        m.ex = m.dx - s.ox;
        m.ey = m.dy - s.oy;
        d[1] = __builtin_fma(m.ex, m.ex, m.ey * m.ey);
    }
    __device__ static void apply(const Src &s, const Tmp &m, const double *inv, Acc &a) {
        const double c = s.g * inv[0];
        a.u = __builtin_fma(-c, m.dy, a.u);
        a.v = __builtin_fma(c, m.dx, a.v);
        const double k = s.g * inv[1];
        a.u = __builtin_fma(k, m.ey, a.u);
        a.v = __builtin_fma(-k, m.ex, a.v);
    }
    // Symmetric evaluation: the direct term is shared between (i, j) and (j, i); each side
    // still gets its own image term (j's image at i, i's image at j).  The two image distances
    // share one reciprocal: |x_i - x_j*|^2 |x_j|^2 = |x_i|^2 |x_j|^2 - 2 R^2 x_i.x_j + R^4 is
    // symmetric in i and j, so 1/|x_j - x_i*|^2 = |x_i|^2 w_j / |x_i - x_j*|^2.  Both
    // differences are still formed from the accurate offsets (no cancellation); only the
    // reciprocal is shared.  Blocks holding a near-centre vortex use the reference's form both
    // ways.
    static constexpr int kSymDenoms = 2;
    using SymScalar = double;
    using SymR = Acc;
    struct SymTgt {
        double x, y, jx, jy, r2;               // jx, jy: delta_i = x_i* - x_i
    };
    struct SymTmp {
        double dx, dy, ex, ey, r2, w, jx, jy;  // x_i - x_j, x_i - x_j*, |x_i|^2, w_j, delta_i
    };
    __device__ static double sym_g(const GeoParams &, double g) { return kInvTwoPi * g; }
    __device__ static SymTgt sym_tgt(const GeoParams &p, double x, double y) {
        const Src s = src(p, x, y, 0.0);
        return {x, y, s.ox, s.oy, __builtin_fma(x, x, y * y)};
    }
    __device__ static void sym_prep(const GeoParams &, const SymTgt &t, const Src &s, SymTmp &m,
                                    double *d) {
        m.dx = t.x - s.x;
        m.dy = t.y - s.y;
        d[0] = __builtin_fma(m.dx, m.dx, m.dy * m.dy);
        m.ex = m.dx - s.ox;
        m.ey = m.dy - s.oy;
        d[1] = __builtin_fma(m.ex, m.ex, m.ey * m.ey);
        m.r2 = t.r2;
        m.w = s.w;
        m.jx = t.jx;
        m.jy = t.jy;
    }
    __device__ static void sym_apply(const Src &s, double gi, const SymTmp &m, const double *inv,
                                     Acc &a, Acc &r) {
        const double c = s.g * inv[0];
        a.u = __builtin_fma(-c, m.dy, a.u);
        a.v = __builtin_fma(c, m.dx, a.v);
        const double k = s.g * inv[1];             // j's image has strength -G_j
        a.u = __builtin_fma(k, m.ey, a.u);
        a.v = __builtin_fma(-k, m.ex, a.v);
        // The reaction: G_i on j and i's image (strength -G_i; G_i |x_i|^2 is the same for
        // every source of the loop) on j.  With x_j - x_i* = -d - delta_i the two combine:
        // ci (dy, -dx) + ki (fy, -fx) = (ci - ki) (dy, -dx) + ki (-delta_iy, delta_ix).
        const double ci = gi * inv[0];
        const double ki = (gi * m.r2) * (inv[1] * m.w);
        const double cm = ci - ki;
        r.u = __builtin_fma(-ki, m.jy, __builtin_fma(cm, m.dy, r.u));
        r.v = __builtin_fma(ki, m.jx, __builtin_fma(-cm, m.dx, r.v));
    }
};

// Mixed disk.  The direct term as in the infinite plane (FP64 difference d = x_i - x_j, FP32
// after).  The image term is the plane-vortex term of -G_j at the image point x_j* (see
// DiskF64), with x_i - x_j* = d - delta_j, where delta_j = x_j* - x_j = x_j (R^2/|x_j|^2 - 1) is
// formed in FP64 once per source and stored in FP32.  The subtraction d - delta_j is safe in
// FP32: x_i is inside the wall and x_j* outside it, so |d - delta_j| is never much smaller than
// |d| + |delta_j| (within a factor ~3), and its rounding stays a few FP32 ulps.  A source within
// 1e-30 R of the centre gets a far, finite offset: its image term then evaluates to 0 (its true
// size is below 1e-30 of the source's own scale) instead of overflowing.
struct DiskMixed {
    struct Src {
        double x, y;
        float g, ix, iy, w;        // g = G/(2 pi); (ix, iy) = delta; w: sym_src
    };
    struct Tgt {
        double x, y;
    };
    using Acc = MixedAcc<Comp>;
    using SymAcc = MixedAcc<BlockSum>;
    static constexpr int kFoldEvery = 1;
    static constexpr bool kBatched = false;
    static constexpr bool kNearTiles = false;
    static constexpr bool kSymNear = true;
    static constexpr int kRingUnroll = 2;
    static constexpr bool kSymTwoPhase = true;
    __device__ static Src src(const GeoParams &p, double x, double y, double g) {
        const double r2 = __builtin_fma(x, x, y * y);
        float ix = 1e30f, iy = 0.f;
        if (r2 >= 1e-60 * p.R2) {
            const double q = p.R2 / r2 - 1.0;
            ix = (float)fmin(fmax(x * q, -1e30), 1e30);
            iy = (float)fmin(fmax(y * q, -1e30), 1e30);
        }
        return {x, y, (float)(kInvTwoPi * g), ix, iy, 0.f};
    }
    __device__ static Tgt tgt(const GeoParams &, double x, double y) { return {x, y}; }
    template <class A> __device__ static void zero(A &a) {
        a.u32 = a.v32 = 0.f;
        a.u.zero();
        a.v.zero();
    }
    // j's image on i; `self` (j == i) leaves the term out: disk_self_image adds it in FP64.
    template <class A>
    __device__ static void image(const Src &s, float dx, float dy, A &a, bool self = false) {
        const float ex = dx - s.ix, ey = dy - s.iy;
        float k = s.g * rcp32(__builtin_fmaf(ex, ex, ey * ey));
        k = self ? 0.f : k;
        a.u32 = __builtin_fmaf(k, ey, a.u32);
        a.v32 = __builtin_fmaf(-k, ex, a.v32);
    }
    template <class A> __device__ static void pair(const GeoParams &, const Tgt &t, const Src &s, A &a) {
        const float dx = (float)(t.x - s.x), dy = (float)(t.y - s.y);
        const float c = s.g * rcp32(__builtin_fmaf(dx, dx, dy * dy));
        a.u32 = __builtin_fmaf(-c, dy, a.u32);
        a.v32 = __builtin_fmaf(c, dx, a.v32);
        image(s, dx, dy, a);
    }
    template <class A> __device__ static void pair_checked(const GeoParams &, const Tgt &t, const Src &s, A &a,
                                        bool self) {
        // The self pair has no direct term, and its own image (d = 0) is added in FP64 where
        // the target's velocity is finalized (disk_self_image), outside the FP32 tile sums.
        const float dx = (float)(t.x - s.x), dy = (float)(t.y - s.y);
        float r2 = __builtin_fmaf(dx, dx, dy * dy);
        r2 = self ? 1.f : r2;
        const float c = self ? 0.f : s.g * rcp32(r2);
        a.u32 = __builtin_fmaf(-c, dy, a.u32);
        a.v32 = __builtin_fmaf(c, dx, a.v32);
        image(s, dx, dy, a, self);
    }
    template <class A> __device__ static void fold(A &a) {
        a.u.add((double)a.u32);
        a.v.add((double)a.v32);
        a.u32 = a.v32 = 0.f;
    }
    template <class A> __device__ static void result(const A &a, double &u, double &v) {
        u = a.u.value();
        v = a.v.value();
    }
    // Symmetric evaluation: d = x_i - x_j once (FP64 difference, FP32 after).  j's image at i is
    // d - delta_j, i's image at j is x_j - x_i* = -d - delta_i (the target carries delta_i).  The
    // two image terms share one FP32 reciprocal, as in DiskF64: 1/|x_j - x_i*|^2 =
    // |x_i|^2 w_j / |x_i - x_j*|^2.  That needs w_j = 1/|x_j|^2 and the product in FP32 range,
    // so blocks holding a vortex within 2^-40 R of the centre (is_sym_near) take the direct form
    // (kDirect), with its own reciprocal for i's image.  The reaction on the source is an FP32
    // sum (over at most 64 x 4 targets) before it joins the FP64 partials.
    static constexpr int kSymDenoms = 0;
    struct SymTgt {
        double x, y;
        float ix, iy;           // delta_i
        float r2;               // |x_i|^2
    };
    struct SymTmp {};
    using SymScalar = float;
    struct SymR {
        float u, v;
    };
    __device__ static bool is_sym_near(const GeoParams &p, double x, double y) {
        return __builtin_fma(x, x, y * y) < 0x1p-80 * p.R2;
    }
    // The symmetric kernel's source record: src() plus w = 1/|x|^2 (0 within 2^-40 R of the
    // centre, where the direct form runs).  Only the single-system symmetric kernel stages it:
    // for the ensemble's small tile pairs the extra reciprocal per staged source cost more
    // than the shared reciprocal saves (mixed ensemble 3.7 % slower), so the ensemble keeps
    // the direct form and src().
    __device__ static Src sym_src(const GeoParams &p, double x, double y, double g) {
        Src s = src(p, x, y, g);
        const double r2 = __builtin_fma(x, x, y * y);
        s.w = r2 >= 0x1p-80 * p.R2 ? rcp32((float)r2) : 0.f;
        return s;
    }
    __device__ static float sym_g(const GeoParams &, double g) { return (float)(kInvTwoPi * g); }
    __device__ static SymTgt sym_tgt(const GeoParams &p, double x, double y) {
        const Src s = src(p, x, y, 0.0);
        return {x, y, s.ix, s.iy, (float)__builtin_fma(x, x, y * y)};
    }
    __device__ static void sym_diff(const GeoParams &, const SymTgt &t, const Src &s, float &dx,
                                    float &dy) {
        dx = (float)(t.x - s.x);
        dy = (float)(t.y - s.y);
    }
    template <bool kDirect, class A>
    __device__ static void sym_rest(const GeoParams &, const SymTgt &t, float gi, const Src &s,
                                    float dx, float dy, A &a, SymR &r) {
        const float inv = rcp32(__builtin_fmaf(dx, dx, dy * dy));
        const float c = s.g * inv;
        a.u32 = __builtin_fmaf(-c, dy, a.u32);
        a.v32 = __builtin_fmaf(c, dx, a.v32);
        const float ci = gi * inv;
        if constexpr (kDirect) {
            r.u = __builtin_fmaf(ci, dy, r.u);
            r.v = __builtin_fmaf(-ci, dx, r.v);
        }
        const float ex = dx - s.ix, ey = dy - s.iy;
        const float inv_e = rcp32(__builtin_fmaf(ex, ex, ey * ey));
        const float k = s.g * inv_e;               // -G_j at x_j*, on i
        a.u32 = __builtin_fmaf(k, ey, a.u32);
        a.v32 = __builtin_fmaf(-k, ex, a.v32);
        if constexpr (kDirect) {
            const float fx = -dx - t.ix, fy = -dy - t.iy;
            const float kf = gi * rcp32(__builtin_fmaf(fx, fx, fy * fy));   // -G_i at x_i*
            r.u = __builtin_fmaf(kf, fy, r.u);
            r.v = __builtin_fmaf(-kf, fx, r.v);
        } else {
            // -G_i at x_i*, on j; G_i |x_i|^2 is the same for every source.  With
            // f = -d - delta_i the direct and image reactions combine:
            // ci (dy, -dx) + kf (fy, -fx) = (ci - kf) (dy, -dx) + kf (-delta_iy, delta_ix).
            const float kf = (gi * t.r2) * (s.w * inv_e);
            const float m = ci - kf;
            r.u = __builtin_fmaf(-kf, t.iy, __builtin_fmaf(m, dy, r.u));
            r.v = __builtin_fmaf(kf, t.ix, __builtin_fmaf(-m, dx, r.v));
        }
    }
};

// The mixed disk's self-image term: vortex i's own image (src/compute.cpp:320-329 with j = i),
// the velocity of -G_i at x_i* on x_i, (G_i / 2 pi) (-delta_y, delta_x) / |delta|^2 with
// delta = x_i* - x_i (image_offset).  Near the wall it is a vortex's largest term (|delta| is
// about twice the gap), and inside an FP32 tile sum every later addition would round at its
// scale (a vortex 5e-4 from the wall: 1.2e-6 relative error in its velocity).  DiskMixed's
// pair_checked therefore leaves it out, and every kernel that finalizes a mixed disk target's
// velocity adds it here, in FP64, once.  A vortex within 1e-30 R of the centre has an image
// term below 1e-30 of its own scale and gets none (src() gives its image a far offset).
__device__ __forceinline__ void disk_self_image(const GeoParams &p, double x, double y, double g,
                                                double &u, double &v) {
    const double r2 = __builtin_fma(x, x, y * y);
    if (!(r2 >= 1e-60 * p.R2))
        return;
    double ox, oy;
    image_offset(p, x, y, r2, ox, oy);
    const double k = kInvTwoPi * g / __builtin_fma(ox, ox, oy * oy);
    u = __builtin_fma(-k, oy, u);
    v = __builtin_fma(k, ox, v);
}
template <class P> constexpr bool kSelfImage = std::is_same_v<P, DiskMixed>;

// ---------------------------------------------------------------- periodic box

// src/compute.cpp:223-249.  With X = k rem(dx, L), Y = k rem(dy, L) in [-pi, pi]:
//   u -= scale G [ sin Y / (cosh X - cos Y) + sum_{m != 0, |m| <= L} sin Y / (cosh(X - 2 pi m) - cos Y) ]
// (and v with X, Y exchanged).  The m = 0 row keeps the reference's cancellation-free
// denominator d0 = 2(sinh^2(X/2) + sin^2(Y/2)).
// Rows m = +-1 in closed form: with c = cosh X, k = cos Y, C = cosh 2 pi, S = sinh 2 pi,
// D+- = cosh(X -+ 2 pi) - k = C c -+ S sinh X - k, so D+ + D- = 2 (C c - k) and
// D+ D- = (C c - k)^2 - S^2 (c^2 - 1) = c^2 - 2 C k c + k^2 + S^2 (C^2 - S^2 = 1).  In the
// doubled variables ch2 = 2c, cs2 = 2k:  t = C ch2 - cs2 = D+ + D-  and
// q = ch2 (ch2 - 2 C cs2) + cs2^2 + 4 S^2 = 4 D+ D-.  q >= 4 (C - cosh pi)^2 while its one
// negative term is below 9% of 4 S^2, so nothing cancels; the two rows add sin Y * 4 t / q and
// join the m = 0 row over the common denominator d0 q.
// Rows |m| >= 2: for |X| < 4 pi the exact expansion
//   sum_{|m| >= 2} ... = 4 sum_{n>=1} G_n sin(nY) cosh(nX),  G_n = sum_{m=2..L} e^{-2 pi n m},
// evaluated with the Chebyshev recurrences for sin(nY) and cosh(nX).  Term n is at most
// 4 n e^{-4 pi n} cosh(n pi) |sin Y| (|sin nY| <= n |sin Y|), and the m = 0 row is at least
// |sin Y| / (cosh pi + 1), so 4 terms leave below 5e-19 of the result (FP64) and 2 terms
// below 4e-11 (mixed).  The reference drops image terms with
// |X - 2 pi m| > 40 (below 5e-18 relative); the closed form and the series keep them.
struct PerF64 {
    struct Src {
        double x, y, gs, pad;
    };
    struct Tgt {
        double x, y;
    };
    struct Acc {
        double u, v;
    };
    using SymAcc = Acc;
    static constexpr int kFoldEvery = 0;
    static constexpr bool kBatched = false;
    static constexpr bool kNearTiles = false;
    static constexpr bool kSymNear = false;
    static constexpr int kRingUnroll = 2;
    __device__ static Src src(const GeoParams &p, double x, double y, double g) {
        return {x, y, p.scale * g, 0.0};
    }
    __device__ static Tgt tgt(const GeoParams &, double x, double y) { return {x, y}; }
    __device__ static void zero(Acc &a) { a.u = a.v = 0.0; }
    __device__ static void terms(const GeoParams &p, const Tgt &t, const Src &s, double &fu,
                                 double &fv, bool self) {
        const double hx = p.hwave * remainder_fast(t.x - s.x, p.L, p.invL);   // X / 2
        const double hy = p.hwave * remainder_fast(t.y - s.y, p.L, p.invL);
        const double hx2 = hx * hx, hy2 = hy * hy;
        double shx, shy, snx, sny;                     // 2 sinh(X/2), 2 sin(X/2)
        sinh_sin2_f64(hx, hx2, shx, snx);
        sinh_sin2_f64(hy, hy2, shy, sny);
        const double sin_x = snx * cos2_f64(hx2);      // 2 sin X
        const double sin_y = sny * cos2_f64(hy2);
        const double snx2 = snx * snx, sny2 = sny * sny;
        const double shx2 = shx * shx, shy2 = shy * shy;
        const double cos2x = 2.0 - snx2, cos2y = 2.0 - sny2;   // 2 cos X
        const double ch2x = 2.0 + shx2, ch2y = 2.0 + shy2;     // 2 cosh X
        const double du = self ? 1.0 : shx2 + sny2;             // 2 d0
        const double dv = self ? 1.0 : shy2 + snx2;
        // Rows |m| >= 2: 2 sin(nY) and 2 cosh(nX) by their recurrences, n = 1..kImageTermsF64
        // (G4 holds G_n, so G_n (2 sin nY)(2 cosh nX) is the row sum's term n).
        double sy0 = 0.0, sy1 = sin_y, sx0 = 0.0, sx1 = sin_x;
        double cx0 = 2.0, cx1 = ch2x, cy0 = 2.0, cy1 = ch2y;
        double au = 0.0, av = 0.0;
#pragma unroll
        for (int n = 1; n <= kImageTermsF64; ++n) {
            au = __builtin_fma(p.G4[n] * sy1, cx1, au);
            av = __builtin_fma(p.G4[n] * sx1, cy1, av);
            if (n < kImageTermsF64) {
                const double sy2 = __builtin_fma(cos2y, sy1, -sy0);
                const double sx2 = __builtin_fma(cos2x, sx1, -sx0);
                const double cx2 = __builtin_fma(ch2x, cx1, -cx0);
                const double cy2 = __builtin_fma(ch2y, cy1, -cy0);
                sy0 = sy1, sy1 = sy2, sx0 = sx1, sx1 = sx2;
                cx0 = cx1, cx1 = cx2, cy0 = cy1, cy1 = cy2;
            }
        }
        // Rows 0 and +-1: sin Y / d0 + 4 sin Y t / q = (2 sin Y) (q + W1 t (2 d0)) / ((2 d0) q),
        // W1 = 2.
        const double tu = __builtin_fma(p.C1, ch2x, -cos2y);
        const double tv = __builtin_fma(p.C1, ch2y, -cos2x);
        const double qu = __builtin_fma(ch2x, __builtin_fma(-p.C2, cos2y, ch2x),
                                        __builtin_fma(cos2y, cos2y, p.S4));
        const double qv = __builtin_fma(ch2y, __builtin_fma(-p.C2, cos2x, ch2y),
                                        __builtin_fma(cos2x, cos2x, p.S4));
        const double nu = __builtin_fma(p.W1 * tu, du, qu);
        const double nv = __builtin_fma(p.W1 * tv, dv, qv);
        fu = __builtin_fma(sin_y * nu, rcp64(du * qu), au);
        fv = __builtin_fma(sin_x * nv, rcp64(dv * qv), av);
    }
    __device__ static void pair(const GeoParams &p, const Tgt &t, const Src &s, Acc &a) {
        double fu, fv;
        terms(p, t, s, fu, fv, false);
        a.u = __builtin_fma(-s.gs, fu, a.u);
        a.v = __builtin_fma(s.gs, fv, a.v);
    }
    __device__ static void pair_checked(const GeoParams &p, const Tgt &t, const Src &s, Acc &a,
                                        bool self) {
        double fu, fv;
        terms(p, t, s, fu, fv, self);
        a.u = __builtin_fma(-s.gs, fu, a.u);
        a.v = __builtin_fma(s.gs, fv, a.v);
    }
    // Symmetric evaluation: with X, Y -> -X, -Y every factor is odd or even, so the term on j
    // from i is the same (fu, fv) with i's circulation and the opposite sign.
    static constexpr int kSymDenoms = 0;
    using SymTgt = Tgt;
    struct SymTmp {};
    using SymScalar = double;
    using SymR = Acc;
    __device__ static double sym_g(const GeoParams &p, double g) { return p.scale * g; }
    __device__ static SymTgt sym_tgt(const GeoParams &p, double x, double y) { return tgt(p, x, y); }
    __device__ static void sym_pair(const GeoParams &p, const SymTgt &t, double gi, const Src &s,
                                    Acc &a, Acc &r) {
        double fu, fv;
        terms(p, t, s, fu, fv, false);
        a.u = __builtin_fma(-s.gs, fu, a.u);
        a.v = __builtin_fma(s.gs, fv, a.v);
        r.u = __builtin_fma(gi, fu, r.u);
        r.v = __builtin_fma(-gi, fv, r.v);
    }
    __device__ static void fold(Acc &) {}
    __device__ static void result(const Acc &a, double &u, double &v) {
        u = a.u;
        v = a.v;
    }
};

// Mixed periodic: the remainder and scaling in FP64, everything after in FP32.
struct PerMixed {
    struct Src {
        double x, y;
        float gs, pad0, pad1, pad2;
    };
    struct Tgt {
        double x, y;
    };
    using Acc = MixedAcc<Comp>;
    using SymAcc = MixedAcc<BlockSum>;
    static constexpr int kFoldEvery = 1;
    static constexpr bool kBatched = false;
    static constexpr bool kNearTiles = false;
    static constexpr bool kSymNear = false;
    static constexpr int kRingUnroll = 1;
    static constexpr bool kSymTwoPhase = false;
    __device__ static Src src(const GeoParams &p, double x, double y, double g) {
        return {x, y, (float)(p.scale * g), 0.f, 0.f, 0.f};
    }
    __device__ static Tgt tgt(const GeoParams &, double x, double y) { return {x, y}; }
    template <class A> __device__ static void zero(A &a) {
        a.u32 = a.v32 = 0.f;
        a.u.zero();
        a.v.zero();
    }
    // The half-angles X/2, Y/2 of the pair: remainder and scaling in FP64, narrowed to FP32.
    __device__ static void half_angles(const GeoParams &p, const Tgt &t, const Src &s, float &hx,
                                       float &hy) {
        hx = (float)(p.hwave * remainder_fast(t.x - s.x, p.L, p.invL));
        hy = (float)(p.hwave * remainder_fast(t.y - s.y, p.L, p.invL));
    }
    __device__ static void terms(const GeoParams &p, const Tgt &t, const Src &s, float &fu,
                                 float &fv, bool self) {
        float hx, hy;
        half_angles(p, t, s, hx, hy);
        terms_h(p, hx, hy, fu, fv, self);
    }
    __device__ static void terms_h(const GeoParams &p, float hx, float hy, float &fu, float &fv,
                                   bool self) {
        const float hx2 = hx * hx, hy2 = hy * hy;
        float shx, shy, snx, sny;
        sinh_sin2_f32(hx, hx2, shx, snx);
        sinh_sin2_f32(hy, hy2, shy, sny);
        const float sin_x = snx * cos2_f32(hx2), sin_y = sny * cos2_f32(hy2);
        const float snx2 = snx * snx, sny2 = sny * sny;
        const float shx2 = shx * shx, shy2 = shy * shy;
        const float cos2x = 2.f - snx2, cos2y = 2.f - sny2;
        const float ch2x = 2.f + shx2, ch2y = 2.f + shy2;
        const float du = self ? 1.f : shx2 + sny2;
        const float dv = self ? 1.f : shy2 + snx2;
        float sy0 = 0.f, sy1 = sin_y, sx0 = 0.f, sx1 = sin_x;
        float cx0 = 2.f, cx1 = ch2x, cy0 = 2.f, cy1 = ch2y;
        float au = 0.f, av = 0.f;
#pragma unroll
        for (int n = 1; n <= kImageTermsF32; ++n) {
            au = __builtin_fmaf(p.G4f[n] * sy1, cx1, au);
            av = __builtin_fmaf(p.G4f[n] * sx1, cy1, av);
            if (n < kImageTermsF32) {
                const float sy2 = __builtin_fmaf(cos2y, sy1, -sy0);
                const float sx2 = __builtin_fmaf(cos2x, sx1, -sx0);
                const float cx2 = __builtin_fmaf(ch2x, cx1, -cx0);
                const float cy2 = __builtin_fmaf(ch2y, cy1, -cy0);
                sy0 = sy1, sy1 = sy2, sx0 = sx1, sx1 = sx2;
                cx0 = cx1, cx1 = cx2, cy0 = cy1, cy1 = cy2;
            }
        }
        const float tu = __builtin_fmaf(p.C1f, ch2x, -cos2y);
        const float tv = __builtin_fmaf(p.C1f, ch2y, -cos2x);
        const float qu = __builtin_fmaf(ch2x, __builtin_fmaf(-p.C2f, cos2y, ch2x),
                                        __builtin_fmaf(cos2y, cos2y, p.S4f));
        const float qv = __builtin_fmaf(ch2y, __builtin_fmaf(-p.C2f, cos2x, ch2y),
                                        __builtin_fmaf(cos2x, cos2x, p.S4f));
        const float nu = __builtin_fmaf(p.W1f * tu, du, qu);
        const float nv = __builtin_fmaf(p.W1f * tv, dv, qv);
        fu = __builtin_fmaf(sin_y * nu, rcp32(du * qu), au);
        fv = __builtin_fmaf(sin_x * nv, rcp32(dv * qv), av);
    }
    template <class A> __device__ static void pair(const GeoParams &p, const Tgt &t, const Src &s, A &a) {
        float fu, fv;
        terms(p, t, s, fu, fv, false);
        a.u32 = __builtin_fmaf(-s.gs, fu, a.u32);
        a.v32 = __builtin_fmaf(s.gs, fv, a.v32);
    }
    template <class A> __device__ static void pair_checked(const GeoParams &p, const Tgt &t, const Src &s, A &a,
                                        bool self) {
        float fu, fv;
        terms(p, t, s, fu, fv, self);
        a.u32 = __builtin_fmaf(-s.gs, fu, a.u32);
        a.v32 = __builtin_fmaf(s.gs, fv, a.v32);
    }
    // Symmetric evaluation as in PerF64 (the term is odd under X, Y -> -X, -Y); the reaction on
    // the source is an FP32 sum (over at most 64 x 4 targets) before it joins the FP64 partials.
    static constexpr int kSymDenoms = 0;
    using SymTgt = Tgt;
    struct SymTmp {};
    using SymScalar = float;
    struct SymR {
        float u, v;
    };
    __device__ static float sym_g(const GeoParams &p, double g) { return (float)(p.scale * g); }
    __device__ static SymTgt sym_tgt(const GeoParams &p, double x, double y) { return tgt(p, x, y); }
    // The pair in two parts as in InfMixed: sym_diff gives the half-angles.
    __device__ static void sym_diff(const GeoParams &p, const SymTgt &t, const Src &s, float &hx,
                                    float &hy) {
        half_angles(p, t, s, hx, hy);
    }
    template <bool kDirect, class A>
    __device__ static void sym_rest(const GeoParams &p, const SymTgt &, float gi, const Src &s,
                                    float hx, float hy, A &a, SymR &r) {
        float fu, fv;
        terms_h(p, hx, hy, fu, fv, false);
        a.u32 = __builtin_fmaf(-s.gs, fu, a.u32);
        a.v32 = __builtin_fmaf(s.gs, fv, a.v32);
        r.u = __builtin_fmaf(gi, fu, r.u);
        r.v = __builtin_fmaf(-gi, fv, r.v);
    }
    template <class A> __device__ static void fold(A &a) {
        a.u.add((double)a.u32);
        a.v.add((double)a.v32);
        a.u32 = a.v32 = 0.f;
    }
    template <class A> __device__ static void result(const A &a, double &u, double &v) {
        u = a.u.value();
        v = a.v.value();
    }
};

} // namespace pv
