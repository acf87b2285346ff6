// pv_sym.hpp -- device pieces of the symmetric pair evaluation (each unordered pair once),
// shared by the single-system kernels (pv_kernels.hip) and the ensemble kernels (pv_batch.hip).
#pragma once
#include "pv_device.hpp"

namespace pv {

// Sum of (r.u, r.v) over the 64 lanes of a wave; every lane gets the same bits (an xor
// butterfly adds the same two values on both partners at every level).
template <class A> __device__ __forceinline__ void wave_sum(A &r) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        r.u += __shfl_xor(r.u, off, 64);
        r.v += __shfl_xor(r.v, off, 64);
    }
}

// DPP wave_ror:1 (control 0x13C): every lane takes the value of its neighbour on one side, the
// whole wave as one ring.  A double moves as two 32-bit halves.
constexpr int kDppWaveRor1 = 0x13C;
__device__ __forceinline__ double wave_ror1(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, kDppWaveRor1, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), kDppWaveRor1, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}
__device__ __forceinline__ float wave_ror1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kDppWaveRor1, 0xF, 0xF, true));
}

// The ring's rotation through the LDS crossbar (ds_bpermute) instead of a DPP move: lane l
// takes lane l + 1's value, so the ring's step is +1.  It issues on the LDS pipe rather than
// the VALU, which pays where the VALU is the bound and the step is short (the mixed infinite
// plane).
__device__ __forceinline__ float lds_ror1(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ int lds_ror1_addr(int lane) { return ((lane + 1) & 63) << 2; }

// A compile-time flag passed as a value (to instantiate a generic lambda both ways).
template <bool B> struct Flag {
    static constexpr bool value = B;
};

// The ring's step in lane units (1 or 63, i.e. -1 mod 64): after k rotations lane l holds what
// lane (l + k dir) mod 64 started with.  Read from the rotation itself, so the ring does not
// depend on which way the hardware turns; the same for every lane.
__device__ __forceinline__ int ring_dir(int lane) {
    const int from = __builtin_amdgcn_mov_dpp(lane, kDppWaveRor1, 0xF, 0xF, true);
    return __builtin_amdgcn_readfirstlane((from - lane) & 63);
}

// A source record for the symmetric ring: the policy's sym_src where it has one (kSymNear).
template <class P>
__device__ __forceinline__ typename P::Src ring_src(const GeoParams &gp, double x, double y,
                                                    double g) {
    if constexpr (P::kSymNear)
        return P::sym_src(gp, x, y, g);
    else
        return P::src(gp, x, y, g);
}

// One source against a lane's targets: the target side into acc[t], the reaction on the source
// into r.  The denominators of all TPT pairs share one reciprocal through the product tree.
// kDirect: the block holds a vortex near the centre (P::kSymNear policies; P::is_sym_near).
template <class P, int TPT, bool kDirect = false>
__device__ __forceinline__ void sym_step(const GeoParams &gp, const typename P::SymTgt *tg,
                                         const typename P::SymScalar *gi,
                                         const typename P::Src &s, typename P::SymAcc *acc,
                                         typename P::SymR &r) {
    if constexpr (P::kSymDenoms > 0) {
        constexpr int ND = P::kSymDenoms;
        typename P::SymTmp m[TPT];
        double d[TPT * ND], inv[TPT * ND];
#pragma unroll
        for (int t = 0; t < TPT; ++t)
            P::sym_prep(gp, tg[t], s, m[t], &d[t * ND]);
        const bool ok = TreeInverse<TPT * ND>::run(d, inv);
        if (tree_fallback(ok)) {
#pragma unroll
            for (int e = 0; e < TPT * ND; ++e)
                inv[e] = rcp64(d[e]);
        }
#pragma unroll
        for (int t = 0; t < TPT; ++t)
            P::sym_apply(s, gi[t], m[t], &inv[t * ND], acc[t], r);
    } else if constexpr (std::is_same_v<typename P::SymScalar, float>) {
        // Mixed policies: the FP64 difference (sym_diff), then the FP32 rest (sym_rest).  With
        // kSymTwoPhase every target's difference comes first and a scheduling barrier keeps the
        // FP64 instructions in one group per step: FP32 instructions that follow an FP64 one in
        // a wave's stream do not issue at the FP32 rate.  That pays on the mixed disk, whose
        // FP32 work per pair is three reciprocal chains; on the infinite plane and the periodic
        // box the interleaved order the compiler picks is faster.
        float dx[TPT], dy[TPT];
        if constexpr (P::kSymTwoPhase) {
#pragma unroll
            for (int t = 0; t < TPT; ++t)
                P::sym_diff(gp, tg[t], s, dx[t], dy[t]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < TPT; ++t)
                P::template sym_rest<kDirect>(gp, tg[t], gi[t], s, dx[t], dy[t], acc[t], r);
        } else {
#pragma unroll
            for (int t = 0; t < TPT; ++t) {
                P::sym_diff(gp, tg[t], s, dx[t], dy[t]);
                P::template sym_rest<kDirect>(gp, tg[t], gi[t], s, dx[t], dy[t], acc[t], r);
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < TPT; ++t)
            P::sym_pair(gp, tg[t], gi[t], s, acc[t], r);
    }
}

// A target's velocity from its partial sums: rows s = 0..nt-1 of (u, v) at p + 2 s stride and
// p + (2 s + 1) stride, added in row order with TwoSum.  The reductions are bound by HBM
// latency (one load pair per row and one target per thread), so the loads of kRowBatch rows are
// issued before their additions; the order of the additions, and so the bits, are those of the
// plain loop.
constexpr int kRowBatch = 8;
__device__ __forceinline__ void sum_rows(const double *p, size_t stride, int nt, Comp &su,
                                         Comp &sv) {
    su.zero();
    sv.zero();
    int s = 0;
    for (; s + kRowBatch <= nt; s += kRowBatch) {
        double ru[kRowBatch], rv[kRowBatch];
#pragma unroll
        for (int q = 0; q < kRowBatch; ++q) {
            ru[q] = __builtin_nontemporal_load(p + (size_t)(2 * (s + q)) * stride);
            rv[q] = __builtin_nontemporal_load(p + (size_t)(2 * (s + q) + 1) * stride);
        }
#pragma unroll
        for (int q = 0; q < kRowBatch; ++q) {
            su.add(ru[q]);
            sv.add(rv[q]);
        }
    }
    for (; s < nt; ++s) {
        su.add(__builtin_nontemporal_load(p + (size_t)(2 * s) * stride));
        sv.add(__builtin_nontemporal_load(p + (size_t)(2 * s + 1) * stride));
    }
}

// Two ring steps whose denominators share one product tree (FP64 policies with kSymDenoms):
// both sources' TPT x ND denominators, one rcp64 for all of them, then the two steps' sums, the
// reaction rotating between them as in two sym_step calls.  Twice the tree amortizes its
// reciprocal, its refinement and its range check over twice the pairs.
template <class P, int TPT>
__device__ __forceinline__ void sym_step_pair(const GeoParams &gp, const typename P::SymTgt *tg,
                                              const typename P::SymScalar *gi,
                                              const typename P::Src &sa, const typename P::Src &sb,
                                              typename P::SymAcc *acc, typename P::SymR &r) {
    constexpr int ND = P::kSymDenoms, K = 2 * TPT * ND;
    typename P::SymTmp ma[TPT], mb[TPT];
    double d[K], inv[K];
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::sym_prep(gp, tg[t], sa, ma[t], &d[t * ND]);
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::sym_prep(gp, tg[t], sb, mb[t], &d[(TPT + t) * ND]);
    const bool ok = TreeInverse<K>::run(d, inv);
    if (tree_fallback(ok)) {
#pragma unroll
        for (int e = 0; e < K; ++e)
            inv[e] = rcp64(d[e]);
    }
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::sym_apply(sa, gi[t], ma[t], &inv[t * ND], acc[t], r);
    r.u = wave_ror1(r.u);
    r.v = wave_ror1(r.v);
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::sym_apply(sb, gi[t], mb[t], &inv[(TPT + t) * ND], acc[t], r);
    r.u = wave_ror1(r.u);
    r.v = wave_ror1(r.v);
}

} // namespace pv
