// pv_kernels.hip -- gfx950 kernels of the velocity backend.
//
// Pair kernel layout (one launch per velocity evaluation):
//   grid.x  = target tiles of kBlock*TPT targets; each lane holds TPT targets in registers,
//             strided by kBlock so that target loads are coalesced;
//   grid.y  = source splits of split_len sources; each block walks its split in LDS tiles of
//             TILE source records (SoA x, y, G in HBM -> derived records in LDS), which every
//             lane of a wave then reads at the same address (an LDS broadcast);
//   splits  > 1 write FP64 partial sums; the last block of a target tile (atomic ticket) adds
//             them in split order, so every target's sum has a fixed order (deterministic).
// The stage arithmetic of RK4 / DOPRI5 is fused into the epilogue of the evaluation that
// produces the stage (StageOp), so an RK4 step is four launches and a DOPRI5 attempt seven.
#include <type_traits>

#include "pv_epilogue.hpp"
#include "pv_sym.hpp"

namespace pv {

__device__ __forceinline__ double epilogue(const PairLaunch &a, const GeoParams &, int i,
                                           double u, double v) {
    return stage_epilogue(a.op, a.op.dt, a.op.atol, a.op.rtol, FailSink{a.st, a.epoch, nullptr},
                          a.x, a.y, i, u, v);
}

template <class P, int TPT, int TILE>
__global__ __launch_bounds__(kBlock) void pair_kernel(PairLaunch a, GeoParams gp) {
    if (a.st->fail_epoch < a.epoch)
        return;
    __shared__ typename P::Src sh[TILE];
    __shared__ int last_block;
    const int tid = threadIdx.x;
    constexpr int TT = kBlock * TPT;
    const int tile = blockIdx.x;
    const int split = blockIdx.y;
    const int t0 = a.tbegin + tile * TT;

    typename P::Tgt tg[TPT];
    typename P::Acc acc[TPT];
    int ti[TPT];
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = t0 + t * kBlock + tid;
        const bool valid = i < a.tend;
        const int ii = valid ? i : a.tbegin;
        const double x = a.x[ii], y = a.y[ii];
        if (valid && split == 0) {
            // src/vortex.h:11-13 and src/compute.cpp:41-48 for every evaluated target.
            if (!finite64(x) || !finite64(y))
                record_failure(a.st, a.epoch, kFailStateNonFinite);
            else if (gp.geometry == 2 && x * x + y * y >= gp.R2)
                record_failure(a.st, a.epoch, kFailOutsideDisk);
        }
        tg[t] = P::tgt(gp, x, y);
        P::zero(acc[t]);
        ti[t] = valid ? i : -1;
    }

    const int s0 = split * a.split_len;
    const int s1 = min(a.n, s0 + a.split_len);
    for (int base = s0; base < s1; base += TILE) {
        const int cnt = min(TILE, s1 - base);
        __syncthreads();
        int near = 0;
        for (int k = tid; k < cnt; k += kBlock) {
            const int j = base + k;
            const double xj = a.x[j], yj = a.y[j];
            sh[k] = P::src(gp, xj, yj, a.g[j]);
            if constexpr (P::kNearTiles)
                near |= P::is_near(gp, xj, yj);
        }
        bool near_tile = false;
        if constexpr (P::kNearTiles)
            near_tile = __syncthreads_or(near) != 0;
        else
            __syncthreads();
        const bool diag = base < t0 + TT && t0 < base + cnt;
        if constexpr (P::kNearTiles) {
            if (near_tile) {
                for (int k = 0; k < cnt; ++k) {
                    const typename P::Src s_a = sh[k];
#pragma unroll
                    for (int t = 0; t < TPT; ++t)
                        P::pair_ref_checked(gp, tg[t], s_a, acc[t], ti[t] == base + k);
                }
                continue;
            }
        }
        if (!diag) {
            int k = 0;
            if constexpr (P::kBatched) {
                // Two sources x TPT targets: all their denominators share one rcp64.
                constexpr int ND = P::kDenoms, K = 2 * TPT * ND;
                for (; k + 2 <= cnt; k += 2) {
                    const typename P::Src s2[2] = {sh[k], sh[k + 1]};
                    typename P::Tmp m[2][TPT];
                    double d[K], inv[K];
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int t = 0; t < TPT; ++t)
                            P::prep(gp, tg[t], s2[q], m[q][t], &d[(q * TPT + t) * ND]);
                    const bool ok = TreeInverse<K>::run(d, inv);
                    if (tree_fallback(ok)) {
#pragma unroll
                        for (int e = 0; e < K; ++e)
                            inv[e] = rcp64(d[e]);
                    }
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int t = 0; t < TPT; ++t)
                            P::apply(s2[q], m[q][t], &inv[(q * TPT + t) * ND], acc[t]);
                }
            }
            for (; k + 2 <= cnt; k += 2) {
                const typename P::Src s_a = sh[k];
                const typename P::Src s_b = sh[k + 1];
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    P::pair(gp, tg[t], s_a, acc[t]);
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    P::pair(gp, tg[t], s_b, acc[t]);
            }
            if (k < cnt) {
                const typename P::Src s_a = sh[k];
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    P::pair(gp, tg[t], s_a, acc[t]);
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                const typename P::Src s_a = sh[k];
                const int j = base + k;
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    P::pair_checked(gp, tg[t], s_a, acc[t], ti[t] == j);
            }
        }
        if (P::kFoldEvery) {
#pragma unroll
            for (int t = 0; t < TPT; ++t)
                P::fold(acc[t]);
        }
    }

    double u[TPT], v[TPT];
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::result(acc[t], u[t], v[t]);

    if (a.nsplit > 1) {
        const size_t n = (size_t)a.n;
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            if (ti[t] >= 0) {
                a.part[(size_t)(2 * split) * n + ti[t]] = u[t];
                a.part[(size_t)(2 * split + 1) * n + ti[t]] = v[t];
            }
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            const unsigned prev = atomicAdd(&a.tickets[tile], 1u);
            last_block = prev == (unsigned)(a.nsplit - 1);
        }
        __syncthreads();
        if (!last_block)
            return;
        __threadfence();
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            if (ti[t] >= 0) {
                Comp su, sv;
                su.zero();
                sv.zero();
                for (int s = 0; s < a.nsplit; ++s) {
                    su.add(__builtin_nontemporal_load(&a.part[(size_t)(2 * s) * n + ti[t]]));
                    sv.add(__builtin_nontemporal_load(&a.part[(size_t)(2 * s + 1) * n + ti[t]]));
                }
                u[t] = su.value();
                v[t] = sv.value();
            }
        }
        if (tid == 0)
            a.tickets[tile] = 0u;
    }

    double err = 0.0;
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        if (ti[t] >= 0) {
            if constexpr (kSelfImage<P>)
                disk_self_image(gp, tg[t].x, tg[t].y, a.g[ti[t]], u[t], v[t]);
            err = fmax(err, epilogue(a, gp, ti[t], u[t], v[t]));
        }
    }
    if (a.op.kind == kOpDopriError) {
        const double m = block_max(err);
        if (tid == 0)
            atomicMax(&a.st->err_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// ------------------------------------------------------------------ symmetric evaluation

// The infinite-plane policies, whose pair is short enough that the ring's data movement is a
// sizeable share of a step: their symmetric kernels also have 8 targets per thread.  The FP64
// infinite and the disk rings read their sources from a doubled window (win), one index add
// per step instead of three; the infinite FP64 diagonal blocks also share one reciprocal over
// two sources x TPT targets, as the split kernel does.  (In the mixed infinite kernel the
// window costs more than it saves: the compiler then loads each source in the step that uses
// it instead of one step ahead.  The mixed disk's two-phase step keeps the load ahead.)
template <class P>
constexpr bool kInfPlane = std::is_same_v<P, InfF64> || std::is_same_v<P, InfMixed>;
template <class P>
constexpr bool kRingWindow =
    std::is_same_v<P, InfF64> || std::is_same_v<P, DiskF64> || std::is_same_v<P, DiskMixed>;
// The diagonal blocks' shared reciprocal over two sources (infinite FP64 only).
template <class P> constexpr bool kDiagShared = std::is_same_v<P, InfF64>;
// The mixed infinite ring at 4 targets per thread rotates its reaction with ds_bpermute
// (lds_ror1), the others with DPP (at 2 targets the bpermute's latency is exposed: 38 % slower).
template <class P, int TPT> constexpr bool kLdsRotate = std::is_same_v<P, InfMixed> && TPT == 4;
// Ring steps in pairs with one product tree over both (sym_step_pair): the FP64 rings with the
// window, up to 4 denominators per step (the infinite plane at 1-4 targets per thread, the disk
// at 1-2; beyond that the tree's registers cost a wave per SIMD).
template <class P, int TPT>
constexpr bool kPairedSteps = kRingWindow<P> && P::kSymDenoms > 0 && TPT * P::kSymDenoms <= 4;

// Slot of source k of a staged chunk: with the window, each group of 64 sources is staged twice
// in a row (group k / 64 at 128 (k / 64)), so that the ring's index needs no wrap.
template <class P> __device__ __forceinline__ int win(int k) {
    if constexpr (kRingWindow<P>)
        return k + (k & ~63);
    else
        return k;
}

// Block -> tile pair (I <= J) in row-major triangular order: b = J (J + 1) / 2 + I.
__device__ __forceinline__ void tile_pair(int b, int &I, int &J) {
    int j = (int)((__builtin_sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= b)
        ++j;
    while (j * (j + 1) / 2 > b)
        --j;
    J = j;
    I = b - j * (j + 1) / 2;
}

// Symmetric blocks of the infinite-plane policies: the nt (nt - 1) / 2 off-diagonal pairs (I < J)
// first, row-major, then the nt diagonal tiles, which take less time there and so fill the end
// of the launch.  The other policies' diagonal blocks take as long as the others (or longer):
// they keep the row-major order over I <= J, which spreads them through the launch.
template <class P> __device__ __forceinline__ void sym_tile_pair(int b, int nt, int &I, int &J) {
    if constexpr (!kInfPlane<P>) {
        tile_pair(b, I, J);
        return;
    }
    const int off = (int)((long long)nt * (nt - 1) / 2);
    if (b >= off) {
        I = J = b - off;
    } else {
        tile_pair(b, I, J);   // b = J (J + 1) / 2 + I with I <= J, so (I, J + 1) is b's pair
        ++J;
    }
}

// At 8 targets per thread the kernel is held to 168 VGPRs (3 waves per SIMD) rather than the
// 180 it would take at 2.
template <class P, int TPT, int TILE>
__global__ __launch_bounds__(sym_block(TPT))
__attribute__((amdgpu_waves_per_eu(TPT >= 8 ? 3 : 1))) void sym_kernel(SymLaunch a, GeoParams gp) {
    if (a.st->fail_epoch < a.epoch)
        return;
    using Acc = typename P::SymAcc;
    __shared__ typename P::Src sh[(kRingWindow<P> ? 2 : 1) * TILE];
    constexpr int B = sym_block(TPT);
    __shared__ double red[B / 64][TILE][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int TT = B * TPT;
    int I, J;
    sym_tile_pair<P>(blockIdx.x, a.nt, I, J);
    const int t0 = I * TT;
    const int s0 = J * TT, s1 = min(a.n, s0 + TT);
    const size_t n = (size_t)a.n;

    double xi[TPT], yi[TPT];
    typename P::SymScalar gi[TPT];
    // Target t of the lane (-1 for padding), recomputed where needed rather than held.
    const auto ti = [&](int t) {
        const int i = t0 + t * B + tid;
        return i < a.n ? i : -1;
    };
    int near = 0, sym_near = 0;
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = t0 + t * B + tid;
        const bool valid = i < a.n;
        const int ii = valid ? i : t0;
        xi[t] = a.x[ii];
        yi[t] = a.y[ii];
        gi[t] = valid ? P::sym_g(gp, a.g[ii]) : 0;       // a padding target adds no reaction
        if constexpr (P::kNearTiles)
            near |= valid && P::is_near(gp, xi[t], yi[t]);
        if constexpr (P::kSymNear)
            sym_near |= valid && P::is_sym_near(gp, xi[t], yi[t]);
    }
    if constexpr (P::kNearTiles) {
        for (int k = tid; k < s1 - s0; k += B)
            near |= P::is_near(gp, a.x[s0 + k], a.y[s0 + k]);
        near = __syncthreads_or(near);
    }
    if constexpr (P::kSymNear) {
        for (int k = tid; k < s1 - s0; k += B)
            sym_near |= P::is_sym_near(gp, a.x[s0 + k], a.y[s0 + k]);
        sym_near = __syncthreads_or(sym_near);
    }

    Acc acc[TPT];
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::zero(acc[t]);

    if (I == J) {
        // Ordered pairs inside the tile, the target side only (self pair skipped).
        typename P::Tgt tg[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t)
            tg[t] = P::tgt(gp, xi[t], yi[t]);
        for (int base = s0; base < s1; base += TILE) {
            const int cnt = min(TILE, s1 - base);
            __syncthreads();
            for (int k = tid; k < cnt; k += B)
                sh[k] = P::src(gp, a.x[base + k], a.y[base + k], a.g[base + k]);
            __syncthreads();
            int k0 = 0;
            if constexpr (kDiagShared<P>) {
                // Q sources x TPT targets share one rcp64 (Q = 2, or 1 at 8 targets per thread).
                // The self pair has zero difference and so no term; its denominator (and a
                // padding lane's) is set to 1 so that the product stays usable.
                constexpr int Q = TPT >= 8 ? 1 : 2, K = Q * TPT;
                for (; k0 + Q <= cnt; k0 += Q) {
                    typename P::Src sq[Q];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        sq[q] = sh[k0 + q];
                    typename P::Tmp m[Q][TPT];
                    double d[K], inv[K];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
#pragma unroll
                        for (int t = 0; t < TPT; ++t) {
                            P::prep(gp, tg[t], sq[q], m[q][t], &d[q * TPT + t]);
                            const int it = ti(t);
                            if (it < 0 || it == base + k0 + q)
                                d[q * TPT + t] = 1.0;
                        }
                    const bool ok = TreeInverse<K>::run(d, inv);
                    if (tree_fallback(ok)) {
#pragma unroll
                        for (int e = 0; e < K; ++e)
                            inv[e] = rcp64(d[e]);
                    }
#pragma unroll
                    for (int q = 0; q < Q; ++q)
#pragma unroll
                        for (int t = 0; t < TPT; ++t)
                            P::apply(sq[q], m[q][t], &inv[q * TPT + t], acc[t]);
                }
            }
            for (int k = k0; k < cnt; ++k) {
                const typename P::Src s_a = sh[k];
#pragma unroll
                for (int t = 0; t < TPT; ++t) {
                    if constexpr (P::kNearTiles) {
                        if (near)
                            P::pair_ref_checked(gp, tg[t], s_a, acc[t], ti(t) == base + k);
                        else
                            P::pair_checked(gp, tg[t], s_a, acc[t], ti(t) == base + k);
                    } else {
                        P::pair_checked(gp, tg[t], s_a, acc[t], ti(t) == base + k);
                    }
                }
            }
            if constexpr (P::kFoldEvery > 0) {
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    P::fold(acc[t]);
            }
        }
    } else {
        typename P::SymTgt tg[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t)
            tg[t] = P::sym_tgt(gp, xi[t], yi[t]);
        const int dir = kLdsRotate<P, TPT> ? 1 : ring_dir(lane);
        const int rot_addr = lds_ror1_addr(lane);
        for (int base = s0; base < s1; base += TILE) {
            const int cnt = min(TILE, s1 - base);
            const int span = (cnt + 63) & ~63;          // <= TILE: TILE is a multiple of 64
            __syncthreads();
            for (int k = tid; k < span; k += B) {
                const typename P::Src v =
                    k < cnt ? ring_src<P>(gp, a.x[base + k], a.y[base + k], a.g[base + k])
                            : ring_src<P>(gp, a.x[base], a.y[base], 0.0);
                sh[win<P>(k)] = v;
                if constexpr (kRingWindow<P>)
                    sh[win<P>(k) + 64] = v;
            }
            __syncthreads();
            bool done = false;
            if constexpr (P::kNearTiles) {
                if (near) {
                    // The reference's form both ways: j on i, and i (as a source) on j.
                    typename P::Tgt tr[TPT];
                    typename P::Src si[TPT];
#pragma unroll
                    for (int t = 0; t < TPT; ++t) {
                        tr[t] = P::tgt(gp, xi[t], yi[t]);
                        si[t] = P::src(gp, xi[t], yi[t], ti(t) >= 0 ? a.g[ti(t)] : 0.0);
                    }
                    for (int k = 0; k < cnt; ++k) {
                        const typename P::Src s_a = sh[win<P>(k)];
                        const typename P::Tgt tj = P::tgt(gp, s_a.x, s_a.y);
                        Acc r;
                        P::zero(r);
#pragma unroll
                        for (int t = 0; t < TPT; ++t) {
                            P::pair_ref_checked(gp, tr[t], s_a, acc[t], false);
                            P::pair_ref_checked(gp, tj, si[t], r, false);
                        }
                        wave_sum(r);
                        if (lane == 0) {
                            red[wave][k][0] = r.u;
                            red[wave][k][1] = r.v;
                        }
                    }
                    done = true;
                }
            }
            // Ring: lane l starts on source g0 + l with that source's reaction in r; after each
            // step r moves one lane (wave_ror:1) and the lane reads the source whose reaction it
            // now holds, (l + dir k) mod 64 after k steps (in the window: entry l + k for dir 1,
            // l + 64 - k for dir 63).  64 steps bring every reaction back to its first lane with
            // the sum over the wave's targets.  Staged pads (strength 0) fill the last group of 64;
            // their reactions are dropped.  `direct` is the near-centre form of kSymNear
            // policies (block-uniform, so each form is its own loop).
            const auto ring = [&](auto direct) {
                constexpr bool kDirect = decltype(direct)::value;
                for (int g0 = 0; g0 < cnt; g0 += 64) {
                    typename P::SymR r;
                    r.u = 0;
                    r.v = 0;
                    const typename P::Src *w = &sh[win<P>(g0) + lane + (dir == 1 ? 0 : 64)];
                    const int step = dir == 1 ? 1 : -1;
                    typename P::Src next = sh[win<P>(g0) + lane];
                    constexpr int kUnroll = TPT >= 8 ? 1 : P::kRingUnroll;
                    if constexpr (kPairedSteps<P, TPT>) {
                        for (int s = 0; s < 64; s += 2) {
                            const typename P::Src s_a = next, s_b = w[step * (s + 1)];
                            next = w[step * (s + 2)];    // at s = 62, the first source again
                            sym_step_pair<P, TPT>(gp, tg, gi, s_a, s_b, acc, r);
                        }
                    } else {
#pragma unroll kUnroll
                        for (int s = 0; s < 64; ++s) {
                            // The next step's source is read one step ahead (at s = 63 this
                            // reads the lane's first source again, unused).
                            const typename P::Src s_a = next;
                            if constexpr (kRingWindow<P>)
                                next = w[step * (s + 1)];
                            else
                                next = sh[g0 + ((lane + dir * (s + 1)) & 63)];
                            sym_step<P, TPT, kDirect>(gp, tg, gi, s_a, acc, r);
                            if constexpr (kLdsRotate<P, TPT>) {
                                r.u = lds_ror1(r.u, rot_addr);
                                r.v = lds_ror1(r.v, rot_addr);
                            } else {
                                r.u = wave_ror1(r.u);
                                r.v = wave_ror1(r.v);
                            }
                        }
                    }
                    if (g0 + lane < cnt) {
                        red[wave][g0 + lane][0] = (double)r.u;
                        red[wave][g0 + lane][1] = (double)r.v;
                    }
                }
            };
            if (!done) {
                if (P::kSymNear && sym_near)
                    ring(Flag<P::kSymNear>{});
                else
                    ring(Flag<false>{});
            }
            if constexpr (P::kFoldEvery > 0) {
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    P::fold(acc[t]);
            }
            __syncthreads();
            // The reaction on this chunk's sources: the block's waves in order, into slot I.
            for (int k = tid; k < cnt; k += B) {
                double ru = red[0][k][0], rv = red[0][k][1];
#pragma unroll
                for (int w = 1; w < B / 64; ++w) {
                    ru += red[w][k][0];
                    rv += red[w][k][1];
                }
                a.part[(size_t)(2 * I) * n + base + k] = ru;
                a.part[(size_t)(2 * I + 1) * n + base + k] = rv;
            }
        }
    }
    // The target side into slot J (slot I on the diagonal, where J == I).
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        if (ti(t) >= 0) {
            double u, v;
            P::result(acc[t], u, v);
            a.part[(size_t)(2 * J) * n + ti(t)] = u;
            a.part[(size_t)(2 * J + 1) * n + ti(t)] = v;
        }
    }
}

// Partials are FP64 in both precisions, so one reduction serves every policy.
__global__ __launch_bounds__(kBlock) void sym_reduce_kernel(SymLaunch a, GeoParams gp) {
    if (a.st->fail_epoch < a.epoch)
        return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double err = 0.0;
    if (i < a.n) {
        const double x = a.x[i], y = a.y[i];
        // src/vortex.h:11-13 and src/compute.cpp:41-48 for every evaluated target.
        if (!finite64(x) || !finite64(y))
            record_failure(a.st, a.epoch, kFailStateNonFinite);
        else if (gp.geometry == 2 && x * x + y * y >= gp.R2)
            record_failure(a.st, a.epoch, kFailOutsideDisk);
        Comp su, sv;
        sum_rows(a.part + i, (size_t)a.n, a.nt, su, sv);
        double u = su.value(), v = sv.value();
        if (a.self_image)
            disk_self_image(gp, x, y, a.g[i], u, v);
        err = stage_epilogue(a.op, a.op.dt, a.op.atol, a.op.rtol,
                             FailSink{a.st, a.epoch, nullptr}, a.x, a.y, i, u, v);
    }
    if (a.op.kind == kOpDopriError) {
        const double m = block_max(err);
        if (threadIdx.x == 0)
            atomicMax(&a.st->err_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// ------------------------------------------------------------------ small kernels

__global__ void reset_status_kernel(DevStatus *st) {
    if (threadIdx.x == 0) {
        st->flags = 0u;
        st->fail_epoch = UINT_MAX;
        st->err_bits = 0ull;
        st->count = 0ull;
        st->barrier = 0ull;
    }
}

// Stage positions from stored stage velocities only (DOPRI5 stage 1).
__global__ void make_stage_kernel(StageOp op, int n, const DevStatus *st, unsigned epoch) {
    if (st->fail_epoch < epoch)
        return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double dx = 0.0, dy = 0.0;
    for (int j = 0; j <= op.cur; ++j) {
        dx += op.coef[j] * op.kx[j][i];
        dy += op.coef[j] * op.ky[j][i];
    }
    op.xn[i] = op.x0[i] + op.dt * dx;
    op.yn[i] = op.y0[i] + op.dt * dy;
}

// Exact zero-denominator test of the reference's loops (src/compute.cpp:119-123, 234-246,
// 307-313), with its non-fused arithmetic.  Runs only after a non-finite result, to tell a
// coincident pair apart from an overflow.
__global__ void classify_kernel(GeoParams gp, const double *x, const double *y, int n,
                                int tbegin, int tend, DevStatus *st) {
#pragma clang fp contract(off)
    const int i = tbegin + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tend)
        return;
    bool hit = false;
    for (int j = 0; j < n && !hit; ++j) {
        if (j == i)
            continue;
        if (gp.geometry == 1) {
            const double sdx = gp.wave * remainder(x[i] - x[j], gp.L);
            const double sdy = gp.wave * remainder(y[i] - y[j], gp.L);
            const double shx = sinh(0.5 * sdx), shy = sinh(0.5 * sdy);
            const double snx = sin(0.5 * sdx), sny = sin(0.5 * sdy);
            const double du = 2.0 * (shx * shx + sny * sny);
            const double dv = 2.0 * (shy * shy + snx * snx);
            hit = du == 0.0 || dv == 0.0;
        } else {
            const double dx = x[i] - x[j], dy = y[i] - y[j];
            double d = dx * dx + dy * dy;
            if (gp.geometry == 0)
                d = d + gp.eps2;
            hit = d == 0.0;
        }
    }
    if (hit)
        atomicOr(&st->flags, (unsigned)kFailCoincident);
}

// Block partial sums of {G, G x, G y, G r^2} (src/compute.cpp:11-27), compensated; each
// block writes 8 doubles (hi, lo pairs) in a fixed order.
__global__ void moments_kernel(const double *x, const double *y, const double *g, int n,
                               double *partials) {
    __shared__ double red[kReduceBlock][8];
    Comp c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        c[q].zero();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double gam = g[i];
        c[0].add(gam);
        c[1].add(gam * x[i]);
        c[2].add(gam * y[i]);
        c[3].add(gam * (x[i] * x[i] + y[i] * y[i]));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        red[threadIdx.x][2 * q] = c[q].hi;
        red[threadIdx.x][2 * q + 1] = c[q].lo;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        Comp s;
        s.zero();
        for (int t = 0; t < (int)blockDim.x; ++t) {
            s.add(red[t][2 * q]);
            s.add(red[t][2 * q + 1]);
        }
        partials[blockIdx.x * 8 + 2 * q] = s.hi;
        partials[blockIdx.x * 8 + 2 * q + 1] = s.lo;
    }
}

// Per-source data of the disk Hamiltonian: the image offset delta_j = x_j* - x_j (image_offset,
// accurate however close x_j is to the wall) and |x_j|^2; zero for a source within 2^-100 R of
// the centre, whose pairs take the reference's form.
struct EnergySrc {
    double ox, oy, r2;
};

__device__ __forceinline__ EnergySrc energy_src(const GeoParams &gp, double x, double y) {
    const double r2 = __builtin_fma(x, x, y * y);
    EnergySrc s{0.0, 0.0, 0.0};
    if (gp.geometry == 2 && r2 >= 0x1p-200 * gp.R2) {
        image_offset(gp, x, y, r2, s.ox, s.oy);
        s.r2 = r2;
    }
    return s;
}

// Pair energy e_ij of the Hamiltonians (symmetric in i, j):
//   infinite  log(r^2 + eps^2)                                      src/compute.cpp:142-151
//   disk      log(|d|^2 R^2 / ((R^2 - x_i.x_j)^2 + (x_i x x_j)^2))   src/compute.cpp:271-282
//   periodic  -X^2/2pi + sum_{|m|<=L} [log D_m - log cosh 2 pi m]   src/compute.cpp:182-199
// with the periodic rows m != 0 summed exactly as -4 sum_n G_n cos(nY) cosh(nX) / n - C_L.
// The disk denominator is |x_j|^2 |x_i - x_j*|^2, the same quantity ((R^2 - a.b)^2 + (a x b)^2
// = R^4 - 2 R^2 a.b + |a|^2 |b|^2), formed from the image offset so that it does not cancel for
// pairs near the wall as R^2 - x_i.x_j does.
__device__ double pair_energy(const GeoParams &gp, double xi, double yi, double xj, double yj,
                              const EnergySrc &sj, bool &zero) {
    if (gp.geometry == 0) {
        const double dx = xi - xj, dy = yi - yj;
        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, gp.eps2));
        zero = r2 == 0.0;
        return log(r2);
    }
    if (gp.geometry == 2) {
        const double dx = xi - xj, dy = yi - yj;
        const double num = __builtin_fma(dx, dx, dy * dy) * gp.R2;
        zero = num == 0.0;
        if (sj.r2 == 0.0) {
            const double dot = __builtin_fma(xi, xj, yi * yj);
            const double cross = __builtin_fma(yi, xj, -(xi * yj));
            const double den = __builtin_fma(gp.R2 - dot, gp.R2 - dot, cross * cross);
            return log(num) - log(den);
        }
        const double ex = dx - sj.ox, ey = dy - sj.oy;
        const double den = sj.r2 * __builtin_fma(ex, ex, ey * ey);
        return log(num) - log(den);
    }
    const double X = gp.wave * remainder_fast(xi - xj, gp.L, gp.invL);
    const double Y = gp.wave * remainder_fast(yi - yj, gp.L, gp.invL);
    const double hx = 0.5 * X, hy = 0.5 * Y;
    const double hx2 = hx * hx, hy2 = hy * hy;
    const double shx = sinh_f64(hx, hx2);
    const double sny = sin_f64(hy, hy2);
    const double shx2 = shx * shx, sny2 = sny * sny;
    const double d0 = 2.0 * (shx2 + sny2);
    zero = d0 == 0.0;
    const double cos2y = __builtin_fma(-4.0, sny2, 2.0);
    const double ch2x = __builtin_fma(4.0, shx2, 2.0);
    double c0 = 1.0, c1 = 0.5 * cos2y, h0 = 1.0, h1 = 0.5 * ch2x, acc = 0.0;
#pragma unroll
    for (int k = 1; k <= kTermsF64; ++k) {
        acc = __builtin_fma(gp.H4[k] * c1, h1, acc);
        const double c2 = __builtin_fma(cos2y, c1, -c0);
        const double h2 = __builtin_fma(ch2x, h1, -h0);
        c0 = c1, c1 = c2, h0 = h1, h1 = h2;
    }
    return -X * X / (2.0 * kPi) + log(d0) - acc - gp.CL;
}

__global__ __launch_bounds__(kReduceBlock) void energy_kernel(GeoParams gp, const double *x,
                                                              const double *y, const double *g,
                                                              int n, double *partials,
                                                              DevStatus *st) {
    __shared__ double sx[kReduceBlock], sy[kReduceBlock], sg[kReduceBlock];
    __shared__ EnergySrc se[kReduceBlock];
    __shared__ double red[kReduceBlock][2];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0;
    Comp e;
    e.zero();
    bool coincident = false;
    for (int base = 0; base < n; base += kReduceBlock) {
        const int cnt = min(kReduceBlock, n - base);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const double xj = x[base + threadIdx.x], yj = y[base + threadIdx.x];
            sx[threadIdx.x] = xj;
            sy[threadIdx.x] = yj;
            sg[threadIdx.x] = g[base + threadIdx.x];
            se[threadIdx.x] = energy_src(gp, xj, yj);
        }
        __syncthreads();
        if (valid) {
            for (int k = 0; k < cnt; ++k) {
                if (base + k == i)
                    continue;
                bool z = false;
                const double eij = pair_energy(gp, xi, yi, sx[k], sy[k], se[k], z);
                coincident |= z;
                e.add(sg[k] * eij);
            }
        }
    }
    double h = 0.0, hl = 0.0;
    if (valid) {
        const double gi = g[i];
        Comp hc;
        hc.zero();
        hc.add(-gi / (8.0 * kPi) * e.hi);
        hc.add(-gi / (8.0 * kPi) * e.lo);
        if (gp.geometry == 2) {
            // Self term G^2 log((R^2 - |x_i|^2)/R) / 4 pi, with R^2 - |x_i|^2 formed without
            // cancellation near the wall (as in image_offset).
            const double ri2 = xi * xi + yi * yi;
            if (ri2 >= gp.R2)
                atomicOr(&st->flags, (unsigned)kFailOutsideDisk);
            hc.add(gi * gi * log(wall_gap(gp, xi, yi) / gp.R) / (4.0 * kPi));
        }
        if (coincident)
            atomicOr(&st->flags, (unsigned)kFailCoincident);
        h = hc.hi;
        hl = hc.lo;
    }
    red[threadIdx.x][0] = h;
    red[threadIdx.x][1] = hl;
    __syncthreads();
    if (threadIdx.x == 0) {
        Comp s;
        s.zero();
        for (int t = 0; t < (int)blockDim.x; ++t) {
            s.add(red[t][0]);
            s.add(red[t][1]);
        }
        partials[2 * blockIdx.x] = s.hi;
        partials[2 * blockIdx.x + 1] = s.lo;
    }
}

// Close opposite-sign pairs (src/dipole.cpp:90-118), with the reference's non-fused
// arithmetic (contraction off) so that candidate sets and their d^2 match it bit for bit.
// One block per pair of 256-vortex tiles (I <= J): targets i of tile I against sources j > i
// of tile J, so the O(N^2 / 2) tests spread evenly over nt (nt + 1) / 2 blocks; the disk wall
// test runs once per vortex, in the diagonal blocks.  Candidates are appended with an atomic
// counter (their order is not significant: the host sorts them).
__global__ __launch_bounds__(kReduceBlock) void dipole_kernel(
    GeoParams gp, int periodic, int disk, double thr2, const double *x, const double *y,
    const double *g, int n, unsigned long long *first, unsigned long long *second, double *d2,
    unsigned long long cap, DevStatus *st) {
#pragma clang fp contract(off)
    __shared__ double sx[kReduceBlock], sy[kReduceBlock], sg[kReduceBlock];
    int I, J;
    tile_pair(blockIdx.x, I, J);
    const int i = I * kReduceBlock + (int)threadIdx.x;
    const bool valid = i < n;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, gi = valid ? g[i] : 0.0;
    const bool active = valid && gi != 0.0;
    const int base = J * kReduceBlock;
    const int cnt = min(kReduceBlock, n - base);
    if ((int)threadIdx.x < cnt) {
        sx[threadIdx.x] = x[base + threadIdx.x];
        sy[threadIdx.x] = y[base + threadIdx.x];
        sg[threadIdx.x] = g[base + threadIdx.x];
    }
    __syncthreads();
    if (active) {
        for (int k = 0; k < cnt; ++k) {
            const int j = base + k;
            if (j <= i)
                continue;
            const double gj = sg[k];
            if (gj == 0.0 || signbit(gi) == signbit(gj))
                continue;
            double dx = xi - sx[k], dy = yi - sy[k];
            if (periodic) {
                dx = remainder_fast(dx, gp.L, gp.invL);
                dy = remainder_fast(dy, gp.L, gp.invL);
            }
            const double dd = dx * dx + dy * dy;
            if (dd < thr2) {
                const unsigned long long slot = atomicAdd(&st->count, 1ull);
                if (slot < cap) {
                    first[slot] = (unsigned long long)i;
                    second[slot] = (unsigned long long)j;
                    d2[slot] = dd;
                }
            }
        }
    }
    if (active && disk && I == J) {
        const double ri2 = xi * xi + yi * yi;
        if (ri2 > 0.0) {
            const double idist = (gp.R2 - ri2) / sqrt(ri2);
            const double dd = idist * idist;
            if (dd < thr2) {
                const unsigned long long slot = atomicAdd(&st->count, 1ull);
                if (slot < cap) {
                    first[slot] = (unsigned long long)i;
                    second[slot] = (unsigned long long)n;
                    d2[slot] = dd;
                }
            }
        }
    }
}

// Plain streaming copy (traffic-counter calibration).
__global__ void copy_kernel(const double4 *__restrict__ src, double4 *__restrict__ dst,
                            size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers

template <class P, int TPT, int TILE>
static void launch_pair_t(const PairLaunch &a, const GeoParams &gp, hipStream_t s) {
    const int targets = a.tend - a.tbegin;
    const int tiles = (targets + kBlock * TPT - 1) / (kBlock * TPT);
    dim3 grid(tiles, a.nsplit);
    hipLaunchKernelGGL((pair_kernel<P, TPT, TILE>), grid, dim3(kBlock), 0, s, a, gp);
}

template <class P, int TILE>
static void launch_pair_tpt(int tpt, const PairLaunch &a, const GeoParams &gp, hipStream_t s) {
    if (tpt >= 4)
        launch_pair_t<P, 4, TILE>(a, gp, s);
    else if (tpt == 2)
        launch_pair_t<P, 2, TILE>(a, gp, s);
    else
        launch_pair_t<P, 1, TILE>(a, gp, s);
}

template <class P>
static void launch_pair_tile(int tpt, int tile, const PairLaunch &a, const GeoParams &gp,
                             hipStream_t s) {
    if (tile == 128)
        launch_pair_tpt<P, 128>(tpt, a, gp, s);
    else
        launch_pair_tpt<P, 256>(tpt, a, gp, s);
}

int pair_tile(int precision) { return precision ? kTileMixed : kTileF64; }

void launch_pair(int geometry, int precision, int tpt, int tile, const PairLaunch &a,
                 const GeoParams &gp, hipStream_t s) {
    if (precision == 0) {
        if (geometry == 0)
            launch_pair_tile<InfF64>(tpt, tile, a, gp, s);
        else if (geometry == 1)
            launch_pair_tile<PerF64>(tpt, tile, a, gp, s);
        else
            launch_pair_tile<DiskF64>(tpt, tile, a, gp, s);
    } else {
        if (geometry == 0)
            launch_pair_tile<InfMixed>(tpt, tile, a, gp, s);
        else if (geometry == 1)
            launch_pair_tile<PerMixed>(tpt, tile, a, gp, s);
        else
            launch_pair_tile<DiskMixed>(tpt, tile, a, gp, s);
    }
}

bool sym_supported(int geometry, int precision) {
    return geometry >= 0 && geometry <= 2 && (precision == 0 || precision == 1);
}

bool sym_tpt_supported(int geometry, int precision, int tpt) {
    if (tpt == 1 || tpt == 2 || tpt == 4)
        return sym_supported(geometry, precision);
    return tpt == 8 && geometry == 0 && (precision == 0 || precision == 1);
}

template <class P, int TILE>
static void launch_sym_tpt(int tpt, const SymLaunch &a, const GeoParams &gp, hipStream_t s) {
    const dim3 grid((unsigned)((size_t)a.nt * (a.nt + 1) / 2));
    if constexpr (kInfPlane<P>) {
        if (tpt >= 8) {
            hipLaunchKernelGGL((sym_kernel<P, 8, TILE>), grid, dim3(sym_block(8)), 0, s, a, gp);
            return;
        }
    }
    if (tpt >= 4)
        hipLaunchKernelGGL((sym_kernel<P, 4, TILE>), grid, dim3(sym_block(4)), 0, s, a, gp);
    else if (tpt == 2)
        hipLaunchKernelGGL((sym_kernel<P, 2, TILE>), grid, dim3(sym_block(2)), 0, s, a, gp);
    else
        hipLaunchKernelGGL((sym_kernel<P, 1, TILE>), grid, dim3(sym_block(1)), 0, s, a, gp);
}

template <class P>
static void launch_sym_p(int tpt, int tile, const SymLaunch &a, const GeoParams &gp,
                         hipStream_t s) {
    if (tile == 128)
        launch_sym_tpt<P, 128>(tpt, a, gp, s);
    else
        launch_sym_tpt<P, 256>(tpt, a, gp, s);
}

void launch_sym(int geometry, int precision, int tpt, int tile, const SymLaunch &a,
                const GeoParams &gp, hipStream_t s) {
    if (precision == 0) {
        if (geometry == 0)
            launch_sym_p<InfF64>(tpt, tile, a, gp, s);
        else if (geometry == 1)
            launch_sym_p<PerF64>(tpt, tile, a, gp, s);
        else
            launch_sym_p<DiskF64>(tpt, tile, a, gp, s);
    } else {
        if (geometry == 0)
            launch_sym_p<InfMixed>(tpt, tile, a, gp, s);
        else if (geometry == 1)
            launch_sym_p<PerMixed>(tpt, tile, a, gp, s);
        else
            launch_sym_p<DiskMixed>(tpt, tile, a, gp, s);
    }
}

void launch_sym_reduce(const SymLaunch &a, const GeoParams &gp, hipStream_t s) {
    const dim3 grid((unsigned)((a.n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(sym_reduce_kernel, grid, dim3(kBlock), 0, s, a, gp);
}

void launch_make_stage(const StageOp &op, int n, const DevStatus *st, unsigned epoch,
                       hipStream_t s) {
    hipLaunchKernelGGL(make_stage_kernel, dim3((n + 255) / 256), dim3(256), 0, s, op, n, st,
                       epoch);
}

void launch_reset_status(DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL(reset_status_kernel, dim3(1), dim3(64), 0, s, st);
}

void launch_classify(const GeoParams &gp, const double *x, const double *y, int n, int tbegin,
                     int tend, DevStatus *st, hipStream_t s) {
    const int cnt = tend - tbegin;
    hipLaunchKernelGGL(classify_kernel, dim3((cnt + 255) / 256), dim3(256), 0, s, gp, x, y, n,
                       tbegin, tend, st);
}

void launch_moments(const double *x, const double *y, const double *g, int n, double *partials,
                    int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(moments_kernel, dim3(nblocks), dim3(kReduceBlock), 0, s, x, y, g, n,
                       partials);
}

void launch_energy(const GeoParams &gp, const double *x, const double *y, const double *g, int n,
                   double *partials, int nblocks, DevStatus *st, hipStream_t s) {
    hipLaunchKernelGGL(energy_kernel, dim3(nblocks), dim3(kReduceBlock), 0, s, gp, x, y, g, n,
                       partials, st);
}

void launch_dipoles(const GeoParams &gp, bool periodic, bool disk, double thr2, const double *x,
                    const double *y, const double *g, int n, unsigned long long *first,
                    unsigned long long *second, double *d2, unsigned long long cap,
                    DevStatus *st, hipStream_t s) {
    const size_t nt = ((size_t)n + kReduceBlock - 1) / kReduceBlock;
    hipLaunchKernelGGL(dipole_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(kReduceBlock), 0,
                       s, gp, (int)periodic, (int)disk, thr2, x, y, g, n, first, second, d2, cap,
                       st);
}

void launch_copy(const double4 *src, double4 *dst, size_t count, hipStream_t s) {
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, src, dst, count);
}

} // namespace pv
