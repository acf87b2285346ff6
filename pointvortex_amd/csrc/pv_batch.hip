// pv_batch.hip -- ensembles of independent vortex systems in one persistent launch per stage.
//
// Layout: every instance's x, y, G live in flat SoA arrays at [offset, offset + n); an
// instance table holds each instance's geometry constants (GeoParams) and range.  Work items
// are 64-target slices of one instance, sorted by decreasing cost on the host.  Each wave of
// the persistent grid takes items from a shared counter until none are left (every wave
// reaches that exit), stages the instance's sources through its own 64-record LDS slice and
// sums them for its 64 targets, one lane per target, in source order -- so each target's
// result does not depend on which wave ran it or when.
//
// Two modes share the pair sums and the fused stage epilogue (pv_epilogue.hpp):
//   fixed     one dt for every instance (RK4 ensembles, captured in hipGraphs);
//   adaptive  per-instance DOPRI5: dt from each instance's InstanceStep, the max-norm error
//             reduced per instance, failures recorded per instance; a controller kernel then
//             applies src/timestep.cpp:104-118 to every instance on the device.
#include "pv_batch.hpp"
#include "pv_epilogue.hpp"
#include "pv_sym.hpp"

#include <type_traits>

namespace pv {

namespace {

constexpr int kWaveTargets = 64;
constexpr int kBatchBlock = 256;

// The pair policies of one precision, by geometry.  A launch carries one set, or two when its
// work items mix FP64 and mixed-precision instances (each instance names its own).
struct F64Set {
    using I = InfF64;
    using P = PerF64;
    using D = DiskF64;
};
struct MixedSet {
    using I = InfMixed;
    using P = PerMixed;
    using D = DiskMixed;
};
struct NoSet {};

constexpr size_t max3(size_t a, size_t b, size_t c) {
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// LDS bytes of one source record, the largest of a set's policies (0 for NoSet).
template <class S> constexpr size_t set_src_bytes() {
    if constexpr (std::is_same<S, NoSet>::value)
        return 0;
    else
        return max3(sizeof(typename S::I::Src), sizeof(typename S::P::Src),
                    sizeof(typename S::D::Src));
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class P, bool ADAPT>
__device__ void batch_item(const BatchLaunch &a, const Instance *ins, const WorkItem &it,
                           typename P::Src *sh, int lane) {
    const GeoParams &gp = ins->gp;
    const int off = ins->offset, n = ins->n;
    const int ti = it.tbegin + lane;
    const bool valid = ti < it.tend;
    const int gi = off + (valid ? ti : it.tbegin);
    InstanceStep *stp = &a.steps[it.instance];
    FailSink fail{a.st, a.epoch, &stp->failure};
    fail.evaluated = stp->evaluated;
    fail.ex = a.x;
    fail.ey = a.y;
    if (!ADAPT) {
        fail.keep = stp->keep;
        fail.x0 = a.op.x0;
        fail.y0 = a.op.y0;
        fail.steps_at_failure = &stp->fail_step;
        fail.steps_done = a.counter + 1;
    }
    const double xt = a.x[gi], yt = a.y[gi];
    if (valid) {
        if (!finite64(xt) || !finite64(yt))
            fail(kFailStateNonFinite);
        else if (gp.geometry == 2 && xt * xt + yt * yt >= gp.R2)
            fail(kFailOutsideDisk);
    }
    const typename P::Tgt tg = P::tgt(gp, xt, yt);
    typename P::Acc acc;
    P::zero(acc);
    for (int base = 0; base < n; base += kWaveTargets) {
        const int cnt = min(kWaveTargets, n - base);
        wave_sync();
        bool near = false;
        if (lane < cnt) {
            const int j = off + base + lane;
            const double xj = a.x[j], yj = a.y[j];
            sh[lane] = P::src(gp, xj, yj, a.g[j]);
            if constexpr (P::kNearTiles)
                near = P::is_near(gp, xj, yj);
        }
        wave_sync();
        const bool diag = base < it.tend && it.tbegin < base + cnt;
        if constexpr (P::kNearTiles) {
            if (__any(near)) {
                for (int k = 0; k < cnt; ++k)
                    P::pair_ref_checked(gp, tg, sh[k], acc, ti == base + k);
                continue;
            }
        }
        if (!diag) {
            int k = 0;
            if constexpr (P::kBatched) {
                constexpr int ND = P::kDenoms, K = 2 * ND;
                for (; k + 2 <= cnt; k += 2) {
                    const typename P::Src s2[2] = {sh[k], sh[k + 1]};
                    typename P::Tmp m[2];
                    double d[K], inv[K];
                    P::prep(gp, tg, s2[0], m[0], &d[0]);
                    P::prep(gp, tg, s2[1], m[1], &d[ND]);
                    const bool ok = TreeInverse<K>::run(d, inv);
                    if (tree_fallback(ok)) {
#pragma unroll
                        for (int e = 0; e < K; ++e)
                            inv[e] = rcp64(d[e]);
                    }
                    P::apply(s2[0], m[0], &inv[0], acc);
                    P::apply(s2[1], m[1], &inv[ND], acc);
                }
            }
            for (; k < cnt; ++k)
                P::pair(gp, tg, sh[k], acc);
        } else {
            for (int k = 0; k < cnt; ++k)
                P::pair_checked(gp, tg, sh[k], acc, ti == base + k);
        }
        if (P::kFoldEvery)
            P::fold(acc);
    }
    double u, v;
    P::result(acc, u, v);
    if (valid) {
        if constexpr (kSelfImage<P>)
            disk_self_image(gp, xt, yt, a.g[gi], u, v);
        const double dt = ADAPT ? stp->dt : a.op.dt;
        const double atol = ADAPT ? stp->atol : a.op.atol;
        const double rtol = ADAPT ? stp->rtol : a.op.rtol;
        const double err = stage_epilogue(a.op, dt, atol, rtol, fail, a.x, a.y, gi, u, v);
        // Every target folds its own error into the instance maximum (errors are >= 0, so the
        // unsigned order of their bits is their numeric order).  No lane-0-only branch here:
        // inside the persistent item loop such a branch lets the compiler specialise the loop
        // header for the other lanes and skip the item fetch they share with lane 0.
        if (ADAPT && a.op.kind == kOpDopriError)
            atomicMax(&stp->err_bits, (unsigned long long)__double_as_longlong(err));
    }
}

template <class F, class M, bool ADAPT>
__global__ __launch_bounds__(kBatchBlock) void batch_kernel(BatchLaunch a) {
    constexpr size_t kSrcBytes = max3(set_src_bytes<F>(), set_src_bytes<M>(), 0);
    __shared__ __align__(16) unsigned char lds[kBatchBlock / 64][kWaveTargets * kSrcBytes];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned char *mine = lds[wave];
    for (;;) {
        int item = 0;
        if (lane == 0)
            item = (int)atomicAdd(a.next, 1u);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= a.nitems)
            break;
        const WorkItem it = a.items[item];
        const Instance *ins = &a.inst[it.instance];
        // An instance that failed in an earlier launch skips the rest, so its failure word
        // holds what the first failing evaluation found and its state stays where it was.
        const InstanceStep &sp = a.steps[it.instance];
        if (sp.failure)
            continue;
        if (ADAPT && (sp.state != kRunning || (a.pass == kPassNeedK0 && sp.fsal)))
            continue;
        // The instance's policy: its geometry, and its precision when the launch carries
        // both sets.
        const int geometry = ins->gp.geometry;
        bool done = false;
        if constexpr (!std::is_same<M, NoSet>::value) {
            if (ins->precision) {
                if (geometry == 0)
                    batch_item<typename M::I, ADAPT>(
                        a, ins, it, reinterpret_cast<typename M::I::Src *>(mine), lane);
                else if (geometry == 1)
                    batch_item<typename M::P, ADAPT>(
                        a, ins, it, reinterpret_cast<typename M::P::Src *>(mine), lane);
                else
                    batch_item<typename M::D, ADAPT>(
                        a, ins, it, reinterpret_cast<typename M::D::Src *>(mine), lane);
                done = true;
            }
        }
        if (!done) {
            if (geometry == 0)
                batch_item<typename F::I, ADAPT>(
                    a, ins, it, reinterpret_cast<typename F::I::Src *>(mine), lane);
            else if (geometry == 1)
                batch_item<typename F::P, ADAPT>(
                    a, ins, it, reinterpret_cast<typename F::P::Src *>(mine), lane);
            else
                batch_item<typename F::D, ADAPT>(
                    a, ins, it, reinterpret_cast<typename F::D::Src *>(mine), lane);
        }
    }
}

// ------------------------------------------------------------------ symmetric mode

// Per-wave LDS of the symmetric item: the 64 TPT source records of tile J, and for the
// near-tile policies the reaction sums of their near path (written by every lane with the same
// value, read back one source per lane, so no lane-0-only store sits inside the persistent
// loop).
template <class P, int TPT, bool RED = P::kNearTiles> struct SymSlice {
    typename P::Src src[64 * TPT];
    double red[64 * TPT][2];
};
template <class P, int TPT> struct SymSlice<P, TPT, false> {
    typename P::Src src[64 * TPT];
};

// The policy of geometry GEO (0..2) in a set.
template <class S, int GEO>
using GeoPolicy = std::conditional_t<GEO == 0, typename S::I,
                                     std::conditional_t<GEO == 1, typename S::P, typename S::D>>;

// Tile widths of the per-instance kernel (TPT = 0 below), which takes each item at its
// instance's width: 1 and 2 targets per lane for every policy, and 4 for the infinite plane,
// whose 4-target code fits the kernel's registers.  The periodic and disk 4-target items run
// in kernels of their own (one geometry, one width: TPT = 4).
template <class P>
constexpr bool kWide4 = std::is_same<P, InfF64>::value || std::is_same<P, InfMixed>::value;

template <class P> constexpr size_t per_instance_slice_bytes() {
    const size_t narrow = sizeof(SymSlice<P, 1>) > sizeof(SymSlice<P, 2>) ? sizeof(SymSlice<P, 1>)
                                                                          : sizeof(SymSlice<P, 2>);
    if constexpr (kWide4<P>)
        return narrow > sizeof(SymSlice<P, 4>) ? narrow : sizeof(SymSlice<P, 4>);
    else
        return narrow;
}

// LDS bytes of one wave's slice: geometry GEO of a set at width TPT, or (TPT = 0, every
// geometry) the largest the per-instance kernel needs; 0 for NoSet.
template <class S, int TPT, int GEO> constexpr size_t set_slice_bytes() {
    if constexpr (std::is_same<S, NoSet>::value)
        return 0;
    else if constexpr (TPT == 0)
        return max3(per_instance_slice_bytes<typename S::I>(),
                    per_instance_slice_bytes<typename S::P>(),
                    per_instance_slice_bytes<typename S::D>());
    else
        return sizeof(SymSlice<GeoPolicy<S, GEO>, TPT>);
}

template <class P, int TPT>
__device__ void batch_sym_item(const BatchLaunch &a, const Instance *ins, const PairItem &it,
                               SymSlice<P, TPT> *sl, int lane, int dir);

// One item with the set's policy of geometry GEO at width TPT, or (TPT = 0, GEO =
// kSymAllGeometries) with its instance's geometry and width.
template <class S, int TPT, int GEO>
__device__ __forceinline__ void sym_item_of(const BatchLaunch &a, const Instance *ins,
                                            const PairItem &it, unsigned char *mine, int lane,
                                            int dir) {
    if constexpr (TPT == 0) {
        const int geometry = ins->gp.geometry;
        if (geometry == 0)
            sym_item_of<S, -1, 0>(a, ins, it, mine, lane, dir);
        else if (geometry == 1)
            sym_item_of<S, -1, 1>(a, ins, it, mine, lane, dir);
        else
            sym_item_of<S, -1, 2>(a, ins, it, mine, lane, dir);
    } else if constexpr (TPT == -1) {
        // The instance's width; the host never gives this kernel a 4-target item of a policy
        // without kWide4 (those run in their own launch), so such an item is skipped here.
        using P = GeoPolicy<S, GEO>;
        const int tpt = ins->tpt;
        if (tpt == 1) {
            batch_sym_item<P, 1>(a, ins, it, reinterpret_cast<SymSlice<P, 1> *>(mine), lane, dir);
        } else if (tpt == 2) {
            batch_sym_item<P, 2>(a, ins, it, reinterpret_cast<SymSlice<P, 2> *>(mine), lane, dir);
        } else if constexpr (kWide4<P>) {
            batch_sym_item<P, 4>(a, ins, it, reinterpret_cast<SymSlice<P, 4> *>(mine), lane, dir);
        }
    } else {
        using P = GeoPolicy<S, GEO>;
        batch_sym_item<P, TPT>(a, ins, it, reinterpret_cast<SymSlice<P, TPT> *>(mine), lane, dir);
    }
}

// One pair of tiles (I, J) of one instance: partial slot J of tile I's targets, and for I < J
// slot I of tile J's sources (the reaction).  I == J: ordered pairs inside the tile.
template <class P, int TPT>
__device__ void batch_sym_item(const BatchLaunch &a, const Instance *ins, const PairItem &it,
                               SymSlice<P, TPT> *sl, int lane, int dir) {
    using Acc = typename P::SymAcc;
    constexpr int TT = 64 * TPT;
    const GeoParams &gp = ins->gp;
    const int off = ins->offset, n = ins->n;
    const size_t nn = (size_t)n;
    double *part = a.part + ins->poff;
    const int t0 = it.I * TT, s0 = it.J * TT;
    const int cnt = min(TT, n - s0);
    const int span = (cnt + 63) & ~63;

    double xi[TPT], yi[TPT];
    typename P::SymScalar gi[TPT];
    int ti[TPT];
    bool near = false;
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = t0 + t * 64 + lane;
        const bool valid = i < n;
        const int ii = off + (valid ? i : t0);
        xi[t] = a.x[ii];
        yi[t] = a.y[ii];
        gi[t] = valid ? P::sym_g(gp, a.g[ii]) : 0;     // a padding target adds no reaction
        ti[t] = valid ? i : -1;
        if constexpr (P::kNearTiles)
            near |= valid && P::is_near(gp, xi[t], yi[t]);
    }
    wave_sync();
    for (int k = lane; k < span; k += 64) {
        const int j = off + s0 + (k < cnt ? k : 0);
        const double xj = a.x[j], yj = a.y[j];
        sl->src[k] = P::src(gp, xj, yj, k < cnt ? a.g[j] : 0.0);
        if constexpr (P::kNearTiles)
            near |= k < cnt && P::is_near(gp, xj, yj);
    }
    wave_sync();
    if constexpr (P::kNearTiles)
        near = __any(near);


    Acc acc[TPT];
#pragma unroll
    for (int t = 0; t < TPT; ++t)
        P::zero(acc[t]);

    if (it.I == it.J) {
        typename P::Tgt tg[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t)
            tg[t] = P::tgt(gp, xi[t], yi[t]);
        for (int k = 0; k < cnt; ++k) {
            const typename P::Src s_a = sl->src[k];
#pragma unroll
            for (int t = 0; t < TPT; ++t) {
                if constexpr (P::kNearTiles) {
                    if (near)
                        P::pair_ref_checked(gp, tg[t], s_a, acc[t], ti[t] == s0 + k);
                    else
                        P::pair_checked(gp, tg[t], s_a, acc[t], ti[t] == s0 + k);
                } else {
                    P::pair_checked(gp, tg[t], s_a, acc[t], ti[t] == s0 + k);
                }
            }
            if constexpr (P::kFoldEvery > 0) {
                if ((k & 63) == 63 || k == cnt - 1) {       // result() reads folded sums only
#pragma unroll
                    for (int t = 0; t < TPT; ++t)
                        P::fold(acc[t]);
                }
            }
        }
    } else {
        bool done = false;
        if constexpr (P::kNearTiles) {
            if (near) {
                // The reference's form both ways: j on i, and i (as a source) on j.
                typename P::Tgt tr[TPT];
                typename P::Src si[TPT];
#pragma unroll
                for (int t = 0; t < TPT; ++t) {
                    tr[t] = P::tgt(gp, xi[t], yi[t]);
                    si[t] = P::src(gp, xi[t], yi[t], ti[t] >= 0 ? a.g[off + ti[t]] : 0.0);
                }
                for (int k = 0; k < cnt; ++k) {
                    const typename P::Src s_a = sl->src[k];
                    const typename P::Tgt tj = P::tgt(gp, s_a.x, s_a.y);
                    Acc r;
                    P::zero(r);
#pragma unroll
                    for (int t = 0; t < TPT; ++t) {
                        P::pair_ref_checked(gp, tr[t], s_a, acc[t], false);
                        P::pair_ref_checked(gp, tj, si[t], r, false);
                    }
                    wave_sum(r);
                    sl->red[k][0] = r.u;
                    sl->red[k][1] = r.v;
                }
                wave_sync();
                for (int k = lane; k < cnt; k += 64) {
                    part[(size_t)(2 * it.I) * nn + s0 + k] = sl->red[k][0];
                    part[(size_t)(2 * it.I + 1) * nn + s0 + k] = sl->red[k][1];
                }
                done = true;
            }
        }
        if (!done) {
            typename P::SymTgt tg[TPT];
#pragma unroll
            for (int t = 0; t < TPT; ++t)
                tg[t] = P::sym_tgt(gp, xi[t], yi[t]);
            // The ring of the single-system symmetric kernel: 64 steps per group of 64
            // sources bring every source's reaction back to its first lane.  kSymNear policies
            // keep their direct form here (their records come from src(), without sym_src's
            // w; see DiskMixed::sym_src).
            {
                constexpr bool kDirect = P::kSymNear;
                for (int g0 = 0; g0 < span; g0 += 64) {
                    typename P::SymR r;
                    r.u = 0;
                    r.v = 0;
                    typename P::Src next = sl->src[g0 + lane];
#pragma unroll P::kRingUnroll
                    for (int s = 0; s < 64; ++s) {
                        const typename P::Src s_a = next;
                        next = sl->src[g0 + ((lane + dir * (s + 1)) & 63)];
                        sym_step<P, TPT, kDirect>(gp, tg, gi, s_a, acc, r);
                        r.u = wave_ror1(r.u);
                        r.v = wave_ror1(r.v);
                    }
                    if (g0 + lane < cnt) {
                        part[(size_t)(2 * it.I) * nn + s0 + g0 + lane] = (double)r.u;
                        part[(size_t)(2 * it.I + 1) * nn + s0 + g0 + lane] = (double)r.v;
                    }
                    if constexpr (P::kFoldEvery > 0) {
#pragma unroll
                        for (int t = 0; t < TPT; ++t)
                            P::fold(acc[t]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        double u, v;
        P::result(acc[t], u, v);
        if (ti[t] >= 0) {
            part[(size_t)(2 * it.J) * nn + ti[t]] = u;
            part[(size_t)(2 * it.J + 1) * nn + ti[t]] = v;
        }
    }
}

// The items of one launch group: TPT = 0 for the per-instance kernel (every geometry, each
// item at its instance's width), or TPT = 4 with GEO its one geometry (the periodic or disk
// 4-target items).  It dispatches on precision when the launch carries both sets.
template <class F, class M, bool ADAPT, int TPT, int GEO>
__global__ __launch_bounds__(kBatchBlock) void batch_sym_kernel(BatchLaunch a) {
    constexpr size_t kSliceBytes =
        max3(set_slice_bytes<F, TPT, GEO>(), set_slice_bytes<M, TPT, GEO>(), 0);
    __shared__ __align__(16) unsigned char lds[kBatchBlock / 64][kSliceBytes];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned char *mine = lds[wave];
    const int dir = ring_dir(lane);
    for (;;) {
        // One chunk of consecutive items per atomic: there are about 4x as many tile-pair
        // items as ordered ones, and one device-scope atomic per item would serialise on the
        // counter.
        int chunk = 0;
        if (lane == 0)
            chunk = (int)atomicAdd(a.next, 1u);
        chunk = __builtin_amdgcn_readfirstlane(__shfl(chunk, 0, 64));
        if (chunk >= a.nchunks)
            break;
        const int first = __builtin_amdgcn_readfirstlane(a.chunks[chunk]);
        const int last = __builtin_amdgcn_readfirstlane(a.chunks[chunk + 1]);
        for (int item = first; item < last; ++item) {
            const PairItem it = a.pitems[item];
            const Instance *ins = &a.inst[it.instance];
            // The item-skip rules of batch_kernel; the reduction applies them per vortex.
            const InstanceStep &sp = a.steps[it.instance];
            if (sp.failure)
                continue;
            if (ADAPT && (sp.state != kRunning || (a.pass == kPassNeedK0 && sp.fsal)))
                continue;
            bool done = false;
            if constexpr (!std::is_same<M, NoSet>::value) {
                if (ins->precision) {
                    sym_item_of<M, TPT, GEO>(a, ins, it, mine, lane, dir);
                    done = true;
                }
            }
            if (!done)
                sym_item_of<F, TPT, GEO>(a, ins, it, mine, lane, dir);
        }
    }
}

// Per flat vortex: the validation and the slot-ordered compensated sum of its instance's
// partials, then the fused stage epilogue (as batch_item's tail).
template <bool ADAPT>
__global__ __launch_bounds__(kBatchBlock) void batch_sym_reduce_kernel(BatchLaunch a) {
    const int v = blockIdx.x * kBatchBlock + threadIdx.x;
    if (v >= a.span)
        return;
    const int k = a.owner[v];
    if (k < 0)
        return;
    InstanceStep *stp = &a.steps[k];
    if (stp->failure)
        return;
    if (ADAPT && (stp->state != kRunning || (a.pass == kPassNeedK0 && stp->fsal)))
        return;
    const Instance ins = a.inst[k];
    FailSink fail{a.st, a.epoch, &stp->failure};
    fail.evaluated = stp->evaluated;
    fail.ex = a.x;
    fail.ey = a.y;
    if (!ADAPT) {
        fail.keep = stp->keep;
        fail.x0 = a.op.x0;
        fail.y0 = a.op.y0;
        fail.steps_at_failure = &stp->fail_step;
        fail.steps_done = a.counter + 1;
    }
    const double xt = a.x[v], yt = a.y[v];
    if (!finite64(xt) || !finite64(yt))
        fail(kFailStateNonFinite);
    else if (ins.gp.geometry == 2 && xt * xt + yt * yt >= ins.gp.R2)
        fail(kFailOutsideDisk);
    Comp su, sv;
    sum_rows(a.part + ins.poff + (v - ins.offset), (size_t)ins.n, ins.nt, su, sv);
    double u = su.value(), w = sv.value();
    if (ins.gp.geometry == 2 && ins.precision == 1)
        disk_self_image(ins.gp, xt, yt, a.g[v], u, w);
    const double dt = ADAPT ? stp->dt : a.op.dt;
    const double atol = ADAPT ? stp->atol : a.op.atol;
    const double rtol = ADAPT ? stp->rtol : a.op.rtol;
    const double err = stage_epilogue(a.op, dt, atol, rtol, fail, a.x, a.y, v, u, w);
    if (ADAPT && a.op.kind == kOpDopriError)
        atomicMax(&stp->err_bits, (unsigned long long)__double_as_longlong(err));
}

__global__ void batch_reset_kernel(unsigned *counter) {
    if (threadIdx.x == 0)
        counter[0] = 0u;
    else if (threadIdx.x <= kSymGroups)
        counter[1 + threadIdx.x] = 0u;
}

// One fixed RK4 step finished (its last stage launch precedes this on the stream).
__global__ void batch_step_done_kernel(unsigned *counter) {
    if (threadIdx.x == 0)
        counter[1] += 1u;
}

// DOPRI5 stage-1 positions with each vortex's own instance dt (owner[i] = instance index).
__global__ void batch_stage1_kernel(StageOp op, const int *owner, const InstanceStep *steps,
                                    int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int k = owner[i];
    if (k < 0 || steps[k].state != kRunning || steps[k].failure)
        return;
    const double dt = steps[k].dt;
    const double dx = op.coef[0] * op.kx[0][i];
    const double dy = op.coef[0] * op.ky[0][i];
    op.xn[i] = op.x0[i] + dt * dx;
    op.yn[i] = op.y0[i] + dt * dy;
}

// src/timestep.cpp:104-118 per instance, then the step landing of src/main.cpp:312-314 and
// the end-of-run test of src/main.cpp:341-344.
__global__ void batch_control_kernel(InstanceStep *steps, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst)
        return;
    InstanceStep &s = steps[k];
    s.accept_now = 0;
    if (s.state != kRunning)
        return;
    s.fsal = 1;
    double error;
    const unsigned long long bits = s.err_bits;
    error = __longlong_as_double((long long)bits);
    s.err_bits = 0ull;
    if (s.failure) {
        s.state = kFailed;
        return;
    }
    const double factor = error == 0.0 ? 5.0 : fmin(fmax(0.9 * pow(error, -0.2), 0.2), 5.0);
    const double suggested = fmin(fmax(s.dt * factor, s.min_dt), s.max_dt);
    if (error <= 1.0) {
        s.accept_now = 1;
        s.t += s.dt;
        s.last_error = error;
        s.sugg = suggested;
        s.accepted += 1;
        s.rejected_now = 0;
        const double slack = 16.0 * 2.220446049250313e-16 * fabs(s.t);
        if (s.t + slack >= s.t_end) {
            s.t = s.t_end;
            s.state = kFinished;
        } else {
            s.dt = fmin(s.sugg, s.t_end - s.t);
        }
        return;
    }
    s.rejected_total += 1;
    if (s.dt <= s.min_dt || ++s.rejected_now > 32) {
        s.failure |= kFailRejected;
        s.state = kFailed;
        return;
    }
    s.dt = fmax(s.min_dt, fmin(suggested, s.dt * 0.9));
}

// Accepted instances take the fifth-order positions and their last stage as the FSAL k0.
__global__ void batch_accept_kernel(const int *owner, const InstanceStep *steps, double *x,
                                    double *y, const double *xn, const double *yn, double *k0x,
                                    double *k0y, const double *k6x, const double *k6y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int k = owner[i];
    if (k < 0 || !steps[k].accept_now)
        return;
    x[i] = xn[i];
    y[i] = yn[i];
    k0x[i] = k6x[i];
    k0y[i] = k6y[i];
}

} // namespace

// precisions: bit 0 set when FP64 instances are among the work items, bit 1 for mixed ones;
// K names the kernel template (batch_kernel or batch_sym_kernel).
#define PV_LAUNCH_SETS(K, precisions, adaptive, a, blocks, s)                                    \
    do {                                                                                        \
        const dim3 grid_(blocks), block_(kBatchBlock);                                         \
        if ((precisions) == 3) {                                                                \
            if (adaptive)                                                                       \
                hipLaunchKernelGGL((K<F64Set, MixedSet, true>), grid_, block_, 0, s, a);        \
            else                                                                                \
                hipLaunchKernelGGL((K<F64Set, MixedSet, false>), grid_, block_, 0, s, a);       \
        } else if ((precisions) == 2) {                                                         \
            if (adaptive)                                                                       \
                hipLaunchKernelGGL((K<MixedSet, NoSet, true>), grid_, block_, 0, s, a);         \
            else                                                                                \
                hipLaunchKernelGGL((K<MixedSet, NoSet, false>), grid_, block_, 0, s, a);        \
        } else {                                                                                \
            if (adaptive)                                                                       \
                hipLaunchKernelGGL((K<F64Set, NoSet, true>), grid_, block_, 0, s, a);           \
            else                                                                                \
                hipLaunchKernelGGL((K<F64Set, NoSet, false>), grid_, block_, 0, s, a);          \
        }                                                                                       \
    } while (0)

void launch_batch(int precisions, bool adaptive, const BatchLaunch &a, int blocks,
                  hipStream_t s) {
    PV_LAUNCH_SETS(batch_kernel, precisions, adaptive, a, blocks, s);
}

#undef PV_LAUNCH_SETS

namespace {

template <class F, class M, int TPT, int GEO>
void launch_sym_group(bool adaptive, const BatchLaunch &a, int blocks, hipStream_t s) {
    const dim3 grid(blocks), block(kBatchBlock);
    if (adaptive)
        hipLaunchKernelGGL((batch_sym_kernel<F, M, true, TPT, GEO>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((batch_sym_kernel<F, M, false, TPT, GEO>), grid, block, 0, s, a);
}

} // namespace

namespace {

template <class F, class M>
void launch_sym_set(int group, bool adaptive, const BatchLaunch &a, int blocks, hipStream_t s) {
    if (group == 1)
        launch_sym_group<F, M, 4, 1>(adaptive, a, blocks, s);
    else if (group == 2)
        launch_sym_group<F, M, 4, 2>(adaptive, a, blocks, s);
    else
        launch_sym_group<F, M, 0, kSymAllGeometries>(adaptive, a, blocks, s);
}

} // namespace

void launch_batch_sym(int precisions, bool adaptive, int group, const BatchLaunch &a, int blocks,
                      hipStream_t s) {
    if (precisions == 3)
        launch_sym_set<F64Set, MixedSet>(group, adaptive, a, blocks, s);
    else if (precisions == 2)
        launch_sym_set<MixedSet, NoSet>(group, adaptive, a, blocks, s);
    else
        launch_sym_set<F64Set, NoSet>(group, adaptive, a, blocks, s);
}

void launch_batch_sym_reduce(bool adaptive, const BatchLaunch &a, hipStream_t s) {
    const dim3 grid((unsigned)((a.span + kBatchBlock - 1) / kBatchBlock)), block(kBatchBlock);
    if (adaptive)
        hipLaunchKernelGGL(batch_sym_reduce_kernel<true>, grid, block, 0, s, a);
    else
        hipLaunchKernelGGL(batch_sym_reduce_kernel<false>, grid, block, 0, s, a);
}

void launch_batch_reset(unsigned *counter, hipStream_t s) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, counter);
}

void launch_batch_step_done(unsigned *counter, hipStream_t s) {
    hipLaunchKernelGGL(batch_step_done_kernel, dim3(1), dim3(64), 0, s, counter);
}

void launch_batch_stage1(const StageOp &op, const int *owner, const InstanceStep *steps, int n,
                         hipStream_t s) {
    hipLaunchKernelGGL(batch_stage1_kernel, dim3((n + 255) / 256), dim3(256), 0, s, op, owner,
                       steps, n);
}

void launch_batch_control(InstanceStep *steps, int ninst, hipStream_t s) {
    hipLaunchKernelGGL(batch_control_kernel, dim3((ninst + 255) / 256), dim3(256), 0, s, steps,
                       ninst);
}

void launch_batch_accept(const int *owner, const InstanceStep *steps, double *x, double *y,
                         const double *xn, const double *yn, double *k0x, double *k0y,
                         const double *k6x, const double *k6y, int n, hipStream_t s) {
    hipLaunchKernelGGL(batch_accept_kernel, dim3((n + 255) / 256), dim3(256), 0, s, owner, steps,
                       x, y, xn, yn, k0x, k0y, k6x, k6y, n);
}

} // namespace pv
