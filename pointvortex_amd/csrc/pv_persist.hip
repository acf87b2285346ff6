// pv_persist.hip -- RK4 steps of a small population in one cooperative launch.
//
// For small N an evaluation is a few microseconds of arithmetic spread over a chain of
// dependent latencies: one launch per stage, the target and source loads, the split partials
// and their ticket, the stage epilogue.  Here every block owns `tb` targets for the whole run
// and the 1024 / tb lane groups of its 1024 threads split the sources among them, so that a
// target's sum is reduced inside its block (in a fixed order) and the stage epilogue runs where
// the sum is.  Stage velocities never leave their block; only the stage positions are
// exchanged, through HBM and a grid-wide barrier between stages.  The stage arithmetic is the
// same StageOp epilogue as the per-stage launches (src/timestep.cpp:19-58), and each stage
// records failures under its own epoch, so the host classifies them exactly as it does for the
// per-stage launches.
#include "pv_epilogue.hpp"

namespace pv {

// Grid-wide barrier.  `target` is the arrival count that ends this barrier (the counter only
// grows within a launch).  The wait is bounded by `limit` ticks of the constant wall clock:
// every block leaves, and false means the wait timed out.
__device__ bool grid_barrier(DevStatus *st, unsigned long long target, unsigned long long limit) {
    __shared__ int ok_s;
    __syncthreads();
    if (threadIdx.x == 0) {
        // Release: this block's stores (ordered before thread 0 by the block barrier).
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        atomicAdd(&st->barrier, 1ull);
        int ok = 1;
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(&st->barrier, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               target) {
            if (wall_clock64() - t0 > limit) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        // Acquire: the other blocks' stores, before this block reads them.
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_s = ok;
    }
    __syncthreads();
    return ok_s != 0;
}

// One target against the sources c0 + k, k = q, q + G, q + 2G, ... < c0 + cnt, held in LDS
// (G = kPersistBlock / tb lane groups).
// Full rounds take two sources per lane.  The self pair (c0 + k == i) can only fall in round
// rs = (first - c0) / (2G), the same round for every target of the block (first is the block's
// first target; tb divides 2G and 2G divides c0), so the other full rounds skip the check.
template <class P>
__device__ __forceinline__ void persist_accumulate(const GeoParams &gp, const typename P::Tgt &tg,
                                                   typename P::Acc &acc,
                                                   const typename P::Src *sh, int cnt, int c0,
                                                   int q, int G, int i, int first, bool near) {
    if constexpr (P::kNearTiles) {
        if (near) {
            for (int k = q; k < cnt; k += G)
                P::pair_ref_checked(gp, tg, sh[k], acc, c0 + k == i);
            return;
        }
    }
    const int rounds = cnt / (2 * G);
    const int rs = (first >= c0 && first < c0 + cnt) ? (first - c0) / (2 * G) : -1;
    const auto round = [&](int r) {
        const int k0 = 2 * G * r + q, k1 = k0 + G;
        if constexpr (P::kBatched) {
            // Both sources' denominators share one rcp64.
            constexpr int ND = P::kDenoms;
            const typename P::Src s0 = sh[k0], s1 = sh[k1];
            typename P::Tmp m0, m1;
            double d[2 * ND], inv[2 * ND];
            P::prep(gp, tg, s0, m0, &d[0]);
            P::prep(gp, tg, s1, m1, &d[ND]);
            const bool ok = TreeInverse<2 * ND>::run(d, inv);
            if (tree_fallback(ok)) {
#pragma unroll
                for (int e = 0; e < 2 * ND; ++e)
                    inv[e] = rcp64(d[e]);
            }
            P::apply(s0, m0, &inv[0], acc);
            P::apply(s1, m1, &inv[ND], acc);
        } else {
            P::pair(gp, tg, sh[k0], acc);
            P::pair(gp, tg, sh[k1], acc);
        }
    };
    // The self round is taken out of the loops, which then carry no check.
    const int r1 = (rs >= 0 && rs < rounds) ? rs : rounds;
#pragma unroll 2
    for (int r = 0; r < r1; ++r)
        round(r);
    if (r1 < rounds) {
        const int k0 = 2 * G * r1 + q, k1 = k0 + G;
        P::pair_checked(gp, tg, sh[k0], acc, c0 + k0 == i);
        P::pair_checked(gp, tg, sh[k1], acc, c0 + k1 == i);
#pragma unroll 2
        for (int r = r1 + 1; r < rounds; ++r)
            round(r);
    }
    // The rest: at most two sources per lane.
    for (int k = 2 * G * rounds + q; k < cnt; k += G)
        P::pair_checked(gp, tg, sh[k], acc, c0 + k == i);
}

// Sum of a target's G lane-group partials red[q tb + t], q = 0..G-1, in a fixed order: 16
// consecutive runs of G/16 (one thread each), then the 16 run sums.  Valid in thread t < tb.
__device__ void persist_reduce(double (*red)[2], double (*runs)[2], int tb, int G, double &u,
                               double &v) {
    const int tid = threadIdx.x;
    const int per = G / 16;
    if (tid < 16 * tb) {
        const int t = tid % tb, h = tid / tb;
        Comp su, sv;
        su.zero();
        sv.zero();
        for (int q = h * per; q < (h + 1) * per; ++q) {
            su.add(red[q * tb + t][0]);
            sv.add(red[q * tb + t][1]);
        }
        runs[tid][0] = su.value();
        runs[tid][1] = sv.value();
    }
    __syncthreads();
    if (tid < tb) {
        Comp su, sv;
        su.zero();
        sv.zero();
        for (int h = 0; h < 16; ++h) {
            su.add(runs[h * tb + tid][0]);
            sv.add(runs[h * tb + tid][1]);
        }
        u = su.value();
        v = sv.value();
    }
}

// The RK4 stage arithmetic of stage_epilogue (src/timestep.cpp:19-30 makeStage, 48-58), for the
// thread that owns target i, with that target's earlier stage velocities (ku, kv) and
// step-start position (x0, y0) kept in registers across the stages instead of in HBM; only the
// stage and step positions, which every block reads, are stored.
struct PersistTarget {
    double x0, y0;
    double ku[3], kv[3];
};

__device__ __forceinline__ void persist_epilogue(const StageOp &op, int c, const FailSink &fail,
                                                 int i, double u, double v, PersistTarget &p) {
    if (!finite64(u) || !finite64(v))
        fail(kFailVelocity);
    const double dt = op.dt;
    if (c < 3) {
        double dx = 0.0, dy = 0.0;
        if (c >= 1) {
            dx += op.coef[0] * p.ku[0];
            dy += op.coef[0] * p.kv[0];
        }
        if (c >= 2) {
            dx += op.coef[1] * p.ku[1];
            dy += op.coef[1] * p.kv[1];
        }
        dx += op.coef[c] * u;
        dy += op.coef[c] * v;
        op.xn[i] = p.x0 + dt * dx;
        op.yn[i] = p.y0 + dt * dy;
        if (c == 0)
            p.ku[0] = u, p.kv[0] = v;
        else if (c == 1)
            p.ku[1] = u, p.kv[1] = v;
        else
            p.ku[2] = u, p.kv[2] = v;
    } else {
        const double xn = p.x0 + dt * (p.ku[0] + 2.0 * p.ku[1] + 2.0 * p.ku[2] + u) / 6.0;
        const double yn = p.y0 + dt * (p.kv[0] + 2.0 * p.kv[1] + 2.0 * p.kv[2] + v) / 6.0;
        op.xn[i] = xn;
        op.yn[i] = yn;
        if (!finite64(xn) || !finite64(yn))
            fail(kFailStateNonFinite);
        p.x0 = xn;
        p.y0 = yn;
    }
}

template <class P>
__global__ __launch_bounds__(kPersistBlock) void rk4_persist_kernel(RkPersistLaunch a,
                                                                    GeoParams gp) {
    if (a.st->fail_epoch < a.epoch0)
        return;
    __shared__ typename P::Src sh[kPersistChunk];
    __shared__ double red[kPersistBlock][2];
    __shared__ double runs[16 * kPersistMaxTb][2];
    const int tid = threadIdx.x;
    const int tb = a.tb, G = kPersistBlock / tb;
    const int t = tid % tb, q = tid / tb;     // tid = q tb + t
    const int first = blockIdx.x * tb;
    const int i = first + t;
    const bool valid = i < a.n;
    unsigned long long arrivals = a.barrier_base;
    PersistTarget own{};
    if (valid && q == 0) {
        own.x0 = a.x[0][i];
        own.y0 = a.y[0][i];
    }

    for (int s = 0; s < a.steps; ++s) {
        for (int c = 0; c < 4; ++c) {
            const unsigned epoch = a.epoch0 + 4u * (unsigned)s + (unsigned)c;
            const StageOp &op = a.ops[(s & 1) * 4 + c];
            // Positions this stage evaluates: x0, then the stage positions (enqueue_rk4).
            const double *xc = c == 0 ? a.x[s & 1] : (c == 2 ? a.xb : a.xa);
            const double *yc = c == 0 ? a.y[s & 1] : (c == 2 ? a.yb : a.ya);

            double xi = 0.0, yi = 0.0;
            if (valid) {
                xi = xc[i];
                yi = yc[i];
                if (q == 0) {
                    // src/vortex.h:11-13 and src/compute.cpp:41-48 for every evaluated target.
                    if (!finite64(xi) || !finite64(yi))
                        record_failure(a.st, epoch, kFailStateNonFinite);
                    else if (gp.geometry == 2 && xi * xi + yi * yi >= gp.R2)
                        record_failure(a.st, epoch, kFailOutsideDisk);
                }
            }
            const typename P::Tgt tg = P::tgt(gp, xi, yi);
            typename P::Acc acc;
            P::zero(acc);
            for (int c0 = 0; c0 < a.n; c0 += kPersistChunk) {
                const int cnt = min(kPersistChunk, a.n - c0);
                __syncthreads();
                int near = 0;
                for (int k = tid; k < cnt; k += kPersistBlock) {
                    const double xj = xc[c0 + k], yj = yc[c0 + k];
                    sh[k] = P::src(gp, xj, yj, a.g[c0 + k]);
                    if constexpr (P::kNearTiles)
                        near |= P::is_near(gp, xj, yj);
                }
                bool near_chunk = false;
                if constexpr (P::kNearTiles)
                    near_chunk = __syncthreads_or(near) != 0;
                else
                    __syncthreads();
                if (valid)
                    persist_accumulate<P>(gp, tg, acc, sh, cnt, c0, q, G, i, first, near_chunk);
                if constexpr (P::kFoldEvery > 0)
                    P::fold(acc);
            }
            double u, v;
            P::result(acc, u, v);
            red[tid][0] = u;
            red[tid][1] = v;
            __syncthreads();
            persist_reduce(red, runs, tb, G, u, v);
            if (valid && q == 0) {
                if constexpr (kSelfImage<P>)
                    disk_self_image(gp, xi, yi, a.g[i], u, v);
                persist_epilogue(op, c, FailSink{a.st, epoch, nullptr}, i, u, v, own);
            }
            if (s == a.steps - 1 && c == 3)
                return;
            arrivals += gridDim.x;
            if (!grid_barrier(a.st, arrivals, a.wait_limit)) {
                if (tid == 0)
                    record_failure(a.st, epoch, kFailBarrier);
                return;
            }
            // Every failure of this stage or earlier was recorded before the barrier and every
            // later one carries a later epoch, so all blocks decide alike.
            if (__hip_atomic_load(&a.st->fail_epoch, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) <= epoch)
                return;
        }
    }
}

template <class P> static const void *persist_fn() {
    return reinterpret_cast<const void *>(&rk4_persist_kernel<P>);
}

static const void *persist_kernel(int geometry, int precision) {
    if (precision == 0)
        return geometry == 0 ? persist_fn<InfF64>()
                             : geometry == 1 ? persist_fn<PerF64>() : persist_fn<DiskF64>();
    return geometry == 0 ? persist_fn<InfMixed>()
                         : geometry == 1 ? persist_fn<PerMixed>() : persist_fn<DiskMixed>();
}

int persistent_capacity(int geometry, int precision, int cu_count) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel(geometry, precision),
                                                     kPersistBlock, 0) != hipSuccess)
        return 0;
    return per_cu * cu_count;
}

hipError_t launch_rk4_persistent(int geometry, int precision, const RkPersistLaunch &a,
                                 const GeoParams &gp, int blocks, hipStream_t s) {
    RkPersistLaunch la = a;
    GeoParams lg = gp;
    void *args[] = {&la, &lg};
    return hipLaunchCooperativeKernel(persist_kernel(geometry, precision), dim3(blocks),
                                      dim3(kPersistBlock), args, 0, s);
}

} // namespace pv
