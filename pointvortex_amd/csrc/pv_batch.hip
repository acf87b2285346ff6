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
//             applies src/timestep.cpp:140-153 to every instance on the device.
#include "pv_batch.hpp"
#include "pv_epilogue.hpp"

namespace pv {

namespace {

constexpr int kWaveTargets = 64;
constexpr int kBatchBlock = 256;

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

template <class PI, class PP, class PD, bool ADAPT>
__global__ __launch_bounds__(kBatchBlock) void batch_kernel(BatchLaunch a) {
    constexpr size_t kSrcBytes =
        sizeof(typename PD::Src) > sizeof(typename PI::Src)
            ? (sizeof(typename PD::Src) > sizeof(typename PP::Src) ? sizeof(typename PD::Src)
                                                                   : sizeof(typename PP::Src))
            : (sizeof(typename PI::Src) > sizeof(typename PP::Src) ? sizeof(typename PI::Src)
                                                                   : sizeof(typename PP::Src));
    __shared__ __align__(16) unsigned char lds[kBatchBlock / 64][kWaveTargets * kSrcBytes];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned char *mine = lds[wave];
    for (;;) {
        int item = 0;
        if (lane == 0)
            item = (int)atomicAdd(a.counter, 1u);
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
        const int geometry = ins->gp.geometry;
        if (geometry == 0)
            batch_item<PI, ADAPT>(a, ins, it, reinterpret_cast<typename PI::Src *>(mine), lane);
        else if (geometry == 1)
            batch_item<PP, ADAPT>(a, ins, it, reinterpret_cast<typename PP::Src *>(mine), lane);
        else
            batch_item<PD, ADAPT>(a, ins, it, reinterpret_cast<typename PD::Src *>(mine), lane);
    }
}

__global__ void batch_reset_kernel(unsigned *counter) {
    if (threadIdx.x == 0)
        *counter = 0u;
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

// src/timestep.cpp:140-153 per instance, then the step landing of src/main.cpp:312-314 and
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

void launch_batch(int precision, bool adaptive, const BatchLaunch &a, int blocks,
                  hipStream_t s) {
    const dim3 grid(blocks), block(kBatchBlock);
    if (precision == 0) {
        if (adaptive)
            hipLaunchKernelGGL((batch_kernel<InfF64, PerF64, DiskF64, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((batch_kernel<InfF64, PerF64, DiskF64, false>), grid, block, 0, s, a);
    } else {
        if (adaptive)
            hipLaunchKernelGGL((batch_kernel<InfMixed, PerMixed, DiskMixed, true>), grid, block, 0,
                               s, a);
        else
            hipLaunchKernelGGL((batch_kernel<InfMixed, PerMixed, DiskMixed, false>), grid, block, 0,
                               s, a);
    }
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
