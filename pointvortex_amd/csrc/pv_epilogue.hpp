// pv_epilogue.hpp -- per-target stage arithmetic fused into the velocity kernels.
#pragma once
#include "pv_kernels.hpp"

#include <climits>

namespace pv {

__device__ __forceinline__ bool finite64(double v) { return __builtin_isfinite(v); }

__device__ __forceinline__ void record_failure(DevStatus *st, unsigned epoch, unsigned kind) {
    atomicOr(&st->flags, kind);
    atomicMin(&st->fail_epoch, epoch);
}

// Where a failure goes: the sticky status of the launch, or (ensembles) the failing
// instance's own word, so that one failing run does not stop the others.  For fixed-step
// ensembles the first failure of an instance also records where the input state of the
// failing step lives (keep = {x0, y0}): the instance is skipped from then on, so that state
// stays intact in one of the two ping-pong buffers.
struct FailSink {
    DevStatus *st;
    unsigned epoch;
    unsigned *word;
    const double **keep = nullptr;     // [2]: where to record {x0, y0}, or null
    const double *x0 = nullptr, *y0 = nullptr;
    unsigned *steps_at_failure = nullptr;   // receives *steps_done at the first failure
    const unsigned *steps_done = nullptr;
    // Ensembles: where to record the positions the failing evaluation read, so that the host
    // can tell a coincident pair from an overflow afterwards (the failing instance is skipped
    // from then on, so those positions stay as they were).
    const double **evaluated = nullptr;
    const double *ex = nullptr, *ey = nullptr;
    __device__ __forceinline__ void operator()(unsigned kind) const {
        if (word) {
            const unsigned old = atomicOr(word, kind);
            if (old == 0u) {
                if (evaluated) {
                    evaluated[0] = ex;
                    evaluated[1] = ey;
                }
                if (keep) {
                    keep[0] = x0;
                    keep[1] = y0;
                    *steps_at_failure = *steps_done;
                }
            }
        } else {
            record_failure(st, epoch, kind);
        }
    }
};

template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Max over the block; result valid in every thread.
__device__ double block_max(double v) {
    __shared__ double red[kBlock / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        red[wave] = v;
    __syncthreads();
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        m = fmax(m, red[w]);
    return m;
}

// Per-target epilogue of an evaluation at positions (xc, yc) with step dt and error tolerances
// (atol, rtol); returns this target's DOPRI5 error contribution (0 otherwise).
__device__ __forceinline__ double stage_epilogue(const StageOp &op, double dt, double atol,
                                                 double rtol, const FailSink &fail,
                                                 const double *xc, const double *yc, int i,
                                                 double u, double v) {
    if (!finite64(u) || !finite64(v))
        fail(kFailVelocity);
    if (op.ku) {
        op.ku[i] = u;
        op.kv[i] = v;
    }
    if (op.kind == kOpStage) {
        // src/timestep.cpp:19-30 (makeStage), coefficients in stage order.
        double dx = 0.0, dy = 0.0;
        for (int j = 0; j < op.cur; ++j) {
            dx += op.coef[j] * op.kx[j][i];
            dy += op.coef[j] * op.ky[j][i];
        }
        dx += op.coef[op.cur] * u;
        dy += op.coef[op.cur] * v;
        op.xn[i] = op.x0[i] + dt * dx;
        op.yn[i] = op.y0[i] + dt * dy;
    } else if (op.kind == kOpRk4Final) {
        // src/timestep.cpp:48-58: combination, then the finiteness check of the new state.
        const double xn = op.x0[i] + dt * (op.kx[0][i] + 2.0 * op.kx[1][i] +
                                              2.0 * op.kx[2][i] + u) / 6.0;
        const double yn = op.y0[i] + dt * (op.ky[0][i] + 2.0 * op.ky[1][i] +
                                              2.0 * op.ky[2][i] + v) / 6.0;
        op.xn[i] = xn;
        op.yn[i] = yn;
        if (!finite64(xn) || !finite64(yn))
            fail(kFailStateNonFinite);
    } else if (op.kind == kOpDopriError) {
        // src/timestep.cpp:83-98: max-norm of the embedded error, scaled per component.
        double ex = 0.0, ey = 0.0;
        for (int s = 0; s < 6; ++s) {
            ex += dt * op.w[s] * op.kx[s][i];
            ey += dt * op.w[s] * op.ky[s][i];
        }
        ex += dt * op.w[6] * u;
        ey += dt * op.w[6] * v;
        const double x5 = xc[i], y5 = yc[i];
        const double scx = atol + rtol * fmax(fabs(op.x0[i]), fabs(x5));
        const double scy = atol + rtol * fmax(fabs(op.y0[i]), fabs(y5));
        if (!finite64(ex) || !finite64(ey) || !finite64(scx) || !finite64(scy) || !(scx > 0.0) ||
            !(scy > 0.0)) {
            fail(kFailErrorEstimate);
            return __builtin_inf();
        }
        return fmax(fabs(ex) / scx, fabs(ey) / scy);
    }
    return 0.0;
}

} // namespace pv
