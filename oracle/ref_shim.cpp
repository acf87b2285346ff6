// ref_shim.cpp -- a C-ABI over the reference's own CPU classes.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the reference's
// src/compute.cpp, src/timestep.cpp, src/dipole.cpp and src/backend_common.cpp, where
// they lie under the reference tree, into oracle/_ref/libpvref.so.  It is used to pin the
// repo's own restatement (oracle/pv_oracle.c), to write the golden vectors under
// tests/golden/, and as the timed CPU baseline.  Nothing here is shipped in the product.
#include "backend.h"
#include "compute.h"
#include "dipole.h"
#include "params.h"
#include "timestep.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
thread_local std::string lastError;

SimParams makeParams(int geometry, double core, double lengthX, double lengthY, int layers,
                     double radius) {
    SimParams p;
    p.boundaryCondition = geometry == 0 ? "infinite" : (geometry == 1 ? "periodic" : "disk");
    p.coreRadius = core;
    p.boxLengthX = lengthX;
    p.boxLengthY = lengthY;
    p.periodicImageLayers = layers;
    p.diskRadius = radius;
    return p;
}

VortexSystem makeSystem(const double *x, const double *y, const double *g, std::size_t n) {
    VortexSystem s(n);
    if (n) {
        std::memcpy(s.x.data(), x, n * sizeof(double));
        std::memcpy(s.y.data(), y, n * sizeof(double));
        std::memcpy(s.circulation.data(), g, n * sizeof(double));
    }
    return s;
}

template <typename F> int guarded(F &&body) {
    try {
        body();
        return 0;
    } catch (const std::invalid_argument &e) {
        lastError = e.what();
        return 1;
    } catch (const std::out_of_range &e) {
        lastError = e.what();
        return 2;
    } catch (const std::logic_error &e) {
        lastError = e.what();
        return 4;
    } catch (const std::exception &e) {
        lastError = e.what();
        return 3;
    }
}
} // namespace

extern "C" {

const char *pvr_last_error() { return lastError.c_str(); }

// VelocityKernel::evaluate (full range + finiteness check) of the reference kernel.
int pvr_evaluate(int geometry, double core, double lengthX, double lengthY, int layers,
                 double radius, const double *x, const double *y, const double *g, std::size_t n,
                 double *u, double *v) {
    return guarded([&] {
        auto kernel = makeReferenceKernel(makeParams(geometry, core, lengthX, lengthY, layers, radius));
        VortexSystem s = makeSystem(x, y, g, n);
        VelocityField vel;
        kernel->evaluate(s, vel);
        for (std::size_t i = 0; i < n; ++i) {
            u[i] = vel.x[i];
            v[i] = vel.y[i];
        }
    });
}

// VelocityKernel::evaluateRange of the reference kernel (entries outside the range are 0).
int pvr_evaluate_range(int geometry, double core, double lengthX, double lengthY, int layers,
                       double radius, const double *x, const double *y, const double *g,
                       std::size_t n, double *u, double *v, std::size_t begin, std::size_t end) {
    return guarded([&] {
        auto kernel = makeReferenceKernel(makeParams(geometry, core, lengthX, lengthY, layers, radius));
        VortexSystem s = makeSystem(x, y, g, n);
        VelocityField vel(n);
        kernel->evaluateRange(s.x, s.y, s.circulation, vel, begin, end);
        for (std::size_t i = 0; i < n; ++i) {
            u[i] = vel.x[i];
            v[i] = vel.y[i];
        }
    });
}

// computeInvariants(vortices, kernel): {circulation, Ix, Iy, angular impulse, Hamiltonian}.
int pvr_invariants(int geometry, double core, double lengthX, double lengthY, int layers,
                   double radius, const double *x, const double *y, const double *g,
                   std::size_t n, double *out) {
    return guarded([&] {
        auto kernel = makeReferenceKernel(makeParams(geometry, core, lengthX, lengthY, layers, radius));
        VortexSystem s = makeSystem(x, y, g, n);
        const Invariants inv = computeInvariants(s, *kernel);
        out[0] = inv.circulation;
        out[1] = inv.linearImpulseX;
        out[2] = inv.linearImpulseY;
        out[3] = inv.angularImpulse;
        out[4] = inv.hamiltonian;
    });
}

// `steps` fixed RK4 steps of size dt, in place.
int pvr_rk4(int geometry, double core, double lengthX, double lengthY, int layers, double radius,
            double *x, double *y, const double *g, std::size_t n, double dt, int steps) {
    return guarded([&] {
        auto kernel = makeReferenceKernel(makeParams(geometry, core, lengthX, lengthY, layers, radius));
        VortexSystem s = makeSystem(x, y, g, n);
        RungeKuttaIntegrator rk(n);
        for (int i = 0; i < steps; ++i)
            rk.rk4Step(s, dt, *kernel);
        std::memcpy(x, s.x.data(), n * sizeof(double));
        std::memcpy(y, s.y.data(), n * sizeof(double));
    });
}

// Adaptive DOPRI5 from t=0 to endTime with step landing min(dt, end-t), as the
// reference's own dopriTest drives it.  Records each accepted step as
// {accepted dt, suggested dt, normalized error, rejected} in `record` (capacity cap steps);
// *taken receives the number of accepted steps.
int pvr_dopri5_run(int geometry, double core, double lengthX, double lengthY, int layers,
                   double radius, double *x, double *y, const double *g, std::size_t n, double dt,
                   double endTime, double atol, double rtol, double minDt, double maxDt,
                   double *record, std::size_t cap, std::size_t *taken) {
    return guarded([&] {
        SimParams p = makeParams(geometry, core, lengthX, lengthY, layers, radius);
        p.absoluteTolerance = atol;
        p.relativeTolerance = rtol;
        p.minimumTimeStep = minDt;
        p.maximumTimeStep = maxDt;
        auto kernel = makeReferenceKernel(p);
        VortexSystem s = makeSystem(x, y, g, n);
        RungeKuttaIntegrator rk(n);
        double time = 0.0;
        std::size_t count = 0;
        while (time < endTime) {
            const StepResult r = rk.dopri5Step(s, std::min(dt, endTime - time), *kernel, p);
            if (count < cap) {
                record[4 * count + 0] = r.acceptedTimeStep;
                record[4 * count + 1] = r.suggestedTimeStep;
                record[4 * count + 2] = r.normalizedError;
                record[4 * count + 3] = r.rejectedSteps;
            }
            ++count;
            time += r.acceptedTimeStep;
            dt = r.suggestedTimeStep;
        }
        *taken = count;
        std::memcpy(x, s.x.data(), n * sizeof(double));
        std::memcpy(y, s.y.data(), n * sizeof(double));
    });
}

// One DipoleManager::process call; the surviving/reinjected state is written back
// (capacity n, the population can only shrink or stay) and *newCount receives its size.
int pvr_dipole_process(int geometry, double lengthX, double lengthY, double radius,
                       double distance, int reinjection, std::uint64_t seed, double *x, double *y,
                       double *g, std::size_t n, std::size_t *newCount, std::size_t *events) {
    return guarded([&] {
        SimParams p = makeParams(geometry, 0.0, lengthX, lengthY, 8, radius);
        p.dipoleRemoval = true;
        p.dipoleRemovalDistance = distance;
        p.dipoleReinjection = reinjection == 0   ? ReinjectionMode::none
                              : reinjection == 1 ? ReinjectionMode::independent
                                                 : ReinjectionMode::paired;
        p.randomSeed = seed;
        DipoleManager manager(p);
        VortexSystem s = makeSystem(x, y, g, n);
        *events = manager.process(s);
        *newCount = s.size();
        for (std::size_t i = 0; i < s.size() && i < n; ++i) {
            x[i] = s.x[i];
            y[i] = s.y[i];
            g[i] = s.circulation[i];
        }
    });
}
}
