// pv_host.hpp -- host-side helpers shared by the C-ABI translation units.
#pragma once
#include "pv_hip.h"
#include "pv_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace pv {
namespace host {

inline thread_local std::string g_error;

struct ApiError {
    pv_status code;
    std::string message;
};

[[noreturn]] inline void raise(pv_status code, const std::string &message) {
    throw ApiError{code, message};
}

inline void hip_check(hipError_t status, const char *operation) {
    if (status != hipSuccess)
        raise(PV_ERR_RUNTIME, std::string(operation) + ": " + hipGetErrorString(status));
}

template <typename F> pv_status guard(F &&body) {
    try {
        body();
        return PV_OK;
    } catch (const ApiError &e) {
        g_error = e.message;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_error = "out of host memory";
        return PV_ERR_RUNTIME;
    } catch (const std::exception &e) {
        g_error = e.what();
        return PV_ERR_RUNTIME;
    }
}

struct DeviceGuard {
    int previous = 0;
    bool changed = false;
    explicit DeviceGuard(int device) {
        hip_check(hipGetDevice(&previous), "query HIP device");
        if (previous != device) {
            hip_check(hipSetDevice(device), "select HIP device");
            changed = true;
        }
    }
    ~DeviceGuard() {
        if (changed)
            (void)hipSetDevice(previous);
    }
};

template <typename T> void dfree(T *&p) {
    if (p)
        (void)hipFree(p);
    p = nullptr;
}

template <typename T> T *dalloc(size_t count, const char *what) {
    void *p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), what);
    return static_cast<T *>(p);
}


inline pv::GeoParams geo_params(const pv_config &c) {
    pv::GeoParams p{};
    p.geometry = c.geometry;
    if (c.geometry == PV_INFINITE) {
        p.eps2 = c.core_radius * c.core_radius;
    } else if (c.geometry == PV_PERIODIC) {
        p.layers = c.image_layers;
        p.L = c.box_length_x;
        p.invL = 1.0 / c.box_length_x;
        p.wave = 2.0 * pv::kPi / c.box_length_x;
        p.hwave = 0.5 * p.wave;
        p.scale = 0.5 / c.box_length_x;
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int n = 1; n <= pv::kMaxTerms; ++n) {
            long double s = 0.0L;
            for (int m = c.image_layers; m >= 2; --m)
                s += expl(-2.0L * pi * n * m);
            p.G4[n] = (double)s;
            p.G4f[n] = (float)s;
            if (c.image_layers >= 1)
                s += expl(-2.0L * pi * n);
            p.H4[n] = (double)(4.0L * s / n);
        }
        const long double ch = coshl(2.0L * pi), sh = sinhl(2.0L * pi);
        p.W1 = c.image_layers >= 1 ? 2.0 : 0.0;
        p.C1 = (double)ch;
        p.C2 = (double)(2.0L * ch);
        p.S4 = (double)(4.0L * sh * sh);
        p.W1f = (float)p.W1;
        p.C1f = (float)ch;
        p.C2f = (float)(2.0L * ch);
        p.S4f = (float)(4.0L * sh * sh);
        long double cl = 0.0L;
        for (int m = c.image_layers; m >= 1; --m)
            cl += log1pl(expl(-4.0L * pi * m));
        p.CL = (double)(2.0L * cl);
    } else {
        p.R = c.disk_radius;
        p.invR = 1.0 / c.disk_radius;
        p.R2 = c.disk_radius * c.disk_radius;
        p.R2lo = std::fma(c.disk_radius, c.disk_radius, -p.R2);
    }
    return p;
}

// Launch geometry of the split pair kernel for `targets` targets against n sources.  tpt,
// splits and tile of 0 are automatic: 2 targets per thread from 32768 targets, the precision's
// LDS tile, and enough source splits for about 4 blocks per CU; the splits are further capped
// at max_splits (a bound on the partial-sum buffer) and at one LDS tile of sources each.
struct SplitShape {
    int tpt, tile;
    size_t tiles, nsplit, split_len;
};

inline SplitShape split_shape(size_t targets, size_t n, int tpt, int splits, int tile,
                              int precision, int cu_count, size_t max_splits = SIZE_MAX) {
    if (tpt <= 0)
        tpt = targets >= 32768 ? 2 : 1;
    tpt = tpt >= 4 ? 4 : (tpt >= 2 ? 2 : 1);
    if (tile != 128 && tile != 256)
        tile = pv::pair_tile(precision);
    SplitShape s{};
    s.tpt = tpt;
    s.tile = tile;
    s.tiles = (targets + (size_t)pv::kBlock * tpt - 1) / ((size_t)pv::kBlock * tpt);
    const size_t max_split = std::max<size_t>(1, (n + tile - 1) / tile);
    size_t nsplit = splits > 0 ? (size_t)splits : 0;
    if (nsplit == 0) {
        const size_t want = (size_t)4 * cu_count;
        nsplit = std::max<size_t>(1, (want + s.tiles - 1) / std::max<size_t>(1, s.tiles));
    }
    nsplit = std::max<size_t>(1, std::min({nsplit, max_split, max_splits}));
    size_t split_len = (n + nsplit - 1) / nsplit;
    split_len = std::max<size_t>(tile, (split_len + tile - 1) / tile * tile);
    s.split_len = split_len;
    s.nsplit = std::max<size_t>(1, (n + split_len - 1) / split_len);
    return s;
}

// Constructor checks of src/compute.cpp:50-56, 165-173, 258-261 and the geometry switch of
// src/backend_common.cpp:4-13.
inline void check_config(const pv_config &c) {
    if (c.geometry == PV_INFINITE) {
        const double r = c.core_radius, r2 = r * r;
        if (!(r >= 0.0) || !std::isfinite(r2) || (r > 0.0 && r2 == 0.0))
            raise(PV_ERR_INVALID_ARGUMENT,
                  "core radius is negative or outside the supported numeric range");
    } else if (c.geometry == PV_PERIODIC) {
        const double lx = c.box_length_x, ly = c.box_length_y;
        if (!(lx > 0.0) || !(ly > 0.0) || !std::isfinite(lx) || !std::isfinite(ly) ||
            !std::isfinite(2.0 * pv::kPi / lx) || c.image_layers < 0 || c.image_layers > 64)
            raise(PV_ERR_INVALID_ARGUMENT, "invalid periodic-box parameters");
        if (std::fabs(lx - ly) > 1e-13 * std::max(lx, ly))
            raise(PV_ERR_INVALID_ARGUMENT,
                  "Weiss-McWilliams kernel requires a square periodic box");
    } else if (c.geometry == PV_DISK) {
        const double r = c.disk_radius, r2 = r * r;
        if (!(r > 0.0) || !std::isfinite(r2) || r2 == 0.0)
            raise(PV_ERR_INVALID_ARGUMENT, "invalid disk parameters");
    } else {
        raise(PV_ERR_INVALID_ARGUMENT, "unsupported boundary condition");
    }
    if (c.precision != PV_FP64 && c.precision != PV_MIXED)
        raise(PV_ERR_INVALID_ARGUMENT, "unsupported precision");
}


} // namespace host
} // namespace pv
