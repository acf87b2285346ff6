// pv_noalloc.cpp -- pv_evaluate_device_noalloc: a velocity evaluation that allocates nothing.
//
// For callers that cannot call hipMalloc (code under graph capture, tight serving loops,
// harnesses that forbid device allocation).  The caller passes device arrays; everything the
// split pair kernel needs besides them -- the sticky status word, the per-tile tickets and the
// split partial sums -- lives in `scratch`, one fixed file-scope device array per GPU, held
// until the process ends.  Host code never uses the array's name as an address (on the host
// it is a stand-in symbol): the first call on a device, with that device current, asks the
// runtime for the array's device address once (hipGetSymbolAddress) and keeps it per device.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind; a call clears
// what must start at zero (status word, tickets) with a small kernel before it launches.
// Calls that overlap on one GPU (other threads, other streams) race on the array.
#include "pv_hip.h"
#include "pv_host.hpp"
#include "pv_kernels.hpp"

#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

namespace {

using namespace pv::host;

constexpr size_t kScratchBytes = PV_NOALLOC_SCRATCH_BYTES;
constexpr size_t kStatusBytes = 256;                 // DevStatus, padded
constexpr size_t kTicketBytes = (size_t)64 << 10;    // one unsigned per target tile
constexpr size_t kPartialOffset = kStatusBytes + kTicketBytes;
constexpr size_t kPartialBytes = kScratchBytes - kPartialOffset;
constexpr size_t kMaxVortices = PV_NOALLOC_MAX_VORTICES;
constexpr int kPrologueBlocks = 64;
constexpr int kPrologueThreads = 256;

static_assert(sizeof(pv::DevStatus) <= kStatusBytes, "status word does not fit its slot");
// The most target tiles a population can have (one target per thread): all need a ticket.
static_assert(kMaxVortices / pv::kBlock * sizeof(unsigned) <= kTicketBytes,
              "tickets do not fit their slot");
static_assert(kPrologueBlocks * 3 * sizeof(double) <= kPartialBytes, "prologue partials");

// The scratch array.  Its name is a host-side stand-in outside device code: host code uses
// only the per-device address that hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
struct DeviceScratch {
    unsigned char *base = nullptr;   // device address of `scratch` on this GPU
    int cu_count = 0;
};
std::mutex g_scratch_mutex;
DeviceScratch g_scratch[kMaxDevices];

// The address of this GPU's scratch array, found once per device with that device current.
DeviceScratch scratch_for(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    DeviceScratch &d = g_scratch[device];
    if (!d.base) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                  "find the address of the no-allocation scratch array");
        int cu = 0;
        hip_check(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device),
                  "query HIP device properties");
        d.cu_count = cu > 0 ? cu : 1;
        d.base = static_cast<unsigned char *>(p);
    }
    return d;
}

// Clears the status word (fail_epoch back to its UINT_MAX sentinel) and `count` tickets.
__global__ void noalloc_clear_kernel(pv::DevStatus *st, unsigned *tickets, int count) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        st->flags = 0u;
        st->fail_epoch = UINT_MAX;
        st->err_bits = 0ull;
        st->count = 0ull;
        st->barrier = 0ull;
    }
    for (int k = t; k < count; k += gridDim.x * blockDim.x)
        tickets[k] = 0u;
}

// Per block: the compensated sums of G and |G| and the number of non-finite G, for the
// validation the host does on host arrays (src/vortex.h:11-13, src/compute.cpp:31-39).
__global__ __launch_bounds__(kPrologueThreads) void noalloc_prologue_kernel(const double *g,
                                                                            int n,
                                                                            double *partials) {
    __shared__ double red[kPrologueThreads][3];
    pv::Comp total, absolute;
    total.zero();
    absolute.zero();
    double bad = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double gi = g[i];
        if (!__builtin_isfinite(gi)) {
            bad += 1.0;
            continue;
        }
        total.add(gi);
        absolute.add(fabs(gi));
    }
    red[threadIdx.x][0] = total.value();
    red[threadIdx.x][1] = absolute.value();
    red[threadIdx.x][2] = bad;
    __syncthreads();
    if (threadIdx.x == 0) {
        pv::Comp s0, s1;
        s0.zero();
        s1.zero();
        double b = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            s0.add(red[t][0]);
            s1.add(red[t][1]);
            b += red[t][2];
        }
        partials[3 * blockIdx.x] = s0.value();
        partials[3 * blockIdx.x + 1] = s1.value();
        partials[3 * blockIdx.x + 2] = b;
    }
}

const char *coincident_message(int geometry) {
    return geometry == PV_INFINITE   ? "coincident vortices in singular point-vortex kernel"
           : geometry == PV_PERIODIC ? "coincident periodic vortices"
                                     : "coincident vortices in disk";
}

void read_status(pv::DevStatus *host, const pv::DevStatus *st, hipStream_t s) {
    hip_check(hipMemcpyAsync(host, st, sizeof(pv::DevStatus), hipMemcpyDeviceToHost, s),
              "read HIP status");
    hip_check(hipStreamSynchronize(s), "synchronize HIP stream");
}

} // namespace

extern "C" pv_status pv_evaluate_device_noalloc(const pv_config *config, const double *x,
                                                const double *y, const double *gamma, size_t n,
                                                double *u, double *v, void *stream) {
    return guard([&] {
        if (!config)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        check_config(*config);
        if (n > kMaxVortices)
            raise(PV_ERR_INVALID_ARGUMENT,
                  "vortex population exceeds the no-allocation limit of " +
                      std::to_string(kMaxVortices) + " vortices");
        if (n == 0)
            return;
        if (!x || !y || !gamma || !u || !v)
            raise(PV_ERR_INVALID_ARGUMENT, "null device array");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "count HIP devices");
        if (config->device < 0 || config->device >= count || config->device >= kMaxDevices)
            raise(PV_ERR_INVALID_ARGUMENT, "invalid HIP device ordinal");
        DeviceGuard dg(config->device);
        const DeviceScratch sc = scratch_for(config->device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        pv::DevStatus *st = reinterpret_cast<pv::DevStatus *>(sc.base);
        unsigned *tickets = reinterpret_cast<unsigned *>(sc.base + kStatusBytes);
        double *part = reinterpret_cast<double *>(sc.base + kPartialOffset);
        const pv::GeoParams gp = geo_params(*config);

        // Circulations: finite, and neutral in a periodic box (the host checks of
        // pv_evaluate, done here on the device arrays).
        hipLaunchKernelGGL(noalloc_prologue_kernel, dim3(kPrologueBlocks), dim3(kPrologueThreads),
                           0, s, gamma, (int)n, part);
        hip_check(hipGetLastError(), "launch HIP validation kernel");
        double sums[3 * kPrologueBlocks];
        hip_check(hipMemcpyAsync(sums, part, sizeof sums, hipMemcpyDeviceToHost, s),
                  "read HIP validation sums");
        hip_check(hipStreamSynchronize(s), "synchronize HIP stream");
        double total = 0.0, absolute = 0.0, bad = 0.0;
        for (int b = 0; b < kPrologueBlocks; ++b) {
            total += sums[3 * b];
            absolute += sums[3 * b + 1];
            bad += sums[3 * b + 2];
        }
        if (bad > 0.0)
            raise(PV_ERR_INVALID_ARGUMENT, "vortex data must be finite");
        if (config->geometry == PV_PERIODIC &&
            std::fabs(total) > 1e-12 * std::max(1.0, absolute))
            raise(PV_ERR_INVALID_ARGUMENT, "periodic box requires zero total circulation");

        // The split pair kernel with the built-in launch shape, its splits capped so that
        // their partial sums fit the scratch array.
        const SplitShape sh = split_shape(n, n, 0, 0, 0, config->precision, sc.cu_count,
                                          kPartialBytes / (2 * n * sizeof(double)));
        hipLaunchKernelGGL(noalloc_clear_kernel, dim3(16), dim3(256), 0, s, st, tickets,
                           (int)sh.tiles);
        hip_check(hipGetLastError(), "launch HIP status reset");
        pv::PairLaunch a{};
        a.x = x, a.y = y, a.g = gamma;
        a.n = (int)n;
        a.tbegin = 0, a.tend = (int)n;
        a.split_len = (int)sh.split_len;
        a.nsplit = (int)sh.nsplit;
        a.part = part;
        a.tickets = tickets;
        a.st = st;
        a.epoch = 1;
        a.op.kind = pv::kOpStore;
        a.op.ku = u, a.op.kv = v;
        pv::launch_pair(config->geometry, config->precision, sh.tpt, sh.tile, a, gp, s);
        hip_check(hipGetLastError(), "launch HIP velocity kernel");

        // The sticky status, turned into the reference's exceptions as pv_evaluate does.
        pv::DevStatus host{};
        read_status(&host, st, s);
        if (host.flags & pv::kFailStateNonFinite)
            raise(PV_ERR_INVALID_ARGUMENT, "vortex data must be finite");
        if (host.flags & pv::kFailOutsideDisk)
            raise(PV_ERR_INVALID_ARGUMENT, "vortex lies on or outside the disk");
        if (host.flags & pv::kFailVelocity) {
            pv::launch_classify(gp, x, y, (int)n, 0, (int)n, st, s);
            hip_check(hipGetLastError(), "launch HIP coincidence check");
            read_status(&host, st, s);
            if (host.flags & pv::kFailCoincident)
                raise(PV_ERR_RUNTIME, coincident_message(config->geometry));
            raise(PV_ERR_RUNTIME, "non-finite velocity; check scales and close encounters");
        }
    });
}
