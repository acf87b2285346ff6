// pv_api.cpp -- the C-ABI of point_vortex_hip (include/pv_hip.h).
//
// A handle owns its device buffers and one HIP stream.  Velocity evaluations, RK stage
// arithmetic, the DOPRI5 error norm, invariants and the dipole search all run on the device;
// the host validates arguments (in the reference's order, with its messages), chooses launch
// shapes, runs the DOPRI5 step-size controller on the one reduced error value per attempt,
// and turns the sticky device status into the reference's exceptions.
#include "pv_hip.h"
#include "pv_host.hpp"
#include "pv_kernels.hpp"

#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using namespace pv::host;

// src/timestep.cpp:69-73 (b5, b4) and src/timestep.h:8-15 (a_{s,j}).
constexpr double kB5[7] = {35.0 / 384.0,     0.0, 500.0 / 1113.0, 125.0 / 192.0,
                           -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
constexpr double kB4[7] = {5179.0 / 57600.0, 0.0,           7571.0 / 16695.0, 393.0 / 640.0,
                           -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
constexpr double kDopriA[7][7] = {
    {0},
    {1.0 / 5.0},
    {3.0 / 40.0, 9.0 / 40.0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
    {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};

// What a launch evaluated, so that a failure can be classified afterwards.
struct EvalRecord {
    const double *x, *y;
    int n, tbegin, tend;
};

} // namespace

struct pv_handle {
    pv_config cfg{};
    // Dipole events: the reinjection generator, src/dipole.cpp:29-38.
    std::mt19937_64 dipole_rng{1234567ULL ^ 0xd1b54a32d192ed03ULL};
    pv::GeoParams gp{};
    int cu_count = 256;
    hipStream_t stream = nullptr;

    size_t n = 0, cap = 0;
    bool state_valid = false;
    bool fsal_valid = false;
    double *x = nullptr, *y = nullptr, *g = nullptr;
    double *xalt = nullptr, *yalt = nullptr;
    double *xa = nullptr, *ya = nullptr, *xb = nullptr, *yb = nullptr;
    double *kx[7] = {}, *ky[7] = {};
    double *u = nullptr, *v = nullptr;

    size_t scap = 0;
    double *sx = nullptr, *sy = nullptr, *sg = nullptr, *su = nullptr, *sv = nullptr;

    double *part = nullptr;
    size_t part_cap = 0;
    unsigned *tickets = nullptr;
    size_t tickets_cap = 0;
    double *red = nullptr;
    size_t red_cap = 0;
    unsigned long long *cand_first = nullptr, *cand_second = nullptr;
    double *cand_d2 = nullptr;
    size_t cand_cap = 0;

    pv::DevStatus *st = nullptr;
    pv::DevStatus *st_host = nullptr;

    int tpt_override = 0, split_override = 0, tile_override = 0, sym_override = -1;
    std::vector<pv_launch_rule> table;   // tuned shapes, first matching rule wins
    bool profiling = false;
    std::vector<hipEvent_t> events;   // pool, two per profiled launch
    size_t events_used = 0;
    unsigned epoch = 0;
    unsigned epoch_base = 1;
    std::vector<EvalRecord> records;

    // Persistent RK4 path (pv_persist.hip): mode -1 automatic, 0 never, 1 whenever it fits.
    int persist_mode = -1, persist_tb = 0;
    double persist_wait_s = 2.0;     // bound on one grid-barrier wait
    bool coop = false;               // the device takes cooperative launches
    double wall_khz = 100000.0;      // rate of the device's constant wall clock
    int persist_cap[2][3] = {{-1, -1, -1}, {-1, -1, -1}};   // co-resident blocks, -1 unknown
    unsigned long long barrier_arrivals = 0;   // since the status word was last reset

    ~pv_handle() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        dfree(x), dfree(y), dfree(g), dfree(xalt), dfree(yalt);
        dfree(xa), dfree(ya), dfree(xb), dfree(yb);
        for (int s = 0; s < 7; ++s)
            dfree(kx[s]), dfree(ky[s]);
        dfree(u), dfree(v);
        dfree(sx), dfree(sy), dfree(sg), dfree(su), dfree(sv);
        dfree(part), dfree(tickets), dfree(red);
        dfree(cand_first), dfree(cand_second), dfree(cand_d2);
        dfree(st);
        for (hipEvent_t e : events)
            (void)hipEventDestroy(e);
        if (st_host)
            (void)hipHostFree(st_host);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace {

void check_population(size_t n) {
    if (n > (size_t)(INT_MAX / 4))
        raise(PV_ERR_INVALID_ARGUMENT, "vortex population exceeds the supported size");
}

void ensure_state(pv_handle *h, size_t n) {
    if (n <= h->cap && h->x)
        return;
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    dfree(h->x), dfree(h->y), dfree(h->g), dfree(h->xalt), dfree(h->yalt);
    dfree(h->xa), dfree(h->ya), dfree(h->xb), dfree(h->yb);
    for (int s = 0; s < 7; ++s)
        dfree(h->kx[s]), dfree(h->ky[s]);
    dfree(h->u), dfree(h->v);
    h->cap = 0;
    h->state_valid = false;
    h->x = dalloc<double>(n, "hipMalloc(x)");
    h->y = dalloc<double>(n, "hipMalloc(y)");
    h->g = dalloc<double>(n, "hipMalloc(circulation)");
    h->xalt = dalloc<double>(n, "hipMalloc(x next)");
    h->yalt = dalloc<double>(n, "hipMalloc(y next)");
    h->xa = dalloc<double>(n, "hipMalloc(stage x)");
    h->ya = dalloc<double>(n, "hipMalloc(stage y)");
    h->xb = dalloc<double>(n, "hipMalloc(stage x)");
    h->yb = dalloc<double>(n, "hipMalloc(stage y)");
    for (int s = 0; s < 7; ++s) {
        h->kx[s] = dalloc<double>(n, "hipMalloc(x stage)");
        h->ky[s] = dalloc<double>(n, "hipMalloc(y stage)");
    }
    h->u = dalloc<double>(n, "hipMalloc(u)");
    h->v = dalloc<double>(n, "hipMalloc(v)");
    h->cap = n;
}

void ensure_scratch(pv_handle *h, size_t n) {
    if (n <= h->scap && h->sx)
        return;
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    dfree(h->sx), dfree(h->sy), dfree(h->sg), dfree(h->su), dfree(h->sv);
    h->scap = 0;
    h->sx = dalloc<double>(n, "hipMalloc(scratch x)");
    h->sy = dalloc<double>(n, "hipMalloc(scratch y)");
    h->sg = dalloc<double>(n, "hipMalloc(scratch circulation)");
    h->su = dalloc<double>(n, "hipMalloc(scratch u)");
    h->sv = dalloc<double>(n, "hipMalloc(scratch v)");
    h->scap = n;
}

void ensure_partials(pv_handle *h, size_t doubles) {
    if (doubles <= h->part_cap && h->part)
        return;
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    dfree(h->part);
    h->part_cap = 0;
    h->part = dalloc<double>(doubles, "hipMalloc(partial sums)");
    h->part_cap = doubles;
}

void ensure_tickets(pv_handle *h, size_t count) {
    if (count <= h->tickets_cap && h->tickets)
        return;
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    dfree(h->tickets);
    h->tickets_cap = 0;
    h->tickets = dalloc<unsigned>(count, "hipMalloc(tile tickets)");
    // Stream-ordered: the handle's stream is non-blocking, so a memset on the null stream
    // would not be ordered before the pair kernel that counts on these tickets.
    hip_check(hipMemsetAsync(h->tickets, 0, std::max<size_t>(count, 1) * sizeof(unsigned),
                             h->stream),
              "clear tile tickets");
    h->tickets_cap = count;
}

void ensure_reduction(pv_handle *h, size_t doubles) {
    if (doubles <= h->red_cap && h->red)
        return;
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    dfree(h->red);
    h->red_cap = 0;
    h->red = dalloc<double>(doubles, "hipMalloc(reduction partials)");
    h->red_cap = doubles;
}

void ensure_candidates(pv_handle *h, size_t cap) {
    if (cap <= h->cand_cap && h->cand_first)
        return;
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    dfree(h->cand_first), dfree(h->cand_second), dfree(h->cand_d2);
    h->cand_cap = 0;
    h->cand_first = dalloc<unsigned long long>(cap, "hipMalloc(dipole candidates)");
    h->cand_second = dalloc<unsigned long long>(cap, "hipMalloc(dipole candidates)");
    h->cand_d2 = dalloc<double>(cap, "hipMalloc(dipole candidates)");
    h->cand_cap = cap;
}

// Starts a synchronised sequence of launches: the status word is cleared (fail_epoch back to
// its UINT_MAX sentinel) and the launch epochs restart from 1, so a sequence's epochs never
// reach the sentinel (sequences are bounded, see kMaxSequenceSteps).
void begin_sequence(pv_handle *h) {
    pv::launch_reset_status(h->st, h->stream);
    hip_check(hipGetLastError(), "launch HIP status reset");
    h->epoch = 0;
    h->epoch_base = 1;
    h->records.clear();
    h->barrier_arrivals = 0;
}

struct Shape {
    int tpt, tile;
    size_t tiles, nsplit, split_len;
    bool sym;      // symmetric path (whole-population evaluations only)
};

// Partial sums of the symmetric path are nt x 2N doubles; beyond this the split path runs.
constexpr size_t kSymPartialLimit = (size_t)8 << 30;

// Launch shape of an evaluation of `targets` targets against n sources.  Precedence: an
// explicit pv_set_launch_shape, then the first matching rule of the handle's launch table,
// then the built-in heuristic (2 targets per thread from 32768 targets; enough source splits
// for about 4 blocks per CU).
Shape resolve_shape(const pv_handle *h, size_t targets, size_t n) {
    const int prec = h->cfg.precision;
    int tpt = 0, splits = 0, tile = 0, sym = 0;
    for (const pv_launch_rule &r : h->table) {
        if (r.geometry == h->cfg.geometry && r.precision == prec && n <= r.n_max) {
            tpt = r.targets_per_thread, splits = r.source_splits, tile = r.tile;
            sym = r.symmetric;
            break;
        }
    }
    if (h->tpt_override > 0)
        tpt = h->tpt_override;
    if (h->split_override > 0)
        splits = h->split_override;
    if (h->tile_override > 0)
        tile = h->tile_override;
    if (h->sym_override >= 0)
        sym = h->sym_override;
    const SplitShape sp = split_shape(targets, n, tpt, splits, tile, prec, h->cu_count);
    Shape s{};
    s.tpt = sp.tpt;
    s.tile = sp.tile;
    s.tiles = sp.tiles;
    s.split_len = sp.split_len;
    s.nsplit = sp.nsplit;
    // The symmetric kernel also has 8 targets per thread on the infinite plane; elsewhere, and
    // on the split path, a request for 8 runs at 4.
    const int sym_tpt =
        tpt >= 8 && pv::sym_tpt_supported(h->cfg.geometry, prec, 8) ? 8 : s.tpt;
    const size_t tt = (size_t)pv::sym_tile(sym_tpt), nt = (n + tt - 1) / tt;
    s.sym = sym == 1 && pv::sym_supported(h->cfg.geometry, prec) && targets == n && n > 0 &&
            nt * 2 * n * sizeof(double) <= kSymPartialLimit;
    if (s.sym)
        s.tpt = sym_tpt;
    return s;
}

// Launch table text: one rule per line, "geometry precision n_max targets_per_thread
// source_splits tile" (geometry infinite|periodic|disk, precision fp64|mixed); '#' starts a
// comment.  A malformed table is an error, not a silent fallback.
std::vector<pv_launch_rule> read_launch_table(const std::string &path) {
    std::ifstream in(path);
    if (!in)
        raise(PV_ERR_INVALID_ARGUMENT, "cannot read launch table " + path);
    std::vector<pv_launch_rule> rules;
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t hash = line.find('#');
        if (hash != std::string::npos)
            line.resize(hash);
        std::istringstream ss(line);
        std::string geo, prec;
        if (!(ss >> geo))
            continue;
        pv_launch_rule r{};
        unsigned long long n_max = 0;
        if (!(ss >> prec >> n_max >> r.targets_per_thread >> r.source_splits >> r.tile))
            raise(PV_ERR_INVALID_ARGUMENT,
                  "launch table " + path + ": line " + std::to_string(lineno) + " is malformed");
        std::string extra;
        if (ss >> extra) {
            if (extra != "0" && extra != "1")
                raise(PV_ERR_INVALID_ARGUMENT, "launch table " + path + ": line " +
                                                   std::to_string(lineno) + " is malformed");
            r.symmetric = extra == "1";
            if (ss >> extra)
                raise(PV_ERR_INVALID_ARGUMENT, "launch table " + path + ": line " +
                                                   std::to_string(lineno) + " is malformed");
        }
        r.geometry = geo == "infinite" ? PV_INFINITE
                     : geo == "periodic" ? PV_PERIODIC
                     : geo == "disk"     ? PV_DISK
                                         : -1;
        r.precision = prec == "fp64" ? PV_FP64 : prec == "mixed" ? PV_MIXED : -1;
        r.n_max = (size_t)n_max;
        if (r.geometry < 0 || r.precision < 0)
            raise(PV_ERR_INVALID_ARGUMENT, "launch table " + path + ": line " +
                                               std::to_string(lineno) +
                                               " names an unknown geometry or precision");
        rules.push_back(r);
    }
    return rules;
}

void check_rules(const pv_launch_rule *rules, size_t count) {
    for (size_t k = 0; k < count; ++k) {
        const pv_launch_rule &r = rules[k];
        if (r.geometry < PV_INFINITE || r.geometry > PV_DISK ||
            (r.precision != PV_FP64 && r.precision != PV_MIXED))
            raise(PV_ERR_INVALID_ARGUMENT, "launch rule names an unknown geometry or precision");
        if (r.targets_per_thread != 0 && r.targets_per_thread != 1 && r.targets_per_thread != 2 &&
            r.targets_per_thread != 4 && r.targets_per_thread != 8)
            raise(PV_ERR_INVALID_ARGUMENT, "targets per thread must be 0, 1, 2, 4 or 8");
        if (r.source_splits < 0)
            raise(PV_ERR_INVALID_ARGUMENT, "source splits must be non-negative");
        if (r.tile != 0 && r.tile != 128 && r.tile != 256)
            raise(PV_ERR_INVALID_ARGUMENT, "source tile must be 0, 128 or 256");
        if (r.symmetric != 0 && r.symmetric != 1)
            raise(PV_ERR_INVALID_ARGUMENT, "symmetric must be 0 or 1");
    }
}

// The default launch table of a new handle: $PV_HIP_LAUNCH_TABLE, else launch_table.txt
// beside this library when present, else none (built-in heuristic).
std::vector<pv_launch_rule> default_launch_table() {
    if (const char *env = std::getenv("PV_HIP_LAUNCH_TABLE")) {
        if (*env) {
            auto rules = read_launch_table(env);
            check_rules(rules.data(), rules.size());
            return rules;
        }
    }
    Dl_info info{};
    if (dladdr((void *)&pv_create, &info) && info.dli_fname) {
        std::string path(info.dli_fname);
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
               "/launch_table.txt";
        std::ifstream probe(path);
        if (probe) {
            auto rules = read_launch_table(path);
            check_rules(rules.data(), rules.size());
            return rules;
        }
    }
    return {};
}

void record_events(pv_handle *h, hipEvent_t &ev0, hipEvent_t &ev1) {
    if (!h->profiling)
        return;
    while (h->events.size() < h->events_used + 2) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "create HIP event");
        h->events.push_back(e);
    }
    ev0 = h->events[h->events_used++];
    ev1 = h->events[h->events_used++];
}

// Whole-population evaluation over unordered tile pairs: the pair kernel writes every
// (slot, vortex) partial once; the reduce kernel sums a vortex's slots in order and applies
// the stage epilogue.  Both launches are bracketed by the profiling events.
void evaluate_symmetric(pv_handle *h, const double *x, const double *y, const double *g,
                        size_t n, const Shape &sh, const pv::StageOp &op) {
    const size_t tt = (size_t)pv::sym_tile(sh.tpt), nt = (n + tt - 1) / tt;
    ensure_partials(h, nt * 2 * n);
    pv::SymLaunch a{};
    a.x = x, a.y = y, a.g = g;
    a.n = (int)n;
    a.nt = (int)nt;
    a.part = h->part;
    a.st = h->st;
    a.epoch = ++h->epoch;
    a.self_image = h->cfg.geometry == PV_DISK && h->cfg.precision == PV_MIXED;
    a.op = op;
    h->records.push_back({x, y, (int)n, 0, (int)n});
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    record_events(h, ev0, ev1);
    if (ev0)
        hip_check(hipEventRecord(ev0, h->stream), "record HIP event");
    pv::launch_sym(h->cfg.geometry, h->cfg.precision, sh.tpt, sh.tile, a, h->gp, h->stream);
    hip_check(hipGetLastError(), "launch HIP velocity kernel");
    pv::launch_sym_reduce(a, h->gp, h->stream);
    hip_check(hipGetLastError(), "launch HIP velocity reduction");
    if (ev1)
        hip_check(hipEventRecord(ev1, h->stream), "record HIP event");
}

// One velocity evaluation of positions (x, y) with circulations g, targets [tb, te).
void evaluate(pv_handle *h, const double *x, const double *y, const double *g, size_t n,
              size_t tb, size_t te, const pv::StageOp &op) {
    const int prec = h->cfg.precision;
    const size_t targets = te - tb;
    const Shape sh = resolve_shape(h, targets, n);
    if (sh.sym && tb == 0 && te == n) {
        evaluate_symmetric(h, x, y, g, n, sh, op);
        return;
    }
    const int tpt = sh.tpt;
    const size_t tiles = sh.tiles, nsplit = sh.nsplit, split_len = sh.split_len;
    if (nsplit > 1)
        ensure_partials(h, nsplit * 2 * n);
    ensure_tickets(h, tiles);

    pv::PairLaunch a{};
    a.x = x, a.y = y, a.g = g;
    a.n = (int)n;
    a.tbegin = (int)tb, a.tend = (int)te;
    a.split_len = (int)split_len;
    a.nsplit = (int)nsplit;
    a.part = h->part;
    a.tickets = h->tickets;
    a.st = h->st;
    a.epoch = ++h->epoch;
    a.op = op;
    h->records.push_back({x, y, (int)n, (int)tb, (int)te});
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    record_events(h, ev0, ev1);
    if (ev0)
        hip_check(hipEventRecord(ev0, h->stream), "record HIP event");
    pv::launch_pair(h->cfg.geometry, prec, tpt, sh.tile, a, h->gp, h->stream);
    hip_check(hipGetLastError(), "launch HIP velocity kernel");
    if (ev1)
        hip_check(hipEventRecord(ev1, h->stream), "record HIP event");
}

pv::StageOp store_op(double *u, double *v) {
    pv::StageOp op{};
    op.kind = pv::kOpStore;
    op.ku = u, op.kv = v;
    return op;
}

const char *coincident_message(int geometry) {
    return geometry == PV_INFINITE   ? "coincident vortices in singular point-vortex kernel"
           : geometry == PV_PERIODIC ? "coincident periodic vortices"
                                     : "coincident vortices in disk";
}

void read_status(pv_handle *h) {
    hip_check(hipMemcpyAsync(h->st_host, h->st, sizeof(pv::DevStatus), hipMemcpyDeviceToHost,
                             h->stream),
              "read HIP status");
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
}

// Waits for the stream, then turns the sticky status into the reference's exception.
// `velocity_must_be_finite` is evaluate() (true) versus evaluateRange() (false).
void finish(pv_handle *h, bool velocity_must_be_finite) {
    read_status(h);
    const unsigned flags = h->st_host->flags;
    if (!flags)
        return;
    if (flags & pv::kFailStateNonFinite)
        raise(PV_ERR_INVALID_ARGUMENT, "vortex data must be finite");
    if (flags & pv::kFailOutsideDisk)
        raise(PV_ERR_INVALID_ARGUMENT, "vortex lies on or outside the disk");
    if (flags & pv::kFailVelocity) {
        const unsigned e = h->st_host->fail_epoch;
        if (e >= h->epoch_base && e - h->epoch_base < h->records.size()) {
            const EvalRecord r = h->records[e - h->epoch_base];
            pv::launch_classify(h->gp, r.x, r.y, r.n, r.tbegin, r.tend, h->st, h->stream);
            hip_check(hipGetLastError(), "launch HIP coincidence check");
            read_status(h);
        }
        if (h->st_host->flags & pv::kFailCoincident)
            raise(PV_ERR_RUNTIME, coincident_message(h->cfg.geometry));
        if (velocity_must_be_finite)
            raise(PV_ERR_RUNTIME, "non-finite velocity; check scales and close encounters");
    }
    if (flags & pv::kFailErrorEstimate)
        raise(PV_ERR_RUNTIME, "non-finite adaptive error estimate or invalid tolerance");
    if (flags & pv::kFailBarrier)
        raise(PV_ERR_RUNTIME, "grid barrier of the persistent RK4 kernel timed out");
}

// src/vortex.h:7-14, src/compute.cpp:31-48 on host arrays.
void validate_host_arrays(const double *x, const double *y, const double *g, size_t n) {
    if (n && (!x || !y || !g))
        raise(PV_ERR_INVALID_ARGUMENT, "vortex arrays have different lengths");
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(g[i]))
            raise(PV_ERR_INVALID_ARGUMENT, "vortex data must be finite");
}

void validate_geometry(const pv_handle *h, const double *x, const double *y, const double *g,
                       size_t n) {
    if (h->cfg.geometry == PV_PERIODIC) {
        double total = 0.0, absolute = 0.0;
        for (size_t i = 0; i < n; ++i) {
            total += g[i];
            absolute += std::fabs(g[i]);
        }
        if (std::fabs(total) > 1e-12 * std::max(1.0, absolute))
            raise(PV_ERR_INVALID_ARGUMENT, "periodic box requires zero total circulation");
    } else if (h->cfg.geometry == PV_DISK) {
        const double r2 = h->cfg.disk_radius * h->cfg.disk_radius;
        for (size_t i = 0; i < n; ++i)
            if (x[i] * x[i] + y[i] * y[i] >= r2)
                raise(PV_ERR_INVALID_ARGUMENT, "vortex lies on or outside the disk");
    }
}

void require_state(const pv_handle *h) {
    if (!h->state_valid)
        raise(PV_ERR_LOGIC, "HIP device state has not been uploaded");
}

void require_dt(double dt) {
    if (!std::isfinite(dt) || !(dt > 0.0))
        raise(PV_ERR_INVALID_ARGUMENT, "timestep must be finite and positive");
}

void upload_to(pv_handle *h, double *dx, double *dy, double *dg, const double *x,
               const double *y, const double *g, size_t n) {
    const size_t bytes = n * sizeof(double);
    hip_check(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, h->stream), "copy x to GPU");
    hip_check(hipMemcpyAsync(dy, y, bytes, hipMemcpyHostToDevice, h->stream), "copy y to GPU");
    if (dg)
        hip_check(hipMemcpyAsync(dg, g, bytes, hipMemcpyHostToDevice, h->stream),
                  "copy circulation to GPU");
}

// Moments {G, G x, G y, G r^2} of device arrays (src/compute.cpp:11-27).
void device_moments(pv_handle *h, const double *x, const double *y, const double *g, size_t n,
                    double out[4]) {
    const int nblocks = (int)std::max<size_t>(
        1, std::min<size_t>(1024, (n + pv::kReduceBlock - 1) / pv::kReduceBlock));
    ensure_reduction(h, (size_t)nblocks * 8);
    pv::launch_moments(x, y, g, (int)n, h->red, nblocks, h->stream);
    hip_check(hipGetLastError(), "launch HIP moments kernel");
    std::vector<double> part((size_t)nblocks * 8);
    hip_check(hipMemcpyAsync(part.data(), h->red, part.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream),
              "read HIP moments");
    hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    for (int q = 0; q < 4; ++q) {
        double s = 0.0, c = 0.0;  // Neumaier over block partials, fixed order
        for (int b = 0; b < nblocks; ++b)
            for (int k = 0; k < 2; ++k) {
                const double xq = part[(size_t)b * 8 + 2 * q + k];
                const double t = s + xq;
                c += std::fabs(s) >= std::fabs(xq) ? (s - t) + xq : (xq - t) + s;
                s = t;
            }
        out[q] = s + c;
    }
    for (int q = 0; q < 4; ++q)
        if (!std::isfinite(out[q]))
            raise(PV_ERR_RUNTIME, "non-finite vortex moment; check input scales");
}

// Hamiltonian of device arrays (src/compute.cpp:85-87, 142-151, 174-202, 262-285).
double device_hamiltonian(pv_handle *h, const double *x, const double *y, const double *g,
                          size_t n) {
    if (n == 0)
        return 0.0;
    const int nblocks = (int)((n + pv::kReduceBlock - 1) / pv::kReduceBlock);
    ensure_reduction(h, (size_t)nblocks * 2);
    begin_sequence(h);
    pv::launch_energy(h->gp, x, y, g, (int)n, h->red, nblocks, h->st, h->stream);
    hip_check(hipGetLastError(), "launch HIP Hamiltonian kernel");
    std::vector<double> part((size_t)nblocks * 2);
    hip_check(hipMemcpyAsync(part.data(), h->red, part.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream),
              "read HIP Hamiltonian");
    read_status(h);
    const unsigned flags = h->st_host->flags;
    if (flags & pv::kFailOutsideDisk)
        raise(PV_ERR_INVALID_ARGUMENT, "vortex outside disk");
    if (flags & pv::kFailCoincident)
        raise(PV_ERR_RUNTIME, h->cfg.geometry == PV_INFINITE   ? "coincident vortices"
                              : h->cfg.geometry == PV_PERIODIC ? "coincident periodic vortices"
                                                               : "coincident vortices in disk");
    double s = 0.0, c = 0.0;
    for (double xq : part) {
        const double t = s + xq;
        c += std::fabs(s) >= std::fabs(xq) ? (s - t) + xq : (xq - t) + s;
        s = t;
    }
    return s + c;
}

void set_geometry(pv_handle *h) { h->gp = geo_params(h->cfg); }

pv_handle *checked(pv_handle *h) {
    if (!h)
        raise(PV_ERR_INVALID_ARGUMENT, "null handle");
    return h;
}

} // namespace

extern "C" {

const char *pv_last_error(void) { return pv::host::g_error.c_str(); }

pv_status pv_create(const pv_config *config, pv_handle **out) {
    return guard([&] {
        if (!config || !out)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        check_config(*config);
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "count HIP devices");
        if (config->device < 0 || config->device >= count)
            raise(PV_ERR_INVALID_ARGUMENT, "invalid HIP device ordinal");
        DeviceGuard dg(config->device);
        auto h = std::make_unique<pv_handle>();
        h->cfg = *config;
        set_geometry(h.get());
        hipDeviceProp_t prop{};
        hip_check(hipGetDeviceProperties(&prop, config->device), "query HIP device properties");
        h->cu_count = std::max(1, prop.multiProcessorCount);
        int coop = 0, wall = 0;
        if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, config->device) ==
            hipSuccess)
            h->coop = coop != 0;
        if (hipDeviceGetAttribute(&wall, hipDeviceAttributeWallClockRate, config->device) ==
                hipSuccess &&
            wall > 0)
            h->wall_khz = wall;
        hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "create HIP stream");
        h->st = dalloc<pv::DevStatus>(1, "hipMalloc(status)");
        hip_check(hipHostMalloc((void **)&h->st_host, sizeof(pv::DevStatus), hipHostMallocDefault),
                  "hipHostMalloc(status)");
        std::memset(h->st_host, 0, sizeof(pv::DevStatus));
        h->table = default_launch_table();
        begin_sequence(h.get());
        hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        *out = h.release();
    });
}

void pv_destroy(pv_handle *h) {
    if (!h)
        return;
    try {
        DeviceGuard dg(h->cfg.device);
        delete h;
    } catch (...) {
        delete h;
    }
}

pv_status pv_set_precision(pv_handle *h, int precision) {
    return guard([&] {
        checked(h);
        if (precision != PV_FP64 && precision != PV_MIXED)
            raise(PV_ERR_INVALID_ARGUMENT, "unsupported precision");
        h->cfg.precision = precision;
    });
}

pv_status pv_set_launch(pv_handle *h, int targets_per_thread, int source_splits) {
    return guard([&] {
        checked(h);
        if (targets_per_thread != 0 && targets_per_thread != 1 && targets_per_thread != 2 &&
            targets_per_thread != 4 && targets_per_thread != 8)
            raise(PV_ERR_INVALID_ARGUMENT, "targets per thread must be 0, 1, 2, 4 or 8");
        if (source_splits < 0)
            raise(PV_ERR_INVALID_ARGUMENT, "source splits must be non-negative");
        h->tpt_override = targets_per_thread;
        h->split_override = source_splits;
        h->tile_override = 0;
        h->sym_override = -1;
    });
}

pv_status pv_set_launch_shape(pv_handle *h, int targets_per_thread, int source_splits,
                              int tile, int symmetric) {
    return guard([&] {
        checked(h);
        if (symmetric < -1 || symmetric > 1)
            raise(PV_ERR_INVALID_ARGUMENT, "symmetric must be -1, 0 or 1");
        const pv_launch_rule r{h->cfg.geometry, h->cfg.precision, 0, targets_per_thread,
                               source_splits, tile, symmetric < 0 ? 0 : symmetric};
        check_rules(&r, 1);
        h->tpt_override = targets_per_thread;
        h->split_override = source_splits;
        h->tile_override = tile;
        h->sym_override = symmetric;
    });
}

static int persistent_blocks(pv_handle *h, size_t n);

pv_status pv_set_persistent(pv_handle *h, int mode, int targets_per_block,
                            double max_wait_seconds) {
    return guard([&] {
        checked(h);
        if (mode < -1 || mode > 1)
            raise(PV_ERR_INVALID_ARGUMENT, "persistent mode must be -1, 0 or 1");
        if (targets_per_block != 0 && targets_per_block != 8 && targets_per_block != 16 &&
            targets_per_block != 32)
            raise(PV_ERR_INVALID_ARGUMENT, "targets per block must be 0, 8, 16 or 32");
        if (!std::isfinite(max_wait_seconds) || max_wait_seconds < 0.0)
            raise(PV_ERR_INVALID_ARGUMENT, "barrier wait must be finite and non-negative");
        h->persist_mode = mode;
        h->persist_tb = targets_per_block;
        h->persist_wait_s = max_wait_seconds > 0.0 ? max_wait_seconds : 2.0;
    });
}

pv_status pv_persistent_blocks(pv_handle *h, size_t n, int *blocks) {
    return guard([&] {
        checked(h);
        if (!blocks)
            raise(PV_ERR_INVALID_ARGUMENT, "null output");
        DeviceGuard dg(h->cfg.device);
        *blocks = persistent_blocks(h, n);
    });
}

pv_status pv_set_launch_table(pv_handle *h, const pv_launch_rule *rules, size_t count) {
    return guard([&] {
        checked(h);
        if (count && !rules)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        check_rules(rules, count);
        h->table.assign(rules, rules + count);
    });
}

pv_status pv_resolve_launch(pv_handle *h, size_t targets, size_t n, int *targets_per_thread,
                            int *source_splits, int *tile, int *symmetric) {
    return guard([&] {
        checked(h);
        if (!targets_per_thread || !source_splits || !tile || !symmetric)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        const Shape s = resolve_shape(h, targets, n);
        *targets_per_thread = s.tpt;
        *source_splits = s.sym ? 0 : (int)s.nsplit;
        *tile = s.tile;
        *symmetric = s.sym ? 1 : 0;
    });
}

static void evaluate_host(pv_handle *h, const double *x, const double *y, const double *g,
                          size_t n, double *u, double *v, size_t begin, size_t end,
                          bool must_be_finite) {
    checked(h);
    check_population(n);
    validate_host_arrays(x, y, g, n);
    if (begin > end || end > n)
        raise(PV_ERR_OUT_OF_RANGE, "invalid target-vortex range");
    validate_geometry(h, x, y, g, n);
    if (begin == end)
        return;
    if (!u || !v)
        raise(PV_ERR_INVALID_ARGUMENT, "null velocity output");
    DeviceGuard dg(h->cfg.device);
    ensure_scratch(h, n);
    begin_sequence(h);
    upload_to(h, h->sx, h->sy, h->sg, x, y, g, n);
    evaluate(h, h->sx, h->sy, h->sg, n, begin, end, store_op(h->su, h->sv));
    const size_t bytes = (end - begin) * sizeof(double);
    hip_check(hipMemcpyAsync(u + begin, h->su + begin, bytes, hipMemcpyDeviceToHost, h->stream),
              "copy u from GPU");
    hip_check(hipMemcpyAsync(v + begin, h->sv + begin, bytes, hipMemcpyDeviceToHost, h->stream),
              "copy v from GPU");
    finish(h, false);
    if (must_be_finite)
        for (size_t i = begin; i < end; ++i)
            if (!std::isfinite(u[i]) || !std::isfinite(v[i]))
                raise(PV_ERR_RUNTIME, "non-finite velocity; check scales and close encounters");
}

pv_status pv_evaluate_range(pv_handle *h, const double *x, const double *y, const double *gamma,
                            size_t n, double *u, double *v, size_t begin, size_t end) {
    return guard([&] { evaluate_host(h, x, y, gamma, n, u, v, begin, end, false); });
}

pv_status pv_evaluate(pv_handle *h, const double *x, const double *y, const double *gamma,
                      size_t n, double *u, double *v) {
    return guard([&] { evaluate_host(h, x, y, gamma, n, u, v, 0, n, true); });
}

pv_status pv_hamiltonian(pv_handle *h, const double *x, const double *y, const double *gamma,
                         size_t n, double *out) {
    return guard([&] {
        checked(h);
        check_population(n);
        if (!out)
            raise(PV_ERR_INVALID_ARGUMENT, "null output");
        validate_host_arrays(x, y, gamma, n);
        DeviceGuard dg(h->cfg.device);
        if (n == 0) {
            *out = 0.0;
            return;
        }
        ensure_scratch(h, n);
        upload_to(h, h->sx, h->sy, h->sg, x, y, gamma, n);
        *out = device_hamiltonian(h, h->sx, h->sy, h->sg, n);
    });
}

pv_status pv_invariants_of(pv_handle *h, const double *x, const double *y, const double *gamma,
                           size_t n, pv_invariants *out) {
    return guard([&] {
        checked(h);
        check_population(n);
        if (!out)
            raise(PV_ERR_INVALID_ARGUMENT, "null output");
        validate_host_arrays(x, y, gamma, n);
        DeviceGuard dg(h->cfg.device);
        double m[4] = {0, 0, 0, 0};
        double ham = 0.0;
        if (n) {
            ensure_scratch(h, n);
            upload_to(h, h->sx, h->sy, h->sg, x, y, gamma, n);
            device_moments(h, h->sx, h->sy, h->sg, n, m);
            ham = device_hamiltonian(h, h->sx, h->sy, h->sg, n);
        }
        if (!std::isfinite(ham))
            raise(PV_ERR_RUNTIME, "non-finite Hamiltonian; check input scales");
        *out = {m[0], m[1], m[2], m[3], ham};
    });
}

pv_status pv_upload(pv_handle *h, const double *x, const double *y, const double *gamma,
                    size_t n) {
    return guard([&] {
        checked(h);
        check_population(n);
        validate_host_arrays(x, y, gamma, n);
        validate_geometry(h, x, y, gamma, n);
        DeviceGuard dg(h->cfg.device);
        h->fsal_valid = false;
        if (n == 0) {
            h->n = 0;
            h->state_valid = true;
            return;
        }
        ensure_state(h, n);
        h->n = n;
        h->state_valid = true;
        upload_to(h, h->x, h->y, h->g, x, y, gamma, n);
        hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
    });
}

pv_status pv_download(pv_handle *h, double *x, double *y, double *gamma, size_t n) {
    return guard([&] {
        checked(h);
        require_state(h);
        if (n != h->n)
            raise(PV_ERR_RUNTIME, "host and HIP vortex populations differ");
        if (n == 0)
            return;
        DeviceGuard dg(h->cfg.device);
        const size_t bytes = n * sizeof(double);
        hip_check(hipMemcpyAsync(x, h->x, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download state x from GPU");
        hip_check(hipMemcpyAsync(y, h->y, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download state y from GPU");
        if (gamma)
            hip_check(hipMemcpyAsync(gamma, h->g, bytes, hipMemcpyDeviceToHost, h->stream),
                      "download circulation from GPU");
        hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(x[i]) || !std::isfinite(y[i]))
                raise(PV_ERR_INVALID_ARGUMENT, "vortex data must be finite");
    });
}

pv_status pv_state_size(pv_handle *h, size_t *n) {
    return guard([&] {
        checked(h);
        require_state(h);
        *n = h->n;
    });
}

pv_status pv_evaluate_device_state(pv_handle *h, double *u, double *v, size_t n) {
    return guard([&] {
        checked(h);
        require_state(h);
        if (n != h->n)
            raise(PV_ERR_RUNTIME, "host and HIP vortex populations differ");
        if (h->n == 0)
            return;
        DeviceGuard dg(h->cfg.device);
        begin_sequence(h);
        evaluate(h, h->x, h->y, h->g, h->n, 0, h->n, store_op(h->u, h->v));
        const size_t bytes = h->n * sizeof(double);
        hip_check(hipMemcpyAsync(u, h->u, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download velocity x from GPU");
        hip_check(hipMemcpyAsync(v, h->v, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download velocity y from GPU");
        finish(h, true);
    });
}

// Enqueues one RK4 step reading (x0, y0) and writing (xn, yn) (src/timestep.cpp:31-61).
// The four stage evaluations of one RK4 step (src/timestep.cpp:36-58): the positions each
// evaluates and its fused epilogue.
struct Rk4Stage {
    const double *x, *y;
    pv::StageOp op;
};

static void rk4_stages(pv_handle *h, double dt, const double *x0, const double *y0, double *xn,
                       double *yn, Rk4Stage st[4]) {
    pv::StageOp op{};
    op.dt = dt;
    op.x0 = x0, op.y0 = y0;
    for (int s = 0; s < 7; ++s)
        op.kx[s] = h->kx[s], op.ky[s] = h->ky[s];
    // k0 at x0 -> stage 1 position x0 + dt (1/2 k0)
    op.kind = pv::kOpStage, op.cur = 0;
    op.coef[0] = 0.5;
    op.ku = h->kx[0], op.kv = h->ky[0], op.xn = h->xa, op.yn = h->ya;
    st[0] = {x0, y0, op};
    // k1 -> stage 2 position x0 + dt (0 k0 + 1/2 k1)
    op.cur = 1, op.coef[0] = 0.0, op.coef[1] = 0.5;
    op.ku = h->kx[1], op.kv = h->ky[1], op.xn = h->xb, op.yn = h->yb;
    st[1] = {h->xa, h->ya, op};
    // k2 -> stage 3 position x0 + dt (0 k0 + 0 k1 + 1 k2)
    op.cur = 2, op.coef[1] = 0.0, op.coef[2] = 1.0;
    op.ku = h->kx[2], op.kv = h->ky[2], op.xn = h->xa, op.yn = h->ya;
    st[2] = {h->xb, h->yb, op};
    // k3 -> x0 + dt (k0 + 2 k1 + 2 k2 + k3) / 6
    op.kind = pv::kOpRk4Final, op.cur = 3;
    op.ku = h->kx[3], op.kv = h->ky[3], op.xn = xn, op.yn = yn;
    st[3] = {h->xa, h->ya, op};
}

static void enqueue_rk4(pv_handle *h, double dt, const double *x0, const double *y0, double *xn,
                        double *yn) {
    Rk4Stage st[4];
    rk4_stages(h, dt, x0, y0, xn, yn, st);
    for (const Rk4Stage &s : st)
        evaluate(h, s.x, s.y, h->g, h->n, 0, h->n, s.op);
}

// Persistent RK4 path: beyond this population the per-stage launches run (automatic mode),
// and nothing larger is ever launched persistently.
constexpr size_t kPersistAutoN = 4096;
constexpr size_t kPersistMaxN = 65536;
constexpr int kPersistTb = 16;

// Blocks of the persistent RK4 launch for n vortices, or 0 when the per-stage launches run:
// automatic mode takes it for n <= kPersistAutoN unless a launch shape was forced; both modes
// need cooperative launches and every block resident at once.
static int persistent_blocks(pv_handle *h, size_t n) {
    if (h->persist_mode == 0 || !h->coop || n == 0 || n > kPersistMaxN)
        return 0;
    if (h->persist_mode < 0 &&
        (n > kPersistAutoN || h->tpt_override || h->split_override || h->tile_override ||
         h->sym_override >= 0))
        return 0;
    const int tb = h->persist_tb ? h->persist_tb : kPersistTb;
    const int blocks = (int)((n + tb - 1) / tb);
    int &cap = h->persist_cap[h->cfg.precision][h->cfg.geometry];
    if (cap < 0)
        cap = pv::persistent_capacity(h->cfg.geometry, h->cfg.precision, h->cu_count);
    return blocks <= cap ? blocks : 0;
}

// `steps` RK4 steps in one cooperative launch; step s reads (bx[s & 1], by[s & 1]) and writes
// the other pair, as the per-stage launches of pv_rk4_advance do, and every stage keeps the
// epoch and evaluation record it would have there.  Returns false, with nothing enqueued and
// the handle unchanged, when the runtime refuses the cooperative launch (only one process can
// hold the device's cooperative resources at a time); the handle then keeps to the per-stage
// launches.
static bool enqueue_rk4_persistent(pv_handle *h, double dt, int steps, int blocks, double *bx[2],
                                   double *by[2]) {
    const size_t n = h->n;
    pv::RkPersistLaunch a{};
    Rk4Stage st[2][4];
    for (int p = 0; p < 2; ++p) {
        rk4_stages(h, dt, bx[p], by[p], bx[1 - p], by[1 - p], st[p]);
        a.x[p] = bx[p], a.y[p] = by[p];
        for (int c = 0; c < 4; ++c)
            a.ops[4 * p + c] = st[p][c].op;
    }
    a.g = h->g;
    a.xa = h->xa, a.ya = h->ya, a.xb = h->xb, a.yb = h->yb;
    a.n = (int)n;
    a.steps = steps;
    a.tb = h->persist_tb ? h->persist_tb : kPersistTb;
    a.st = h->st;
    a.epoch0 = h->epoch + 1;
    a.barrier_base = h->barrier_arrivals;
    a.wait_limit = (unsigned long long)(h->persist_wait_s * h->wall_khz * 1e3);
    const size_t events_used = h->events_used;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    record_events(h, ev0, ev1);
    if (ev0)
        hip_check(hipEventRecord(ev0, h->stream), "record HIP event");
    const hipError_t e = pv::launch_rk4_persistent(h->cfg.geometry, h->cfg.precision, a, h->gp,
                                                   blocks, h->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        h->coop = false;
        h->events_used = events_used;
        return false;
    }
    if (ev1)
        hip_check(hipEventRecord(ev1, h->stream), "record HIP event");
    h->epoch += 4u * (unsigned)steps;
    h->barrier_arrivals += (unsigned long long)blocks * (4ull * (unsigned)steps - 1ull);
    h->records.reserve(h->records.size() + 4 * (size_t)steps);
    for (int s = 0; s < steps; ++s)
        for (int c = 0; c < 4; ++c)
            h->records.push_back({st[s & 1][c].x, st[s & 1][c].y, (int)n, 0, (int)n});
    return true;
}

pv_status pv_rk4_step(pv_handle *h, double dt) {
    return guard([&] {
        checked(h);
        require_dt(dt);
        require_state(h);
        if (h->n == 0)
            return;
        DeviceGuard dg(h->cfg.device);
        h->fsal_valid = false;
        begin_sequence(h);
        double *bx[2] = {h->x, h->xalt}, *by[2] = {h->y, h->yalt};
        const int blocks = persistent_blocks(h, h->n);
        if (!blocks || !enqueue_rk4_persistent(h, dt, 1, blocks, bx, by))
            enqueue_rk4(h, dt, h->x, h->y, h->xalt, h->yalt);
        finish(h, true);
        std::swap(h->x, h->xalt);
        std::swap(h->y, h->yalt);
    });
}

// RK4 steps per synchronised sequence of pv_rk4_advance: four epochs (and evaluation
// records) per step, far below the 2^32 epoch range.
constexpr int kMaxSequenceSteps = 1 << 16;

pv_status pv_rk4_advance(pv_handle *h, double dt, int steps) {
    return guard([&] {
        checked(h);
        require_dt(dt);
        require_state(h);
        if (steps < 0)
            raise(PV_ERR_INVALID_ARGUMENT, "step count must be non-negative");
        if (h->n == 0 || steps == 0)
            return;
        DeviceGuard dg(h->cfg.device);
        h->fsal_valid = false;
        for (int done = 0; done < steps;) {
            const int count = std::min(steps - done, kMaxSequenceSteps);
            begin_sequence(h);
            double *bx[2] = {h->x, h->xalt}, *by[2] = {h->y, h->yalt};
            const int blocks = persistent_blocks(h, h->n);
            if (!blocks || !enqueue_rk4_persistent(h, dt, count, blocks, bx, by)) {
                for (int s = 0; s < count; ++s)
                    enqueue_rk4(h, dt, bx[s & 1], by[s & 1], bx[(s + 1) & 1], by[(s + 1) & 1]);
            }
            try {
                finish(h, true);
            } catch (ApiError &e) {
                // Keep the state of the last good step: the input of the failing one.
                const unsigned e0 = h->st_host->fail_epoch;
                const int failed = (int)((e0 - h->epoch_base) / 4);
                h->x = bx[failed & 1], h->xalt = bx[(failed + 1) & 1];
                h->y = by[failed & 1], h->yalt = by[(failed + 1) & 1];
                e.message += " (RK4 step " + std::to_string(done + failed + 1) + " of " +
                             std::to_string(steps) + ")";
                throw;
            }
            h->x = bx[count & 1], h->xalt = bx[(count + 1) & 1];
            h->y = by[count & 1], h->yalt = by[(count + 1) & 1];
            done += count;
        }
    });
}

pv_status pv_dopri5_step(pv_handle *h, double dt, double atol, double rtol, double min_dt,
                         double max_dt, pv_step_result *out) {
    return guard([&] {
        checked(h);
        require_dt(dt);
        require_state(h);
        if (!out)
            raise(PV_ERR_INVALID_ARGUMENT, "null step result");
        const size_t n = h->n;
        unsigned rejected = 0;
        DeviceGuard dg(h->cfg.device);
        if (n == 0) {
            // An empty population has zero error: src/timestep.cpp:104-114.
            const double suggested = std::clamp(dt * 5.0, min_dt, max_dt);
            h->fsal_valid = true;
            *out = {dt, suggested, 0.0, 0u};
            return;
        }
        // src/timestep.cpp:75-77: FSAL derivative at the current state.
        if (!h->fsal_valid) {
            begin_sequence(h);
            evaluate(h, h->x, h->y, h->g, n, 0, n, store_op(h->kx[0], h->ky[0]));
            finish(h, true);
        }
        for (;;) {
            begin_sequence(h);
            pv::StageOp op{};
            op.dt = dt;
            op.x0 = h->x, op.y0 = h->y;
            op.atol = atol, op.rtol = rtol;
            for (int s = 0; s < 7; ++s) {
                op.kx[s] = h->kx[s], op.ky[s] = h->ky[s];
                op.w[s] = kB5[s] - kB4[s];
            }
            // Stage positions alternate between the two stage buffers; stage 6 lands in B,
            // which becomes the state on acceptance.
            double *px[7] = {nullptr, h->xa, h->xb, h->xa, h->xb, h->xa, h->xb};
            double *py[7] = {nullptr, h->ya, h->yb, h->ya, h->yb, h->ya, h->yb};
            op.kind = pv::kOpStage, op.cur = 0;
            op.coef[0] = kDopriA[1][0];
            op.xn = px[1], op.yn = py[1];
            pv::launch_make_stage(op, (int)n, h->st, h->epoch, h->stream);
            hip_check(hipGetLastError(), "launch HIP Runge--Kutta stage");
            for (int s = 1; s < 7; ++s) {
                op.cur = s;
                op.ku = h->kx[s], op.kv = h->ky[s];
                if (s < 6) {
                    op.kind = pv::kOpStage;
                    for (int j = 0; j <= s; ++j)
                        op.coef[j] = kDopriA[s + 1][j];
                    op.xn = px[s + 1], op.yn = py[s + 1];
                } else {
                    op.kind = pv::kOpDopriError;
                    op.xn = nullptr, op.yn = nullptr;
                }
                evaluate(h, px[s], py[s], h->g, n, 0, n, op);
            }
            finish(h, true);
            double error;
            std::memcpy(&error, &h->st_host->err_bits, sizeof error);
            // src/timestep.cpp:104-118 (controller and acceptance).
            const double factor =
                error == 0.0 ? 5.0 : std::clamp(0.9 * std::pow(error, -0.2), 0.2, 5.0);
            const double suggested = std::clamp(dt * factor, min_dt, max_dt);
            if (error <= 1.0) {
                std::swap(h->x, h->xb);
                std::swap(h->y, h->yb);
                std::swap(h->kx[0], h->kx[6]);
                std::swap(h->ky[0], h->ky[6]);
                h->fsal_valid = true;
                *out = {dt, suggested, error, rejected};
                return;
            }
            if (dt <= min_dt || ++rejected > 32)
                raise(PV_ERR_RUNTIME, "adaptive integrator could not satisfy tolerance");
            dt = std::max(min_dt, std::min(suggested, dt * 0.9));
        }
    });
}

void pv_invalidate_fsal(pv_handle *h) {
    if (h)
        h->fsal_valid = false;
}

pv_status pv_state_invariants(pv_handle *h, pv_invariants *out) {
    return guard([&] {
        checked(h);
        require_state(h);
        if (!out)
            raise(PV_ERR_INVALID_ARGUMENT, "null output");
        DeviceGuard dg(h->cfg.device);
        double m[4] = {0, 0, 0, 0};
        double ham = 0.0;
        if (h->n) {
            device_moments(h, h->x, h->y, h->g, h->n, m);
            ham = device_hamiltonian(h, h->x, h->y, h->g, h->n);
        }
        if (!std::isfinite(ham))
            raise(PV_ERR_RUNTIME, "non-finite Hamiltonian; check input scales");
        *out = {m[0], m[1], m[2], m[3], ham};
    });
}

pv_status pv_dipole_candidates(pv_handle *h, double distance, uint64_t *first, uint64_t *second,
                               double *distance_squared, size_t cap, size_t *count) {
    return guard([&] {
        checked(h);
        require_state(h);
        if (!count)
            raise(PV_ERR_INVALID_ARGUMENT, "null count");
        *count = 0;
        if (h->n == 0)
            return;
        DeviceGuard dg(h->cfg.device);
        const size_t dcap = std::max<size_t>(cap, 1);
        ensure_candidates(h, dcap);
        begin_sequence(h);
        const double thr2 = distance * distance;
        pv::launch_dipoles(h->gp, h->cfg.geometry == PV_PERIODIC, h->cfg.geometry == PV_DISK,
                           thr2, h->x, h->y, h->g, (int)h->n, h->cand_first, h->cand_second,
                           h->cand_d2, (unsigned long long)cap, h->st, h->stream);
        hip_check(hipGetLastError(), "launch HIP dipole search");
        read_status(h);
        const size_t found = (size_t)h->st_host->count;
        const size_t m = std::min(found, cap);
        if (m) {
            if (!first || !second || !distance_squared)
                raise(PV_ERR_INVALID_ARGUMENT, "null candidate output");
            hip_check(hipMemcpyAsync(first, h->cand_first, m * sizeof(uint64_t),
                                     hipMemcpyDeviceToHost, h->stream),
                      "read dipole candidates");
            hip_check(hipMemcpyAsync(second, h->cand_second, m * sizeof(uint64_t),
                                     hipMemcpyDeviceToHost, h->stream),
                      "read dipole candidates");
            hip_check(hipMemcpyAsync(distance_squared, h->cand_d2, m * sizeof(double),
                                     hipMemcpyDeviceToHost, h->stream),
                      "read dipole candidates");
            hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        }
        *count = found;
    });
}

// ---------------------------------------------------------------- dipole events
//
// DipoleManager::process (src/dipole.cpp:42-125) on the resident state: the candidate search
// on the device, then on the host the order, the closest-first disjoint selection, the
// compaction and the reinjection, with the handle's std::mt19937_64 drawn in the reference's
// order (uniform_real_distribution<double>(0, 1), one call per draw), so removals and
// reinjected positions are those of the reference for the same generator state.

namespace {

struct DipoleCandidate {
    uint64_t first, second;
    double d2;
};

// pi rounded to double: the same value as C++20's std::numbers::pi, which src/dipole.cpp uses.
constexpr double kPi = 3.141592653589793238462643383279502884;

// All candidates of the resident state, unordered (the device search appends them).
std::vector<DipoleCandidate> device_candidates(pv_handle *h, double distance) {
    size_t cap = std::max<size_t>(h->cand_cap, 1024);
    for (;;) {
        ensure_candidates(h, cap);
        begin_sequence(h);
        pv::launch_dipoles(h->gp, h->cfg.geometry == PV_PERIODIC, h->cfg.geometry == PV_DISK,
                           distance * distance, h->x, h->y, h->g, (int)h->n, h->cand_first,
                           h->cand_second, h->cand_d2, (unsigned long long)cap, h->st,
                           h->stream);
        hip_check(hipGetLastError(), "launch HIP dipole search");
        read_status(h);
        const size_t found = (size_t)h->st_host->count;
        if (found > cap) {
            cap = found;
            continue;
        }
        std::vector<uint64_t> a(found), b(found);
        std::vector<double> d(found);
        if (found) {
            hip_check(hipMemcpyAsync(a.data(), h->cand_first, found * sizeof(uint64_t),
                                     hipMemcpyDeviceToHost, h->stream),
                      "read dipole candidates");
            hip_check(hipMemcpyAsync(b.data(), h->cand_second, found * sizeof(uint64_t),
                                     hipMemcpyDeviceToHost, h->stream),
                      "read dipole candidates");
            hip_check(hipMemcpyAsync(d.data(), h->cand_d2, found * sizeof(double),
                                     hipMemcpyDeviceToHost, h->stream),
                      "read dipole candidates");
            hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        }
        std::vector<DipoleCandidate> out(found);
        for (size_t k = 0; k < found; ++k)
            out[k] = {a[k], b[k], d[k]};
        return out;
    }
}

struct HostState {
    std::vector<double> x, y, g;
    void push(double a, double b, double c) {
        x.push_back(a);
        y.push_back(b);
        g.push_back(c);
    }
};

double unit_draw(std::mt19937_64 &rng) {
    std::uniform_real_distribution<double> unit(0.0, 1.0);
    return unit(rng);
}

// src/dipole.cpp:127-133: a single vortex uniformly in the disk.
void inject_single(const pv_config &c, std::mt19937_64 &rng, HostState &s, double g) {
    const double radius = c.disk_radius * std::sqrt(unit_draw(rng));
    const double angle = 2.0 * kPi * unit_draw(rng);
    s.push(radius * std::cos(angle), radius * std::sin(angle), g);
}

// src/dipole.cpp:135-176: a pair at independent uniform positions, or (paired) a first
// position and its partner at the mean spacing in a uniform direction.
void inject_pair(const pv_config &c, int mode, std::mt19937_64 &rng, HostState &s, double g1,
                 double g2, size_t population) {
    const bool periodic = c.geometry == PV_PERIODIC;
    const double area = periodic ? c.box_length_x * c.box_length_y
                                 : kPi * c.disk_radius * c.disk_radius;
    const double spacing = std::sqrt(area / (double)population);
    auto position = [&]() {
        if (periodic) {
            const double px = (unit_draw(rng) - 0.5) * c.box_length_x;
            const double py = (unit_draw(rng) - 0.5) * c.box_length_y;
            return std::pair{px, py};
        }
        const double radius = c.disk_radius * std::sqrt(unit_draw(rng));
        const double angle = 2.0 * kPi * unit_draw(rng);
        return std::pair{radius * std::cos(angle), radius * std::sin(angle)};
    };
    auto first = position();
    auto second = position();
    if (mode == PV_REINJECT_PAIRED) {
        bool placed = false;
        for (size_t attempt = 0; attempt < 100000; ++attempt) {
            first = position();
            const double angle = 2.0 * kPi * unit_draw(rng);
            second = {first.first + spacing * std::cos(angle),
                      first.second + spacing * std::sin(angle)};
            if (periodic) {
                second.first = std::remainder(second.first, c.box_length_x);
                second.second = std::remainder(second.second, c.box_length_y);
                placed = true;
                break;
            }
            if (second.first * second.first + second.second * second.second <
                c.disk_radius * c.disk_radius) {
                placed = true;
                break;
            }
        }
        if (!placed)
            raise(PV_ERR_RUNTIME, "could not find a valid paired reinjection position");
    }
    s.push(first.first, first.second, g1);
    s.push(second.first, second.second, g2);
}

} // namespace

pv_status pv_dipole_set_random(pv_handle *h, uint64_t seed, const char *state) {
    return guard([&] {
        checked(h);
        std::mt19937_64 rng(seed ^ 0xd1b54a32d192ed03ULL);
        if (state && *state) {
            std::istringstream in(state);
            if (!(in >> rng))
                raise(PV_ERR_RUNTIME, "invalid random-generator state in checkpoint");
        }
        h->dipole_rng = rng;
    });
}

pv_status pv_dipole_random_state(pv_handle *h, char *buffer, size_t length, size_t *needed) {
    return guard([&] {
        checked(h);
        std::ostringstream out;
        out << h->dipole_rng;
        const std::string text = out.str();
        if (needed)
            *needed = text.size() + 1;
        if (buffer && length) {
            std::strncpy(buffer, text.c_str(), length - 1);
            buffer[length - 1] = '\0';
        }
    });
}

pv_status pv_dipole_process(pv_handle *h, double distance, int reinjection, size_t *events) {
    return guard([&] {
        checked(h);
        require_state(h);
        if (!events)
            raise(PV_ERR_INVALID_ARGUMENT, "null event count");
        if (!std::isfinite(distance) || distance < 0.0)
            raise(PV_ERR_INVALID_ARGUMENT, "dipole removal distance must be finite and non-negative");
        if (reinjection != PV_REINJECT_NONE && reinjection != PV_REINJECT_INDEPENDENT &&
            reinjection != PV_REINJECT_PAIRED)
            raise(PV_ERR_INVALID_ARGUMENT, "unknown dipole reinjection mode");
        *events = 0;
        const size_t n = h->n;
        if (n == 0)
            return;
        DeviceGuard dg(h->cfg.device);
        std::vector<DipoleCandidate> cand = device_candidates(h, distance);
        if (cand.empty())
            return;
        // src/dipole.cpp:79-84: closest first, ties by index.
        std::sort(cand.begin(), cand.end(), [](const DipoleCandidate &a, const DipoleCandidate &b) {
            if (a.d2 != b.d2)
                return a.d2 < b.d2;
            return std::pair{a.first, a.second} < std::pair{b.first, b.second};
        });
        HostState cur;
        cur.x.resize(n), cur.y.resize(n), cur.g.resize(n);
        const size_t bytes = n * sizeof(double);
        hip_check(hipMemcpyAsync(cur.x.data(), h->x, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download state x from GPU");
        hip_check(hipMemcpyAsync(cur.y.data(), h->y, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download state y from GPU");
        hip_check(hipMemcpyAsync(cur.g.data(), h->g, bytes, hipMemcpyDeviceToHost, h->stream),
                  "download circulation from GPU");
        hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        // src/dipole.cpp:86-100: disjoint selection; a wall candidate (second == n) removes
        // the real vortex only.
        struct Event {
            double g1, g2;
            bool wall;
        };
        std::vector<char> selected(n, 0);
        std::vector<Event> ev;
        for (const DipoleCandidate &c : cand) {
            if (selected[c.first] || (c.second < n && selected[c.second]))
                continue;
            selected[c.first] = 1;
            if (c.second < n) {
                selected[c.second] = 1;
                ev.push_back({cur.g[c.first], cur.g[c.second], false});
            } else {
                ev.push_back({cur.g[c.first], 0.0, true});
            }
        }
        // src/dipole.cpp:104-111: survivors in index order.
        HostState next;
        for (size_t i = 0; i < n; ++i)
            if (!selected[i])
                next.push(cur.x[i], cur.y[i], cur.g[i]);
        // src/dipole.cpp:114-122.
        if (reinjection != PV_REINJECT_NONE)
            for (const Event &e : ev) {
                if (e.wall)
                    inject_single(h->cfg, h->dipole_rng, next, e.g1);
                else
                    inject_pair(h->cfg, reinjection, h->dipole_rng, next, e.g1, e.g2, n);
            }
        const size_t m = next.x.size();
        h->fsal_valid = false;
        h->n = m;
        if (m) {
            ensure_state(h, m);
            upload_to(h, h->x, h->y, h->g, next.x.data(), next.y.data(), next.g.data(), m);
            hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        }
        h->state_valid = true;
        *events = ev.size();
    });
}

pv_status pv_runtime_details(int device, char *buffer, size_t length) {
    return guard([&] {
        hipDeviceProp_t prop{};
        hip_check(hipGetDeviceProperties(&prop, device), "query HIP device properties");
        int runtime = 0;
        hip_check(hipRuntimeGetVersion(&runtime), "query HIP runtime version");
        const std::string s = "hip_device " + std::to_string(device) + "\nhip_device_name \"" +
                              std::string(prop.name) + "\"\nhip_arch " +
                              std::string(prop.gcnArchName) + "\nhip_compute_units " +
                              std::to_string(prop.multiProcessorCount) + "\nhip_runtime " +
                              std::to_string(runtime);
        if (buffer && length) {
            std::strncpy(buffer, s.c_str(), length - 1);
            buffer[length - 1] = '\0';
        }
    });
}

pv_status pv_evaluate_resident(pv_handle *h, int repeats) {
    return guard([&] {
        checked(h);
        require_state(h);
        if (repeats > kMaxSequenceSteps)
            raise(PV_ERR_INVALID_ARGUMENT, "too many repeats in one resident evaluation");
        if (h->n == 0 || repeats <= 0)
            return;
        DeviceGuard dg(h->cfg.device);
        begin_sequence(h);
        for (int r = 0; r < repeats; ++r)
            evaluate(h, h->x, h->y, h->g, h->n, 0, h->n, store_op(h->u, h->v));
    });
}

pv_status pv_synchronize(pv_handle *h) {
    return guard([&] {
        checked(h);
        DeviceGuard dg(h->cfg.device);
        finish(h, true);
    });
}

pv_status pv_profile_begin(pv_handle *h) {
    return guard([&] {
        checked(h);
        h->profiling = true;
        h->events_used = 0;
    });
}

pv_status pv_profile_end(pv_handle *h, double *kernel_ms, int *launches) {
    return guard([&] {
        checked(h);
        DeviceGuard dg(h->cfg.device);
        hip_check(hipStreamSynchronize(h->stream), "synchronize HIP stream");
        double total = 0.0;
        for (size_t k = 0; k + 1 < h->events_used; k += 2) {
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, h->events[k], h->events[k + 1]),
                      "read HIP event time");
            total += ms;
        }
        if (kernel_ms)
            *kernel_ms = total;
        if (launches)
            *launches = (int)(h->events_used / 2);
        h->profiling = false;
        h->events_used = 0;
    });
}

pv_status pv_calibration_copy(int device, size_t bytes, int repeats, double *ms_per_copy) {
    return guard([&] {
        if (bytes < 32 || repeats <= 0)
            raise(PV_ERR_INVALID_ARGUMENT, "copy size must be >= 32 bytes and repeats > 0");
        DeviceGuard dg(device);
        const size_t count = bytes / 32;
        double4 *src = nullptr, *dst = nullptr;
        hipStream_t stream = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (src)
                (void)hipFree(src);
            if (dst)
                (void)hipFree(dst);
            if (e0)
                (void)hipEventDestroy(e0);
            if (e1)
                (void)hipEventDestroy(e1);
            if (stream)
                (void)hipStreamDestroy(stream);
        };
        try {
            hip_check(hipMalloc((void **)&src, count * 32), "hipMalloc(copy source)");
            hip_check(hipMalloc((void **)&dst, count * 32), "hipMalloc(copy destination)");
            hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "create HIP stream");
            hip_check(hipMemsetAsync(src, 0, count * 32, stream), "clear copy source");
            hip_check(hipEventCreate(&e0), "create HIP event");
            hip_check(hipEventCreate(&e1), "create HIP event");
            pv::launch_copy(src, dst, count, stream);
            hip_check(hipEventRecord(e0, stream), "record HIP event");
            for (int r = 0; r < repeats; ++r)
                pv::launch_copy(src, dst, count, stream);
            hip_check(hipEventRecord(e1, stream), "record HIP event");
            hip_check(hipEventSynchronize(e1), "wait for HIP event");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "read HIP event time");
            if (ms_per_copy)
                *ms_per_copy = ms / repeats;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

} // extern "C"
