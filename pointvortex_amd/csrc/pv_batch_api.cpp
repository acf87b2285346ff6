// pv_batch_api.cpp -- C-ABI of the ensemble engine (pv_batch_* in include/pv_hip.h).
//
// All instances share flat SoA device arrays (x, y, G, next positions, two stage position
// buffers, seven stage velocities, u, v).  Fixed-step ensembles are laid out back to back by
// pv_batch_load; adaptive runs (serving) take vortex ranges from a first-fit allocator and an
// instance slot each, so runs can be admitted and released while others are in flight.
#include "pv_batch.hpp"
#include "pv_hip.h"
#include "pv_host.hpp"

#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

using namespace pv::host;

namespace {

// Work counters: [0] next ordered item, [1] RK4 steps since the load, [2 + g] next chunk of
// symmetric group g.
constexpr size_t kCounterWords = 2 + pv::kSymGroups;

// Ensemble launch table text: one rule per line, "geometry precision n_max targets_per_lane
// chunk_items" (geometry infinite|periodic|disk, precision fp64|mixed); '#' starts a comment.
// A malformed table is an error, not a silent fallback.
std::vector<pv_batch_rule> read_batch_table(const std::string &path) {
    std::ifstream in(path);
    if (!in)
        raise(PV_ERR_INVALID_ARGUMENT, "cannot read ensemble launch table " + path);
    std::vector<pv_batch_rule> rules;
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t hash = line.find('#');
        if (hash != std::string::npos)
            line.resize(hash);
        std::istringstream ss(line);
        std::string geo, prec, extra;
        if (!(ss >> geo))
            continue;
        pv_batch_rule r{};
        unsigned long long n_max = 0;
        const std::string where =
            "ensemble launch table " + path + ": line " + std::to_string(lineno);
        if (!(ss >> prec >> n_max >> r.targets_per_lane >> r.chunk_items) || (ss >> extra))
            raise(PV_ERR_INVALID_ARGUMENT, where + " is malformed");
        r.geometry = geo == "infinite" ? PV_INFINITE
                     : geo == "periodic" ? PV_PERIODIC
                     : geo == "disk"     ? PV_DISK
                                         : -1;
        r.precision = prec == "fp64" ? PV_FP64 : prec == "mixed" ? PV_MIXED : -1;
        r.n_max = (size_t)n_max;
        if (r.geometry < 0 || r.precision < 0)
            raise(PV_ERR_INVALID_ARGUMENT, where + " names an unknown geometry or precision");
        rules.push_back(r);
    }
    return rules;
}

void check_batch_rules(const pv_batch_rule *rules, size_t count) {
    for (size_t k = 0; k < count; ++k) {
        const pv_batch_rule &r = rules[k];
        if (r.geometry < PV_INFINITE || r.geometry > PV_DISK ||
            (r.precision != PV_FP64 && r.precision != PV_MIXED))
            raise(PV_ERR_INVALID_ARGUMENT, "launch rule names an unknown geometry or precision");
        if (r.targets_per_lane != 0 && r.targets_per_lane != 1 && r.targets_per_lane != 2 &&
            r.targets_per_lane != 4)
            raise(PV_ERR_INVALID_ARGUMENT, "targets per lane must be 0, 1, 2 or 4");
        if (r.chunk_items < 0 || r.chunk_items > 64)
            raise(PV_ERR_INVALID_ARGUMENT, "chunk items must be between 0 and 64");
    }
}

// The default table of a new ensemble: $PV_HIP_ENSEMBLE_TABLE, else ensemble_table.txt beside
// this library when present, else none (2 targets per lane, chunks of 4).
std::vector<pv_batch_rule> default_batch_table() {
    std::string path;
    if (const char *env = std::getenv("PV_HIP_ENSEMBLE_TABLE"); env && *env) {
        path = env;
    } else {
        Dl_info info{};
        if (!dladdr((void *)&pv_batch_create, &info) || !info.dli_fname)
            return {};
        path = info.dli_fname;
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
               "/ensemble_table.txt";
        if (!std::ifstream(path))
            return {};
    }
    auto rules = read_batch_table(path);
    check_batch_rules(rules.data(), rules.size());
    return rules;
}

// src/timestep.h:8-15 and src/timestep.cpp:69-73.
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

constexpr size_t kDoublesPerVortex = 25;  // x y G, next x y, 2 stage x y, 7 k x y, u v

pv_config to_config(const pv_instance_config &c, int precision) {
    pv_config p{};
    p.geometry = c.geometry;
    p.precision = precision;
    p.core_radius = c.core_radius;
    p.box_length_x = p.box_length_y = c.box_length;
    p.image_layers = c.image_layers;
    p.disk_radius = c.disk_radius;
    return p;
}

// Relative cost of one pair by geometry, used only to order work items and cut chunks: the
// ensemble's RK4 step times of four N=65536 instances at 2 targets per lane, against the
// infinite plane's (FP64 periodic 6.3x, disk 1.75x; mixed 4.1x and 2.0x).
double pair_weight(int geometry, int precision) {
    if (geometry == PV_PERIODIC)
        return precision ? 4.1 : 6.3;
    if (geometry == PV_DISK)
        return precision ? 2.0 : 1.75;
    return 1.0;
}

void validate_instance(const pv_instance_config &c, const double *x, const double *y,
                       const double *g, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(g[i]))
            raise(PV_ERR_INVALID_ARGUMENT, "vortex data must be finite");
    if (c.geometry == PV_PERIODIC) {
        double total = 0.0, absolute = 0.0;
        for (size_t i = 0; i < n; ++i) {
            total += g[i];
            absolute += std::fabs(g[i]);
        }
        if (std::fabs(total) > 1e-12 * std::max(1.0, absolute))
            raise(PV_ERR_INVALID_ARGUMENT, "periodic box requires zero total circulation");
    } else if (c.geometry == PV_DISK) {
        const double r2 = c.disk_radius * c.disk_radius;
        for (size_t i = 0; i < n; ++i)
            if (x[i] * x[i] + y[i] * y[i] >= r2)
                raise(PV_ERR_INVALID_ARGUMENT, "vortex lies on or outside the disk");
    }
}

} // namespace

struct pv_batch {
    int device = 0;
    int precision = 0;                 // of the instances loaded or admitted next
    int cu_count = 256;
    hipStream_t stream = nullptr;
    // Symmetric groups run side by side: group g on `stream` (g = 0) or side[g - 1], forked
    // from and joined back to `stream` by events, so that one group's tail overlaps the
    // others' work (eager launches and captured graphs alike).
    static constexpr int kSides = pv::kSymGroups - 1;
    hipStream_t side[kSides] = {};
    hipEvent_t fork = nullptr, join[kSides] = {};
    size_t capacity = 0, max_inst = 0;
    // The per-vortex arrays below hold vcap <= capacity vortices; they grow as instances are
    // loaded or admitted.  byte_limit (0: none) bounds everything the ensemble allocates on
    // the device: an admission that would exceed it is refused.
    size_t vcap = 0;
    size_t byte_limit = 0;
    int item_precisions = 1;           // bit 0: FP64 instances among the work items, bit 1 mixed

    double *x = nullptr, *y = nullptr, *g = nullptr, *xn = nullptr, *yn = nullptr;
    double *xa = nullptr, *ya = nullptr, *xb = nullptr, *yb = nullptr;
    double *kx[7] = {}, *ky[7] = {};
    double *u = nullptr, *v = nullptr;
    int *owner = nullptr;

    pv::Instance *inst_d = nullptr;
    std::vector<pv::Instance> inst_h;
    pv::InstanceStep *steps_d = nullptr;
    pv::InstanceStep *steps_h = nullptr;   // pinned mirror
    pv::WorkItem *items_d = nullptr;
    size_t items_cap = 0;
    int nitems = 0;
    // symmetric mode (pv::PairItem work items, per-instance partial sums)
    int sym_mode = -1;                 // -1: fixed-step ensembles only; 0: never; 1: always
    bool sym = false;                  // the mode the current work items were built for
    pv::PairItem *pitems_d = nullptr;
    size_t pitems_cap = 0;
    int *chunks_d = nullptr;           // chunk starts of every group (chunks_cap of them)
    size_t chunks_cap = 0;
    int nchunks = 0;
    // The launch groups of the current symmetric items (pv::sym_group: 0 every item at its
    // instance's width, 1 and 2 the periodic and disk 4-target items), each with its items at
    // pitems_d + item_off and its chunk starts (relative to item_off) at chunks_d + chunk_off.
    struct Group {
        int group, precisions;
        int item_off, nitems, chunk_off, nchunks;
    };
    std::vector<Group> groups;
    std::vector<pv_batch_rule> table;  // tile and chunk rules of the symmetric decomposition
    double *part = nullptr;
    size_t part_cap = 0;               // doubles
    int span = 0;                      // flat vortices the reduction covers
    std::vector<int> item_instances;   // instances of the current work items
    unsigned *counter = nullptr;
    pv::DevStatus *st = nullptr;
    pv::DevStatus *st_host = nullptr;
    unsigned epoch = 0;

    // fixed-step layout
    size_t total = 0;
    size_t nfixed = 0;
    unsigned long long steps_done = 0;   // RK4 steps since pv_batch_load
    // Two captured RK4 steps per (dt, buffer the pair starts from): a graph holds the pointers
    // it was captured with, and an odd eager tail swaps x and xn.
    std::map<std::pair<double, const double *>, hipGraphExec_t> graphs;

    // adaptive layout
    std::vector<char> used;
    std::vector<char> in_items;                           // slots the work items cover
    std::vector<std::pair<size_t, size_t>> free_ranges;   // [begin, end), sorted
    bool items_dirty = true;
    bool adaptive = false;

    ~pv_batch() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        for (auto &kv : graphs)
            (void)hipGraphExecDestroy(kv.second);
        dfree(x), dfree(y), dfree(g), dfree(xn), dfree(yn);
        dfree(xa), dfree(ya), dfree(xb), dfree(yb);
        for (int s = 0; s < 7; ++s)
            dfree(kx[s]), dfree(ky[s]);
        dfree(u), dfree(v), dfree(owner);
        dfree(inst_d), dfree(steps_d), dfree(items_d), dfree(counter), dfree(st);
        dfree(pitems_d), dfree(chunks_d), dfree(part);
        if (steps_h)
            (void)hipHostFree(steps_h);
        if (st_host)
            (void)hipHostFree(st_host);
        for (int k = 0; k < kSides; ++k) {
            if (side[k])
                (void)hipStreamSynchronize(side[k]);
            if (side[k])
                (void)hipStreamDestroy(side[k]);
            if (join[k])
                (void)hipEventDestroy(join[k]);
        }
        if (fork)
            (void)hipEventDestroy(fork);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace {

pv_batch *checked(pv_batch *b) {
    if (!b)
        raise(PV_ERR_INVALID_ARGUMENT, "null ensemble");
    return b;
}

void reset_status(pv_batch *b) {
    pv::launch_reset_status(b->st, b->stream);
    hip_check(hipGetLastError(), "launch HIP status reset");
}

int batch_blocks(const pv_batch *b, int nitems) {
    const int want = (nitems + 3) / 4;
    return std::max(1, std::min(want, 8 * b->cu_count));
}

void launch(pv_batch *b, const double *px, const double *py, const pv::StageOp &op,
            bool adaptive, int pass) {
    pv::launch_batch_reset(b->counter, b->stream);
    hip_check(hipGetLastError(), "launch HIP ensemble reset");
    pv::BatchLaunch a{};
    a.inst = b->inst_d;
    a.items = b->items_d;
    a.nitems = b->nitems;
    a.counter = b->counter;
    a.next = b->counter;
    a.x = px, a.y = py, a.g = b->g;
    a.st = b->st;
    a.epoch = ++b->epoch;
    a.op = op;
    a.steps = b->steps_d;
    a.pass = pass;
    if (b->sym) {
        // One launch per group, each with its own chunk counter; they write disjoint partial
        // slots, and the reduction after them adds every vortex's slots.
        a.part = b->part;
        a.owner = b->owner;
        a.span = b->span;
        const int ngroups = (int)b->groups.size();
        const int sides = std::min(ngroups - 1, pv_batch::kSides);
        if (sides > 0) {
            hip_check(hipEventRecord(b->fork, b->stream), "record HIP event");
            for (int k = 0; k < sides; ++k)
                hip_check(hipStreamWaitEvent(b->side[k], b->fork, 0), "wait on HIP event");
        }
        for (int g = 0; g < ngroups; ++g) {
            const pv_batch::Group &gr = b->groups[g];
            const int lane = g;     // at most kSides + 1 groups
            a.pitems = b->pitems_d + gr.item_off;
            a.chunks = b->chunks_d + gr.chunk_off;
            a.nchunks = gr.nchunks;
            a.nitems = gr.nitems;
            a.next = b->counter + 2 + gr.group;
            // Each group has a persistent grid of its own: blocks of a group whose chunks are
            // done exit and free their slots for the others' (measured on the 4096-instance
            // ensemble with three groups: grids split in proportion to the groups' estimated
            // costs were 9 % slower in FP64, the groups one after another 4 % slower mixed).
            pv::launch_batch_sym(gr.precisions, adaptive, gr.group, a,
                                 batch_blocks(b, gr.nitems),
                                 lane == 0 ? b->stream : b->side[lane - 1]);
            hip_check(hipGetLastError(), "launch HIP ensemble pair kernel");
        }
        for (int k = 0; k < sides; ++k) {
            hip_check(hipEventRecord(b->join[k], b->side[k]), "record HIP event");
            hip_check(hipStreamWaitEvent(b->stream, b->join[k], 0), "wait on HIP event");
        }
        pv::launch_batch_sym_reduce(adaptive, a, b->stream);
        hip_check(hipGetLastError(), "launch HIP ensemble reduction");
        return;
    }
    pv::launch_batch(b->item_precisions, adaptive, a, batch_blocks(b, b->nitems), b->stream);
    hip_check(hipGetLastError(), "launch HIP ensemble kernel");
}

bool wants_sym(const pv_batch *b) {
    return b->sym_mode == 1 || (b->sym_mode < 0 && !b->adaptive);
}

// Symmetric launch shape of an instance: targets per lane (tiles of 64 tpt vortices) and the
// most items a wave takes per counter increment.  The first rule of the ensemble's table that
// covers (geometry, precision, n) decides; otherwise 2 targets per lane and chunks of 4.
struct SymShape {
    int tpt, chunk;
};
constexpr int kChunkItemsDefault = 4;

SymShape sym_shape(const pv_batch *b, int geometry, int precision, size_t n) {
    for (const pv_batch_rule &r : b->table)
        if (r.geometry == geometry && r.precision == precision && n <= r.n_max)
            return {r.targets_per_lane ? r.targets_per_lane : pv::kSymTptDefault,
                    r.chunk_items ? r.chunk_items : kChunkItemsDefault};
    return {pv::kSymTptDefault, kChunkItemsDefault};
}

// Symmetric work items: every tile pair (I <= J) of every listed instance, each instance's
// partial sums placed after the previous one's.  Items are grouped by pv::sym_group, one
// launch per group, longest items first within a group.
void build_pair_items(pv_batch *b, const std::vector<int> &instances) {
    struct Keyed {
        double cost;
        pv::PairItem item;
        int chunk;          // the instance's chunk limit
    };
    std::vector<Keyed> per[pv::kSymGroups];
    int group_precisions[pv::kSymGroups] = {};
    size_t poff = 0;
    for (int k : instances) {
        pv::Instance &in = b->inst_h[k];
        const SymShape sh = sym_shape(b, in.gp.geometry, in.precision, (size_t)in.n);
        const int TT = 64 * sh.tpt;
        const int nt = (in.n + TT - 1) / TT;
        in.tpt = sh.tpt;
        in.nt = nt;
        in.poff = poff;
        poff += (size_t)nt * 2 * (size_t)in.n;
        const int g = pv::sym_group(sh.tpt, in.gp.geometry);
        group_precisions[g] |= in.precision ? 2 : 1;
        const double w = pair_weight(in.gp.geometry, in.precision);
        for (int J = 0; J < nt; ++J) {
            const double rj = std::min(TT, in.n - J * TT);
            for (int I = 0; I <= J; ++I) {
                const double ri = std::min(TT, in.n - I * TT);
                per[g].push_back({w * ri * rj * (I == J ? 1.0 : 0.5), {k, I, J, 0}, sh.chunk});
            }
        }
    }
    // Within a group: chunks of at most the first item's instance limit of consecutive items
    // (every item has a fixed fetch and staging latency besides its pairs) and of at most the
    // group's total / kChunksWanted cost (an item costlier than that is a chunk by itself).
    constexpr double kChunksWanted = 16384.0;
    std::vector<pv::PairItem> items;
    std::vector<int> chunks;
    std::vector<pv_batch::Group> groups;
    for (int g = 0; g < pv::kSymGroups; ++g) {
        std::vector<Keyed> &v = per[g];
        if (v.empty())
            continue;
        std::stable_sort(v.begin(), v.end(),
                         [](const Keyed &x, const Keyed &y) { return x.cost > y.cost; });
        double total_cost = 0.0;
        for (const Keyed &k : v)
            total_cost += k.cost;
        const double quota = total_cost / kChunksWanted;
        pv_batch::Group gr{};
        gr.group = g;
        gr.precisions = group_precisions[g];
        gr.item_off = (int)items.size();
        gr.nitems = (int)v.size();
        gr.chunk_off = (int)chunks.size();
        size_t start = 0;
        double acc = 0.0;
        chunks.push_back(0);
        for (size_t i = 0; i < v.size(); ++i) {
            const size_t held = i - start;
            if (held > 0 && (held >= (size_t)v[start].chunk || acc + v[i].cost > quota)) {
                chunks.push_back((int)i);
                start = i;
                acc = 0.0;
            }
            acc += v[i].cost;
        }
        chunks.push_back((int)v.size());
        gr.nchunks = (int)chunks.size() - gr.chunk_off - 1;
        for (const Keyed &k : v)
            items.push_back(k.item);
        groups.push_back(gr);
    }
    if (items.size() > b->pitems_cap || chunks.size() > b->chunks_cap || poff > b->part_cap) {
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        if (items.size() > b->pitems_cap) {
            dfree(b->pitems_d);
            b->pitems_cap = 0;
            b->pitems_d = dalloc<pv::PairItem>(items.size(), "hipMalloc(work items)");
            b->pitems_cap = items.size();
        }
        if (chunks.size() > b->chunks_cap) {
            dfree(b->chunks_d);
            b->chunks_cap = 0;
            b->chunks_d = dalloc<int>(chunks.size(), "hipMalloc(work chunks)");
            b->chunks_cap = chunks.size();
        }
        if (poff > b->part_cap) {
            dfree(b->part);
            b->part_cap = 0;
            b->part = dalloc<double>(poff, "hipMalloc(partial sums)");
            b->part_cap = poff;
        }
    }
    // The instance table is host-authoritative: the device never writes it.
    hip_check(hipMemcpyAsync(b->inst_d, b->inst_h.data(), b->max_inst * sizeof(pv::Instance),
                             hipMemcpyHostToDevice, b->stream),
              "upload instances");
    if (!items.empty()) {
        hip_check(hipMemcpyAsync(b->pitems_d, items.data(), items.size() * sizeof(pv::PairItem),
                                 hipMemcpyHostToDevice, b->stream),
                  "upload work items");
        hip_check(hipMemcpyAsync(b->chunks_d, chunks.data(), chunks.size() * sizeof(int),
                                 hipMemcpyHostToDevice, b->stream),
                  "upload work chunks");
    }
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    b->nitems = (int)items.size();
    b->nchunks = (int)(chunks.size() - groups.size());
    b->groups = std::move(groups);
}

void build_items(pv_batch *b, const std::vector<int> &instances);

// Cost-sorted work items of the given instances (longest first), ordered or symmetric.
void build_ordered_items(pv_batch *b, const std::vector<int> &instances) {
    struct Keyed {
        double cost;
        pv::WorkItem item;
    };
    std::vector<Keyed> v;
    for (int k : instances) {
        const int n = b->inst_h[k].n;
        const double w = pair_weight(b->inst_h[k].gp.geometry, b->inst_h[k].precision) * n;
        for (int tb = 0; tb < n; tb += 64)
            v.push_back({w, {k, tb, std::min(n, tb + 64), 0}});
    }
    std::stable_sort(v.begin(), v.end(),
                     [](const Keyed &a, const Keyed &c) { return a.cost > c.cost; });
    if (v.size() > b->items_cap) {
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        dfree(b->items_d);
        b->items_cap = 0;
        b->items_d = dalloc<pv::WorkItem>(v.size(), "hipMalloc(work items)");
        b->items_cap = v.size();
    }
    std::vector<pv::WorkItem> items(v.size());
    for (size_t i = 0; i < v.size(); ++i)
        items[i] = v[i].item;
    if (!items.empty()) {
        hip_check(hipMemcpyAsync(b->items_d, items.data(), items.size() * sizeof(pv::WorkItem),
                                 hipMemcpyHostToDevice, b->stream),
                  "upload work items");
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    }
    b->nitems = (int)items.size();
}

void build_items(pv_batch *b, const std::vector<int> &instances) {
    b->item_instances = instances;
    b->sym = wants_sym(b);
    int mask = 0;
    for (int k : instances)
        mask |= b->inst_h[k].precision ? 2 : 1;
    b->item_precisions = mask ? mask : 1;
    if (b->sym)
        build_pair_items(b, instances);
    else
        build_ordered_items(b, instances);
}

void read_status(pv_batch *b) {
    hip_check(hipMemcpyAsync(b->st_host, b->st, sizeof(pv::DevStatus), hipMemcpyDeviceToHost,
                             b->stream),
              "read HIP status");
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
}

// A non-finite velocity of an instance is a coincident pair when one of the reference's
// denominators is exactly zero (src/compute.cpp:119-123, 234-246, 307-313); the single-system
// path asks classify_kernel after a failure, and so does this, on the positions the failing
// evaluation read (kept intact: a failed instance is skipped from then on).  The result goes
// into the instance's failure word, once per failure, with its geometry for the message.
void classify_failures(pv_batch *b, size_t ninst) {
    for (size_t k = 0; k < ninst; ++k) {
        pv::InstanceStep &s = b->steps_h[k];
        if (!(s.failure & pv::kFailVelocity) || (s.failure & pv::kFailClassified) ||
            !s.evaluated[0])
            continue;
        const pv::Instance &in = b->inst_h[k];
        reset_status(b);
        pv::launch_classify(in.gp, s.evaluated[0] + in.offset, s.evaluated[1] + in.offset, in.n,
                            0, in.n, b->st, b->stream);
        hip_check(hipGetLastError(), "launch HIP coincidence check");
        read_status(b);
        s.failure |= pv::kFailClassified;
        if (b->st_host->flags & pv::kFailCoincident)
            s.failure |= pv::kFailCoincident |
                         ((unsigned)in.gp.geometry << pv::kFailGeometryShift);
        hip_check(hipMemcpyAsync(&b->steps_d[k].failure, &s.failure, sizeof(unsigned),
                                 hipMemcpyHostToDevice, b->stream),
                  "upload instance failure");
    }
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
}

// Fixed-step calls fail per instance (the instance stops; the others go on): refresh the
// host mirror of the instance words once the stream is done.
void finish_fixed(pv_batch *b) {
    if (b->nfixed)
        hip_check(hipMemcpyAsync(b->steps_h, b->steps_d, b->nfixed * sizeof(pv::InstanceStep),
                                 hipMemcpyDeviceToHost, b->stream),
                  "read instance states");
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    classify_failures(b, b->nfixed);
}

// Every instance slot idle, no vortex owned, the whole capacity free.
void clear_instances(pv_batch *b) {
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    for (size_t k = 0; k < b->max_inst; ++k) {
        std::memset(&b->steps_h[k], 0, sizeof(pv::InstanceStep));
        b->steps_h[k].state = pv::kIdle;
    }
    hip_check(hipMemcpyAsync(b->steps_d, b->steps_h, b->max_inst * sizeof(pv::InstanceStep),
                             hipMemcpyHostToDevice, b->stream),
              "upload instance states");
    if (b->vcap) {
        std::vector<int> minus(b->vcap, -1);
        hip_check(hipMemcpyAsync(b->owner, minus.data(), b->vcap * sizeof(int),
                                 hipMemcpyHostToDevice, b->stream),
                  "clear owners");
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    }
    b->used.assign(b->max_inst, 0);
    b->in_items.assign(b->max_inst, 0);
    b->free_ranges = {{0, b->capacity}};
    b->total = 0;
    b->nfixed = 0;
    b->nitems = 0;
    b->items_dirty = true;
    b->item_instances.clear();
    b->steps_done = 0;
    hip_check(hipMemsetAsync(b->counter, 0, kCounterWords * sizeof(unsigned), b->stream),
              "clear step count");
}

// One RK4 step of every instance from (x0, y0) into (x1, y1) (src/timestep.cpp:31-61).
void enqueue_rk4(pv_batch *b, double dt, const double *x0, const double *y0, double *x1,
                 double *y1) {
    pv::StageOp op{};
    op.dt = dt;
    op.x0 = x0, op.y0 = y0;
    for (int s = 0; s < 7; ++s)
        op.kx[s] = b->kx[s], op.ky[s] = b->ky[s];
    op.kind = pv::kOpStage, op.cur = 0, op.coef[0] = 0.5;
    op.ku = b->kx[0], op.kv = b->ky[0], op.xn = b->xa, op.yn = b->ya;
    launch(b, x0, y0, op, false, pv::kPassAll);
    op.cur = 1, op.coef[0] = 0.0, op.coef[1] = 0.5;
    op.ku = b->kx[1], op.kv = b->ky[1], op.xn = b->xb, op.yn = b->yb;
    launch(b, b->xa, b->ya, op, false, pv::kPassAll);
    op.cur = 2, op.coef[1] = 0.0, op.coef[2] = 1.0;
    op.ku = b->kx[2], op.kv = b->ky[2], op.xn = b->xa, op.yn = b->ya;
    launch(b, b->xb, b->yb, op, false, pv::kPassAll);
    op.kind = pv::kOpRk4Final, op.cur = 3;
    op.ku = b->kx[3], op.kv = b->ky[3], op.xn = x1, op.yn = y1;
    launch(b, b->xa, b->ya, op, false, pv::kPassAll);
    pv::launch_batch_step_done(b->counter, b->stream);
    hip_check(hipGetLastError(), "launch HIP ensemble step count");
}

void upload_slot(pv_batch *b, int k) {
    hip_check(hipMemcpyAsync(b->inst_d + k, &b->inst_h[k], sizeof(pv::Instance),
                             hipMemcpyHostToDevice, b->stream),
              "upload instance");
    hip_check(hipMemcpyAsync(b->steps_d + k, b->steps_h + k, sizeof(pv::InstanceStep),
                             hipMemcpyHostToDevice, b->stream),
              "upload instance state");
}

void set_owner(pv_batch *b, size_t begin, size_t n, int value) {
    std::vector<int> o(n, value);
    hip_check(hipMemcpyAsync(b->owner + begin, o.data(), n * sizeof(int), hipMemcpyHostToDevice,
                             b->stream),
              "upload owners");
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
}

void require_fixed(pv_batch *b) {
    if (b->adaptive)
        raise(PV_ERR_LOGIC, "ensemble holds admitted runs; fixed-step calls need pv_batch_load");
}

// The 25 per-vortex FP64 arrays.
std::vector<double **> vortex_arrays(pv_batch *b) {
    std::vector<double **> v = {&b->x, &b->y, &b->g, &b->xn, &b->yn, &b->xa, &b->ya, &b->xb,
                                &b->yb, &b->u, &b->v};
    for (int s = 0; s < 7; ++s) {
        v.push_back(&b->kx[s]);
        v.push_back(&b->ky[s]);
    }
    return v;
}

constexpr size_t kVortexBytes = kDoublesPerVortex * sizeof(double) + sizeof(int);

// Device bytes that do not depend on the vortices: instance tables, counters, status.
size_t fixed_bytes(const pv_batch *b) {
    return b->max_inst * (sizeof(pv::Instance) + sizeof(pv::InstanceStep)) +
           kCounterWords * sizeof(unsigned) + sizeof(pv::DevStatus);
}

// Everything the ensemble holds on the device now.
size_t device_bytes(const pv_batch *b) {
    return fixed_bytes(b) + b->vcap * kVortexBytes + b->part_cap * sizeof(double) +
           b->pitems_cap * sizeof(pv::PairItem) + b->chunks_cap * sizeof(int) +
           b->items_cap * sizeof(pv::WorkItem);
}

// Work-item and partial-sum bytes one running instance of n vortices needs in `sym` mode (its
// tile from the ensemble's launch rules).  The chunk starts are counted as one per item plus
// one, which bounds what the groups need.
size_t instance_bytes(const pv_batch *b, int geometry, int precision, size_t n, bool sym) {
    if (!sym)
        return (n + 63) / 64 * sizeof(pv::WorkItem);
    const size_t tile = 64 * (size_t)sym_shape(b, geometry, precision, n).tpt;
    const size_t nt = (n + tile - 1) / tile;
    return nt * (nt + 1) / 2 * (sizeof(pv::PairItem) + sizeof(int)) + sizeof(int) +
           nt * 2 * n * sizeof(double);
}

// The per-vortex arrays grown to newcap vortices (contents kept, new owners -1), one array at
// a time so that at most one old array is held beside the new ones.  Captured graphs hold the
// old pointers and are dropped.
void grow_vortices(pv_batch *b, size_t newcap) {
    if (newcap <= b->vcap)
        return;
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    for (auto &kv : b->graphs)
        (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    const size_t old = b->vcap;
    for (double **p : vortex_arrays(b)) {
        double *q = dalloc<double>(newcap, "hipMalloc(ensemble state)");
        if (*p && old)
            hip_check(hipMemcpyAsync(q, *p, old * sizeof(double), hipMemcpyDeviceToDevice,
                                     b->stream),
                      "copy ensemble state");
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        dfree(*p);
        *p = q;
    }
    int *o = dalloc<int>(newcap, "hipMalloc(owners)");
    if (b->owner && old)
        hip_check(hipMemcpyAsync(o, b->owner, old * sizeof(int), hipMemcpyDeviceToDevice,
                                 b->stream),
                  "copy owners");
    std::vector<int> minus(newcap - old, -1);
    hip_check(hipMemcpyAsync(o + old, minus.data(), minus.size() * sizeof(int),
                             hipMemcpyHostToDevice, b->stream),
              "clear owners");
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    dfree(b->owner);
    b->owner = o;
    b->vcap = newcap;
}

// Frees the work-item and partial-sum buffers; the next attempt rebuilds them at the size the
// running instances need.
void trim_items(pv_batch *b) {
    hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    dfree(b->part), dfree(b->pitems_d), dfree(b->chunks_d), dfree(b->items_d);
    b->part_cap = b->pitems_cap = b->chunks_cap = b->items_cap = 0;
    b->nitems = b->nchunks = 0;
    b->groups.clear();
    b->items_dirty = true;
}

} // namespace

extern "C" {

size_t pv_batch_bytes_per_vortex(void) { return kDoublesPerVortex * sizeof(double) + 8; }

pv_status pv_batch_set_byte_limit(pv_batch *b, size_t bytes) {
    return guard([&] {
        checked(b);
        b->byte_limit = bytes;
    });
}

pv_status pv_batch_device_bytes(pv_batch *b, size_t *bytes) {
    return guard([&] {
        checked(b);
        if (!bytes)
            raise(PV_ERR_INVALID_ARGUMENT, "null output");
        *bytes = device_bytes(b);
    });
}

pv_status pv_batch_instance_bytes(pv_batch *b, int geometry, int precision, size_t n, int sym,
                                  size_t *bytes) {
    return guard([&] {
        checked(b);
        if (!bytes)
            raise(PV_ERR_INVALID_ARGUMENT, "null output");
        if (geometry < PV_INFINITE || geometry > PV_DISK ||
            (precision != PV_FP64 && precision != PV_MIXED))
            raise(PV_ERR_INVALID_ARGUMENT, "unknown geometry or precision");
        *bytes = instance_bytes(b, geometry, precision, n, sym != 0);
    });
}

pv_status pv_batch_set_launch_table(pv_batch *b, const pv_batch_rule *rules, size_t count) {
    return guard([&] {
        checked(b);
        if (count && !rules)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        check_batch_rules(rules, count);
        DeviceGuard dg(b->device);
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        b->table.assign(rules, rules + count);
        for (auto &kv : b->graphs)
            (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        if (b->adaptive)
            b->items_dirty = true;
        else if (!b->item_instances.empty())
            build_items(b, std::vector<int>(b->item_instances));
    });
}

pv_status pv_batch_resolve_launch(pv_batch *b, int geometry, int precision, size_t n,
                                  int *targets_per_lane, int *chunk_items) {
    return guard([&] {
        checked(b);
        if (!targets_per_lane || !chunk_items)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        const SymShape sh = sym_shape(b, geometry, precision, n);
        *targets_per_lane = sh.tpt;
        *chunk_items = sh.chunk;
    });
}

pv_status pv_batch_create(int device, int precision, size_t capacity, size_t max_instances,
                          pv_batch **out) {
    return guard([&] {
        if (!out)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if (precision != PV_FP64 && precision != PV_MIXED)
            raise(PV_ERR_INVALID_ARGUMENT, "unsupported precision");
        if (capacity == 0 || capacity > (size_t)INT_MAX / 2 || max_instances == 0 ||
            max_instances > (size_t)INT_MAX / 2)
            raise(PV_ERR_INVALID_ARGUMENT, "invalid ensemble capacity");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "count HIP devices");
        if (device < 0 || device >= count)
            raise(PV_ERR_INVALID_ARGUMENT, "invalid HIP device ordinal");
        DeviceGuard dg(device);
        auto b = std::make_unique<pv_batch>();
        b->device = device;
        b->precision = precision;
        b->capacity = capacity;
        b->max_inst = max_instances;
        hipDeviceProp_t prop{};
        hip_check(hipGetDeviceProperties(&prop, device), "query HIP device properties");
        b->cu_count = std::max(1, prop.multiProcessorCount);
        b->table = default_batch_table();
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "create HIP stream");
        for (int k = 0; k < pv_batch::kSides; ++k) {
            hip_check(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking),
                      "create HIP stream");
            hip_check(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming),
                      "create HIP event");
        }
        hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "create HIP event");
        b->inst_d = dalloc<pv::Instance>(max_instances, "hipMalloc(instances)");
        b->steps_d = dalloc<pv::InstanceStep>(max_instances, "hipMalloc(instance states)");
        hip_check(hipHostMalloc((void **)&b->steps_h, max_instances * sizeof(pv::InstanceStep),
                                hipHostMallocDefault),
                  "hipHostMalloc(instance states)");
        b->inst_h.assign(max_instances, pv::Instance{});
        for (size_t k = 0; k < max_instances; ++k) {
            std::memset(&b->steps_h[k], 0, sizeof(pv::InstanceStep));
            b->steps_h[k].state = pv::kIdle;
        }
        // All initialisation is ordered on the ensemble's own (non-blocking) stream.
        hip_check(hipMemcpyAsync(b->steps_d, b->steps_h, max_instances * sizeof(pv::InstanceStep),
                                 hipMemcpyHostToDevice, b->stream),
                  "upload instance states");
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        b->counter = dalloc<unsigned>(kCounterWords, "hipMalloc(work counter)");
        hip_check(hipMemsetAsync(b->counter, 0, kCounterWords * sizeof(unsigned), b->stream),
                  "clear work counter");
        b->st = dalloc<pv::DevStatus>(1, "hipMalloc(status)");
        hip_check(hipHostMalloc((void **)&b->st_host, sizeof(pv::DevStatus), hipHostMallocDefault),
                  "hipHostMalloc(status)");
        b->used.assign(max_instances, 0);
        b->in_items.assign(max_instances, 0);
        b->free_ranges = {{0, capacity}};
        reset_status(b.get());
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        *out = b.release();
    });
}

void pv_batch_destroy(pv_batch *b) {
    if (!b)
        return;
    try {
        DeviceGuard dg(b->device);
        delete b;
    } catch (...) {
        delete b;
    }
}

pv_status pv_batch_set_precision(pv_batch *b, int precision) {
    return guard([&] {
        checked(b);
        if (precision != PV_FP64 && precision != PV_MIXED)
            raise(PV_ERR_INVALID_ARGUMENT, "unsupported precision");
        b->precision = precision;
        if (b->adaptive)
            return;     // admitted runs keep theirs; the next admissions take this one
        DeviceGuard dg(b->device);
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        for (auto &kv : b->graphs)
            (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        for (size_t k = 0; k < b->nfixed; ++k)
            b->inst_h[k].precision = precision;
        if (!b->item_instances.empty())
            build_items(b, std::vector<int>(b->item_instances));
    });
}

pv_status pv_batch_set_symmetric(pv_batch *b, int mode) {
    return guard([&] {
        checked(b);
        if (mode < -1 || mode > 1)
            raise(PV_ERR_INVALID_ARGUMENT, "symmetric mode must be -1, 0 or 1");
        if (mode == b->sym_mode)
            return;
        DeviceGuard dg(b->device);
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        b->sym_mode = mode;
        for (auto &kv : b->graphs)
            (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        if (b->adaptive)
            b->items_dirty = true;
        else if (!b->item_instances.empty())
            build_items(b, std::vector<int>(b->item_instances));
    });
}

int pv_batch_symmetric(const pv_batch *b) { return b && b->sym ? 1 : 0; }

pv_status pv_batch_load(pv_batch *b, const pv_instance_config *configs, const size_t *counts,
                        size_t ninst, const double *x, const double *y, const double *gamma) {
    return guard([&] {
        checked(b);
        if (ninst > b->max_inst)
            raise(PV_ERR_INVALID_ARGUMENT, "more instances than the ensemble was created for");
        size_t total = 0;
        for (size_t k = 0; k < ninst; ++k)
            total += counts[k];
        if (total > b->capacity)
            raise(PV_ERR_INVALID_ARGUMENT, "more vortices than the ensemble was created for");
        if (b->byte_limit) {
            size_t need = fixed_bytes(b) + std::max(total, b->vcap) * kVortexBytes;
            for (size_t k = 0; k < ninst; ++k)
                need += instance_bytes(b, configs[k].geometry, b->precision, counts[k],
                                       b->sym_mode != 0);
            if (need > b->byte_limit)
                raise(PV_ERR_RUNTIME, "ensemble would exceed its device byte limit");
        }
        DeviceGuard dg(b->device);
        clear_instances(b);
        grow_vortices(b, total);
        size_t off = 0;
        std::vector<int> owner(total);
        std::vector<int> all;
        for (size_t k = 0; k < ninst; ++k) {
            const pv_config c = to_config(configs[k], b->precision);
            check_config(c);
            validate_instance(configs[k], x + off, y + off, gamma + off, counts[k]);
            b->inst_h[k].gp = geo_params(c);
            b->inst_h[k].offset = (int)off;
            b->inst_h[k].n = (int)counts[k];
            b->inst_h[k].precision = b->precision;
            b->steps_h[k].state = pv::kRunning;
            std::fill(owner.begin() + off, owner.begin() + off + counts[k], (int)k);
            if (counts[k])
                all.push_back((int)k);
            off += counts[k];
        }
        const size_t bytes = total * sizeof(double);
        hip_check(hipMemcpyAsync(b->x, x, bytes, hipMemcpyHostToDevice, b->stream), "copy x");
        hip_check(hipMemcpyAsync(b->y, y, bytes, hipMemcpyHostToDevice, b->stream), "copy y");
        hip_check(hipMemcpyAsync(b->g, gamma, bytes, hipMemcpyHostToDevice, b->stream),
                  "copy circulation");
        hip_check(hipMemcpyAsync(b->owner, owner.data(), total * sizeof(int),
                                 hipMemcpyHostToDevice, b->stream),
                  "copy owners");
        hip_check(hipMemcpyAsync(b->inst_d, b->inst_h.data(), ninst * sizeof(pv::Instance),
                                 hipMemcpyHostToDevice, b->stream),
                  "copy instances");
        hip_check(hipMemcpyAsync(b->steps_d, b->steps_h, ninst * sizeof(pv::InstanceStep),
                                 hipMemcpyHostToDevice, b->stream),
                  "upload instance states");
        b->total = total;
        b->nfixed = ninst;
        b->adaptive = false;
        b->span = (int)total;
        build_items(b, all);
        for (auto &kv : b->graphs)
            (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        reset_status(b);
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    });
}

pv_status pv_batch_rk4(pv_batch *b, double dt, int steps, int use_graph) {
    return guard([&] {
        checked(b);
        require_fixed(b);
        if (!std::isfinite(dt) || !(dt > 0.0))
            raise(PV_ERR_INVALID_ARGUMENT, "timestep must be finite and positive");
        if (steps < 0)
            raise(PV_ERR_INVALID_ARGUMENT, "step count must be non-negative");
        if (steps == 0 || b->nitems == 0)
            return;
        DeviceGuard dg(b->device);
        reset_status(b);
        int done = 0;
        if (use_graph && steps >= 2) {
            // Two steps per replay: x -> xn -> x, so the captured pointers stay valid.  Keyed
            // on the buffer x is in now: after an odd step count x and xn have traded places,
            // and a graph captured the other way round would step from the stale buffer.
            const auto key = std::make_pair(dt, static_cast<const double *>(b->x));
            auto it = b->graphs.find(key);
            if (it == b->graphs.end()) {
                hipGraph_t graph = nullptr;
                hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
                          "begin HIP graph capture");
                try {
                    enqueue_rk4(b, dt, b->x, b->y, b->xn, b->yn);
                    enqueue_rk4(b, dt, b->xn, b->yn, b->x, b->y);
                } catch (...) {
                    (void)hipStreamEndCapture(b->stream, &graph);
                    if (graph)
                        (void)hipGraphDestroy(graph);
                    throw;
                }
                hip_check(hipStreamEndCapture(b->stream, &graph), "end HIP graph capture");
                hipGraphExec_t exec = nullptr;
                const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                hip_check(e, "instantiate HIP graph");
                it = b->graphs.emplace(key, exec).first;
            }
            for (; done + 2 <= steps; done += 2)
                hip_check(hipGraphLaunch(it->second, b->stream), "launch HIP graph");
        }
        for (; done < steps; ++done) {
            enqueue_rk4(b, dt, b->x, b->y, b->xn, b->yn);
            std::swap(b->x, b->xn);
            std::swap(b->y, b->yn);
        }
        b->steps_done += (unsigned long long)steps;
        finish_fixed(b);
    });
}

pv_status pv_batch_read(pv_batch *b, double *x, double *y, size_t total) {
    return guard([&] {
        checked(b);
        require_fixed(b);
        if (total != b->total)
            raise(PV_ERR_RUNTIME, "host and HIP vortex populations differ");
        DeviceGuard dg(b->device);
        hip_check(hipMemcpyAsync(x, b->x, total * sizeof(double), hipMemcpyDeviceToHost,
                                 b->stream),
                  "download x");
        hip_check(hipMemcpyAsync(y, b->y, total * sizeof(double), hipMemcpyDeviceToHost,
                                 b->stream),
                  "download y");
        // A failed instance was left untouched from its failing step on: its state is the
        // input of that step, in whichever position buffer the step read.
        for (size_t k = 0; k < b->nfixed; ++k) {
            const pv::InstanceStep &s = b->steps_h[k];
            if (!s.failure || !s.keep[0] || s.keep[0] == b->x)
                continue;
            const size_t off = (size_t)b->inst_h[k].offset, n = (size_t)b->inst_h[k].n;
            hip_check(hipMemcpyAsync(x + off, s.keep[0] + off, n * sizeof(double),
                                     hipMemcpyDeviceToHost, b->stream),
                      "download x");
            hip_check(hipMemcpyAsync(y + off, s.keep[1] + off, n * sizeof(double),
                                     hipMemcpyDeviceToHost, b->stream),
                      "download y");
        }
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    });
}

pv_status pv_batch_failures(pv_batch *b, unsigned *failures, unsigned long long *steps_done,
                            size_t ninst) {
    return guard([&] {
        checked(b);
        require_fixed(b);
        if (ninst != b->nfixed || (ninst && !failures))
            raise(PV_ERR_INVALID_ARGUMENT, "instance count differs from the loaded ensemble");
        for (size_t k = 0; k < ninst; ++k) {
            const pv::InstanceStep &s = b->steps_h[k];
            failures[k] = s.failure;
            if (steps_done)
                steps_done[k] = s.failure ? s.fail_step : b->steps_done;
        }
    });
}

pv_status pv_batch_evaluate(pv_batch *b, double *u, double *v, size_t total) {
    return guard([&] {
        checked(b);
        require_fixed(b);
        if (total != b->total)
            raise(PV_ERR_RUNTIME, "host and HIP vortex populations differ");
        if (total == 0)
            return;
        DeviceGuard dg(b->device);
        reset_status(b);
        pv::StageOp op{};
        op.kind = pv::kOpStore;
        op.ku = b->u, op.kv = b->v;
        op.x0 = b->x, op.y0 = b->y;   // where a failing instance's state stays
        launch(b, b->x, b->y, op, false, pv::kPassAll);
        hip_check(hipMemcpyAsync(u, b->u, total * sizeof(double), hipMemcpyDeviceToHost,
                                 b->stream),
                  "download u");
        hip_check(hipMemcpyAsync(v, b->v, total * sizeof(double), hipMemcpyDeviceToHost,
                                 b->stream),
                  "download v");
        finish_fixed(b);
    });
}

pv_status pv_batch_admit(pv_batch *b, const pv_instance_config *config, const double *x,
                         const double *y, const double *gamma, size_t n, double t_end,
                         double dt0, double atol, double rtol, double min_dt, double max_dt,
                         int *slot) {
    return guard([&] {
        checked(b);
        if (!config || !slot)
            raise(PV_ERR_INVALID_ARGUMENT, "null argument");
        *slot = -1;
        const pv_config c = to_config(*config, b->precision);
        check_config(c);
        if (n == 0)
            raise(PV_ERR_INVALID_ARGUMENT, "an admitted run needs at least one vortex");
        validate_instance(*config, x, y, gamma, n);
        if (!std::isfinite(dt0) || !(dt0 > 0.0) || !std::isfinite(t_end) || t_end < 0.0)
            raise(PV_ERR_INVALID_ARGUMENT, "timestep must be finite and positive");
        if (!b->adaptive) {
            // First admission turns a loaded fixed-step ensemble into an adaptive one.
            DeviceGuard dg(b->device);
            clear_instances(b);
            b->adaptive = true;
        }
        int k = -1;
        for (size_t s = 0; s < b->max_inst; ++s)
            if (!b->used[s]) {
                k = (int)s;
                break;
            }
        if (k < 0)
            return;
        size_t r = b->free_ranges.size();
        for (size_t q = 0; q < b->free_ranges.size(); ++q)
            if (b->free_ranges[q].second - b->free_ranges[q].first >= n) {
                r = q;
                break;
            }
        if (r == b->free_ranges.size())
            return;
        const size_t off = b->free_ranges[r].first;
        // Admission by device footprint: the per-vortex arrays grown to cover the run, the work
        // items and partial sums every admitted run needs, and (while the arrays grow) one old
        // array held beside the new ones, all within the byte limit.
        DeviceGuard dg(b->device);
        const size_t need = off + n;
        size_t newcap = b->vcap;
        if (need > b->vcap)
            newcap = std::min(b->capacity,
                              std::max({need, b->vcap + b->vcap / 2, (size_t)1 << 16}));
        if (b->byte_limit) {
            const bool sym = b->sym_mode == 1;
            size_t items = instance_bytes(b, config->geometry, b->precision, n, sym);
            for (size_t q = 0; q < b->max_inst; ++q)
                if (b->used[q])
                    items += instance_bytes(b, b->inst_h[q].gp.geometry, b->inst_h[q].precision,
                                            (size_t)b->inst_h[q].n, sym);
            const size_t held = device_bytes(b) - fixed_bytes(b) - b->vcap * kVortexBytes;
            auto projected = [&](size_t cap, size_t buffers) {
                const size_t growing = cap > b->vcap ? b->vcap * sizeof(double) : 0;
                return fixed_bytes(b) + cap * kVortexBytes + growing + buffers;
            };
            if (projected(newcap, std::max(items, held)) > b->byte_limit)
                newcap = std::max(need, b->vcap);
            if (projected(newcap, std::max(items, held)) > b->byte_limit) {
                if (projected(newcap, items) > b->byte_limit)
                    return;                 // does not fit now
                trim_items(b);              // buffers still sized for runs that have left
            }
        }
        grow_vortices(b, newcap);
        b->free_ranges[r].first += n;
        if (b->free_ranges[r].first == b->free_ranges[r].second)
            b->free_ranges.erase(b->free_ranges.begin() + r);
        const size_t bytes = n * sizeof(double);
        hip_check(hipMemcpyAsync(b->x + off, x, bytes, hipMemcpyHostToDevice, b->stream), "copy x");
        hip_check(hipMemcpyAsync(b->y + off, y, bytes, hipMemcpyHostToDevice, b->stream), "copy y");
        hip_check(hipMemcpyAsync(b->g + off, gamma, bytes, hipMemcpyHostToDevice, b->stream),
                  "copy circulation");
        b->inst_h[k].gp = geo_params(c);
        b->inst_h[k].offset = (int)off;
        b->inst_h[k].n = (int)n;
        b->inst_h[k].precision = b->precision;
        pv::InstanceStep &s = b->steps_h[k];
        std::memset(&s, 0, sizeof s);
        // src/main.cpp:162-165 (initial dt) and 312-314 (landing on the end time).
        const double dt = std::clamp(dt0, min_dt, max_dt);
        s.t = 0.0;
        s.t_end = t_end;
        s.sugg = dt;
        s.dt = std::min(dt, t_end);
        s.atol = atol, s.rtol = rtol, s.min_dt = min_dt, s.max_dt = max_dt;
        s.state = t_end > 0.0 ? pv::kRunning : pv::kFinished;
        s.fsal = 0;
        upload_slot(b, k);
        set_owner(b, off, n, k);
        b->used[k] = 1;
        b->items_dirty = true;
        *slot = k;
    });
}

// One DOPRI5 attempt of every running run, enqueued on the ensemble's stream.
static void enqueue_attempt(pv_batch *b) {
    if (b->items_dirty) {
        std::vector<int> running;
        b->in_items.assign(b->max_inst, 0);
        for (size_t k = 0; k < b->max_inst; ++k)
            if (b->used[k] && b->steps_h[k].state == pv::kRunning) {
                running.push_back((int)k);
                b->in_items[k] = 1;
            }
        build_items(b, running);
        b->items_dirty = false;
    }
    if (b->nitems == 0)
        return;
    // Per-vortex passes only need to reach the end of the highest occupied range.
    size_t high = 0;
    for (size_t k = 0; k < b->max_inst; ++k)
        if (b->used[k])
            high = std::max(high, (size_t)b->inst_h[k].offset + (size_t)b->inst_h[k].n);
    const int n = (int)high;
    b->span = n;
    pv::StageOp op{};
    for (int s = 0; s < 7; ++s) {
        op.kx[s] = b->kx[s], op.ky[s] = b->ky[s];
        op.w[s] = kB5[s] - kB4[s];
    }
    op.x0 = b->x, op.y0 = b->y;
    // FSAL derivative for runs that have none yet (src/timestep.cpp:75-77).
    op.kind = pv::kOpStore;
    op.ku = b->kx[0], op.kv = b->ky[0];
    launch(b, b->x, b->y, op, true, pv::kPassNeedK0);
    // Stage positions: 1 -> A, 2 -> B, 3 -> A, 4 -> B, 5 -> A, 6 (fifth order) -> next.
    double *px[7] = {nullptr, b->xa, b->xb, b->xa, b->xb, b->xa, b->xn};
    double *py[7] = {nullptr, b->ya, b->yb, b->ya, b->yb, b->ya, b->yn};
    op.kind = pv::kOpStage, op.cur = 0, op.coef[0] = kDopriA[1][0];
    op.xn = px[1], op.yn = py[1];
    pv::launch_batch_stage1(op, b->owner, b->steps_d, n, b->stream);
    hip_check(hipGetLastError(), "launch HIP ensemble stage");
    for (int s = 1; s < 7; ++s) {
        op.cur = s;
        op.ku = b->kx[s], op.kv = b->ky[s];
        if (s < 6) {
            op.kind = pv::kOpStage;
            for (int j = 0; j <= s; ++j)
                op.coef[j] = kDopriA[s + 1][j];
            op.xn = px[s + 1], op.yn = py[s + 1];
        } else {
            op.kind = pv::kOpDopriError;
            op.xn = nullptr, op.yn = nullptr;
        }
        launch(b, px[s], py[s], op, true, pv::kPassAll);
    }
    pv::launch_batch_control(b->steps_d, (int)b->max_inst, b->stream);
    hip_check(hipGetLastError(), "launch HIP ensemble controller");
    pv::launch_batch_accept(b->owner, b->steps_d, b->x, b->y, b->xn, b->yn, b->kx[0],
                            b->ky[0], b->kx[6], b->ky[6], n, b->stream);
    hip_check(hipGetLastError(), "launch HIP ensemble accept");
}

pv_status pv_batch_attempts(pv_batch *b, int count) {
    return guard([&] {
        checked(b);
        if (!b->adaptive)
            raise(PV_ERR_LOGIC, "no admitted runs in this ensemble");
        if (count < 1)
            raise(PV_ERR_INVALID_ARGUMENT, "attempt count must be positive");
        DeviceGuard dg(b->device);
        // Runs that finish or fail part-way sit out the rest: every kernel of an attempt skips
        // a run that is not running.  The work items are rebuilt at the next call.
        for (int c = 0; c < count; ++c)
            enqueue_attempt(b);
        hip_check(hipMemcpyAsync(b->steps_h, b->steps_d, b->max_inst * sizeof(pv::InstanceStep),
                                 hipMemcpyDeviceToHost, b->stream),
                  "read instance states");
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
        // Before the run can be released and its vortex range reused.
        classify_failures(b, b->max_inst);
        for (size_t k = 0; k < b->max_inst; ++k)
            if (b->in_items[k] && b->steps_h[k].state != pv::kRunning)
                b->items_dirty = true;
    });
}

pv_status pv_batch_attempt(pv_batch *b) { return pv_batch_attempts(b, 1); }

pv_status pv_batch_poll(pv_batch *b, pv_slot_state *states, size_t cap, size_t *count) {
    return guard([&] {
        checked(b);
        if (!count)
            raise(PV_ERR_INVALID_ARGUMENT, "null count");
        size_t m = 0;
        for (size_t k = 0; k < b->max_inst; ++k) {
            if (!b->used[k])
                continue;
            if (m < cap && states) {
                const pv::InstanceStep &s = b->steps_h[k];
                pv_slot_state &o = states[m];
                o.slot = (int)k;
                o.state = s.state;
                o.t = s.t;
                o.dt = s.dt;
                o.last_error = s.last_error;
                o.accepted = s.accepted;
                o.rejected = s.rejected_total;
                o.failure = s.failure;
                o.n = (size_t)b->inst_h[k].n;
            }
            ++m;
        }
        *count = m;
    });
}

pv_status pv_batch_fetch(pv_batch *b, int slot, double *x, double *y, size_t n) {
    return guard([&] {
        checked(b);
        if (slot < 0 || (size_t)slot >= b->max_inst || !b->used[slot])
            raise(PV_ERR_INVALID_ARGUMENT, "no run in this slot");
        if (n != (size_t)b->inst_h[slot].n)
            raise(PV_ERR_RUNTIME, "host and HIP vortex populations differ");
        DeviceGuard dg(b->device);
        const size_t off = (size_t)b->inst_h[slot].offset;
        hip_check(hipMemcpyAsync(x, b->x + off, n * sizeof(double), hipMemcpyDeviceToHost,
                                 b->stream),
                  "download x");
        hip_check(hipMemcpyAsync(y, b->y + off, n * sizeof(double), hipMemcpyDeviceToHost,
                                 b->stream),
                  "download y");
        hip_check(hipStreamSynchronize(b->stream), "synchronize HIP stream");
    });
}

pv_status pv_batch_release(pv_batch *b, int slot) {
    return guard([&] {
        checked(b);
        if (slot < 0 || (size_t)slot >= b->max_inst || !b->used[slot])
            raise(PV_ERR_INVALID_ARGUMENT, "no run in this slot");
        DeviceGuard dg(b->device);
        const size_t off = (size_t)b->inst_h[slot].offset, n = (size_t)b->inst_h[slot].n;
        std::memset(&b->steps_h[slot], 0, sizeof(pv::InstanceStep));
        b->steps_h[slot].state = pv::kIdle;
        upload_slot(b, slot);
        set_owner(b, off, n, -1);
        b->used[slot] = 0;
        // Return [off, off + n) to the free list, merging neighbours.
        auto &fr = b->free_ranges;
        auto pos = std::lower_bound(fr.begin(), fr.end(), std::make_pair(off, off + n));
        pos = fr.insert(pos, {off, off + n});
        if (pos + 1 != fr.end() && pos->second == (pos + 1)->first) {
            pos->second = (pos + 1)->second;
            fr.erase(pos + 1);
        }
        if (pos != fr.begin() && (pos - 1)->second == pos->first) {
            (pos - 1)->second = pos->second;
            fr.erase(pos);
        }
        b->items_dirty = true;
    });
}

const char *pv_batch_failure_message(unsigned failure) {
    if (failure & pv::kFailStateNonFinite)
        return "vortex data must be finite";
    if (failure & pv::kFailOutsideDisk)
        return "vortex lies on or outside the disk";
    if (failure & pv::kFailCoincident) {
        // The single-system messages (src/compute.cpp:131-132, 254-255, 334-335).
        const unsigned geometry = (failure >> pv::kFailGeometryShift) & 3u;
        return geometry == PV_INFINITE   ? "coincident vortices in singular point-vortex kernel"
               : geometry == PV_PERIODIC ? "coincident periodic vortices"
                                         : "coincident vortices in disk";
    }
    if (failure & pv::kFailVelocity)
        return "non-finite velocity; check scales and close encounters";
    if (failure & pv::kFailErrorEstimate)
        return "non-finite adaptive error estimate or invalid tolerance";
    if (failure & pv::kFailRejected)
        return "adaptive integrator could not satisfy tolerance";
    return "";
}

} // extern "C"
