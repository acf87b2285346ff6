// pv_batch.hpp -- batched ensembles: many independent vortex systems of uneven N advanced
// together, one persistent launch per RK stage.
#pragma once
#include "pv_kernels.hpp"

namespace pv {

// One instance of the ensemble: its vortices are [offset, offset + n) of the flat SoA arrays.
// Symmetric mode: the instance's tiles are 64 tpt vortices, and its partial sums are nt tile
// rows x 2n doubles at part + poff.
struct Instance {
    GeoParams gp;
    int offset;
    int n;
    int nt;
    int precision;   // the instance's pair arithmetic: 0 FP64, 1 mixed
    size_t poff;
    int tpt;         // symmetric mode: targets per lane (1, 2 or 4)
    int pad;
};

// One unit of work: 64 targets [tbegin, tend) (instance-local) of one instance against all
// of that instance's sources.  Items are sorted by decreasing cost on the host, and waves take
// them from a shared counter, so the largest items start first (longest-processing-time
// order); every target's sum is still computed by one wave in a fixed order.
struct WorkItem {
    int instance;
    int tbegin, tend;
    int pad;
};

// Per-instance state resident on the device: the adaptive DOPRI5 run (src/timestep.cpp:62-119
// and the step landing of src/main.cpp:312-314), and for both modes the instance's failure
// word.  A fixed-step instance that fails keeps the input state of its failing step at
// keep[0..1] (one of the two ping-pong position buffers).
struct InstanceStep {
    double t, dt, sugg, t_end;        // time, trial step, carried suggestion, horizon
    double atol, rtol, min_dt, max_dt;
    double last_error;
    unsigned long long err_bits;      // max-norm error of the current attempt
    int state;                        // kRunning, kFinished, kFailed, kIdle
    int fsal;                         // k0 holds f(x) at the current state
    int accept_now;                   // set by the controller for the accept pass
    unsigned accepted, rejected_total, rejected_now;
    unsigned failure;                 // FailureKind bits (kFailRejected: too many rejections)
    unsigned fail_step;               // fixed-step mode: steps completed before the failing one
    const double *keep[2];            // fixed-step mode: {x, y} before the failing step
    const double *evaluated[2];       // {x, y} the first failing evaluation read
};

// Symmetric mode: one unit of work is a pair of tiles (I <= J) of one instance, 64 tpt
// vortices each, tpt its instance's targets per lane.  One wave computes each of its unordered
// pairs once and writes the reaction on tile J into partial slot I and tile I's side into slot
// J; every (slot, vortex) is written by exactly one item, and a reduction adds a vortex's slots
// in slot order.  One launch (group 0) takes every item at its instance's width, except the
// periodic and disk 4-target items, which run in launches of their own (groups 1 and 2: their
// 4-target code needs more registers than the shared kernel should hold).
constexpr int kSymTptDefault = 2;
constexpr int kSymGroups = 3;
constexpr int kSymAllGeometries = 3;
inline int sym_group(int tpt, int geometry) { return tpt >= 4 && geometry != 0 ? geometry : 0; }
// A wave takes one chunk of consecutive items per counter increment: the host cuts the
// cost-sorted items into runs of about equal cost (the largest items alone), so the counter
// sees far fewer atomics than items and the tail stays balanced.
struct PairItem {
    int instance;
    int I, J;
    int pad;
};

enum InstanceState : int { kRunning = 0, kFinished = 1, kFailed = 2, kIdle = 3 };
constexpr unsigned kFailRejected = 32u;
// Set by the host once a non-finite velocity of the instance has been classified (the exact
// zero-denominator test); with kFailCoincident, bits 8-9 carry the instance's geometry so the
// message names it as the single-system path does.
constexpr unsigned kFailClassified = 64u;
constexpr int kFailGeometryShift = 8;

enum BatchPass : int {
    kPassAll = 0,        // every item
    kPassNeedK0 = 1,     // only instances whose FSAL derivative is missing (adaptive)
};

struct BatchLaunch {
    const Instance *inst;
    const WorkItem *items;
    int nitems;
    unsigned *counter;         // [0] next ordered item, [2 + g] next chunk of symmetric group g
                               // (zeroed before every stage); [1] RK4 steps completed since
                               // the ensemble was loaded
    unsigned *next;            // this launch's item (ordered) or chunk (symmetric) counter
    const double *x, *y, *g;   // positions evaluated (flat) and circulations
    DevStatus *st;
    unsigned epoch;
    StageOp op;
    InstanceStep *steps;       // adaptive mode: per-instance dt, error and failure sink
    int pass;
    // symmetric mode
    const PairItem *pitems;    // work items (nitems of them) instead of `items`
    const int *chunks;         // chunk c is items [chunks[c], chunks[c + 1]); nchunks chunks
    int nchunks;
    double *part;              // partial sums, Instance::poff per instance
    const int *owner;          // instance of every flat vortex (-1: unused)
    int span;                  // flat vortices the reduction covers
};

// Launchers (pv_batch.hip).  `precisions` says which pair arithmetic the work items' instances
// use: bit 0 FP64, bit 1 mixed (both: one launch dispatches on each instance's precision).
void launch_batch(int precisions, bool adaptive, const BatchLaunch &a, int blocks, hipStream_t s);
// Zeroes counter[0] and the kSymGroups chunk counters.
void launch_batch_reset(unsigned *counter, hipStream_t s);
// Symmetric mode: the pair kernel over the PairItems of one launch group (sym_group), then
// (once per stage) the per-vortex reduction with the stage epilogue.
void launch_batch_sym(int precisions, bool adaptive, int group, const BatchLaunch &a, int blocks,
                      hipStream_t s);
void launch_batch_sym_reduce(bool adaptive, const BatchLaunch &a, hipStream_t s);
void launch_batch_step_done(unsigned *counter, hipStream_t s);
void launch_batch_stage1(const StageOp &op, const int *owner, const InstanceStep *steps, int n,
                         hipStream_t s);
void launch_batch_control(InstanceStep *steps, int ninst, hipStream_t s);
void launch_batch_accept(const int *owner, const InstanceStep *steps, double *x, double *y,
                         const double *xn, const double *yn, double *k0x, double *k0y,
                         const double *k6x, const double *k6y, int n, hipStream_t s);

} // namespace pv
