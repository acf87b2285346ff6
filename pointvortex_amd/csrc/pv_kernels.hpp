// pv_kernels.hpp -- launch-side declarations shared by the kernels (pv_kernels.hip) and the
// C-ABI (pv_api.cpp).
#pragma once
#include "pv_device.hpp"

namespace pv {

// Sticky device status word, one per handle.  Kernels of a launch whose epoch is later than
// fail_epoch return at once, so the buffers stay as they were when the failure happened.
struct DevStatus {
    unsigned flags;                // FailureKind bits
    unsigned fail_epoch;           // epoch of the first failing launch (UINT_MAX: none)
    unsigned long long err_bits;   // DOPRI5 max-norm error (bit pattern of a double >= 0)
    unsigned long long count;      // dipole candidates found
    unsigned long long barrier;    // grid-barrier arrivals of the persistent RK4 kernel
    unsigned long long pad[3];
};

enum FailureKind : unsigned {
    kFailStateNonFinite = 1u,    // "vortex data must be finite"
    kFailOutsideDisk = 2u,       // "vortex lies on or outside the disk"
    kFailVelocity = 4u,          // non-finite velocity (coincident pairs end up here too)
    kFailErrorEstimate = 8u,     // "non-finite adaptive error estimate or invalid tolerance"
    kFailCoincident = 16u,       // set by the classification pass
    kFailBarrier = 32u,          // a grid barrier of the persistent RK4 kernel timed out
};

enum OpKind : int { kOpStore = 0, kOpStage = 1, kOpRk4Final = 2, kOpDopriError = 3 };

// Fused per-target epilogue of a velocity evaluation (src/timestep.cpp:19-30, 48-57, 62-119).
struct StageOp {
    int kind;
    int cur;                   // stage index of this evaluation
    double dt;
    double coef[7];            // kOpStage: next position = x0 + dt * sum_{j<=cur} coef[j] k_j
    double w[7];               // kOpDopriError: b5 - b4 (the epilogue scales by dt)
    double atol, rtol;
    double *ku, *kv;           // where this evaluation's velocity is stored (may be null)
    const double *kx[7], *ky[7];  // earlier stage velocities
    const double *x0, *y0;     // positions at the start of the step
    double *xn, *yn;           // next positions
};

struct PairLaunch {
    const double *x, *y, *g;   // positions evaluated and circulations
    int n;                     // population; sources are [0, n)
    int tbegin, tend;          // targets
    int split_len;             // sources per split (multiple of the LDS tile)
    int nsplit;
    double *part;              // [nsplit][2][n] partial sums when nsplit > 1
    unsigned *tickets;         // one per target tile, zero between launches
    DevStatus *st;
    unsigned epoch;
    StageOp op;
};

// Symmetric evaluation of a whole population (every geometry, FP64 and mixed): one block per
// unordered pair of tiles (I <= J) of sym_tile(TPT) vortices.  Off-diagonal blocks compute each
// pair once and write the target side (into tile I's vortices, slot J) and the reaction (into
// tile J's vortices, slot I); diagonal blocks compute the ordered pairs inside their tile (slot
// I).  Every (slot, vortex) entry of `part` is written exactly once; the reduce kernel then sums
// a vortex's nt slots in slot order and applies the stage epilogue.
struct SymLaunch {
    const double *x, *y, *g;
    int n;
    int nt;                    // tiles of sym_tile(TPT) vortices
    double *part;              // [nt][2][n]
    DevStatus *st;
    unsigned epoch;
    int self_image;            // 1 for the mixed disk: the reduction adds disk_self_image
    StageOp op;
};

constexpr int kBlock = 256;

// Threads of a symmetric block: kBlock, except 128 at 8 targets per thread, so that a tile
// stays at 1024 vortices and N = 65536 still gives 2080 blocks.  8 targets per thread serve the
// infinite-plane policies only (sym_tpt_supported).
constexpr int sym_block(int tpt) { return tpt >= 8 ? 128 : kBlock; }
constexpr int sym_tile(int tpt) { return sym_block(tpt) * tpt; }
constexpr int kTileF64 = 256;
constexpr int kTileMixed = 128;

// RK4 steps of a whole population in one cooperative launch (pv_persist.hip): block b of
// kPersistBlock threads owns targets [b tb, b tb + tb), tb in {8, 16, 32}; its
// kPersistBlock / tb lane groups split the sources, staged in LDS kPersistChunk at a time.
// ops[4 p + c] is stage c of a step of parity p (p = step & 1), exactly as the per-stage
// launches would be given it; step s evaluates stage 0 at (x[p], y[p]) and writes the new
// state to (x[1-p], y[1-p]).
constexpr int kPersistBlock = 1024;
constexpr int kPersistChunk = 1024;
constexpr int kPersistMaxTb = 32;
struct RkPersistLaunch {
    const double *x[2], *y[2];
    const double *g;
    const double *xa, *ya, *xb, *yb;   // stage positions (written through ops[].xn)
    int n, steps, tb;
    DevStatus *st;
    unsigned epoch0;                   // epoch of stage 0 of step 0; stage c of step s: +4s+c
    unsigned long long barrier_base;   // arrivals counted before this launch
    unsigned long long wait_limit;     // barrier wait bound, in wall-clock ticks
    StageOp ops[8];
};

// Launchers (pv_kernels.hip).  `tpt` is 1, 2 or 4 (launch_sym: also 8 where sym_tpt_supported).
// tile: sources staged in LDS per step, 128 or 256 (pair_tile() gives each precision's default).
void launch_pair(int geometry, int precision, int tpt, int tile, const PairLaunch &a,
                 const GeoParams &gp, hipStream_t s);
int pair_tile(int precision);
// Symmetric path (every geometry and precision): two launches, the pairs then the
// reduce + epilogue.
bool sym_supported(int geometry, int precision);
bool sym_tpt_supported(int geometry, int precision, int tpt);
void launch_sym(int geometry, int precision, int tpt, int tile, const SymLaunch &a,
                const GeoParams &gp, hipStream_t s);
void launch_sym_reduce(const SymLaunch &a, const GeoParams &gp, hipStream_t s);
void launch_make_stage(const StageOp &op, int n, const DevStatus *st, unsigned epoch,
                       hipStream_t s);
void launch_reset_status(DevStatus *st, hipStream_t s);
void launch_classify(const GeoParams &gp, const double *x, const double *y, int n, int tbegin,
                     int tend, DevStatus *st, hipStream_t s);
void launch_moments(const double *x, const double *y, const double *g, int n, double *partials,
                    int nblocks, hipStream_t s);
void launch_energy(const GeoParams &gp, const double *x, const double *y, const double *g, int n,
                   double *partials, int nblocks, DevStatus *st, hipStream_t s);
void launch_dipoles(const GeoParams &gp, bool periodic, bool disk, double thr2, const double *x,
                    const double *y, const double *g, int n, unsigned long long *first,
                    unsigned long long *second, double *d2, unsigned long long cap,
                    DevStatus *st, hipStream_t s);

void launch_copy(const double4 *src, double4 *dst, size_t count, hipStream_t s);

// Blocks of the persistent RK4 kernel that fit on the device at once (0 if unknown).
int persistent_capacity(int geometry, int precision, int cu_count);
hipError_t launch_rk4_persistent(int geometry, int precision, const RkPersistLaunch &a,
                                 const GeoParams &gp, int blocks, hipStream_t s);

constexpr int kReduceBlock = 256;

} // namespace pv
