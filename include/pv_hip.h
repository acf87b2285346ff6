/*
 * pv_hip.h -- C-ABI of point_vortex_hip, the MI355X (gfx950) velocity backend for
 * PointVortex.
 *
 * Plain C types only: pointers to host double arrays, sizes, and small POD structs.
 * Every call returns a pv_status; on failure pv_last_error() holds the message the
 * reference would have thrown (same text where the reference has one).
 *
 * Status codes map one-to-one onto the reference's exception types:
 *   PV_ERR_INVALID_ARGUMENT -> std::invalid_argument
 *   PV_ERR_OUT_OF_RANGE     -> std::out_of_range
 *   PV_ERR_RUNTIME          -> std::runtime_error   (includes HIP API failures, reported
 *                                                     as "<operation>: <hipGetErrorString>")
 *   PV_ERR_LOGIC            -> std::logic_error
 *
 * A handle is used by one host thread at a time (as the reference's VelocityKernel
 * objects are); different handles may be used concurrently.  Each handle owns its
 * device buffers and one HIP stream.
 */
#ifndef PV_HIP_H
#define PV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pv_status {
    PV_OK = 0,
    PV_ERR_INVALID_ARGUMENT = 1,
    PV_ERR_OUT_OF_RANGE = 2,
    PV_ERR_RUNTIME = 3,
    PV_ERR_LOGIC = 4
} pv_status;

/* Boundary conditions of src/params.h:52 (boundaryCondition). */
typedef enum pv_geometry { PV_INFINITE = 0, PV_PERIODIC = 1, PV_DISK = 2 } pv_geometry;

/* Arithmetic of the pair sums.  PV_FP64 is the drop-in (matches the reference's FP64 CPU
 * path to rounding).  PV_MIXED forms pair differences in FP64, evaluates the pair terms in
 * FP32 and folds per-tile FP32 partial sums into FP64 with compensated (TwoSum)
 * accumulation (in the symmetric path: a plain FP64 sum of a block's few tile sums, then
 * TwoSum over the blocks); the disk's self-image term is added in FP64.  It has no
 * counterpart in the reference. */
typedef enum pv_precision { PV_FP64 = 0, PV_MIXED = 1 } pv_precision;

/* Kernel parameters: the geometry fields of SimParams (src/params.h:44-57) that
 * makeReferenceKernel (src/backend_common.cpp:4-13) reads. */
typedef struct pv_config {
    int geometry;          /* pv_geometry */
    int precision;         /* pv_precision */
    double core_radius;    /* coreRadius, infinite plane only */
    double box_length_x;   /* boxLengthX, periodic only */
    double box_length_y;   /* boxLengthY, periodic only (must equal box_length_x) */
    int image_layers;      /* periodicImageLayers, 0..64 */
    double disk_radius;    /* diskRadius, disk only */
    int device;            /* HIP device ordinal */
} pv_config;

/* StepResult (src/compute.h:11-16). */
typedef struct pv_step_result {
    double accepted_time_step;
    double suggested_time_step;
    double normalized_error;
    unsigned rejected_steps;
} pv_step_result;

/* Invariants (src/compute.h:78-81). */
typedef struct pv_invariants {
    double circulation;
    double linear_impulse_x;
    double linear_impulse_y;
    double angular_impulse;
    double hamiltonian;
} pv_invariants;

typedef struct pv_handle pv_handle;

/* Message of the last failed call on this thread (never NULL). */
const char *pv_last_error(void);

/* Replaces makeBackendKernel(params) (src/backend.h:16, src/backend_cuda.cu:550-552): the
 * constructor checks of InfinitePlaneKernel/PeriodicBoxKernel/DiskKernel
 * (src/compute.cpp:50-56, 165-173, 258-261) with the same messages. */
pv_status pv_create(const pv_config *config, pv_handle **out);
void pv_destroy(pv_handle *h);

/* Switch the pair-sum arithmetic of an existing handle (no reference counterpart). */
pv_status pv_set_precision(pv_handle *h, int precision);

/* VelocityKernel::evaluateRange (src/compute.h:24-26, src/compute.cpp:88-133, 203-256,
 * 286-336): pure evaluation of host arrays; u[begin..end), v[begin..end) are written,
 * other entries are left untouched.  Does not disturb the uploaded device state (unlike
 * src/backend_cuda.cu:229-236). */
pv_status pv_evaluate_range(pv_handle *h, const double *x, const double *y, const double *gamma,
                            size_t n, double *u, double *v, size_t begin, size_t end);

/* VelocityKernel::evaluate (src/compute.cpp:60-69): full range plus the non-finite check. */
pv_status pv_evaluate(pv_handle *h, const double *x, const double *y, const double *gamma,
                      size_t n, double *u, double *v);

/* VelocityKernel::hamiltonian (src/compute.cpp:85-87, 174-202, 262-285) of host arrays,
 * computed on the device. */
pv_status pv_hamiltonian(pv_handle *h, const double *x, const double *y, const double *gamma,
                         size_t n, double *out);

/* computeInvariants(vortices, kernel) (src/compute.cpp:157-163) of host arrays. */
pv_status pv_invariants_of(pv_handle *h, const double *x, const double *y, const double *gamma,
                           size_t n, pv_invariants *out);

/* uploadDeviceState (src/compute.h:30, src/backend_cuda.cu:249-269). */
pv_status pv_upload(pv_handle *h, const double *x, const double *y, const double *gamma,
                    size_t n);

/* downloadDeviceState (src/compute.h:31, src/backend_cuda.cu:270-282): n must equal the
 * uploaded population.  gamma may be NULL. */
pv_status pv_download(pv_handle *h, double *x, double *y, double *gamma, size_t n);

/* Number of vortices in the device state. */
pv_status pv_state_size(pv_handle *h, size_t *n);

/* evaluateDeviceState (src/compute.h:32, src/backend_cuda.cu:283-298). */
pv_status pv_evaluate_device_state(pv_handle *h, double *u, double *v, size_t n);

/* deviceRk4Step (src/compute.h:33, src/backend_cuda.cu:299-323): one classical RK4 step of
 * the device state, validated before it returns. */
pv_status pv_rk4_step(pv_handle *h, double dt);

/* Enqueue `steps` RK4 steps of the device state with one synchronisation at the end (no
 * reference counterpart; the reference synchronises after every evaluation).  On failure
 * the message names the first failing step; the state is that of the last good step. */
pv_status pv_rk4_advance(pv_handle *h, double dt, int steps);

/* deviceDopri5Step (src/compute.h:34, src/backend_cuda.cu:324-357, src/timestep.cpp:62-119):
 * FSAL, max-norm error estimate on the device, the reference's controller and rejection
 * rule. */
pv_status pv_dopri5_step(pv_handle *h, double dt, double atol, double rtol, double min_dt,
                         double max_dt, pv_step_result *out);

/* invalidateDeviceDerivative (src/compute.h:35, src/backend_cuda.cu:358). */
void pv_invalidate_fsal(pv_handle *h);

/* computeInvariants of the device state (moments and the FP64 Hamiltonian log-sum on the
 * device; replaces the CPU Hamiltonian of src/backend_cuda.cu:245-247). */
pv_status pv_state_invariants(pv_handle *h, pv_invariants *out);

/* The O(N^2) search of DipoleManager::process (src/dipole.cpp:87-119) on the device state:
 * opposite-sign nonzero pairs i<j closer than `distance` (periodic: minimum image), plus in
 * a disk the real/image wall candidate written as (i, n).  Candidates come back unordered;
 * the caller sorts them (src/dipole.cpp:120-125).  *count receives the number found, which
 * may exceed cap (then only cap entries are written). */
pv_status pv_dipole_candidates(pv_handle *h, double distance, uint64_t *first, uint64_t *second,
                               double *distance_squared, size_t cap, size_t *count);

/* ReinjectionMode (src/params.h:10). */
enum pv_reinjection { PV_REINJECT_NONE = 0, PV_REINJECT_INDEPENDENT = 1, PV_REINJECT_PAIRED = 2 };

/* DipoleManager::process (src/dipole.cpp:42-125) on the resident state: the candidate search
 * above, then on the host the closest-first disjoint selection, the compaction (survivors in
 * index order) and the reinjection (src/dipole.cpp:127-176) with the handle's generator; the
 * new state (its size can change) replaces the resident one and the FSAL derivative is
 * invalidated.  *events receives the number of removed pairs (wall events included).  The
 * driver would call it where it calls DipoleManager::process on host vectors. */
pv_status pv_dipole_process(pv_handle *h, double distance, int reinjection, size_t *events);

/* The reinjection generator: DipoleManager's constructors (src/dipole.cpp:29-38): seeded with
 * seed ^ 0xd1b54a32d192ed03, or restored from `state` (the text std::mt19937_64 writes, as
 * stored in checkpoints, src/checkpoint.cpp:66-68) when it is non-empty.
 * pv_dipole_random_state writes that text (DipoleManager::state, src/dipole.cpp:178-182);
 * *needed receives its length plus one. */
pv_status pv_dipole_set_random(pv_handle *h, uint64_t seed, const char *state);
pv_status pv_dipole_random_state(pv_handle *h, char *buffer, size_t length, size_t *needed);

/* backendRuntimeDetails (src/backend.h:14): newline-separated "key value" lines. */
pv_status pv_runtime_details(int device, char *buffer, size_t length);

/* Device-resident velocity evaluation for benchmarks: evaluates the uploaded state into
 * device buffers and does not copy back or synchronise.  pv_synchronize waits and reports
 * any sticky failure. */
pv_status pv_evaluate_resident(pv_handle *h, int repeats);
pv_status pv_synchronize(pv_handle *h);

/* Launch shapes of the pair kernels (no reference counterpart; the CUDA backend used one
 * fixed shape, src/backend_cuda.cu:402-416).  A shape is targets per thread (1, 2 or 4; 8 on the
 * symmetric path of the infinite plane, 4 elsewhere; 0 = automatic), source splits (0 = automatic), the LDS source tile (128 or 256; 0 = the
 * precision's default) and whether whole-population evaluations use the symmetric path (each
 * unordered pair computed once; 0 or 1; both precisions).  Precedence: pv_set_launch /
 * pv_set_launch_shape (symmetric -1 = not forced), then the first matching rule of the handle's
 * launch table, then a built-in heuristic (split path).  A new handle's table is read from the
 * file named by $PV_HIP_LAUNCH_TABLE, else from launch_table.txt beside libpvhip.so when present
 * (lines "geometry precision n_max tpt splits tile [symmetric]", as written by the autotuner,
 * pointvortex_amd/autotune.py).  pv_resolve_launch reports the shape an evaluation of
 * `targets` targets against n sources would use. */
typedef struct pv_launch_rule {
    int geometry;            /* pv_geometry */
    int precision;           /* pv_precision */
    size_t n_max;            /* applies to populations of at most n_max vortices */
    int targets_per_thread;
    int source_splits;
    int tile;
    int symmetric;
} pv_launch_rule;

pv_status pv_set_launch(pv_handle *h, int targets_per_thread, int source_splits);
pv_status pv_set_launch_shape(pv_handle *h, int targets_per_thread, int source_splits,
                              int tile, int symmetric);
pv_status pv_set_launch_table(pv_handle *h, const pv_launch_rule *rules, size_t count);
pv_status pv_resolve_launch(pv_handle *h, size_t targets, size_t n, int *targets_per_thread,
                            int *source_splits, int *tile, int *symmetric);

/* Persistent RK4 path for small populations (no reference counterpart; it serves
 * deviceRk4Step, src/compute.h:33, and pv_rk4_advance).  pv_rk4_step and pv_rk4_advance then
 * run all their stages in one cooperative launch: each block owns targets_per_block targets
 * (8, 16 or 32; 0 = 16) for the whole run and the blocks exchange stage positions through HBM
 * and a grid barrier, instead of one launch per stage.  Results agree with the per-stage
 * launches to rounding (each target's sum runs in another order); failures, their messages
 * and the state kept after one are the same.  mode -1 (default): used for N <= 4096 unless a
 * launch shape was forced with pv_set_launch / pv_set_launch_shape; 1: used whenever every
 * block fits on the device at once (up to N = 65536); 0: never.  When the runtime refuses the
 * cooperative launch (another process holds the device's cooperative resources) the handle
 * keeps to the per-stage launches.  A barrier wait longer than
 * max_wait_seconds (0 = 2 s) ends the launch with PV_ERR_RUNTIME instead of waiting on.
 * pv_persistent_blocks gives the blocks a step of n vortices would launch (0: per-stage
 * launches). */
pv_status pv_set_persistent(pv_handle *h, int mode, int targets_per_block,
                            double max_wait_seconds);
pv_status pv_persistent_blocks(pv_handle *h, size_t n, int *blocks);

/* Kernel timing for benchmarks (no reference counterpart): between begin and end every
 * velocity-kernel launch of the handle is bracketed by HIP events recorded on the handle's own
 * stream; end waits and returns the summed kernel time and the number of launches. */
pv_status pv_profile_begin(pv_handle *h);
pv_status pv_profile_end(pv_handle *h, double *kernel_ms, int *launches);

/* Measurement helper (no reference counterpart): copies `bytes` bytes between two device
 * buffers with a plain grid-stride copy kernel, `repeats` times, and returns the average
 * kernel time.  Used to calibrate the HBM-traffic counters (FETCH_SIZE, WRITE_SIZE). */
pv_status pv_calibration_copy(int device, size_t bytes, int repeats, double *ms_per_copy);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the evaluation of
 * pv_evaluate (VelocityKernel::evaluate, src/compute.cpp:60-69) for a population already in
 * device memory, allocating nothing.  x, y, gamma (n each) are read and u, v (n each) written;
 * all five are device pointers.  The split pair kernel's status word, tile tickets and partial
 * sums live in a fixed device array of PV_NOALLOC_SCRATCH_BYTES that each GPU the library runs
 * on holds until the process ends.  n is at most PV_NOALLOC_MAX_VORTICES (larger populations
 * are rejected with PV_ERR_INVALID_ARGUMENT); source splits are capped so that their partial
 * sums fit.  Runs on `stream` (a hipStream_t; NULL is the null stream) and waits for it.
 * Validation and messages are those of pv_evaluate.  Calls on one GPU share its array:
 * calls that overlap on one GPU race. */
#define PV_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
#define PV_NOALLOC_MAX_VORTICES ((size_t)1 << 22)
pv_status pv_evaluate_device_noalloc(const pv_config *config, const double *x, const double *y,
                                     const double *gamma, size_t n, double *u, double *v,
                                     void *stream);

/* ------------------------------------------------------------------------------------
 * Ensembles (no reference counterpart; each instance is what the reference runs as one
 * process).  Instances of uneven N and any geometry share flat device arrays and advance
 * together, one persistent launch per RK stage.
 * ---------------------------------------------------------------------------------- */

typedef struct pv_batch pv_batch;

/* Geometry of one instance (the SimParams fields of pv_config). */
typedef struct pv_instance_config {
    int geometry;
    int image_layers;
    double core_radius;
    double box_length;
    double disk_radius;
} pv_instance_config;

/* Device-side state of one adaptive instance, as read back by pv_batch_poll. */
typedef struct pv_slot_state {
    int slot;
    int state;             /* 0 running, 1 finished, 2 failed, 3 idle */
    double t, dt, last_error;
    unsigned accepted, rejected;
    unsigned failure;      /* failure kind bits; see pv_batch_failure_message */
    size_t n;
} pv_slot_state;

/* Room for up to `capacity` vortices and `max_instances` instances.  The per-vortex arrays
 * (208 bytes of HBM per vortex) are allocated as instances are loaded or admitted. */
pv_status pv_batch_create(int device, int precision, size_t capacity, size_t max_instances,
                          pv_batch **out);
void pv_batch_destroy(pv_batch *b);
/* The pair arithmetic of instances.  A fixed-step ensemble switches all its instances; in an
 * ensemble of admitted runs it applies to the runs admitted from then on, each run keeping its
 * own, so FP64 and mixed runs share one ensemble and advance in the same launches. */
pv_status pv_batch_set_precision(pv_batch *b, int precision);
size_t pv_batch_bytes_per_vortex(void);
/* Admission by device footprint (no reference counterpart): with a limit of `bytes` (0: none),
 * pv_batch_admit places a run only if everything the ensemble then holds on the device -- its
 * per-vortex arrays, the work items and partial sums of every admitted run, and the tables --
 * stays within it (otherwise *slot is -1 and the run can be offered again later), and
 * pv_batch_load fails with PV_ERR_RUNTIME if the loaded ensemble would exceed it.
 * pv_batch_device_bytes reports what the ensemble holds on the device now. */
pv_status pv_batch_set_byte_limit(pv_batch *b, size_t bytes);
pv_status pv_batch_device_bytes(pv_batch *b, size_t *bytes);

/* How an ensemble evaluates its pairs (no reference counterpart; the reference runs one
 * system per process).  mode 1: symmetric -- work items are pairs of tiles of one instance
 * (64, 128 or 256 vortices, per pv_batch_set_launch_table), each unordered pair computed once,
 * with per-instance partial sums (16 B per vortex per tile of its instance, on top of the
 * 208 B per vortex) reduced in a fixed order; mode 0: ordered -- 64-target items against all
 * of the instance's sources; mode -1 (default): symmetric for fixed-step ensembles
 * (pv_batch_load), ordered for admitted runs, whose admission budget counts 208 B per vortex.
 * Both give every instance a result that does not depend on the rest of the ensemble.
 * pv_batch_symmetric reports the mode the current work items use. */
pv_status pv_batch_set_symmetric(pv_batch *b, int mode);
int pv_batch_symmetric(const pv_batch *b);

/* Launch shapes of the symmetric decomposition (no reference counterpart): per (geometry,
 * precision, N-bucket), the targets each lane holds (1, 2 or 4: tiles of 64, 128 or 256
 * vortices; 0 = 2) and the most work items a wave takes per counter increment (1..64; 0 = 4).
 * The first rule covering an instance (n <= n_max) decides; instances no rule covers take the
 * defaults.  The items of each (tile, geometry) group run in a launch of their own.  A new
 * ensemble reads its rules from $PV_HIP_ENSEMBLE_TABLE, else from ensemble_table.txt beside
 * libpvhip.so when present (lines "geometry precision n_max targets_per_lane chunk_items", as
 * written by pointvortex_amd/autotune.py --ensemble).  Results depend on an instance's own
 * rule only: a run gives the same bits alone and in a mix.  pv_batch_resolve_launch reports
 * the shape an instance would take; pv_batch_instance_bytes the device bytes one admitted run
 * adds beyond n x pv_batch_bytes_per_vortex() (work items, and with sym != 0 the symmetric
 * items and partial sums at its tile) -- what admission by byte limit counts. */
typedef struct pv_batch_rule {
    int geometry;            /* pv_geometry */
    int precision;           /* pv_precision */
    size_t n_max;            /* applies to instances of at most n_max vortices */
    int targets_per_lane;
    int chunk_items;
} pv_batch_rule;

pv_status pv_batch_set_launch_table(pv_batch *b, const pv_batch_rule *rules, size_t count);
pv_status pv_batch_resolve_launch(pv_batch *b, int geometry, int precision, size_t n,
                                  int *targets_per_lane, int *chunk_items);
pv_status pv_batch_instance_bytes(pv_batch *b, int geometry, int precision, size_t n, int sym,
                                  size_t *bytes);

/* Fixed-step ensembles: replace the contents with `ninst` instances laid out one after the
 * other in x, y, gamma (counts[k] vortices each), advance them with RK4 (optionally as
 * replays of a captured hipGraph), read positions back, or evaluate velocities.
 * Failures are per instance: where the reference's single run throws
 * (src/compute.cpp:60-69, src/timestep.cpp:54-58; backend_cuda.cu:387-393), the failing
 * instance stops at the input state of its failing step and the others go on.
 * pv_batch_failures reports each instance's failure bits (0 = healthy; see
 * pv_batch_failure_message) and, if steps_done is not null, the RK4 steps it completed
 * since pv_batch_load; pv_batch_read returns a failed instance's stopped state and
 * pv_batch_evaluate leaves its velocities unspecified. */
pv_status pv_batch_load(pv_batch *b, const pv_instance_config *configs, const size_t *counts,
                        size_t ninst, const double *x, const double *y, const double *gamma);
pv_status pv_batch_rk4(pv_batch *b, double dt, int steps, int use_graph);
pv_status pv_batch_read(pv_batch *b, double *x, double *y, size_t total);
pv_status pv_batch_evaluate(pv_batch *b, double *u, double *v, size_t total);
pv_status pv_batch_failures(pv_batch *b, unsigned *failures, unsigned long long *steps_done,
                            size_t ninst);

/* Adaptive ensembles with admission while others run (serving).  pv_batch_admit places one
 * run (DOPRI5 from t=0 to t_end) if its vortices fit; *slot is -1 when they do not.
 * pv_batch_attempt performs one DOPRI5 attempt for every running instance with the
 * step-size controllers on the device; pv_batch_attempts performs `count` of them back to back
 * with one synchronisation at the end (a run that finishes or fails part-way sits out the
 * rest); pv_batch_poll reads every occupied slot's state as of the last attempt;
 * pv_batch_fetch copies a slot's positions; pv_batch_release frees the slot. */
pv_status pv_batch_admit(pv_batch *b, const pv_instance_config *config, const double *x,
                         const double *y, const double *gamma, size_t n, double t_end,
                         double dt0, double atol, double rtol, double min_dt, double max_dt,
                         int *slot);
pv_status pv_batch_attempt(pv_batch *b);
pv_status pv_batch_attempts(pv_batch *b, int count);
pv_status pv_batch_poll(pv_batch *b, pv_slot_state *states, size_t cap, size_t *count);
pv_status pv_batch_fetch(pv_batch *b, int slot, double *x, double *y, size_t n);
pv_status pv_batch_release(pv_batch *b, int slot);
const char *pv_batch_failure_message(unsigned failure);

#ifdef __cplusplus
}
#endif

#endif /* PV_HIP_H */
