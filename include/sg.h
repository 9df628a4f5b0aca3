/*
 * sg.h — C-ABI of libsg, the MI355X-native parallel-tempering engine for the
 * 3D Edwards–Anderson ±J spin glass.
 *
 * The reference (Marco-Mendivil-Carboni/Spin-Glass) exposes its hot path
 * through the C++ class `eamsim` (Program/src/eamsim.cuh:66-110) driven by
 * `main` (Program/src/main.cu:8-84).  There is no FFI in the reference; each
 * entry point below replaces one member of that class, as cited.  No C++
 * exception and no torch type crosses this boundary: every call returns an
 * sg_status and sg_last_error() gives the message (the reference throws
 * `error`, Program/src/util.cuh:46-54, caught once in main.cu:70-75).
 *
 * Threading: one handle is used by one host thread at a time; different
 * handles may run concurrently (each owns its own HIP stream).
 *
 * Canonical algorithm (the product and the CPU oracle under oracle/ both
 * follow it bit for bit): 64 disorder samples per uint64 word (bit l =
 * sample l of a "group"), spin bit 1 = +1, coupling bit 1 = J=-1; Philox4x32-10
 * counters keyed by (group, replica, temperature slot, sweep, site); integer
 * Metropolis thresholds floor(2^32 exp(-beta dE)); PT swaps of adjacent
 * temperature slots, even pairs on even PT steps and odd pairs on odd steps.
 */
#ifndef SG_H
#define SG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sg_status {
  SG_OK = 0,
  SG_ERR_INVALID = 1,   /* bad parameter or argument */
  SG_ERR_DEVICE = 2,    /* HIP runtime error */
  SG_ERR_NOMEM = 3,     /* allocation failed */
  SG_ERR_FORMAT = 4,    /* checkpoint size / magic / checksum mismatch */
  SG_ERR_STATE = 5      /* call out of order (e.g. run before init) */
} sg_status;

typedef struct sg_handle sg_handle;

/* Run parameters.  They replace the reference's compile-time constants
 * (Program/src/eamsim.cuh:13-25) and its constructor argument H
 * (eamsim.cuh:73).  Zero-initialise, then fill; sg_params_default() fills the
 * reference's own values (L=16, 2048 samples, 2 copies, 24 temperatures of
 * the ladder beta_i = 2*2^(-4i/23), PT every 16, measurement every 1024). */
typedef struct sg_params {
  int32_t L;             /* even, 2..48 */
  int32_t n_samples;     /* disorder samples, packed 64 per group */
  int32_t n_temps;       /* temperature slots */
  int32_t n_replicas;    /* real replicas per sample ("copies", >= 2 for q) */
  const double *betas;   /* n_temps inverse temperatures, slot order; NULL: use T_min/T_max */
  double T_min, T_max;   /* geometric temperature ladder (slot 0 = T_min) if betas == NULL */
  double H;              /* external field, 0 <= H <= 4 (eamsim.cu:520) */
  uint64_t coupling_seed;
  uint64_t spin_seed;    /* initial spins, Metropolis and PT random numbers */
  int32_t pt_every;      /* sweeps between PT steps (0 = no PT) */
  int32_t meas_every;    /* sweeps between measurements (0 = none) */
  int32_t device;        /* HIP device ordinal */
  int32_t form;          /* SG_FORM_* */
  int32_t kernel;        /* SG_KERNEL_*: how the lattices are held during sweeps */
  int32_t tile;          /* threads-per-work-group variant of the shape (sg_tune_tiles):
                            0 = the tuner's pick with SG_KERNEL_AUTO, else the first;
                            t >= 1 = tile t - 1 (SG_ERR_INVALID if the shape lacks it).
                            Batch-engine jobs run their class's tile: the tuner's
                            cached one for SG_KERNEL_AUTO, else the first */
} sg_params;

enum {
  SG_FORM_PLAIN = 0,      /* per-interval launches, one group at a time */
  SG_FORM_BATCHED = 1     /* all groups per launch, interval captured in a hipGraph */
};

enum {
  SG_KERNEL_AUTO = 0,     /* chosen per L (and lattice count) from measured rates */
  SG_KERNEL_RESIDENT = 1, /* lattice pair resident in one work-group for a PT interval */
  SG_KERNEL_STREAMED = 2, /* one launch per colour, lattices streamed in chunks (L >= 12) */
  SG_KERNEL_RESIDENT_SINGLE = 3, /* one lattice per work-group (twice the work-groups of
                                   RESIDENT), PT in a launch of its own; the same as
                                   RESIDENT at L = 22..26, whose kernel holds one lattice */
  SG_KERNEL_SPLIT = 4     /* L = 32 and 48: a lattice held on chip by 4
                             work-groups for a PT interval, exchanging boundary
                             planes (all work-groups of a launch run at once: 64
                             lattices a launch, more as consecutive launches); PT
                             in a launch of its own.  Not used by the batch
                             engine */
};

/* Per-measurement callback: rec holds n_groups*n_temps*64 records of nfield
 * int32 each (layout [group][temp][lane][field]); see sg_record_layout. */
typedef void (*sg_obs_sink)(void *user, int64_t sweep, const int32_t *rec,
                            int64_t n_records, int32_t nfield);

/* replaces the reference's constants and ladder, eamsim.cuh:13-25 and
 * eamsim::init_repib (eamsim.cu:725-744): L = 16, 2048 samples, 2 copies,
 * the 24-temperature ladder, PT every 16, a record every 1024, the batched
 * form (every group in one launch per PT interval; SG_FORM_PLAIN gives the
 * same bits one group at a time) and SG_KERNEL_AUTO */
void sg_params_default(sg_params *p);

/* replaces eamsim::eamsim(float H) (eamsim.cu:516-543): validates, allocates
 * the device state, builds threshold tables */
sg_status sg_create(const sg_params *p, sg_handle **out);

/* replaces eamsim::~eamsim() (eamsim.cu:546-559) */
sg_status sg_destroy(sg_handle *h);

/* message of the last failed call on this handle (or of the last failed
 * sg_create when h == NULL) */
const char *sg_last_error(const sg_handle *h);

/* replaces eamsim::init_lattice() (eamsim.cu:622-652): couplings from
 * coupling_seed (shared by all replicas of a sample), hot-start spins from
 * spin_seed; resets the sweep counter */
sg_status sg_init_lattice(sg_handle *h);

/* replaces eamsim::run_simulation(std::ofstream&) (eamsim.cu:679-722):
 * n_sweeps sweeps with PT and measurements on the configured cadence; each
 * measurement goes to sink (may be NULL) and into the handle's FP64
 * accumulators */
sg_status sg_run(sg_handle *h, int64_t n_sweeps, sg_obs_sink sink, void *user);

/* no counterpart in the reference (whose launches are not captured): builds
 * every captured launch sequence (hipGraph) that a call sg_run(h, n_sweeps,
 * ...) starting at sweep from_sweep (the current sweep when negative) would
 * replay, and runs nothing, so that a timed sg_run afterwards includes no
 * capture or instantiation.  sg_graph_count gives the number of instantiated
 * graphs the handle holds (-1 for NULL). */
sg_status sg_prepare(sg_handle *h, int64_t from_sweep, int64_t n_sweeps);
int32_t sg_graph_count(const sg_handle *h);

/* replaces eamsim::save_checkpoint / load_checkpoint (eamsim.cu:655-676).
 * The image carries a versioned header (L, sizes, ladder, seeds, sweep
 * counter, accumulators) and a checksum; load rejects a size, magic,
 * parameter or checksum mismatch.  save with buf == NULL returns the size. */
sg_status sg_save_checkpoint(sg_handle *h, void *buf, size_t *size);
sg_status sg_load_checkpoint(sg_handle *h, const void *buf, size_t size);

/* --- inspection (no counterpart in the reference: used by the CLI, tests
 *     and the analysis writer) ------------------------------------------- */

int32_t sg_nfield(const sg_handle *h);            /* 2R + 1 + 3L */
int32_t sg_n_groups(const sg_handle *h);
int64_t sg_sweep_count(const sg_handle *h);
sg_status sg_get_betas(const sg_handle *h, double *betas);            /* [T] */
/* spins in natural site order s = x + L(y + Lz), [group][replica][temp][N] */
sg_status sg_get_spins(sg_handle *h, uint64_t *out);
sg_status sg_set_spins(sg_handle *h, const uint64_t *in);
/* couplings [group][3][N], bit 1 = antiferromagnetic bond s -> s + e_d */
sg_status sg_get_couplings(sg_handle *h, uint64_t *out);
sg_status sg_set_couplings(sg_handle *h, const uint64_t *in);
/* per-lane unsatisfied-bond count U and up-spin count K of every lattice
 * [group][replica][temp][64] */
sg_status sg_get_energies(sg_handle *h, int32_t *U, int32_t *K);
/* accepted / attempted PT swaps per adjacent pair (t, t+1), summed over
 * groups, replicas and lanes, [T-1] each */
sg_status sg_get_swap_counts(sg_handle *h, uint64_t *accepted, uint64_t *attempted);
/* FP64 per-temperature results over all measurements so far and all active
 * samples: [T][SG_NRES] with the columns below */
enum {
  SG_RES_BETA = 0, SG_RES_E, SG_RES_M, SG_RES_Q, SG_RES_Q2, SG_RES_Q4,
  SG_RES_BINDER, SG_RES_CHI0, SG_RES_CHI1, SG_RES_XI_L, SG_RES_NMEAS,
  SG_NRES
};
sg_status sg_get_results(sg_handle *h, double *res);
/* device time of sg_run calls since creation (ms, HIP events on the handle's
 * stream) and the spin-flip attempts they performed */
sg_status sg_get_timing(sg_handle *h, double *kernel_ms, double *flips);
/* per-kernel timing: with profiling on, every sweep-kernel launch is
 * bracketed by HIP events on the handle's stream (the batched form then
 * launches directly instead of through its hipGraph).  kind 0 = sweep kernel,
 * 1 = PT exchange, 2 = measurement.  Returns launches, summed device ms and
 * the spin-flip attempts of those launches. */
sg_status sg_set_profiling(sg_handle *h, int32_t on);
sg_status sg_get_kernel_stats(sg_handle *h, int32_t kind, int64_t *launches,
                              double *total_ms, double *flips);
/* the autotuner behind SG_KERNEL_AUTO (no counterpart in the reference): the
 * work-group shape (SG_KERNEL_RESIDENT, _RESIDENT_SINGLE or _STREAMED) it
 * picks for n_lattices lattices of size L on a device, with the measured
 * device time of the lattice-pair and the streamed shapes over the same 64
 * sweeps (0 where the L has no such shape).  The first call for a (device,
 * L, field, power-of-two lattice count) measures every shape the L has;
 * later ones are answered from the process-wide cache.  Errors:
 * sg_last_error(NULL). */
sg_status sg_tune_kernel(int32_t device, int32_t L, int32_t n_lattices, int32_t field,
                         int32_t *kernel, double *resident_ms, double *streamed_ms);
/* the same decision with the time of every shape: ms[k] for k =
 * SG_KERNEL_RESIDENT, _STREAMED, _RESIDENT_SINGLE, _SPLIT (ms[0] unused; 0
 * where the L has no such shape) */
sg_status sg_tune_shapes(int32_t device, int32_t L, int32_t n_lattices, int32_t field,
                         int32_t *kernel, double ms[5]);
/* the same decision down to the tile: each shape is timed at each of its
 * tiles (threads per work-group: the streamed kernels' 1 or 2 quads or
 * columns per thread, the split kernel's 1024 or 512 threads at L = 32);
 * *kernel and *tile (0-based, as sg_params.tile - 1) receive the fastest,
 * ms[k * 4 + t] the device time of shape k at tile t over 64 sweeps (0
 * where the L has no such shape or tile).  Cached with sg_tune_shapes. */
sg_status sg_tune_tiles(int32_t device, int32_t L, int32_t n_lattices, int32_t field,
                        int32_t *kernel, int32_t *tile, double ms[20]);
/* the shape and 0-based tile a handle sweeps with */
sg_status sg_get_shape(const sg_handle *h, int32_t *kernel, int32_t *tile);

/* No counterpart in the reference (it has no entry point of this kind): for
 * callers that cannot allocate device memory — code under HIP graph capture,
 * serving loops and harnesses that forbid hipMalloc.  The sweeps with PT of
 * sg_run (same parameters, same bits; measurements are not taken, so
 * p->meas_every must be 0) over lattices the caller holds in device memory,
 * allocating nothing: the working buffers (coupling site records,
 * descriptors, thresholds, task lists, sweep counter) live in one static
 * device array of SG_STATIC_SCRATCH_BYTES per device, held until the process
 * ends.  Sizes whose buffers exceed it are rejected (SG_ERR_INVALID); that is
 * G * 6 * L^3 * 8 bytes of site records plus about 100 bytes per lattice (L =
 * 16: about 110 groups of 64 samples at 24 temperatures and 2 replicas).
 *   spins       in/out, device, [group][replica][temp][N] uint64 in the
 *               kernels' order: colour c = (x+y+z)&1 at c*N/2, then
 *               (x>>1) + (L/2)(y + L z)  (sg_get_spins gives natural order)
 *   couplings   in, device, [group][3][N] natural order (as sg_get_couplings)
 *   sweep0      the sweep counter at entry (Philox counters and PT parity)
 *   energies    out, device, [group][replica][temp][128] int32: U in 0..63,
 *               K in 64..127 of the final states (as sg_get_energies)
 *   swaps_accepted out, device, [T-1] uint64: accepted swaps of this call
 *               (may be NULL when n_temps == 1)
 *   stream      hipStream_t to queue the work on (NULL: the null stream);
 *               the call returns once the work is queued.
 * p->coupling_seed, form and kernel are not used; p->device selects the
 * device (made current).  Calls that overlap on one device race: they share
 * its array.  Errors: sg_last_error(NULL). */
#define SG_STATIC_SCRATCH_BYTES (32u << 20)
sg_status sg_run_static(const sg_params *p, uint64_t *spins, const uint64_t *couplings,
                        int64_t sweep0, int64_t n_sweeps, int32_t *energies,
                        uint64_t *swaps_accepted, void *stream);

/* tuner measurements run in this process so far (each sg_tune_* call for an
 * unmeasured (device, L, field, bucket), and each sg_create with
 * SG_KERNEL_AUTO for one).  The batch engine never measures: its classes take
 * the cached result, or for an unmeasured bucket the rule the measurements
 * support (pairs from 512 lattices at L <= 16 and 256 at L = 18, 20, single
 * lattices below; the one-lattice kernel at L = 22..26; streamed above). */
int64_t sg_tune_count(void);

/* VALU issue-rate probe (no counterpart in the reference; measurement only,
 * for the roofline of the sweep kernels): one launch that fills every CU
 * (2048 work-groups of 256 threads) with 16 independent chains per thread of
 * one instruction — kind 0 v_bitop3_b32, kind 1 v_mad_u64_u32 — for iters
 * steps (even, >= 2), after one untimed launch.  Returns the lane-operations
 * the timed launch issued and its HIP-event duration in ms.  Errors:
 * sg_probe_last_error(). */
sg_status sg_probe_valu(int32_t device, int32_t kind, int32_t iters, double *lane_ops,
                        double *ms);
const char *sg_probe_last_error(void);

/* --- batch and serving engine -------------------------------------------
 * No counterpart in the reference, which runs one fixed-size simulation per
 * process (Program/src/main.cu:8-84).  An engine holds many jobs (each a
 * simulation as above, of any L, ladder, seeds and sweep count) on one
 * device, their state resident in a device arena between steps.  Jobs are
 * admitted in submission order while their footprint fits the arena.  The
 * jobs of one (L, field) class share a stream and, per PT interval, one
 * sweep-kernel launch (SG_FORM_BATCHED) or one launch per job
 * (SG_FORM_PLAIN); each step queues slice_pt_steps PT intervals of every
 * class whose stream holds fewer than max_inflight unfinished slices, so
 * classes of short intervals run ahead of slower ones.  A job's results are bit-identical
 * to those of the same parameters run alone through sg_run: every random
 * number is keyed by the job's own seeds and counters.  One host thread per
 * engine. */
typedef struct sg_batch sg_batch;

typedef struct sg_batch_params {
  int32_t device;
  int32_t pt_every;        /* PT cadence of every job, sweeps (> 0) */
  int32_t slice_pt_steps;  /* PT intervals per sg_step (>= 1) */
  int32_t form;            /* SG_FORM_BATCHED or SG_FORM_PLAIN, as above */
  int64_t arena_bytes;     /* device memory for job state */
  int32_t max_inflight;    /* slices queued per class on the device ahead of the host (>= 1) */
  int32_t pad;
} sg_batch_params;

enum { SG_JOB_QUEUED = 0, SG_JOB_RUNNING = 1, SG_JOB_DONE = 2 };

void sg_batch_params_default(sg_batch_params *bp);
sg_status sg_batch_create(const sg_batch_params *bp, sg_batch **out);
sg_status sg_batch_destroy(sg_batch *b);
/* message of the last failed call on this engine (or of the last failed
 * sg_batch_create when b == NULL) */
const char *sg_batch_last_error(const sg_batch *b);
/* device bytes a job with these parameters takes from the arena */
int64_t sg_job_footprint(const sg_params *p);
/* queues a job of n_sweeps sweeps.  p->pt_every must be 0 or the engine's;
 * n_sweeps and p->meas_every must be multiples of the engine's pt_every;
 * p->form and p->device are ignored.  SG_ERR_NOMEM when the footprint
 * exceeds the whole arena.  *id >= 0 names the job from then on. */
sg_status sg_submit(sg_batch *b, const sg_params *p, int64_t n_sweeps, int64_t *id);
/* one step: retire finished jobs, admit queued ones, queue one slice of
 * every class with room (waiting for a slot when none has).  Returns the
 * counts after the step (either may be NULL). */
sg_status sg_step(sg_batch *b, int32_t *n_running, int32_t *n_queued);
/* steps until every submitted job is done */
sg_status sg_drain(sg_batch *b);
sg_status sg_poll(sg_batch *b, int64_t id, int32_t *state);
/* [T][SG_NRES] results and [T-1] accepted swap counts (may be NULL) of a
 * finished job; the job is released and its id becomes invalid */
sg_status sg_result(sg_batch *b, int64_t id, double *res, uint64_t *swaps_accepted);
/* host wall-clock seconds since the engine's creation: submitted (or the
 * arrival time given to sg_serve), admitted, finished (seen done by the host) */
sg_status sg_job_times(sg_batch *b, int64_t id, double *t_submit, double *t_admit,
                       double *t_done);
/* spin-flip attempts launched, steps run, arena bytes in use and capacity */
sg_status sg_batch_stats(sg_batch *b, double *flips, int64_t *steps, int64_t *arena_in_use,
                         int64_t *arena_capacity);
/* batched form: class PT intervals launched directly and through replays of
 * captured hipGraphs, and the graphs instantiated.  A class captures its
 * intervals once its task tables have served 8 intervals unchanged; the
 * graphs of a class are dropped when a job joins or leaves it. */
sg_status sg_batch_launch_stats(sg_batch *b, int64_t *intervals_direct,
                                int64_t *intervals_graphed, int64_t *graphs_built);
/* Host-side allocations of the engine beyond its arena: growths of the
 * classes' task tables (each a hipMalloc of twice the size; outgrown tables
 * are freed with the batch, never while it runs) and pinned blocks allocated
 * for the jobs' final copies (retired jobs give theirs back for reuse).  No
 * counterpart in the reference. */
sg_status sg_batch_alloc_stats(sg_batch *b, int64_t *table_growths, int64_t *pinned_blocks);

/* Serving loop: jobs[i] arrives arrival_s[i] seconds after the call starts
 * (non-decreasing); each is submitted when the host clock reaches its
 * arrival, sg_step runs while any job is queued or running, and the loop
 * returns once all are done.  ids[i] receives the job id (results stay
 * available through sg_result); wall_s the loop's duration. */
sg_status sg_serve(sg_batch *b, const sg_params *jobs, const int64_t *n_sweeps,
                   const double *arrival_s, int32_t n_jobs, int64_t *ids, double *wall_s);

#ifdef __cplusplus
}
#endif
#endif
