// sg_kernels.h — host-side launch entry points of the HIP kernels of libsg.
#pragma once

#include "sg_common.h"

namespace sg {

// a sweep kernel exists for every even L in [2, 48]: k_block (lattice in
// LDS, couplings in VGPRs) up to L = 16, k_mid (two lattices in LDS,
// couplings from the site records) at L = 18, 20, k_big (in place in global
// memory) above
bool sweep_kernel_supported(int L);

// One work-group per task (one lattice, or two lattices of one group: a PT
// pair of slots t and t+1 of one replica, or two lattices left without a
// partner at this PT parity); a task's buffers are those of its job,
// jt[task.job], so one launch may cover many jobs of one L.  With do_pt, the
// task's PT pair is decided and exchanged first; then nsweeps sweeps;
// want_energy = 1 writes the per-lane U of the final states to the energy
// slots, 2 writes U and K.
hipError_t launch_block(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks,
                        int ntasks, int do_pt, int nsweeps, int want_energy, hipStream_t st);

// ---- single-lattice resident sweeps (k_one): L = 22, 24, 26, where one
// lattice fits in LDS and two do not.  Tasks: lat[0] = lattice descriptor,
// lat[1] unused, job; no PT (the caller runs launch_pt_chunks first); then
// nsweeps sweeps and, with want_energy, the energies as in launch_block.
bool single_kernel_supported(int L);
hipError_t launch_single(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int ntasks,
                         int nsweeps, int want_energy, hipStream_t st);

// Measures every (group, temperature) of md: folds the measurement into the
// ObsAcc sums and, with write_records, writes the int32 records to out.
hipError_t launch_measure(int L, int R, const uint64_t *spins, const int32_t *energy,
                          const MeasDesc *md, int nmd, int32_t *out, int nfield, ObsAcc *acc,
                          const double *cs, int write_records, hipStream_t st);

hipError_t launch_init_spins(int L, uint64_t *spins, const LatDesc *descs,
                             int ndesc, hipStream_t st);
hipError_t launch_init_couplings(int L, uint64_t *J, uint64_t j_off, uint32_t g,
                                 uint32_t key0, uint32_t key1, hipStream_t st);

// ---- streamed lattices: one launch per colour (k_phase), PT in its own launch
constexpr int kPhaseNT = 256;     // Philox quads per chunk (per task)
// Tiles of the streamed kernels, picked by the tuner: tile t runs 2^t quads
// (k_phase) or columns (k_slab) per thread, so kPhaseNT / 2^t threads per
// k_phase work-group; the tasks are the same for every tile.
constexpr int kPhaseTiles = 2;
constexpr int kPtWords = 1024;    // words of each lattice per PT work-group
enum : int { kPhUpdate = 1, kPhEnergy = 2, kPhEnergyK = 4, kPhZero = 8 };

inline int phase_chunks(int L) { return (L * L * L / 4 + kPhaseNT - 1) / kPhaseNT; }
inline int pt_chunks(int L) { return (L * L * L + kPtWords - 1) / kPtWords; }
bool phase_kernel_supported(int L);

// Tasks: lat[0] = lattice descriptor, lat[1] = chunk, job.  c = colour, it =
// sweep within the interval (the Philox counter is the job's sweep + it),
// mode = kPh* bits (energy bits with c = 1 only).
hipError_t launch_phase(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int ntasks,
                        int c, int it, int mode, int tile, hipStream_t st);
// ---- slab sweeps of streamed lattices: one launch per sweep (k_slab)
constexpr int kSlabParts = 4;     // slab pairs per lattice (tasks per lattice)
bool slab_kernel_supported(int L);
// Tasks: lat[0] = lattice descriptor, lat[1] = slab pair (>= kSlabParts:
// padding, no work), job.  Reads spins (dir = 0) or spins2 (dir = 1) and
// writes the other buffer; it = sweep within the interval; mode = kPh* bits
// (kPhZero in an earlier launch than the energy bits).
hipError_t launch_slab(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int ntasks,
                       int dir, int it, int mode, int tile, hipStream_t st);

// ---- split sweeps: a lattice (L = 32, 48) held on chip by P work-groups
// for a whole PT interval, exchanging boundary planes through jt's xbuf /
// xflag (k_split).  split_slabs: the P used for n_lattices lattices of size
// L (0: no split kernel for this L); P > 1 needs all n_lattices * P
// work-groups of a launch resident at once (one per CU), so a launch takes
// at most split_batch(P) lattices (more run as consecutive launches) and
// callers never run two split launches at the same time.
int split_slabs(int L, int n_lattices);
constexpr int split_batch(int P) { return 256 / P; }
constexpr int kSplitTicketWords = 8;   // one ticket per slab pair (P <= 7) and a done count
// Tasks: lat[0] = lattice descriptor, one per lattice; ticket:
// kSplitTicketWords zeroed uint32 counters kept for this task list (the
// launch leaves them zeroed);
// nsweeps >= 1; want_energy as in launch_block.  PT runs before the launch.
// split_tiles(L): the tiles (threads per work-group) the split kernel has
// at L (0: none).
int split_tiles(int L);
hipError_t launch_split(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int nlat,
                        int P, uint32_t *ticket, int nsweeps, int want_energy, int tile,
                        hipStream_t st);

// Tasks: lat[0] = PairDesc index, lat[1] = chunk, job.
hipError_t launch_pt_chunks(const JobPtrs *jt, const TaskDesc *tasks, int ntasks, hipStream_t st);
// coupling site records of one group from its [3][N] couplings
hipError_t launch_build_jsite(int L, uint64_t *Js, const uint64_t *J, uint64_t j_off,
                              hipStream_t st);

// p[0..n) = v (a kernel, so that it may sit in a captured launch sequence)
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, int n, hipStream_t st);

// adds nsweeps to the sweep counter of each of the njobs jobs of jt
hipError_t launch_advance(const JobPtrs *jt, int njobs, int nsweeps, hipStream_t st);

}  // namespace sg
