// sg_kernels.h — host-side launch entry points of the HIP kernels of libsg.
#pragma once

#include "sg_common.h"

namespace sg {

// Per-job device state read by the kernels, so that a launch sequence can be
// captured once in a hipGraph and replayed while the counters advance.
struct JobState {
  uint32_t sweep;     // sweeps completed
  uint32_t pt_every;  // PT cadence (0 = none)
  uint32_t pad[2];
};

// a sweep kernel exists for every even L in [2, 48]: k_block (lattice in
// LDS, couplings in VGPRs) up to L = 16, k_big above
bool sweep_kernel_supported(int L);

// One work-group per task (one lattice, or the lattices of slots t and t+1
// of one group and replica): with do_pt, the task's PT pair is decided and
// exchanged first; then nsweeps sweeps; want_energy = 1 writes the per-lane
// U of the final states to the energy slots, 2 writes U and K.
hipError_t launch_block(int L, bool field, uint64_t *spins, const uint64_t *J,
                        const LatDesc *descs, const TaskDesc *tasks, int ntasks,
                        const PairDesc *pairs, const ThrRec *thr, const JobState *jobs,
                        int do_pt, int nsweeps, int want_energy, int32_t *energy,
                        unsigned long long *counters, hipStream_t st);

// PT step of every pair whose parity matches the job's current step; pairs
// of the other parity return at once.
hipError_t launch_pt(uint64_t *spins, int32_t *energy, const PairDesc *pairs,
                     int npairs, const JobState *jobs,
                     unsigned long long *counters, hipStream_t st);

// Measures every (group, temperature) of md: folds the measurement into the
// ObsAcc sums and, with write_records, writes the int32 records to out.
hipError_t launch_measure(int L, int R, const uint64_t *spins, const uint64_t *J,
                          const MeasDesc *md, int nmd, int32_t *out, int nfield, ObsAcc *acc,
                          const double *cs, int write_records, hipStream_t st);

hipError_t launch_init_spins(int L, uint64_t *spins, const LatDesc *descs,
                             int ndesc, hipStream_t st);
hipError_t launch_init_couplings(int L, uint64_t *J, uint64_t j_off, uint32_t g,
                                 uint32_t key0, uint32_t key1, hipStream_t st);

hipError_t launch_advance(JobState *jobs, int njobs, int nsweeps, hipStream_t st);

}  // namespace sg
