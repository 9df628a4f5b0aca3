// sg_kernels.h — host-side launch entry points of the HIP kernels of libsg.
#pragma once

#include "sg_common.h"

namespace sg {

// a sweep kernel exists for every even L in [2, 48]: k_block (lattice in
// LDS, couplings in VGPRs) up to L = 16, k_big above
bool sweep_kernel_supported(int L);

// One work-group per task (one lattice, or the lattices of slots t and t+1
// of one group and replica); a task's buffers are those of its job,
// jt[task.job], so one launch may cover many jobs of one L.  With do_pt, the
// task's PT pair is decided and exchanged first; then nsweeps sweeps;
// want_energy = 1 writes the per-lane U of the final states to the energy
// slots, 2 writes U and K.
hipError_t launch_block(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks,
                        int ntasks, int do_pt, int nsweeps, int want_energy, hipStream_t st);

// Measures every (group, temperature) of md: folds the measurement into the
// ObsAcc sums and, with write_records, writes the int32 records to out.
hipError_t launch_measure(int L, int R, const uint64_t *spins, const uint64_t *J,
                          const MeasDesc *md, int nmd, int32_t *out, int nfield, ObsAcc *acc,
                          const double *cs, int write_records, hipStream_t st);

hipError_t launch_init_spins(int L, uint64_t *spins, const LatDesc *descs,
                             int ndesc, hipStream_t st);
hipError_t launch_init_couplings(int L, uint64_t *J, uint64_t j_off, uint32_t g,
                                 uint32_t key0, uint32_t key1, hipStream_t st);

// adds nsweeps to the sweep counter of each of the njobs jobs of jt
hipError_t launch_advance(const JobPtrs *jt, int njobs, int nsweeps, hipStream_t st);

}  // namespace sg
