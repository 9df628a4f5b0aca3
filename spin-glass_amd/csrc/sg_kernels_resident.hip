// sg_kernels_resident.hip — the resident sweep kernels (k_block, k_mid,
// k_big, k_one) and their launchers, compiled from sg_kernels.hip as a
// translation unit of their own (the three parts of sg_kernels.hip build in
// parallel).
#define SG_RESIDENT_PART 1
#include "sg_kernels.hip"
