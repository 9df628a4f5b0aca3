// The resident sweep kernels (k_block, k_mid, k_big, k_one) and their
// launchers, in a translation unit of their own so that they are compiled with
// LLVM's max-ILP machine scheduler (build_sg.py adds the flag for this file).
// At L = 16 (2048 samples, 24 T, 2 R) it takes a k_block launch from 169.1 to
// 167.5 us and the mixed-L batch from 23.5k to 24.2k flips/ns, with the same
// register count (126 VGPRs, no spills); the streamed kernels (k_slab, k_phase)
// measured 4 % slower under it and stay with the default scheduler in
// sg_kernels.hip.
#define SG_RESIDENT_PART 1
#include "sg_kernels.hip"
