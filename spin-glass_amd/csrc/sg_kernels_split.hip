// sg_kernels_split.hip — the split sweep kernel (k_split: a lattice held on
// chip by several work-groups that exchange boundary planes) and its
// launcher, compiled from sg_kernels.hip as a translation unit of their own.
#define SG_SPLIT_PART 1
#include "sg_kernels.hip"
