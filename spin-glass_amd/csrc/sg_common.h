// sg_common.h — definitions shared by the HIP kernels and the host runtime
// of libsg: Philox4x32-10, the deterministic exp used for every acceptance
// threshold, the bit layout and the launch descriptors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {

// ------------------------------------------------------------------ layout
// A lattice of one (group, replica, temperature slot) is N = L^3 uint64
// words in checkerboard-split order: colour c = (x+y+z)&1 at word offset
// c*N/2, then h = (x>>1) + (L/2)*(y + L*z).  Bit l of a word is sample l of
// the group.  Couplings of a group are [3][N] words in natural order
// s = x + L*(y + L*z), bit 1 = antiferromagnetic bond s -> s+e_d.
constexpr int kMaxL = 48;

// Philox counter word 3: tag<<28 | (rho & 0xFFF)<<16 | (t & 0xFFFF)
enum : uint32_t { kTagCoup = 1, kTagSpin = 2, kTagMC = 3, kTagPT = 4 };

__host__ __device__ inline uint32_t stream_word(uint32_t tag, uint32_t rho, uint32_t t) {
  return (tag << 28) | ((rho & 0xFFFu) << 16) | (t & 0xFFFFu);
}

// Philox4x32-10 (Salmon et al., SC'11).  One call yields four 32-bit words.
struct u32x4 { uint32_t v[4]; };

__host__ __device__ inline u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2,
                                        uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  u32x4 r;
  r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
  return r;
}

// exp(x) for x <= 0 as a fixed sequence of IEEE double operations, so the
// host and the GPU produce the same bits: reduction x = k ln2 + r with ln2
// split into a head with trailing zero bits and a tail, Taylor series of e^r
// to degree 14 by Horner, exact scaling by 2^k from the exponent field.
// Below -40 it returns 0 (floor(2^32 e^x) is 0 there), so no subnormal
// result is ever formed.  Contraction into FMA is disabled.
__host__ __device__ inline double det_exp(double x) {
#pragma clang fp contract(off)
  if (!(x <= 0.0)) return 1.0;
  if (x < -40.0) return 0.0;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double kd = floor(x * inv_ln2 + 0.5);
  const double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double p = 1.0 / 87178291200.0;
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  const int k = (int)kd;
  const uint64_t bits = (uint64_t)(k + 1023) << 52;
  double two_k;
  __builtin_memcpy(&two_k, &bits, sizeof(double));
  return p * two_k;
}

// Swap acceptance of one lane: x = (beta_a - beta_b) * dE; accept iff x >= 0
// or r < floor(2^32 e^x).
__host__ __device__ inline bool pt_accept(double beta_a, double beta_b, int32_t Ua,
                                          int32_t Ub, int32_t Ka, int32_t Kb, double H,
                                          uint32_t r) {
#pragma clang fp contract(off)
  const double dE = (double)(2 * (Ua - Ub)) - H * (double)(2 * (Ka - Kb));
  const double x = (beta_a - beta_b) * dE;
  if (x >= 0.0) return true;
  return (uint64_t)r < (uint64_t)(det_exp(x) * 4294967296.0);
}

// Metropolis thresholds of one temperature.  For spin bit sig and u
// unsatisfied bonds, dE = 12 - 4u + 2 H sigma; a flip is accepted iff
// r < T = floor(2^32 exp(-beta dE)), and always when dE <= 0.  Because dE
// falls with u, the entries with dE > 0 form a prefix u < nfin[sig], and
// T rises with u.  beta > 0 keeps every finite T below 2^32.
struct ThrRec {
  uint32_t t[2][8];
  uint32_t nfin[2];
  uint32_t pad[2];
};

// One lattice processed by a sweep launch.
struct LatDesc {
  uint64_t spin_off;   // word offset of the lattice
  uint64_t j_off;      // word offset of its group's couplings [3][N]
  uint32_t key0, key1; // spin_seed
  uint32_t g;          // group id in the Philox counter
  uint32_t rt;         // (rho << 16) | t
  uint32_t thr_idx;    // ThrRec index
  uint32_t e_idx;      // energy slot: U at [e_idx*128 + l], K at [e_idx*128 + 64 + l]
  uint32_t pad[2];
};

// One PT pair (t, t+1) of one (group, replica).
struct PairDesc {
  uint64_t a_off, b_off;  // word offsets of the lattices at slots t and t+1
  double beta_a, beta_b;
  double H;
  uint32_t e_a, e_b;      // energy slots
  uint32_t key0, key1;
  uint32_t g, rt;         // rt = (rho << 16) | t
  uint32_t pad;
  uint32_t cnt_idx;       // accepted-swap counter index
  uint32_t n_words;       // N
  uint32_t parity;        // t & 1: the pair is active on steps with b & 1 == parity
};

// Per-job device state read by the kernels, so that a launch sequence can be
// captured once in a hipGraph and replayed while the counters advance.
struct JobState {
  uint32_t sweep;     // sweeps completed; a PT step before sweep n has b = n / pt_every
  uint32_t pt_every;  // PT cadence (0 = none)
  uint32_t pad[2];
};

// The device buffers of one job (one simulation: one L, its groups,
// replicas and temperatures).  Descriptor offsets and indices are relative
// to these bases.
struct JobPtrs {
  uint64_t *spins;
  const uint64_t *J;
  const uint64_t *Js;  // coupling site records [group][colour][h][6]
  const LatDesc *descs;
  const PairDesc *pairs;
  const ThrRec *thr;
  JobState *js;
  int32_t *energy;
  unsigned long long *counters;
  uint64_t *spins2;    // slab sweeps: the second spin buffer (same layout), or null
  // split sweeps (a lattice held by P work-groups): boundary-plane exchange
  // buffer [lattice][P][parity][half][side][L^2/2], the phase flag of each
  // (lattice, slab) and an error word set by a wait that timed out; null
  // where the job has no split shape
  uint64_t *xbuf;
  uint32_t *xflag;
  uint32_t *xerr;
};

// One work-group task of the sweep kernel: lattice descriptor indices (the
// second is 0xFFFFFFFF for a single lattice), when the two form a PT pair
// (slots t and t+1) the PairDesc index, and the job's entry in the launch's
// JobPtrs table.
struct TaskDesc {
  uint32_t lat[2];
  uint32_t pair;
  uint32_t job;
};

// One (group, temperature) measured across its replicas.
struct MeasDesc {
  uint64_t spin_off;      // replica 0 lattice
  uint64_t rho_stride;    // words between replicas
  uint64_t e_idx;         // energy slot of replica 0 (U, K of the current states)
  uint64_t e_stride;      // energy slots between replicas
  uint64_t out_off;       // int32 offset of the 64 records
  uint64_t acc_off;       // ObsAcc index of lane 0
  uint64_t cs_off;        // double offset of cos[L], sin[L] of 2 pi c / L
};

// Observable sums of one (group, temperature, lane) over all measurements:
// exact integers (Q = N - 2W is the overlap times N; Q^4 needs 128 bits at
// L = 48) and one FP64 sum for the k = 2 pi / L overlap.
struct ObsAcc {
  int64_t sU, sK;           // unsatisfied bonds, up spins: summed over replicas
  int64_t sQ, sAbsQ;
  uint64_t sQ2;
  uint64_t sQ4lo, sQ4hi;
  double sq1;               // sum of (|q1_x|^2 + |q1_y|^2 + |q1_z|^2) / 3
};

}  // namespace sg
