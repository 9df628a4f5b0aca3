// sg_kernels.hip — CDNA4 (gfx950) kernels of libsg.
//
// Hot path: checkerboard Metropolis sweeps on multispin words (64 samples per
// uint64), with the lattice resident in LDS for the whole launch and the
// couplings of each thread's sites held in VGPRs.  The local-field count is
// a bit-sliced full-adder tree and the accept decision a comparison of one
// Philox word per site against integer thresholds, so no float enters the
// update.  Reference counterpart: perform_MC_steps, Program/src/eamsim.cu:
// 112-208 (one curand() and one LUT read per spin there).
#include "sg_kernels.h"

// This file is compiled three times: as itself (everything but the resident
// and split sweep kernels), through sg_kernels_resident.hip
// (SG_RESIDENT_PART) and through sg_kernels_split.hip (SG_SPLIT_PART).
#if !defined(SG_RESIDENT_PART) && !defined(SG_SPLIT_PART)
#define SG_MAIN_PART 1
#endif

namespace sg {

__device__ __forceinline__ uint64_t maj3(uint64_t a, uint64_t b, uint64_t c) {
  return (a & b) | (c & (a | b));
}

// v_bitop3_b32: any 3-input boolean function in one VALU op.  The 8-bit
// immediate is the function applied to the operand patterns a = 0xF0,
// b = 0xCC, c = 0xAA (bit 4a+2b+c of the immediate is f(a, b, c)).
constexpr uint32_t kA = 0xF0, kB = 0xCC, kC = 0xAA;
constexpr uint32_t kXor3 = (kA ^ kB ^ kC) & 0xFF;                      // a^b^c
constexpr uint32_t kMaj3 = ((kA & kB) | (kA & kC) | (kB & kC)) & 0xFF;  // majority
constexpr uint32_t kSel = ((kA & kB) | (~kA & kC)) & 0xFF;              // a ? b : c
constexpr uint32_t kXorOr = (kA ^ (kB | kC)) & 0xFF;                    // a ^ (b|c)
#define SG_BITOP3(a, b, c, imm) __builtin_amdgcn_bitop3_b32((a), (b), (c), (imm))

// One Metropolis update of a 32-lane half word at H = 0: s is the centre
// half, n[k]/j[k] the six neighbour and coupling halves.  A lane flips iff
// its count u of unsatisfied bonds is at least k, where k = 3 - n and n is the
// number of thresholds T(u = 2) >= T(u = 1) >= T(u = 0) the site's random
// number is below (u >= 3 always flips).  u >= k is the carry out of
// u + (5 + n) >= 8, so the bias 5 + n enters the carry-save tree as extra
// inputs and only carries are formed.  With B0 = n odd and B1 = n >= 2
// (all-ones masks) the bias bits are ~B0 (weight 1), B0 ^ B1 (weight 2),
// ~(B0 & B1) (weight 4) and B0 & B1 (weight 8): at weight 1, FA(b0..b2) +
// FA(b3..b5) + ~B0 gives carries k1, k2, k3; at weight 2, FA(k1, k2, k3)
// gives t and m1; the weight-8 carry with the weight-8 bias bit is then
// m1 | (t & (B0 ^ B1)) | (B0 & B1) = m1 | maj(t, B0, B1).  15 VALU ops.
constexpr uint32_t kMaj3NotC = ((kA & kB) | (kA & ~kC) | (kB & ~kC)) & 0xFF;  // maj(a, b, ~c)
__device__ __forceinline__ uint32_t metropolis32(uint32_t s, const uint32_t (&n)[6],
                                                 const uint32_t (&j)[6], uint32_t B0,
                                                 uint32_t B1) {
  const uint32_t b0 = SG_BITOP3(s, n[0], j[0], kXor3);
  const uint32_t b1 = SG_BITOP3(s, n[1], j[1], kXor3);
  const uint32_t b2 = SG_BITOP3(s, n[2], j[2], kXor3);
  const uint32_t b3 = SG_BITOP3(s, n[3], j[3], kXor3);
  const uint32_t b4 = SG_BITOP3(s, n[4], j[4], kXor3);
  const uint32_t b5 = SG_BITOP3(s, n[5], j[5], kXor3);
  const uint32_t s1 = SG_BITOP3(b0, b1, b2, kXor3), k1 = SG_BITOP3(b0, b1, b2, kMaj3);
  const uint32_t s2 = SG_BITOP3(b3, b4, b5, kXor3), k2 = SG_BITOP3(b3, b4, b5, kMaj3);
  const uint32_t k3 = SG_BITOP3(s1, s2, B0, kMaj3NotC);
  const uint32_t t = SG_BITOP3(k1, k2, k3, kXor3), m1 = SG_BITOP3(k1, k2, k3, kMaj3);
  const uint32_t g = SG_BITOP3(t, B0, B1, kMaj3);
  return SG_BITOP3(s, m1, g, kXorOr);                 // s ^ (m1 | g)
}

// The same at H != 0: a lane flips iff rr < T[sigma][u], sigma its spin bit
// and u its unsatisfied bonds; u = 5, 6 always flip, since dE = 2H sigma' - 8
// <= 0 for H <= 4.  T[sigma][u] rises with u, so for each sigma the lane
// flips iff u >= k_sigma, and u >= k is the carry out of u + (8 - k) >= 8 as
// at H = 0, with a bias 8 - k in 3..8 whose bits now depend on the lane's
// spin: C[sigma] = {bit 0, bit 1, bit 2 | bit 3, bit 3} as all-ones masks,
// each picked by s with one op.  The weight-8 carry is maj(m1, t & bit 1,
// bit 2), and bit 3 (k = 0: always) forces the flip.  20 VALU ops.
__device__ __forceinline__ uint32_t metropolis32_field(uint32_t s, const uint32_t (&n)[6],
                                                       const uint32_t (&j)[6],
                                                       const uint32_t (&C)[2][4]) {
  const uint32_t b0 = SG_BITOP3(s, n[0], j[0], kXor3);
  const uint32_t b1 = SG_BITOP3(s, n[1], j[1], kXor3);
  const uint32_t b2 = SG_BITOP3(s, n[2], j[2], kXor3);
  const uint32_t b3 = SG_BITOP3(s, n[3], j[3], kXor3);
  const uint32_t b4 = SG_BITOP3(s, n[4], j[4], kXor3);
  const uint32_t b5 = SG_BITOP3(s, n[5], j[5], kXor3);
  const uint32_t s1 = SG_BITOP3(b0, b1, b2, kXor3), k1 = SG_BITOP3(b0, b1, b2, kMaj3);
  const uint32_t s2 = SG_BITOP3(b3, b4, b5, kXor3), k2 = SG_BITOP3(b3, b4, b5, kMaj3);
  const uint32_t c0 = SG_BITOP3(s, C[1][0], C[0][0], kSel);
  const uint32_t k3 = SG_BITOP3(s1, s2, c0, kMaj3);
  const uint32_t t = SG_BITOP3(k1, k2, k3, kXor3), m1 = SG_BITOP3(k1, k2, k3, kMaj3);
  const uint32_t c1 = SG_BITOP3(s, C[1][1], C[0][1], kSel);
  const uint32_t m2 = t & c1;
  const uint32_t d2 = SG_BITOP3(s, C[1][2], C[0][2], kSel);
  const uint32_t w8 = SG_BITOP3(m1, m2, d2, kMaj3);
  const uint32_t c3 = SG_BITOP3(s, C[1][3], C[0][3], kSel);
  return SG_BITOP3(s, w8, c3, kXorOr);                 // s ^ (w8 | c3)
}

// Philox4x32-10 with the key XORs as single 3-input ops (same numbers as
// sg::philox).
__device__ __forceinline__ u32x4 philox_dev(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = SG_BITOP3((uint32_t)(p1 >> 32), c1, k0, kXor3);
    const uint32_t n2 = SG_BITOP3((uint32_t)(p0 >> 32), c3, k1, kXor3);
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

// The same numbers for a counter (j, n, g, w) in which only j differs
// between the threads: the first round's product of g and the second round's
// product of the first output word (both the same in every thread) are
// taken as 32-bit scalar multiplies, off the vector unit.
__device__ __forceinline__ u32x4 philox_sweep(uint32_t j, uint32_t n, uint32_t g, uint32_t w,
                                              uint32_t k0, uint32_t k1) {
  // round 1: M1 g is uniform, M0 j depends on the thread only (not the sweep)
  const uint64_t p0 = (uint64_t)0xD2511F53u * j;
  uint32_t c0 = __umulhi(0xCD9E8D57u, g) ^ n ^ k0;
  uint32_t c1 = 0xCD9E8D57u * g;
  uint32_t c2 = SG_BITOP3((uint32_t)(p0 >> 32), w, k1, kXor3);
  uint32_t c3 = (uint32_t)p0;
  k0 += 0x9E3779B9u;
  k1 += 0xBB67AE85u;
  // round 2: c0 is uniform
  {
    const uint32_t q0hi = __umulhi(0xD2511F53u, c0), q0lo = 0xD2511F53u * c0;
    const uint64_t q1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = SG_BITOP3((uint32_t)(q1 >> 32), c1, k0, kXor3);
    const uint32_t n2 = SG_BITOP3(q0hi, c3, k1, kXor3);
    c1 = (uint32_t)q1;
    c3 = q0lo;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
#pragma unroll
  for (int round = 2; round < 10; ++round) {
    const uint64_t p0r = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1r = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = SG_BITOP3((uint32_t)(p1r >> 32), c1, k0, kXor3);
    const uint32_t n2 = SG_BITOP3((uint32_t)(p0r >> 32), c3, k1, kXor3);
    c1 = (uint32_t)p1r;
    c3 = (uint32_t)p0r;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  u32x4 r;
  r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
  return r;
}

// Word i & 3 of a Philox block by selects: an indexed read of the block would
// place it in memory (LDS or scratch).
__device__ __forceinline__ uint32_t pick_word(const u32x4 &r, int i) {
  const uint32_t lo = (i & 1) ? r.v[1] : r.v[0];
  const uint32_t hi = (i & 1) ? r.v[3] : r.v[2];
  return (i & 2) ? hi : lo;
}

// Address spaces of what a kernel reaches through its job table.  The table
// holds generic pointers; accessed as such, every load and store would be a
// flat access and no uniform load could become a scalar load.  Descriptors,
// thresholds and counters are read-only during a launch and uniform across
// the work-group: they are read through the constant address space (s_load
// into SGPRs).  Spins and couplings are global.
#define SG_GLOBAL __attribute__((address_space(1)))
#define SG_CONST __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ T ld_uniform(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized records only");
  const SG_CONST uint32_t *q = (const SG_CONST uint32_t *)p;
  T v;
  uint32_t *d = (uint32_t *)&v;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(q[i]);
  return v;
}

template <class T>
__device__ __forceinline__ SG_GLOBAL T *to_global(T *p) {
  return (SG_GLOBAL T *)p;
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// A 32-bit byte offset the compiler cannot see through (a pure asm, free to
// move): loads from a uniform base at such an offset take the SGPR-base form
// and no 64-bit address per item is formed or kept.
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
  asm("" : "+v"(v));
  return v;
}
__device__ __forceinline__ uint64_t gload(const SG_GLOBAL uint64_t *base, int idx) {
  return *(const SG_GLOBAL uint64_t *)((const SG_GLOBAL char *)base + opaque((uint32_t)idx * 8u));
}

// site record idx of Jsb without the offset barrier: the compiler may issue
// these loads as early as it likes (the items that use them are independent)
__device__ __forceinline__ void site_record_free(const SG_GLOBAL char *__restrict__ Jsb, int idx,
                                                 uint64_t (&jw)[6]) {
  const SG_GLOBAL u64x2 *r = (const SG_GLOBAL u64x2 *)(Jsb + opaque((uint32_t)idx * 48u));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const u64x2 v = r[k];
    jw[2 * k] = v.x;
    jw[2 * k + 1] = v.y;
  }
}

template <int L>
struct Geo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int Q = N / 4;          // Philox quads: one per thread per sweep
  static constexpr int LH = L / 2;
  static constexpr int NT = ((Q + 63) / 64) * 64;
  static constexpr int NW = NT / 64;
};

// word index of (x, y, z) in the checkerboard-split layout
template <int L>
__device__ __forceinline__ int cb_index(int x, int y, int z) {
  return ((x + y + z) & 1) * Geo<L>::NH + (x >> 1) + Geo<L>::LH * (y + L * z);
}

// Bit-sliced add: accumulator planes A (NP planes) += 3-bit count (u2 u1 u0).
template <int NP>
__device__ __forceinline__ void vc_add3(uint64_t (&A)[NP], uint64_t u0, uint64_t u1,
                                        uint64_t u2) {
  uint64_t c = A[0] & u0;
  A[0] ^= u0;
  uint64_t s = A[1] ^ u1 ^ c;
  c = maj3(A[1], u1, c);
  A[1] = s;
  s = A[2] ^ u2 ^ c;
  c = maj3(A[2], u2, c);
  A[2] = s;
#pragma unroll
  for (int p = 3; p < NP; ++p) {
    const uint64_t t = A[p] & c;
    A[p] ^= c;
    c = t;
  }
}

template <int NP>
__device__ __forceinline__ void vc_add1(uint64_t (&A)[NP], uint64_t w) {
  uint64_t c = w;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const uint64_t t = A[p] & c;
    A[p] ^= c;
    c = t;
  }
}

// DPP source fetch of a 64-bit value (two 32-bit DPP moves); lanes whose
// source is outside the row, or whose row is not in ROW_MASK, read 0.
// GFX9 DPP controls: row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROW_MASK, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROW_MASK, 0xF, true);
  return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

// A (NP bit planes) += the DPP-fetched A of another lane; NP + 1 planes out.
template <int NP, int CTRL, int ROW_MASK, int NA>
__device__ __forceinline__ void dpp_add(uint64_t (&A)[NA]) {
  static_assert(NP < NA, "plane array too small");
  uint64_t B[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) B[p] = dpp64<CTRL, ROW_MASK>(A[p]);
  uint64_t carry = 0;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const uint64_t a = A[p];
    A[p] = a ^ B[p] ^ carry;
    carry = maj3(a, B[p], carry);
  }
  A[NP] = carry;
}

// Sum the bit-sliced per-thread counters of a work-group into one integer per
// lane (sample).  A0 holds P0 planes.  Within a wave, a DPP tree (row_shr 1,
// 2, 4, 8, then row_bcast 15 and 31) leaves the wave total in lane 63 (P0+6
// planes); the wave totals meet in wave 0 through LDS, a second row_shr tree
// leaves the work-group total in lane 15 (P0+10 planes), and every lane
// reads those planes back (readlane) and extracts its own bit.  Result valid
// in wave 0.  red must hold NW*24 words; NW <= 16.  Every thread calls it.
template <int P0, int NW>
__device__ int32_t wg_sum_planes(const uint64_t (&A0)[P0], uint64_t *red) {
  static_assert(NW <= 16, "at most 16 waves");
  constexpr int PW = P0 + 6, PT = P0 + 10;
  uint64_t A[PT];
#pragma unroll
  for (int p = 0; p < P0; ++p) A[p] = A0[p];
  dpp_add<P0 + 0, 0x111, 0xF>(A);
  dpp_add<P0 + 1, 0x112, 0xF>(A);
  dpp_add<P0 + 2, 0x114, 0xF>(A);
  dpp_add<P0 + 3, 0x118, 0xF>(A);
  dpp_add<P0 + 4, 0x142, 0xA>(A);
  dpp_add<P0 + 5, 0x143, 0xC>(A);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63) {
#pragma unroll
    for (int p = 0; p < PW; ++p) red[wave * 24 + p] = A[p];
  }
  __syncthreads();
  int32_t v = 0;
  if (wave == 0) {
#pragma unroll
    for (int p = 0; p < PW; ++p) A[p] = lane < NW ? red[lane * 24 + p] : 0;
    dpp_add<PW + 0, 0x111, 0xF>(A);
    dpp_add<PW + 1, 0x112, 0xF>(A);
    dpp_add<PW + 2, 0x114, 0xF>(A);
    dpp_add<PW + 3, 0x118, 0xF>(A);
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)A[p], 15);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(A[p] >> 32), 15);
      const uint64_t tot = (uint64_t)lo | ((uint64_t)hi << 32);
      v |= (int32_t)((tot >> lane) & 1) << p;
    }
  }
  __syncthreads();
  return v;
}

// Bit-sliced S (NS planes) += B (NB planes), ripple carry; the caller sizes
// NS so that the sum cannot overflow.
template <int NS, int NA, int NB>
__device__ __forceinline__ void vc_add(uint64_t (&S)[NA], const uint64_t (&B)[NB]) {
  static_assert(NS <= NA && NB <= NS, "plane counts");
  uint64_t c = 0;
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    const uint64_t a = S[p];
    if (p < NB) {
      S[p] = a ^ B[p] ^ c;
      c = maj3(a, B[p], c);
    } else {
      S[p] = a ^ c;
      c = a & c;
    }
  }
}

constexpr int ceil_log2(int n) { return n <= 1 ? 0 : 1 + ceil_log2((n + 1) / 2); }

// 32-bit DPP fetch / bit-sliced add (see dpp64 / dpp_add)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, true);
}

template <int NP, int CTRL, int ROW_MASK, int NA>
__device__ __forceinline__ void dpp_add32(uint32_t (&A)[NA]) {
  static_assert(NP < NA, "plane array too small");
  uint32_t B[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) B[p] = dpp32<CTRL, ROW_MASK>(A[p]);
  uint32_t carry = 0;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const uint32_t a = A[p];
    A[p] = SG_BITOP3(a, B[p], carry, kXor3);
    carry = SG_BITOP3(a, B[p], carry, kMaj3);
  }
  A[NP] = carry;
}

// Work-group total, per lane (sample), of the bit-sliced per-thread counts A
// of NL lattices (P0 planes each).  A 64-bit DPP tree level costs ~6 VALU ops
// per plane on every thread, so:
//  1. the counts are folded through LDS onto the first TW <= 4 waves (one per
//     SIMD), each folded thread adding K entries with ripple adds;
//  2. v_permlane32_swap splits every 64-bit plane: thread t < 32 keeps
//     samples 0..31 (its low dword plus thread t+32's), thread t >= 32
//     samples 32..63, so the rest of the tree runs on 32-bit planes;
//  3. a 32-bit DPP tree within each half (row_shr 1, 2, 4, 8, row_bcast 15)
//     leaves the half totals in lanes 31 and 63;
//  4. each lane reads its half's planes (readlane) and extracts its own bit:
//     one integer per lane per tree wave, which wave 0 adds.
// fold holds the first PS of the NL*P0 planes (NT words each) and fold2 the
// rest; red holds TW*NL*64 int32; all LDS.  Every thread calls it; out is
// valid in wave 0 (lane = sample).  `during` runs in every thread between the
// two barriers, after the tree in the tree waves: work that touches neither
// fold nor red (a lattice's write-out) overlaps the tree there.
struct NoWork {
  __device__ void operator()() const {}
};
template <int NT, int PS>
__device__ __forceinline__ uint64_t *fold_plane(uint64_t *fold, uint64_t *fold2, int idx) {
  return idx < PS ? fold + idx * NT : fold2 + (idx - PS) * NT;
}

template <int P0, int NT, int NL, int PS = NL * P0, class Fn = NoWork>
__device__ void wg_count_lanes(const uint64_t (&A)[NL][P0], uint64_t *fold, int32_t *red,
                               int32_t (&out)[NL], uint64_t *fold2 = nullptr,
                               const Fn &during = Fn()) {
  constexpr int NW = NT / 64;
  constexpr int TW = NW < 4 ? NW : 4;
  constexpr int FT = TW * 64;
  constexpr int K = (NT + FT - 1) / FT;
  constexpr int PF = P0 + ceil_log2(K);   // planes after the fold
  constexpr int PH = PF + 1;              // after the 32-lane split
  constexpr int PW = PH + 5;              // after the half-wave tree
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int a = 0; a < NL; ++a)
#pragma unroll
    for (int p = 0; p < P0; ++p) fold_plane<NT, PS>(fold, fold2, a * P0 + p)[tid] = A[a][p];
  __syncthreads();
  if (wave < TW) {
    uint32_t H[NL][PW];
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      uint64_t S[PF];
#pragma unroll
      for (int p = 0; p < PF; ++p) S[p] = p < P0 ? A[a][p] : 0;
#pragma unroll
      for (int k = 1; k < K; ++k) {
        const int i = tid + k * FT;
        if (i < NT) {
          uint64_t B[P0];
#pragma unroll
          for (int p = 0; p < P0; ++p) B[p] = fold_plane<NT, PS>(fold, fold2, a * P0 + p)[i];
          vc_add<PF>(S, B);
        }
      }
      // split: x = own low dwords, y = own high dwords; after the swap,
      // lanes < 32 hold (own low, low of lane + 32) and lanes >= 32
      // (high of lane - 32, own high)
      uint32_t carry = 0;
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const auto r = __builtin_amdgcn_permlane32_swap((uint32_t)S[p], (uint32_t)(S[p] >> 32),
                                                        false, false);
        const uint32_t x = r[0], y = r[1];
        H[a][p] = SG_BITOP3(x, y, carry, kXor3);
        carry = SG_BITOP3(x, y, carry, kMaj3);
      }
      H[a][PF] = carry;
#pragma unroll
      for (int p = PH; p < PW; ++p) H[a][p] = 0;
    }
#pragma unroll
    for (int a = 0; a < NL; ++a) dpp_add32<PH + 0, 0x111, 0xF>(H[a]);
#pragma unroll
    for (int a = 0; a < NL; ++a) dpp_add32<PH + 1, 0x112, 0xF>(H[a]);
#pragma unroll
    for (int a = 0; a < NL; ++a) dpp_add32<PH + 2, 0x114, 0xF>(H[a]);
#pragma unroll
    for (int a = 0; a < NL; ++a) dpp_add32<PH + 3, 0x118, 0xF>(H[a]);
#pragma unroll
    for (int a = 0; a < NL; ++a) dpp_add32<PH + 4, 0x142, 0xA>(H[a]);
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      int32_t v = 0;
#pragma unroll
      for (int p = 0; p < PW; ++p) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)H[a][p], 31);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)H[a][p], 63);
        const uint32_t word = lane < 32 ? lo : hi;
        v |= (int32_t)((word >> (lane & 31)) & 1u) << p;
      }
      red[(wave * NL + a) * 64 + lane] = v;
    }
  }
  during();
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NL; ++a) {
    int32_t v = 0;
    if (wave == 0)
#pragma unroll
      for (int w = 0; w < TW; ++w) v += red[(w * NL + a) * 64 + lane];
    out[a] = v;
  }
}

// ------------------------------------------------------------- sweep kernel
// One work-group per task: one lattice, or two lattices of adjacent
// temperature slots t, t+1 of the same (group, replica).  Both share the
// couplings and neighbour indices held in VGPRs.  When the launch starts at
// a PT step and the task is that step's pair, the group first decides the
// 64 lane swaps (wave 0, one Philox word per lane) and exchanges the masked
// bits in LDS; then both lattices are swept nsweeps times in LDS; then the
// per-lane U (and K) of the final states are reduced for the next PT step.
// Thread j owns the Philox quad j of each lattice: black and white sites at
// h = j and h = j + N/4; per sweep it draws one Philox block per lattice
// (r0, r1 for its black sites, r2, r3 for its white sites).
// Reference counterparts: perform_MC_steps and perform_PT_shuffle,
// Program/src/eamsim.cu:112-330.
template <int L, bool FIELD>
__device__ __forceinline__ void update_site(uint64_t &s, const uint64_t (&nv)[6],
                                            const uint64_t (&jw)[6], uint32_t rr,
                                            const ThrRec &tr) {
  if (!FIELD) {
    // u >= 3 always flips; u = 2, 1, 0 flip iff rr < T(u).  The host keeps
    // T(0) <= T(1) <= T(2), so below T(0) implies below T(1) and T(2).
    const bool a0 = rr < tr.t[0][0], a1 = rr < tr.t[0][1], a2 = rr < tr.t[0][2];
    const uint32_t B0 = (a0 ^ a1 ^ a2) ? ~0u : 0u;   // the count of a's is odd
    const uint32_t B1 = a1 ? ~0u : 0u;               // ... is 2 or 3
    uint32_t nlo[6], nhi[6], jlo[6], jhi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      nlo[k] = (uint32_t)nv[k];
      nhi[k] = (uint32_t)(nv[k] >> 32);
      jlo[k] = (uint32_t)jw[k];
      jhi[k] = (uint32_t)(jw[k] >> 32);
    }
    const uint32_t lo = metropolis32((uint32_t)s, nlo, jlo, B0, B1);
    const uint32_t hi = metropolis32((uint32_t)(s >> 32), nhi, jhi, B0, B1);
    s = (uint64_t)lo | ((uint64_t)hi << 32);
  } else {
    // a[u] = rr is below T[sigma][u] (thresholds past nfin are 2^32: always
    // below; the host keeps nfin <= 5 and T rising in u): a thermometer code,
    // a[u] implies a[u + 1], of m = a[0] + ... + a[4] = 5 - k.  The bias
    // 8 - k = 3 + m has bit 0 = m even, bit 1 = m in {0, 3, 4}, bit 2 | bit 3
    // = m >= 1 and bit 3 = m == 5.
    uint32_t C[2][4];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      bool a[5];
#pragma unroll
      for (int u = 0; u < 5; ++u) a[u] = (uint32_t)u >= tr.nfin[g] || rr < tr.t[g][u];
      C[g][0] = (a[0] ^ a[1] ^ a[2] ^ a[3] ^ a[4]) ? 0u : ~0u;
      C[g][1] = (a[4] == a[2] && !a[0]) ? ~0u : 0u;
      C[g][2] = a[4] ? ~0u : 0u;
      C[g][3] = a[0] ? ~0u : 0u;
    }
    uint32_t nlo[6], nhi[6], jlo[6], jhi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      nlo[k] = (uint32_t)nv[k];
      nhi[k] = (uint32_t)(nv[k] >> 32);
      jlo[k] = (uint32_t)jw[k];
      jhi[k] = (uint32_t)(jw[k] >> 32);
    }
    const uint32_t lo = metropolis32_field((uint32_t)s, nlo, jlo, C);
    const uint32_t hi = metropolis32_field((uint32_t)(s >> 32), nhi, jhi, C);
    s = (uint64_t)lo | ((uint64_t)hi << 32);
  }
}

// bit-sliced unsatisfied-bond count (u2 u1 u0) of one site
__device__ __forceinline__ void bond_count(uint64_t s, const uint64_t (&nv)[6],
                                           const uint64_t (&jw)[6], uint64_t &u0,
                                           uint64_t &u1, uint64_t &u2) {
  const uint64_t b0 = s ^ nv[0] ^ jw[0];
  const uint64_t b1 = s ^ nv[1] ^ jw[1];
  const uint64_t b2 = s ^ nv[2] ^ jw[2];
  const uint64_t b3 = s ^ nv[3] ^ jw[3];
  const uint64_t b4 = s ^ nv[4] ^ jw[4];
  const uint64_t b5 = s ^ nv[5] ^ jw[5];
  const uint64_t s1 = b0 ^ b1 ^ b2, c1 = maj3(b0, b1, b2);
  const uint64_t s2 = b3 ^ b4 ^ b5, c2 = maj3(b3, b4, b5);
  const uint64_t c3 = s1 & s2;
  u0 = s1 ^ s2;
  u1 = c1 ^ c2 ^ c3;
  u2 = maj3(c1, c2, c3);
}

constexpr uint32_t kNone = 0xFFFFFFFFu;

template <int L, bool FIELD>
__global__ __launch_bounds__(Geo<L>::NT) void k_block(
    const JobPtrs *__restrict__ jt, const TaskDesc *__restrict__ tasks, int do_pt,
    int nsweeps, int want_energy) {
  using G = Geo<L>;
  // colour 0 (black) and colour 1 (white) halves of the two lattices as
  // separate arrays: a phase writes one and reads the other, so the compiler
  // may reorder the loads and stores of a phase freely
  __shared__ uint64_t latB[2][G::NH], latW[2][G::NH];
  __shared__ uint64_t fold[2 * 4 * G::NT];   // energy reduction scratch
  __shared__ int32_t redi[4 * 2 * 64];

  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  SG_GLOBAL uint64_t *__restrict__ spins = to_global(jp.spins);
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy);
  unsigned long long *__restrict__ counters = jp.counters;
  const bool two = tk.lat[1] != kNone;
  const LatDesc d0 = ld_uniform(jp.descs + tk.lat[0]);
  const LatDesc d1 = ld_uniform(jp.descs + (two ? tk.lat[1] : tk.lat[0]));
  const int tid = threadIdx.x, lane = tid & 63;
  const bool active = tid < G::Q;

  // ---- PT step of the pair (t, t+1) at b = sweep / pt_every.  Every wave
  // decides the 64 lanes (samples) itself, so the mask needs no LDS round
  // trip; the energy loads go out first and the decision overlaps the
  // lattice loads.
  const bool pt_here = do_pt && two && tk.pair != kNone;
  int32_t ua = 0, ub = 0, ka = 0, kb = 0;
  PairDesc pd{};
  uint32_t pt_b = 0;
  if (pt_here) {
    pd = ld_uniform(jp.pairs + tk.pair);
    const JobState js = ld_uniform(jp.js);
    pt_b = js.sweep / js.pt_every;
    const SG_GLOBAL int32_t *Ea = energy + (size_t)pd.e_a * 128, *Eb = energy + (size_t)pd.e_b * 128;
    ua = Ea[lane]; ub = Eb[lane]; ka = Ea[64 + lane]; kb = Eb[64 + lane];
  }

  // the task's lattices, at most two words per colour per thread (NH <= 2 NT)
  static_assert(G::NH <= 2 * G::NT, "two load rounds");
  const SG_GLOBAL uint64_t *g0 = spins + d0.spin_off, *g1 = spins + d1.spin_off;
  uint64_t w[2][4];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int i = tid + k * G::NT;
    if (i < G::NH) {
      w[k][0] = g0[i];
      w[k][1] = g0[G::NH + i];
      if (two) {
        w[k][2] = g1[i];
        w[k][3] = g1[G::NH + i];
      }
    }
  }

  // neighbour indices within the opposite colour's half (y, z: same for both
  // colours; x: per colour) and the 6 couplings of each owned site
  // (couplings from the site records: three 16-byte loads per site,
  // contiguous across the wave).  One x neighbour of a site is the
  // opposite-colour site of the same index h, which this thread owns: its
  // value comes from a register, and its coupling goes first (the bond
  // count is symmetric in the six (neighbour, coupling) pairs); xo is the
  // index of the other x neighbour.
  const SG_GLOBAL char *Js = (const SG_GLOBAL char *)(to_global(jp.Js) + 2 * d0.j_off);
  int yz[2][4], xo[2][2];
  uint64_t jw[2][2][6];
  if (active) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int h = tid + q * G::Q;
      const int xh = h % G::LH, y = (h / G::LH) % L, z = h / (G::LH * L);
      const int yp = (y + 1) % L, ym = (y + L - 1) % L;
      const int zp = (z + 1) % L, zm = (z + L - 1) % L;
      yz[q][0] = xh + G::LH * (yp + L * z);
      yz[q][1] = xh + G::LH * (ym + L * z);
      yz[q][2] = xh + G::LH * (y + L * zp);
      yz[q][3] = xh + G::LH * (y + L * zm);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        // x even: x + 1 is (xh, y, z) of the other colour, x - 1 the other
        // neighbour; x odd: the reverse
        const bool odd = ((y + z + c) & 1) != 0;
        const int x = 2 * xh + (odd ? 1 : 0);
        const int xn = odd ? (x + 1) % L : (x + L - 1) % L;
        xo[q][c] = (xn >> 1) + G::LH * (y + L * z);
        uint64_t r[6];
        site_record_free(Js, c * G::NH + h, r);
        jw[q][c][0] = odd ? r[1] : r[0];
        jw[q][c][1] = odd ? r[0] : r[1];
#pragma unroll
        for (int k = 2; k < 6; ++k) jw[q][c][k] = r[k];
      }
      // the two colours' sites of index h are x neighbours: one bond, one
      // register pair
      jw[q][1][0] = jw[q][0][0];
    }
  }

  uint64_t m = 0;
  if (pt_here) {
    const uint32_t sw3 = (kTagPT << 28) | (pd.rt & 0x0FFFFFFFu);
    const u32x4 r = philox_dev((uint32_t)(lane >> 2), pt_b, pd.g, sw3, pd.key0, pd.key1);
    const bool acc = pt_accept(pd.beta_a, pd.beta_b, ua, ub, ka, kb, pd.H, pick_word(r, lane));
    m = __ballot(acc);
  }
  // accepted lanes trade their bits on the way into LDS
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int i = tid + k * G::NT;
    if (i < G::NH) {
      if (two) {
        const uint64_t db = (w[k][0] ^ w[k][2]) & m, dw = (w[k][1] ^ w[k][3]) & m;
        latB[0][i] = w[k][0] ^ db;
        latW[0][i] = w[k][1] ^ dw;
        latB[1][i] = w[k][2] ^ db;
        latW[1][i] = w[k][3] ^ dw;
      } else {
        latB[0][i] = w[k][0];
        latW[0][i] = w[k][1];
      }
    }
  }
  ThrRec tr0{}, tr1{};
  uint32_t n0 = 0;
  if (nsweeps > 0) {
    tr0 = ld_uniform(jp.thr + d0.thr_idx);
    tr1 = ld_uniform(jp.thr + d1.thr_idx);
    n0 = ld_uniform(jp.js).sweep;
  }
  __syncthreads();

  // the thread's own sites of both colours, held in registers for the launch
  uint64_t own[2][2][2] = {};   // [lattice][colour][q]
  if (active) {
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (a == 1 && !two) break;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        own[a][0][q] = latB[a][tid + q * G::Q];
        own[a][1][q] = latW[a][tid + q * G::Q];
      }
    }
  }
  const uint32_t sw0 = (kTagMC << 28) | (d0.rt & 0x0FFFFFFFu);
  const uint32_t sw1 = (kTagMC << 28) | (d1.rt & 0x0FFFFFFFu);
  for (int it = 0; it < nsweeps; ++it) {
    const uint32_t n = n0 + (uint32_t)it;
    u32x4 rb[2];   // the sweep's Philox block of each lattice
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (active) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          if (a == 1 && !two) break;
          uint64_t *mine = c == 0 ? latB[a] : latW[a];
          const uint64_t *opp = c == 0 ? latW[a] : latB[a];
          const ThrRec &tr = a == 0 ? tr0 : tr1;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            uint64_t nv[6];
            nv[0] = own[a][1 - c][q];
            nv[1] = opp[xo[q][c]];
            nv[2] = opp[yz[q][0]];
            nv[3] = opp[yz[q][1]];
            nv[4] = opp[yz[q][2]];
            nv[5] = opp[yz[q][3]];
            // at colour 0 the lattice's Philox block is drawn while the first
            // site's neighbour reads are in flight
            if (c == 0 && q == 0) {
              const LatDesc &d = a == 0 ? d0 : d1;
              rb[a] = philox_sweep((uint32_t)tid, n, d.g, a == 0 ? sw0 : sw1, d.key0, d.key1);
            }
            uint64_t s = own[a][c][q];
            update_site<L, FIELD>(s, nv, jw[q][c], rb[a].v[2 * c + q], tr);
            mine[tid + q * G::Q] = s;
            own[a][c][q] = s;
          }
        }
      }
      __syncthreads();
    }
  }

  // energies of the final states: every bond touches exactly one black site
  // (colour 0).  The bit-sliced counts are formed first (registers), the
  // lattices leave for global memory, then the counts are reduced.
  uint64_t AU[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  uint64_t AK[2][3] = {{0, 0, 0}, {0, 0, 0}};
  if (want_energy && active) {
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (a == 1 && !two) break;
      const uint64_t *opp = latW[a];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const uint64_t s = own[a][0][q];
        uint64_t nv[6];
        nv[0] = own[a][1][q];
        nv[1] = opp[xo[q][0]];
        nv[2] = opp[yz[q][0]];
        nv[3] = opp[yz[q][1]];
        nv[4] = opp[yz[q][2]];
        nv[5] = opp[yz[q][3]];
        uint64_t u0, u1, u2;
        bond_count(s, nv, jw[q][0], u0, u1, u2);
        vc_add3(AU[a], u0, u1, u2);
        if (want_energy > 1) {
          vc_add1(AK[a], s);
          vc_add1(AK[a], own[a][1][q]);
        }
      }
    }
  }

  // the lattices leave for global memory; with energies, while the tree
  // waves of the reduction run (the write-out reads only the lattices)
  auto write_out = [&]() {
    SG_GLOBAL uint64_t *g0 = spins + d0.spin_off, *g1 = spins + d1.spin_off;
    for (int i = tid; i < G::NH; i += G::NT) {
      g0[i] = latB[0][i];
      g0[G::NH + i] = latW[0][i];
    }
    if (two)
      for (int i = tid; i < G::NH; i += G::NT) {
        g1[i] = latB[1][i];
        g1[G::NH + i] = latW[1][i];
      }
  };
  if (!want_energy) write_out();

  if (want_energy) {
    int32_t U[2];
    wg_count_lanes<4, G::NT, 2>(AU, fold, redi, U, nullptr, write_out);
    if (tid < 64) {
      energy[(size_t)d0.e_idx * 128 + tid] = U[0];
      if (two) energy[(size_t)d1.e_idx * 128 + tid] = U[1];
    }
    if (want_energy > 1) {
      __syncthreads();  // fold and redi are reused
      int32_t K[2];
      wg_count_lanes<3, G::NT, 2>(AK, fold, redi, K);
      if (tid < 64) {
        energy[(size_t)d0.e_idx * 128 + 64 + tid] = K[0];
        if (two) energy[(size_t)d1.e_idx * 128 + 64 + tid] = K[1];
      }
    }
  }
  if (pt_here && tid == 0) atomicAdd(&counters[pd.cnt_idx], (unsigned long long)__popcll(m));
}

// ------------------------------------------------------- large-lattice kernel
// L >= 22: two lattices no longer fit in LDS, so each thread walks several
// quads (j = tid, tid + NT, ...), reads the couplings of each site from L2,
// draws the quad's Philox block in each phase, and updates the lattices in
// place in global memory (L2/MALL-resident), which is coherent within the
// work-group because all its waves share one CU.  Same task / PT / energy
// semantics as k_block.
template <int L>
struct BigGeo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int Q = N / 4;
  static constexpr int LH = L / 2;
  static constexpr int NT = 1024;
  static constexpr int NW = NT / 64;
  static constexpr int QPT = (Q + NT - 1) / NT;   // quads per thread
  // bits of a per-thread count of up to 12 * QPT (U) or 4 * QPT (K)
  static constexpr int PU = QPT <= 1 ? 4 : QPT <= 2 ? 5 : QPT <= 5 ? 6 : QPT <= 10 ? 7 : QPT <= 21 ? 8 : 9;
  static constexpr int PK = QPT <= 1 ? 3 : QPT <= 3 ? 4 : QPT <= 7 ? 5 : QPT <= 15 ? 6 : 7;
};

template <int L>
__device__ __forceinline__ void site_geometry(int h, int c, int (&nb)[6], int &s_nat,
                                              int (&jn)[6]) {
  using G = BigGeo<L>;
  const int xh = h % G::LH, y = (h / G::LH) % L, z = h / (G::LH * L);
  const int x = 2 * xh + ((y + z + c) & 1);
  const int xp = (x + 1) % L, xm = (x + L - 1) % L;
  const int yp = (y + 1) % L, ym = (y + L - 1) % L;
  const int zp = (z + 1) % L, zm = (z + L - 1) % L;
  // indices into the opposite colour's half
  nb[0] = (xp >> 1) + G::LH * (y + L * z);
  nb[1] = (xm >> 1) + G::LH * (y + L * z);
  nb[2] = xh + G::LH * (yp + L * z);
  nb[3] = xh + G::LH * (ym + L * z);
  nb[4] = xh + G::LH * (y + L * zp);
  nb[5] = xh + G::LH * (y + L * zm);
  s_nat = x + L * (y + L * z);
  // natural indices of the six bonds' owners: +x, -x, +y, -y, +z, -z
  jn[0] = s_nat;
  jn[1] = xm + L * (y + L * z);
  jn[2] = s_nat + G::N;
  jn[3] = x + L * (ym + L * z) + G::N;
  jn[4] = s_nat + 2 * G::N;
  jn[5] = x + L * (y + L * zm) + 2 * G::N;
}

template <int L, bool FIELD>
__global__ __launch_bounds__(1024) void k_big(
    const JobPtrs *__restrict__ jt, const TaskDesc *__restrict__ tasks, int do_pt,
    int nsweeps, int want_energy) {
  using G = BigGeo<L>;
  __shared__ uint64_t red[G::NW * 24];
  __shared__ unsigned long long smask;

  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  SG_GLOBAL uint64_t *__restrict__ spins = to_global(jp.spins);
  const SG_GLOBAL uint64_t *__restrict__ Jg = to_global(jp.J);
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy);
  unsigned long long *__restrict__ counters = jp.counters;
  const bool two = tk.lat[1] != kNone;
  const LatDesc d0 = ld_uniform(jp.descs + tk.lat[0]);
  const LatDesc d1 = ld_uniform(jp.descs + (two ? tk.lat[1] : tk.lat[0]));
  const int tid = threadIdx.x;
  uint64_t *lat[2] = {(uint64_t *)(spins + d0.spin_off), (uint64_t *)(spins + d1.spin_off)};
  const SG_GLOBAL uint64_t *J = Jg + d0.j_off;
  ThrRec tr0{}, tr1{};
  uint32_t n0 = 0;
  if (nsweeps > 0) {
    tr0 = ld_uniform(jp.thr + d0.thr_idx);
    tr1 = ld_uniform(jp.thr + d1.thr_idx);
    n0 = ld_uniform(jp.js).sweep;
  }
  __syncthreads();

  if (do_pt && two && tk.pair != kNone) {
    const PairDesc p = ld_uniform(jp.pairs + tk.pair);
    const JobState js = ld_uniform(jp.js);
    if (tid < 64) {
      const uint32_t b = js.sweep / js.pt_every;
      const int l = tid;
      const uint32_t sw3 = (kTagPT << 28) | (p.rt & 0x0FFFFFFFu);
      const u32x4 r = philox_dev((uint32_t)(l >> 2), b, p.g, sw3, p.key0, p.key1);
      const uint32_t rr = pick_word(r, l);
      const SG_GLOBAL int32_t *Ea = energy + (size_t)p.e_a * 128, *Eb = energy + (size_t)p.e_b * 128;
      const bool acc = pt_accept(p.beta_a, p.beta_b, Ea[l], Eb[l], Ea[64 + l], Eb[64 + l], p.H, rr);
      const unsigned long long m = __ballot(acc);
      if (l == 0) {
        smask = m;
        atomicAdd(&counters[p.cnt_idx], (unsigned long long)__popcll(m));
      }
    }
    __syncthreads();
    const uint64_t m = smask;
    if (m)
      for (int i = tid; i < G::N; i += G::NT) {
        const uint64_t a = lat[0][i], bb = lat[1][i], dd = (a ^ bb) & m;
        lat[0][i] = a ^ dd;
        lat[1][i] = bb ^ dd;
      }
    __syncthreads();
  }

  const uint32_t sw[2] = {(kTagMC << 28) | (d0.rt & 0x0FFFFFFFu),
                          (kTagMC << 28) | (d1.rt & 0x0FFFFFFFu)};
  for (int it = 0; it < nsweeps; ++it) {
    const uint32_t n = n0 + (uint32_t)it;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll 1
      for (int jq0 = tid; jq0 < G::Q; jq0 += G::NT) {
        // opaque to the optimiser: when NT is a whole number of planes (L = 32)
        // it would otherwise hoist ~20 per-thread 64-bit neighbour addresses
        // out of the sweep loop and spill them
        int jq = jq0;
        asm volatile("" : "+v"(jq));
        const u32x4 r0 = philox_sweep((uint32_t)jq, n, d0.g, sw[0], d0.key0, d0.key1);
        u32x4 r1 = r0;
        if (two) r1 = philox_sweep((uint32_t)jq, n, d1.g, sw[1], d1.key0, d1.key1);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int h = jq + q * G::Q;
          int nb[6], s_nat, jn[6];
          site_geometry<L>(h, c, nb, s_nat, jn);
          uint64_t jw[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) jw[k] = J[jn[k]];
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            if (a == 1 && !two) break;
            const uint64_t *opp = lat[a] + (1 - c) * G::NH;
            uint64_t nv[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) nv[k] = opp[nb[k]];
            uint64_t s = lat[a][c * G::NH + h];
            update_site<L, FIELD>(s, nv, jw, (a == 0 ? r0 : r1).v[2 * c + q], a == 0 ? tr0 : tr1);
            lat[a][c * G::NH + h] = s;
          }
        }
      }
      __syncthreads();
    }
  }

  if (want_energy) {
    // one lattice at a time keeps the live bit planes small
#pragma unroll 1
    for (int a = 0; a < 2; ++a) {
      if (a == 1 && !two) break;
      const uint64_t *la = a == 0 ? lat[0] : lat[1];
      uint64_t AU[G::PU];
#pragma unroll
      for (int p = 0; p < G::PU; ++p) AU[p] = 0;
#pragma unroll 1
      for (int jq = tid; jq < G::Q; jq += G::NT) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int h = jq + q * G::Q;
          int nb[6], s_nat, jn[6];
          site_geometry<L>(h, 0, nb, s_nat, jn);
          uint64_t jw[6], nv[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            jw[k] = J[jn[k]];
            nv[k] = la[G::NH + nb[k]];
          }
          uint64_t u0, u1, u2;
          bond_count(la[h], nv, jw, u0, u1, u2);
          vc_add3(AU, u0, u1, u2);
        }
      }
      const LatDesc &d = a == 0 ? d0 : d1;
      const int32_t U = wg_sum_planes<G::PU, G::NW>(AU, red);
      if (tid < 64) energy[(size_t)d.e_idx * 128 + tid] = U;
      if (want_energy > 1) {
        uint64_t AK[G::PK];
#pragma unroll
        for (int p = 0; p < G::PK; ++p) AK[p] = 0;
#pragma unroll 1
        for (int jq = tid; jq < G::Q; jq += G::NT) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int h = jq + q * G::Q;
            vc_add1(AK, la[h]);
            vc_add1(AK, la[G::NH + h]);
          }
        }
        const int32_t K = wg_sum_planes<G::PK, G::NW>(AK, red);
        if (tid < 64) energy[(size_t)d.e_idx * 128 + 64 + tid] = K;
      }
    }
  }
}

// ------------------------------------------------- mid-size resident kernel
// L = 18, 20: the two lattices of a pair task still fit in LDS (4 NH words,
// 125 KB at L = 20) but their couplings no longer fit in VGPRs beside them.
// They are read from the coupling site records (6 words per site, three
// 16-byte loads, L2-resident), one site ahead: the record of the next site
// (across phases too) is requested before the current site is updated in
// both lattices, and the barrier between phases waits for LDS traffic only,
// so the load stays in flight across it.  A thread owns the Philox quads
// j = tid and j = tid + NT (two quads per thread, NT = 768 at L = 18 and
// 1024 at L = 20); each quad's block is drawn once per sweep and lattice,
// words 0, 1 for colour 0 and 2, 3 for colour 1, as in k_block.  The quad's
// sites h = j and h = j + N/4 share a column (x, y); their neighbours are
// kept as offsets from h.  Same task, PT and energy semantics as k_block.
template <int L>
struct MidGeo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int Q = N / 4;
  static constexpr int LH = L / 2;
  static constexpr int PL = LH * L;                    // words per z-plane, one colour
  static constexpr int NT = ((Q / 2 + 63) / 64) * 64;  // two quads per thread
  static constexpr int NLOAD = (NH + NT - 1) / NT;
  static constexpr int SP = LH & 1;                    // site 1 parity flip
};

// Work-group barrier that waits for this wave's LDS operations only: global
// loads in flight stay in flight across it.  The empty asm statements keep
// the compiler from moving memory operations across the barrier.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt not waited for
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}


// The six bond words of record idx of the record array at Jsb (a group's
// block, or its colour-c half): a uniform base and a 32-bit byte offset, so
// the loads take the SGPR-base form and no 64-bit address per site is kept.
__device__ __forceinline__ void site_record(const SG_GLOBAL char *__restrict__ Jsb, int idx,
                                            bool valid, uint64_t (&jw)[6]) {
  if (!valid) return;
  uint32_t off = (uint32_t)idx * 48u;
  asm volatile("" : "+v"(off));   // keeps base + zext(off)
  const SG_GLOBAL u64x2 *r = (const SG_GLOBAL u64x2 *)(Jsb + off);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const u64x2 v = r[k];
    jw[2 * k] = v.x;
    jw[2 * k + 1] = v.y;
  }
}

template <int L, bool FIELD>
__global__ __launch_bounds__(MidGeo<L>::NT) void k_mid(
    const JobPtrs *__restrict__ jt, const TaskDesc *__restrict__ tasks, int do_pt,
    int nsweeps, int want_energy) {
  using G = MidGeo<L>;
  static_assert(G::NT <= G::Q && G::Q <= 2 * G::NT, "two quads per thread");
  __shared__ uint64_t latB[2][G::NH], latW[2][G::NH];
  __shared__ int32_t redi[4 * 64];

  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  SG_GLOBAL uint64_t *__restrict__ spins = to_global(jp.spins);
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy);
  unsigned long long *__restrict__ counters = jp.counters;
  const bool two = tk.lat[1] != kNone;
  const LatDesc d0 = ld_uniform(jp.descs + tk.lat[0]);
  const LatDesc d1 = ld_uniform(jp.descs + (two ? tk.lat[1] : tk.lat[0]));
  const SG_GLOBAL char *__restrict__ Js = (const SG_GLOBAL char *)(to_global(jp.Js) + 2 * d0.j_off);
  const int tid = threadIdx.x, lane = tid & 63;
  const bool act1 = tid + G::NT < G::Q;   // second quad present

  const bool pt_here = do_pt && two && tk.pair != kNone;
  int32_t ua = 0, ub = 0, ka = 0, kb = 0;
  PairDesc pd{};
  uint32_t pt_b = 0;
  if (pt_here) {
    pd = ld_uniform(jp.pairs + tk.pair);
    const JobState js = ld_uniform(jp.js);
    pt_b = js.sweep / js.pt_every;
    const SG_GLOBAL int32_t *Ea = energy + (size_t)pd.e_a * 128, *Eb = energy + (size_t)pd.e_b * 128;
    ua = Ea[lane]; ub = Eb[lane]; ka = Ea[64 + lane]; kb = Eb[64 + lane];
  }

  const SG_GLOBAL uint64_t *g0 = spins + d0.spin_off, *g1 = spins + d1.spin_off;
  uint64_t w[G::NLOAD][4];
#pragma unroll
  for (int k = 0; k < G::NLOAD; ++k) {
    const int i = tid + k * G::NT;
    if (i < G::NH) {
      w[k][0] = g0[i];
      w[k][1] = g0[G::NH + i];
      if (two) {
        w[k][2] = g1[i];
        w[k][3] = g1[G::NH + i];
      }
    }
  }

  // site (q, s): h = tid + q NT + s Q; the record of the first site
  uint64_t jc[6], jn[6];
  site_record(Js, tid, true, jc);

  // neighbour offsets from h in the opposite colour's half.  x: one of the
  // two x-neighbours shares the index h, the other is h +- 1 (wrapping within
  // the row), depending on the site's parity (y + z + c) & 1; xo[q][c] holds
  // the offsets (+x, -x) of site 0 at colour c, and site 1 at colour c uses
  // those of colour c ^ SP.  y: dy[q] (+y, -y).  z: site 0 (z < L/2) has +z
  // at +PL and -z at dz[q][0]; site 1 (z >= L/2) +z at dz[q][1], -z at -PL.
  int xo[2][2][2], dy[2][2], dz[2][2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = tid + q * G::NT;
    const int xh = j % G::LH, y = (j / G::LH) % L, z0 = j / G::PL;
    const int dxp = xh == G::LH - 1 ? -(G::LH - 1) : 1;
    const int dxm = xh == 0 ? G::LH - 1 : -1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const bool odd = ((y + z0 + c) & 1) != 0;
      xo[q][c][0] = odd ? dxp : 0;
      xo[q][c][1] = odd ? 0 : dxm;
    }
    dy[q][0] = y == L - 1 ? -(L - 1) * G::LH : G::LH;
    dy[q][1] = y == 0 ? (L - 1) * G::LH : -G::LH;
    dz[q][0] = z0 == 0 ? (L - 1) * G::PL : -G::PL;
    dz[q][1] = z0 + G::LH == L - 1 ? -(L - 1) * G::PL : G::PL;
  }

  uint64_t m = 0;
  if (pt_here) {
    const uint32_t sw3 = (kTagPT << 28) | (pd.rt & 0x0FFFFFFFu);
    const u32x4 r = philox_dev((uint32_t)(lane >> 2), pt_b, pd.g, sw3, pd.key0, pd.key1);
    const bool acc = pt_accept(pd.beta_a, pd.beta_b, ua, ub, ka, kb, pd.H, pick_word(r, lane));
    m = __ballot(acc);
  }
#pragma unroll
  for (int k = 0; k < G::NLOAD; ++k) {
    const int i = tid + k * G::NT;
    if (i < G::NH) {
      if (two) {
        const uint64_t db = (w[k][0] ^ w[k][2]) & m, dw = (w[k][1] ^ w[k][3]) & m;
        latB[0][i] = w[k][0] ^ db;
        latW[0][i] = w[k][1] ^ dw;
        latB[1][i] = w[k][2] ^ db;
        latW[1][i] = w[k][3] ^ dw;
      } else {
        latB[0][i] = w[k][0];
        latW[0][i] = w[k][1];
      }
    }
  }
  ThrRec tr0{}, tr1{};
  uint32_t n0 = 0;
  if (nsweeps > 0) {
    tr0 = ld_uniform(jp.thr + d0.thr_idx);
    tr1 = ld_uniform(jp.thr + d1.thr_idx);
    n0 = ld_uniform(jp.js).sweep;
  }
  lds_barrier();

  // the six neighbours of site (q, s) at colour c, from the half opp
#define MID_NEIGHBOURS(opp, q, s, c, h, nv)                                               \
  do {                                                                                   \
    const int cs_ = (c) ^ ((s) & G::SP);                                                 \
    nv[0] = opp[h + xo[q][cs_][0]];                                                      \
    nv[1] = opp[h + xo[q][cs_][1]];                                                      \
    nv[2] = opp[h + dy[q][0]];                                                           \
    nv[3] = opp[h + dy[q][1]];                                                           \
    nv[4] = opp[h + ((s) == 0 ? G::PL : dz[q][1])];                                      \
    nv[5] = opp[h + ((s) == 0 ? dz[q][0] : -G::PL)];                                     \
  } while (0)

  const uint32_t sw0 = (kTagMC << 28) | (d0.rt & 0x0FFFFFFFu);
  const uint32_t sw1 = (kTagMC << 28) | (d1.rt & 0x0FFFFFFFu);
  for (int it = 0; it < nsweeps; ++it) {
    const uint32_t n = n0 + (uint32_t)it;
    uint32_t r1[2] = {0, 0}, r23[2][2][2];   // [q][lattice][word 2, 3]
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          // the next site's record: site 1 of this quad, site 0 of quad 1, or
          // the next phase's first site
          if (s == 0) {
            site_record(Js + c * G::NH * 48, tid + q * G::NT + G::Q, q == 0 || act1, jn);
          } else if (q == 0) {
            const int nidx = act1 ? c * G::NH + tid + G::NT : (1 - c) * G::NH + tid;
            site_record(Js, nidx, true, jn);
          } else {
            site_record(Js + (1 - c) * G::NH * 48, tid, true, jn);
          }
          if (q == 0 || act1) {
            const int h = tid + q * G::NT + s * G::Q;
            uint32_t rw[2];
            if (c == 0 && s == 0) {
              const uint32_t j = (uint32_t)(tid + q * G::NT);
              const u32x4 ra = philox_sweep(j, n, d0.g, sw0, d0.key0, d0.key1);
              u32x4 rb = ra;
              if (two) rb = philox_sweep(j, n, d1.g, sw1, d1.key0, d1.key1);
              rw[0] = ra.v[0]; r1[0] = ra.v[1]; r23[q][0][0] = ra.v[2]; r23[q][0][1] = ra.v[3];
              rw[1] = rb.v[0]; r1[1] = rb.v[1]; r23[q][1][0] = rb.v[2]; r23[q][1][1] = rb.v[3];
            } else if (c == 0) {
              rw[0] = r1[0];
              rw[1] = r1[1];
            } else {
              rw[0] = r23[q][0][s];
              rw[1] = r23[q][1][s];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
              if (a == 1 && !two) break;
              uint64_t *mine = c == 0 ? latB[a] : latW[a];
              const uint64_t *opp = c == 0 ? latW[a] : latB[a];
              uint64_t nv[6];
              MID_NEIGHBOURS(opp, q, s, c, h, nv);
              uint64_t sv = mine[h];
              update_site<L, FIELD>(sv, nv, jc, rw[a], a == 0 ? tr0 : tr1);
              mine[h] = sv;
            }
          }
#pragma unroll
          for (int e = 0; e < 6; ++e) jc[e] = jn[e];
        }
      }
      lds_barrier();
    }
  }

  // energies of the final states (every bond has one black end); jc holds
  // the record of colour 0, site (0, 0)
  uint64_t AU[2][1][5] = {}, AK[2][1][4] = {};
  if (want_energy) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (s == 0)
          site_record(Js, tid + q * G::NT + G::Q, q == 0 || act1, jn);
        else if (q == 0)
          site_record(Js, tid + G::NT, act1, jn);
        if (q == 0 || act1) {
          const int h = tid + q * G::NT + s * G::Q;
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            if (a == 1 && !two) break;
            uint64_t nv[6];
            MID_NEIGHBOURS(latW[a], q, s, 0, h, nv);
            const uint64_t sv = latB[a][h];
            uint64_t u0, u1, u2;
            bond_count(sv, nv, jc, u0, u1, u2);
            vc_add3(AU[a][0], u0, u1, u2);
            if (want_energy > 1) {
              vc_add1(AK[a][0], sv);
              vc_add1(AK[a][0], latW[a][h]);
            }
          }
        }
#pragma unroll
        for (int e = 0; e < 6; ++e) jc[e] = jn[e];
      }
    }
  }
#undef MID_NEIGHBOURS

  {
    SG_GLOBAL uint64_t *o0 = spins + d0.spin_off, *o1 = spins + d1.spin_off;
    for (int i = tid; i < G::NH; i += G::NT) {
      o0[i] = latB[0][i];
      o0[G::NH + i] = latW[0][i];
    }
    if (two)
      for (int i = tid; i < G::NH; i += G::NT) {
        o1[i] = latB[1][i];
        o1[G::NH + i] = latW[1][i];
      }
  }

  if (want_energy) {
    // the lattices have left: their black halves serve as the fold scratch
    // (5 NT <= 2 NH words), one lattice and one count at a time
    uint64_t *fold = &latB[0][0];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (a == 1 && !two) break;
      const LatDesc &d = a == 0 ? d0 : d1;
      __syncthreads();
      int32_t U[1];
      wg_count_lanes<5, G::NT, 1>(AU[a], fold, redi, U);
      if (tid < 64) energy[(size_t)d.e_idx * 128 + tid] = U[0];
      if (want_energy > 1) {
        __syncthreads();
        int32_t K[1];
        wg_count_lanes<4, G::NT, 1>(AK[a], fold, redi, K);
        if (tid < 64) energy[(size_t)d.e_idx * 128 + 64 + tid] = K[0];
      }
    }
  }
  if (pt_here && tid == 0) atomicAdd(&counters[pd.cnt_idx], (unsigned long long)__popcll(m));
}

// ------------------------------------------- single-lattice resident kernel
// L = 22, 24, 26: one lattice fits in LDS (N words, 137 KB at L = 26), two
// do not, so a task is one lattice and its PT step runs before the launch
// (k_pt_chunk over its pairs).  Thread t owns the column col = t mod PL
// (x/2, y fixed) and the z-run of QZ quads z0 = zg QZ + k, zg = t / PL: its x
// and y neighbour offsets are fixed per thread, the site parity alternates
// with k, and only the quads at z0 = 0 and z0 = L/2 - 1 have z wraps.  Per
// sweep each quad's Philox block is drawn once (words 2, 3 kept for colour
// 1); the coupling records of the next quad (or of the next phase's first
// quad) are requested before the current quad is updated, so one quad's
// records are in flight while another's are used.  Energies as in k_block; the fold
// scratch of the reduction spans both colour halves.
template <int L>
struct OneGeo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int Q = N / 4;
  static constexpr int LH = L / 2;
  static constexpr int PL = LH * L;                           // columns
  // z-runs per column: as many as 1024 threads allow, except at L = 24,
  // where three runs (864 threads, 14 waves) are held to 128 VGPRs and
  // spill; two runs (576 threads, up to 168 VGPRs) measure faster there
  static constexpr int ZG = L == 24 ? 2 : (1024 / PL < LH ? 1024 / PL : LH);
  static constexpr int QZ = (LH + ZG - 1) / ZG;               // quads per thread
  static constexpr int NA = PL * ZG;                          // threads with quads
  static constexpr int NT = ((NA + 63) / 64) * 64;
  static constexpr int NLOAD = (NH + NT - 1) / NT;
  static constexpr int SP = LH & 1;                           // site 1 parity flip
  static constexpr int PU = ceil_log2(12 * QZ + 1);           // per-thread U bits
  static constexpr int PK = ceil_log2(4 * QZ + 1);            // per-thread K bits
};

template <int L, bool FIELD>
__global__ __launch_bounds__(OneGeo<L>::NT) void k_one(const JobPtrs *__restrict__ jt,
                                                       const TaskDesc *__restrict__ tasks,
                                                       int nsweeps, int want_energy) {
  using G = OneGeo<L>;
  static_assert(((G::PU + 1) / 2) * G::NT <= G::NH && ((G::PK + 1) / 2) * G::NT <= G::NH,
                "fold scratch within the colour halves");
  __shared__ uint64_t latB[G::NH], latW[G::NH];
  __shared__ int32_t redi[4 * 64];

  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  SG_GLOBAL uint64_t *__restrict__ spins = to_global(jp.spins);
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy);
  const LatDesc d = ld_uniform(jp.descs + tk.lat[0]);
  const SG_GLOBAL char *__restrict__ Js = (const SG_GLOBAL char *)(to_global(jp.Js) + 2 * d.j_off);
  const int tid = threadIdx.x;
  const bool act = tid < G::NA;
  const int col = act ? tid % G::PL : 0, zg = act ? tid / G::PL : 0;

  SG_GLOBAL uint64_t *g = spins + d.spin_off;
  uint64_t w[G::NLOAD][2];
#pragma unroll
  for (int k = 0; k < G::NLOAD; ++k) {
    const int i = tid + k * G::NT;
    if (i < G::NH) {
      w[k][0] = g[i];
      w[k][1] = g[G::NH + i];
    }
  }

  // quad k of the thread: j = col + PL z0, valid while z0 < L/2 (k < nq);
  // its sites h = j + s Q, s = 0, 1
  const int jb = col + G::PL * zg * G::QZ;
  const int nq = act ? G::LH - zg * G::QZ : 0;
#define quad_ok(k) ((k) < nq)
  // records of both sites of quad k at colour c
#define ONE_RECORDS(c, k, jw)                                                             \
  do {                                                                                   \
    site_record(Js + (c) * G::NH * 48, jb + (k) * G::PL, quad_ok(k), jw[0]);             \
    site_record(Js + (c) * G::NH * 48, jb + (k) * G::PL + G::Q, quad_ok(k), jw[1]);      \
  } while (0)
  uint64_t jc[2][6], jn[2][6];
  ONE_RECORDS(0, 0, jc);

  // x offsets by site parity (1: +x at h + dxp, -x at h; 0: +x at h, -x at
  // h + dxm), XO[b] for parity p0 ^ b, p0 the parity of the column's first
  // site at colour 0; y offsets; z offsets of the wrapping quads
  const int xh = col % G::LH, y = col / G::LH;
  const int dxp = xh == G::LH - 1 ? -(G::LH - 1) : 1;
  const int dxm = xh == 0 ? G::LH - 1 : -1;
  const bool p0 = ((y + zg * G::QZ) & 1) != 0;
  int XO[2][2];
  XO[0][0] = p0 ? dxp : 0;
  XO[0][1] = p0 ? 0 : dxm;
  XO[1][0] = p0 ? 0 : dxp;
  XO[1][1] = p0 ? dxm : 0;
  const int dyp = y == L - 1 ? -(L - 1) * G::LH : G::LH;
  const int dym = y == 0 ? (L - 1) * G::LH : -G::LH;
  const int dz_first = zg == 0 ? (L - 1) * G::PL : -G::PL;    // -z of site 0, quad 0

#pragma unroll
  for (int k = 0; k < G::NLOAD; ++k) {
    const int i = tid + k * G::NT;
    if (i < G::NH) {
      latB[i] = w[k][0];
      latW[i] = w[k][1];
    }
  }
  ThrRec tr{};
  uint32_t n0 = 0;
  if (nsweeps > 0) {
    tr = ld_uniform(jp.thr + d.thr_idx);
    n0 = ld_uniform(jp.js).sweep;
  }
  lds_barrier();

  // the six neighbours of site s of quad k at colour c, from the half opp
#define ONE_NEIGHBOURS(opp, k, s, c, h, nv)                                               \
  do {                                                                                   \
    const int b_ = ((k) ^ (c) ^ ((s) & G::SP)) & 1;                                      \
    nv[0] = opp[h + XO[b_][0]];                                                          \
    nv[1] = opp[h + XO[b_][1]];                                                          \
    nv[2] = opp[h + dyp];                                                                \
    nv[3] = opp[h + dym];                                                                \
    if ((s) == 0) {                                                                      \
      nv[4] = opp[h + G::PL];                                                            \
      nv[5] = opp[h + ((k) == 0 ? dz_first : -G::PL)];                                   \
    } else {                                                                             \
      nv[4] = opp[h + (zg * G::QZ + (k) == G::LH - 1 ? -(L - 1) * G::PL : G::PL)];       \
      nv[5] = opp[h - G::PL];                                                            \
    }                                                                                    \
  } while (0)

  const uint32_t sw = (kTagMC << 28) | (d.rt & 0x0FFFFFFFu);
  for (int it = 0; it < nsweeps; ++it) {
    const uint32_t n = n0 + (uint32_t)it;
    uint32_t r23[G::QZ][2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      uint64_t *mine = c == 0 ? latB : latW;
      const uint64_t *opp = c == 0 ? latW : latB;
#pragma unroll
      for (int k = 0; k < G::QZ; ++k) {
        // the next quad's records (the next phase's first quad after the last)
        if (k + 1 < G::QZ)
          ONE_RECORDS(c, k + 1, jn);
        else
          ONE_RECORDS(1 - c, 0, jn);
        if (quad_ok(k)) {
          const int j = jb + k * G::PL;
          uint32_t rw[2];
          if (c == 0) {
            const u32x4 r = philox_sweep((uint32_t)j, n, d.g, sw, d.key0, d.key1);
            rw[0] = r.v[0];
            rw[1] = r.v[1];
            r23[k][0] = r.v[2];
            r23[k][1] = r.v[3];
          } else {
            rw[0] = r23[k][0];
            rw[1] = r23[k][1];
          }
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            const int h = j + s * G::Q;
            uint64_t nv[6];
            ONE_NEIGHBOURS(opp, k, s, c, h, nv);
            uint64_t sv = mine[h];
            update_site<L, FIELD>(sv, nv, jc[s], rw[s], tr);
            mine[h] = sv;
          }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
          for (int e = 0; e < 6; ++e) jc[s][e] = jn[s][e];
      }
      lds_barrier();
    }
  }

  // energies of the final state; jc holds the records of colour 0, quad 0
  uint64_t AU[1][G::PU] = {}, AK[1][G::PK] = {};
  if (want_energy) {
#pragma unroll
    for (int k = 0; k < G::QZ; ++k) {
      if (k + 1 < G::QZ) ONE_RECORDS(0, k + 1, jn);
      if (quad_ok(k)) {
        const int j = jb + k * G::PL;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int h = j + s * G::Q;
          uint64_t nv[6];
          ONE_NEIGHBOURS(latW, k, s, 0, h, nv);
          const uint64_t sv = latB[h];
          uint64_t u0, u1, u2;
          bond_count(sv, nv, jc[s], u0, u1, u2);
          vc_add3(AU[0], u0, u1, u2);
          if (want_energy > 1) {
            vc_add1(AK[0], sv);
            vc_add1(AK[0], latW[h]);
          }
        }
      }
      if (k + 1 < G::QZ)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
          for (int e = 0; e < 6; ++e) jc[s][e] = jn[s][e];
    }
  }
#undef ONE_NEIGHBOURS
#undef ONE_RECORDS

  for (int i = tid; i < G::NH; i += G::NT) {
    g[i] = latB[i];
    g[G::NH + i] = latW[i];
  }

  if (want_energy) {
    // the lattice has left: its halves serve as the fold scratch
    constexpr int PSU = (G::PU + 1) / 2, PSK = (G::PK + 1) / 2;
    __syncthreads();
    int32_t U[1];
    wg_count_lanes<G::PU, G::NT, 1, PSU>(AU, latB, redi, U, latW);
    if (tid < 64) energy[(size_t)d.e_idx * 128 + tid] = U[0];
    if (want_energy > 1) {
      __syncthreads();
      int32_t K[1];
      wg_count_lanes<G::PK, G::NT, 1, PSK>(AK, latB, redi, K, latW);
      if (tid < 64) energy[(size_t)d.e_idx * 128 + 64 + tid] = K[0];
    }
  }
#undef quad_ok
}

// ---------------------------------------------------------- measure kernel
// One work-group (4 waves) per (group, temperature); lane l of every wave is
// sample l.  U and K of every replica are the per-lane energy slots, which
// the sweep launch ending at a measurement fills (U and K).  The overlap
// w = s1 XOR s2 of replicas 0 and 1 is staged in LDS a chunk of z-planes at
// a time with coalesced loads; each wave then walks a quarter of the chunk's
// (y, z) rows reading the words as LDS broadcasts, and each lane counts its
// own bit into W and the x-, y- and z-plane profiles.  Reference
// counterpart: compute_q (eamsim.cu:451-511) plus the E, M scan of
// perform_PT_shuffle.
template <int L>
struct MeasGeo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int LH = L / 2;
  static constexpr int PL = LH * L;                   // words per colour plane
  // z-planes per chunk: the largest divisor of L whose staging fits 48 KiB
  static constexpr int zc(int z) { return z <= 1 ? 1 : (L % z == 0 && 16 * PL * z <= 49152 ? z : zc(z - 1)); }
  static constexpr int ZC = zc(L);
};

template <int L>
__global__ __launch_bounds__(256) void k_measure(const uint64_t *__restrict__ spins,
                                                 const int32_t *__restrict__ energy,
                                                 const MeasDesc *__restrict__ md, int R,
                                                 int32_t *__restrict__ out, int nfield,
                                                 ObsAcc *__restrict__ acc,
                                                 const double *__restrict__ cs,
                                                 int write_records) {
  using G = MeasGeo<L>;
  constexpr int N = G::N;
  __shared__ uint64_t wbuf[2][G::ZC][G::PL];
  __shared__ int32_t sx[L][64], sy[L][64], sz[L][64];
  __shared__ int32_t sW[64];
  const MeasDesc d = md[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (R >= 2) {
    for (int i = tid; i < L * 64; i += 256) {
      (&sx[0][0])[i] = 0;
      (&sy[0][0])[i] = 0;
      (&sz[0][0])[i] = 0;
    }
    if (tid < 64) sW[tid] = 0;
    const uint64_t *S0 = spins + d.spin_off, *S1 = S0 + d.rho_stride;
    int32_t wx[L];
#pragma unroll
    for (int x = 0; x < L; ++x) wx[x] = 0;
    int32_t wt = 0;
    for (int z0 = 0; z0 < L; z0 += G::ZC) {
      __syncthreads();   // the previous chunk has been read
      constexpr int W = 2 * G::ZC * G::PL;
      for (int i = tid; i < W; i += 256) {
        const int c = i / (G::ZC * G::PL), r = i % (G::ZC * G::PL);
        const int k = c * G::NH + z0 * G::PL + r;     // word: colour c, plane z0 + r / PL
        (&wbuf[0][0][0])[i] = S0[k] ^ S1[k];
      }
      __syncthreads();
      for (int row = wave; row < G::ZC * L; row += 4) {
        const int zz = row / L, y = row % L, z = z0 + zz;
        int32_t rs = 0;
#pragma unroll
        for (int x = 0; x < L; ++x) {
          const int c = (x + y + z) & 1;
          const int32_t wb = (int32_t)((wbuf[c][zz][(x >> 1) + G::LH * y] >> lane) & 1);
          wx[x] += wb;
          rs += wb;
        }
        atomicAdd(&sy[y][lane], rs);
        atomicAdd(&sz[z][lane], rs);
        wt += rs;
      }
    }
#pragma unroll
    for (int x = 0; x < L; ++x) atomicAdd(&sx[x][lane], wx[x]);
    atomicAdd(&sW[lane], wt);
  }
  __syncthreads();
  if (tid < 64) {
    // fold this measurement into the lane's running sums (one owner per lane)
    ObsAcc a = acc[d.acc_off + tid];
    for (int rho = 0; rho < R; ++rho) {
      const int32_t *e = energy + (size_t)(d.e_idx + (uint64_t)rho * d.e_stride) * 128;
      a.sU += e[tid];
      a.sK += e[64 + tid];
    }
    if (R >= 2) {
#pragma clang fp contract(off)
      const int64_t Q = (int64_t)N - 2 * (int64_t)sW[tid];
      const uint64_t Q2 = (uint64_t)(Q * Q);
      const unsigned __int128 Q4 = (unsigned __int128)Q2 * Q2;
      const unsigned __int128 s4 =
          (((unsigned __int128)a.sQ4hi << 64) | a.sQ4lo) + Q4;
      a.sQ += Q;
      a.sAbsQ += Q < 0 ? -Q : Q;
      a.sQ2 += Q2;
      a.sQ4lo = (uint64_t)s4;
      a.sQ4hi = (uint64_t)(s4 >> 64);
      const double *cw = cs + d.cs_off, *sw = cw + L;
      double q1 = 0.0;
      for (int dir = 0; dir < 3; ++dir) {
        double re = 0.0, im = 0.0;
        for (int c = 0; c < L; ++c) {
          const double w = (double)(dir == 0 ? sx[c][tid] : dir == 1 ? sy[c][tid] : sz[c][tid]);
          re += cw[c] * w;
          im += sw[c] * w;
        }
        re *= -2.0 / N;
        im *= -2.0 / N;
        q1 += re * re + im * im;
      }
      a.sq1 += q1 / 3.0;
    }
    acc[d.acc_off + tid] = a;
  }
  if (!write_records) return;
  int32_t *o = out + d.out_off;
  for (int i = tid; i < 64 * nfield; i += 256) {
    const int l = i / nfield, f = i % nfield;
    int32_t v;
    if (f < 2 * R) {
      const int rho = f < R ? f : f - R;
      v = energy[(size_t)(d.e_idx + (uint64_t)rho * d.e_stride) * 128 + (f < R ? 0 : 64) + l];
    } else if (f == 2 * R) {
      v = R >= 2 ? sW[l] : 0;
    } else {
      const int g = f - 2 * R - 1;
      const int dir = g / L, c = g % L;
      v = R < 2 ? 0 : (dir == 0 ? sx[c][l] : dir == 1 ? sy[c][l] : sz[c][l]);
    }
    o[i] = v;
  }
}

// ------------------------------------------------------------ init kernels
template <int L>
__global__ __launch_bounds__(256) void k_init_spins(uint64_t *__restrict__ spins,
                                                    const LatDesc *__restrict__ descs) {
  constexpr int N = L * L * L;
  const LatDesc d = descs[blockIdx.y];
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= N) return;
  const int x = s % L, y = (s / L) % L, z = s / (L * L);
  const uint32_t sw3 = (kTagSpin << 28) | (d.rt & 0x0FFFFFFFu);
  const u32x4 r = philox((uint32_t)s, 0u, d.g, sw3, d.key0, d.key1);
  spins[d.spin_off + cb_index<L>(x, y, z)] = (uint64_t)r.v[0] | ((uint64_t)r.v[1] << 32);
}

#ifdef SG_MAIN_PART  // non-template kernels: defined once, in sg_kernels.hip
__global__ __launch_bounds__(256) void k_init_couplings(uint64_t *__restrict__ J, uint64_t j_off,
                                                        int N, uint32_t g, uint32_t key0,
                                                        uint32_t key1) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int dir = blockIdx.y;
  if (s >= N) return;
  const u32x4 r = philox((uint32_t)s, (uint32_t)dir, g, kTagCoup << 28, key0, key1);
  J[j_off + (uint64_t)dir * N + s] = (uint64_t)r.v[0] | ((uint64_t)r.v[1] << 32);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_advance(const JobPtrs *jt, int njobs, int nsweeps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < njobs) jt[i].js->sweep += (uint32_t)nsweeps;
}

#endif

// ------------------------------------------------------- coupling site records
// Site record of colour c, checkerboard index h: the six bond words of that
// site (+x, -x, +y, -y, +z, -z) at word (c*NH + h)*6 of the group's block
// (6N words, at 2*j_off).  One site's couplings are then three 16-byte loads
// and a wave reads 3 KiB contiguous, where the natural [3][N] layout gives
// six 8-byte loads at stride 2.
template <int L>
__global__ __launch_bounds__(256) void k_build_jsite(uint64_t *__restrict__ Js,
                                                     const uint64_t *__restrict__ J,
                                                     uint64_t j_off) {
  constexpr int N = L * L * L, NH = N / 2, LH = L / 2;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const int c = i / NH, h = i % NH;
  const int xh = h % LH, y = (h / LH) % L, z = h / (LH * L);
  const int x = 2 * xh + ((y + z + c) & 1);
  const int xm = (x + L - 1) % L, ym = (y + L - 1) % L, zm = (z + L - 1) % L;
  const uint64_t *Jx = J + j_off, *Jy = Jx + N, *Jz = Jy + N;
  const int s = x + L * (y + L * z);
  uint64_t *o = Js + 2 * j_off + (uint64_t)i * 6;
  o[0] = Jx[s];
  o[1] = Jx[xm + L * (y + L * z)];
  o[2] = Jy[s];
  o[3] = Jy[x + L * (ym + L * z)];
  o[4] = Jz[s];
  o[5] = Jz[x + L * (y + L * zm)];
}

// ---------------------------------------------------------- per-phase kernel
// One launch updates one colour of every lattice of its task list, the
// lattices streamed from HBM/L2 in chunks of kPhaseNT Philox quads: a task
// is (lattice, chunk), a thread owns quad j and updates the sites h = j and
// h = j + N/4 of the launch's colour (words 2c and 2c+1 of the quad's Philox
// block).  The kernel boundary orders the phases, so a lattice spreads over
// as many work-groups as it has chunks (108 at L = 48) with no waiting
// between them.  mode bits: update the spins; zero the lattice's energy
// slots (chunk 0; before any accumulation, in an earlier launch); accumulate
// U (and K) of the states after this launch (colour 1 only: every bond has
// one white end, every site index h one white and one black word) into the
// slots with per-lane atomics.
template <int L, bool FIELD, bool ENG, int QPT>
__global__ __launch_bounds__(kPhaseNT / QPT) void k_phase(const JobPtrs *__restrict__ jt,
                                                          const TaskDesc *__restrict__ tasks,
                                                          int c, int it, int mode) {
  using G = Geo<L>;
  constexpr int NT = kPhaseNT / QPT;        // QPT quads per thread: a tile of the tuner
  constexpr int PU = ceil_log2(12 * QPT + 1), PK = ceil_log2(4 * QPT + 1);
  static_assert(kPhaseNT % QPT == 0 && NT % 64 == 0, "whole waves");
  __shared__ uint64_t fold[PU * NT];
  __shared__ int32_t redi[4 * 64];
  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  const LatDesc d = ld_uniform(jp.descs + tk.lat[0]);
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy) + (size_t)d.e_idx * 128;
  const int tid = threadIdx.x;
  if ((mode & kPhZero) && tk.lat[1] == 0)
    for (int i = tid; i < 128; i += NT) energy[i] = 0;
  SG_GLOBAL uint64_t *__restrict__ lat = to_global(jp.spins) + d.spin_off;
  SG_GLOBAL uint64_t *__restrict__ mine = lat + c * G::NH;
  const SG_GLOBAL uint64_t *__restrict__ opp = lat + (1 - c) * G::NH;
  const SG_GLOBAL uint64_t *__restrict__ Js =
      to_global(jp.Js) + 2 * d.j_off + (uint64_t)c * G::NH * 6;
  const bool upd = (mode & kPhUpdate) != 0;
  const bool eng = ENG && (mode & (kPhEnergy | kPhEnergyK)) != 0;
  constexpr int PL = G::LH * L;  // words per plane of one colour
  // quad j = xh + (L/2)(y + L z) with z < L/2; its sites are (xh, y, z) and
  // (xh, y, z + L/2): one column, one set of divisions.  Every load of the
  // thread's quads is issued before the Philox rounds, which then hide
  // their latency.
  int jq[QPT], hq[QPT][2];
  bool act[QPT];
  uint64_t jw[QPT][2][6], nv[QPT][2][6], sq[QPT][2], bq[QPT][2];
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    const int j = (int)tk.lat[1] * kPhaseNT + tid + k * NT;
    jq[k] = j;
    act[k] = j < G::Q && (upd || eng);
    const int xh = j % G::LH, y = (j / G::LH) % L, z0 = j / (G::LH * L);
    const int ry = G::LH * y;
    const int ryp = G::LH * (y + 1 == L ? 0 : y + 1), rym = G::LH * (y == 0 ? L - 1 : y - 1);
    bq[k][0] = bq[k][1] = 0;
    if (act[k]) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int z = z0 + q * (L / 2);
        const int pz = PL * z;
        const int pzp = PL * (z + 1 == L ? 0 : z + 1), pzm = PL * (z == 0 ? L - 1 : z - 1);
        const int x = 2 * xh + ((y + z + c) & 1);
        const int xp = x + 1 == L ? 0 : x + 1, xm = x == 0 ? L - 1 : x - 1;
        const int h = xh + ry + pz;
        hq[k][q] = h;
        const SG_GLOBAL uint64_t *jr = Js + (size_t)h * 6;
#pragma unroll
        for (int e = 0; e < 6; ++e) jw[k][q][e] = jr[e];
        nv[k][q][0] = opp[(xp >> 1) + ry + pz];
        nv[k][q][1] = opp[(xm >> 1) + ry + pz];
        nv[k][q][2] = opp[xh + ryp + pz];
        nv[k][q][3] = opp[xh + rym + pz];
        nv[k][q][4] = opp[xh + ry + pzp];
        nv[k][q][5] = opp[xh + ry + pzm];
        sq[k][q] = mine[h];
        if (ENG && (mode & kPhEnergyK)) bq[k][q] = opp[h];
      }
    }
  }
  // each quad's Philox block, drawn by both colour launches of the sweep:
  // words 2c and 2c + 1 (recomputing it costs less than passing two words
  // per quad through memory from one launch to the next)
  uint32_t rw[QPT][2];
  ThrRec tr{};
  if (upd) tr = ld_uniform(jp.thr + d.thr_idx);
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    rw[k][0] = rw[k][1] = 0;
    if (upd && jq[k] < G::Q) {
      const uint32_t n = ld_uniform(jp.js).sweep + (uint32_t)it;
      const u32x4 r = philox_sweep((uint32_t)jq[k], n, d.g, (kTagMC << 28) | (d.rt & 0x0FFFFFFFu),
                                   d.key0, d.key1);
      rw[k][0] = c ? r.v[2] : r.v[0];
      rw[k][1] = c ? r.v[3] : r.v[1];
    }
  }
  uint64_t AU[1][PU] = {}, AK[1][PK] = {};
#pragma unroll
  for (int k = 0; k < QPT; ++k) {
    if (!act[k]) continue;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      uint64_t s = sq[k][q];
      if (upd) {
        update_site<L, FIELD>(s, nv[k][q], jw[k][q], rw[k][q], tr);
        mine[hq[k][q]] = s;
      }
      if (eng) {
        uint64_t u0, u1, u2;
        bond_count(s, nv[k][q], jw[k][q], u0, u1, u2);
        vc_add3(AU[0], u0, u1, u2);
        if (ENG && (mode & kPhEnergyK)) {
          vc_add1(AK[0], s);
          vc_add1(AK[0], bq[k][q]);
        }
      }
    }
  }
  if (eng) {
    int32_t U[1];
    wg_count_lanes<PU, NT, 1>(AU, fold, redi, U);
    if (tid < 64) atomicAdd((int32_t *)(energy + tid), U[0]);
    if (ENG && (mode & kPhEnergyK)) {
      __syncthreads();
      int32_t K[1];
      wg_count_lanes<PK, NT, 1>(AK, fold, redi, K);
      if (tid < 64) atomicAdd((int32_t *)(energy + 64 + tid), K[0]);
    }
  }
}

// ------------------------------------------------------------- slab kernel
// Streamed lattices, one launch per sweep.  A task is (lattice, slab pair p):
// the owned planes are z in [za, za + S) and their images z + L/2 (za = pS,
// S = L/8, four slab pairs per lattice), so that each owned Philox quad's
// four words are used by the task that draws them.  Phase 1 updates colour
// 0 on the owned planes and one halo plane on each side, into LDS; the halo
// updates recompute, bit for bit, what the neighbouring task computes for
// its own planes (same counter-based random numbers, same inputs).  Phase 2
// updates colour 1 on the owned planes from the new colour 0 in LDS.  The
// sweep reads one spin buffer and writes the other (dir selects which), so
// no task reads a word that another task of the launch writes.  The
// colour-1 planes the task reads (owned plus two halo planes each side) are
// staged in LDS first; the kernel exists where both colours' planes fit
// (L = 32: 112 KB).  Where they do not (L = 40, 48), reading colour 1 from
// L2 in the item loop measured slower than the colour launches (k_phase),
// which remain the streamed kernel there.  Energies as in k_phase: zeroed by
// one launch (kPhZero, task p = 0), accumulated with per-lane atomics by a
// later one (colour-1 sites: every bond once).
template <int L, int CPT_>
struct SlabGeo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int LH = L / 2;
  static constexpr int PL = LH * L;                     // columns (words per colour plane)
  static constexpr int NP = kSlabParts;                 // slab pairs per lattice
  static constexpr int S = LH / NP;                     // owned planes per half
  static constexpr int E = S + 2;                       // colour-0 planes per half, with halo
  static constexpr int CPT = CPT_;                      // columns per thread: a tuner tile
  static constexpr int NT = PL / CPT;                   // every thread has its columns
  static constexpr int C1W = 2 * (S + 4) * PL;          // staged colour-1 words
  static constexpr int PU = ceil_log2(12 * S * CPT + 1);
  static constexpr int PK = ceil_log2(4 * S * CPT + 1);
  static_assert(LH % NP == 0 && PL % CPT == 0 && NT % 64 == 0 && NT <= 1024, "slab shape");
  static_assert((2 * E * PL + C1W) * 8 <= 150 * 1024, "both colours' planes in LDS");
};



// ENG: the launch also accumulates the energies of the new states (mode
// carries kPhEnergy, and kPhEnergyK for K).  The items of both phases are
// straight-line code (no per-item branches), so the compiler can issue the
// loads of later items early.
template <int L, bool FIELD, bool ENG, int CPT>
__global__ __launch_bounds__((SlabGeo<L, CPT>::NT)) void k_slab(const JobPtrs *__restrict__ jt,
                                                         const TaskDesc *__restrict__ tasks,
                                                         int dir, int it, int mode) {
  using G = SlabGeo<L, CPT>;
  static_assert(G::PU * G::NT <= 2 * G::E * G::PL, "fold scratch within the colour-0 planes");
  __shared__ uint64_t c0[2][G::E][G::PL];
  __shared__ uint64_t c1s[G::C1W];
  __shared__ int32_t redi[4 * 64];
  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  if (tk.lat[1] >= (uint32_t)G::NP) return;   // padding task: the whole work-group
  const JobPtrs jp = ld_uniform(jt + tk.job);
  const LatDesc d = ld_uniform(jp.descs + tk.lat[0]);
  const SG_GLOBAL uint64_t *src = to_global(dir ? jp.spins2 : jp.spins) + d.spin_off;
  SG_GLOBAL uint64_t *dst = to_global(dir ? jp.spins : jp.spins2) + d.spin_off;
  const SG_GLOBAL char *__restrict__ Js = (const SG_GLOBAL char *)(to_global(jp.Js) + 2 * d.j_off);
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy) + (size_t)d.e_idx * 128;
  const int tid = threadIdx.x;
  const int za = (int)tk.lat[1] * G::S;
  if ((mode & kPhZero) && tk.lat[1] == 0 && tid < 128) energy[tid] = 0;

  // colour-1 planes za - 2 ... za + S + 1 of both halves, staged in LDS
  auto c1 = [&](int half, int e1, int col) -> uint64_t & {
    return c1s[(half * (G::S + 4) + e1) * G::PL + col];
  };
  {
    constexpr int W = G::C1W, NL = (W + G::NT - 1) / G::NT;
    uint64_t v[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int i = tid + k * G::NT;
      if (i < W) {
        const int col = i % G::PL, r = i / G::PL;
        const int half = r / (G::S + 4), e1 = r % (G::S + 4);
        const int z = (za - 2 + e1 + half * G::LH + 2 * L) % L;
        v[k] = src[G::NH + col + G::PL * z];
      }
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int i = tid + k * G::NT;
      if (i < W) c1s[i] = v[k];
    }
  }

  // per column: x offsets by site parity (XE[t] for parity (y + t) & 1,
  // t = (z + c) & 1), y offsets
  int cols[G::CPT], XE[G::CPT][2][2], dyp[G::CPT], dym[G::CPT];
#pragma unroll
  for (int i = 0; i < G::CPT; ++i) {
    const int col = tid + i * G::NT;
    cols[i] = col;
    const int xh = col % G::LH, y = col / G::LH;
    const int dxp = xh == G::LH - 1 ? -(G::LH - 1) : 1;
    const int dxm = xh == 0 ? G::LH - 1 : -1;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const bool odd = ((y + t) & 1) != 0;
      XE[i][t][0] = odd ? dxp : 0;
      XE[i][t][1] = odd ? 0 : dxm;
    }
    dyp[i] = y == L - 1 ? -(L - 1) * G::LH : G::LH;
    dym[i] = y == 0 ? (L - 1) * G::LH : -G::LH;
  }
  const ThrRec tr = ld_uniform(jp.thr + d.thr_idx);
  const uint32_t n = ld_uniform(jp.js).sweep + (uint32_t)it;
  const uint32_t sw = (kTagMC << 28) | (d.rt & 0x0FFFFFFFu);
  __syncthreads();

  // ---- phase 1: colour 0 on planes za - 1 ... za + S of both halves
  uint32_t r23[G::S][G::CPT][2];
#pragma unroll
  for (int e = 0; e < G::E; ++e) {
    const int pl = (za - 1 + e + L) % L;               // lower plane (uniform)
    const int ph = (pl + G::LH) % L;
    const int z0 = pl % G::LH;                         // the quad's plane
    const bool slo = pl >= G::LH;                      // site index of the lower plane
#pragma unroll
    for (int i = 0; i < G::CPT; ++i) {
      const int col = cols[i];
      const u32x4 r = philox_sweep((uint32_t)(col + G::PL * z0), n, d.g, sw, d.key0, d.key1);
      if (e >= 1 && e <= G::S) {
        r23[e - 1][i][0] = r.v[2];
        r23[e - 1][i][1] = r.v[3];
      }
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int z = half ? ph : pl;
        const int h = col + G::PL * z;
        const int t = z & 1;
        uint64_t nv[6];
        nv[0] = c1(half, e + 1, col + XE[i][t][0]);
        nv[1] = c1(half, e + 1, col + XE[i][t][1]);
        nv[2] = c1(half, e + 1, col + dyp[i]);
        nv[3] = c1(half, e + 1, col + dym[i]);
        nv[4] = c1(half, e + 2, col);
        nv[5] = c1(half, e, col);
        uint64_t jw[6];
        site_record_free(Js, h, jw);
        uint64_t sv = gload(src, h);
        update_site<L, FIELD>(sv, nv, jw, (half ? !slo : slo) ? r.v[1] : r.v[0], tr);
        c0[half][e][col] = sv;
      }
    }
  }
  __syncthreads();

  // ---- phase 2: colour 1 on the owned planes; both colours leave
  uint64_t AU[1][G::PU] = {}, AK[1][G::PK] = {};
#pragma unroll
  for (int e = 1; e <= G::S; ++e) {
    const int pl = za - 1 + e;                         // < L/2: site 0 of its quad
#pragma unroll
    for (int i = 0; i < G::CPT; ++i) {
      const int col = cols[i];
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int z = pl + half * G::LH;
        const int h = col + G::PL * z;
        const int t = (z + 1) & 1;
        uint64_t nv[6];
        nv[0] = c0[half][e][col + XE[i][t][0]];
        nv[1] = c0[half][e][col + XE[i][t][1]];
        nv[2] = c0[half][e][col + dyp[i]];
        nv[3] = c0[half][e][col + dym[i]];
        nv[4] = c0[half][e + 1][col];
        nv[5] = c0[half][e - 1][col];
        uint64_t jw[6];
        site_record_free(Js + G::NH * 48, h, jw);
        uint64_t sv = c1(half, e + 1, col);
        update_site<L, FIELD>(sv, nv, jw, r23[e - 1][i][half], tr);
        const uint64_t b = c0[half][e][col];
        dst[G::NH + h] = sv;
        dst[h] = b;
        if (ENG) {
          uint64_t u0, u1, u2;
          bond_count(sv, nv, jw, u0, u1, u2);
          vc_add3(AU[0], u0, u1, u2);
          vc_add1(AK[0], sv);
          vc_add1(AK[0], b);
        }
      }
    }
  }
  if (ENG) {
    uint64_t *fold = &c0[0][0][0];
    __syncthreads();
    int32_t U[1];
    wg_count_lanes<G::PU, G::NT, 1>(AU, fold, redi, U);
    if (tid < 64) atomicAdd((int32_t *)(energy + tid), U[0]);
    if (mode & kPhEnergyK) {
      __syncthreads();
      int32_t K[1];
      wg_count_lanes<G::PK, G::NT, 1>(AK, fold, redi, K);
      if (tid < 64) atomicAdd((int32_t *)(energy + 64 + tid), K[0]);
    }
  }
}

#ifdef SG_MAIN_PART  // non-template kernels: defined once, in sg_kernels.hip
// PT step of streamed lattices: a task is (pair, chunk of kPtWords words);
// every work-group decides the 64 lanes itself (same energies, same Philox
// words, so the same mask) and exchanges its chunk; chunk 0 counts.
__global__ __launch_bounds__(256) void k_pt_chunk(const JobPtrs *__restrict__ jt,
                                                  const TaskDesc *__restrict__ tasks) {
  const TaskDesc tk = ld_uniform(tasks + blockIdx.x);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  const PairDesc p = ld_uniform(jp.pairs + tk.lat[0]);
  const JobState js = ld_uniform(jp.js);
  const int tid = threadIdx.x, lane = tid & 63;
  // the chunk's words go out before the decision
  SG_GLOBAL uint64_t *A = to_global(jp.spins) + p.a_off, *B = to_global(jp.spins) + p.b_off;
  const uint32_t i0 = tk.lat[1] * kPtWords;
  uint64_t wa[kPtWords / 256], wb[kPtWords / 256];
#pragma unroll
  for (int k = 0; k < kPtWords / 256; ++k) {
    const uint32_t i = i0 + (uint32_t)(k * 256 + tid);
    wa[k] = i < p.n_words ? A[i] : 0;
    wb[k] = i < p.n_words ? B[i] : 0;
  }
  const uint32_t b = js.sweep / js.pt_every;
  const SG_GLOBAL int32_t *E = to_global(jp.energy);
  const SG_GLOBAL int32_t *Ea = E + (size_t)p.e_a * 128, *Eb = E + (size_t)p.e_b * 128;
  const u32x4 r = philox_dev((uint32_t)(lane >> 2), b, p.g, (kTagPT << 28) | (p.rt & 0x0FFFFFFFu),
                             p.key0, p.key1);
  const bool acc = pt_accept(p.beta_a, p.beta_b, Ea[lane], Eb[lane], Ea[64 + lane], Eb[64 + lane],
                             p.H, pick_word(r, lane));
  const uint64_t m = __ballot(acc);
  if (m) {
#pragma unroll
    for (int k = 0; k < kPtWords / 256; ++k) {
      const uint32_t i = i0 + (uint32_t)(k * 256 + tid);
      if (i < p.n_words) {
        const uint64_t dd = (wa[k] ^ wb[k]) & m;
        A[i] = wa[k] ^ dd;
        B[i] = wb[k] ^ dd;
      }
    }
  }
  if (tk.lat[1] == 0 && tid == 0) atomicAdd(&jp.counters[p.cnt_idx], (unsigned long long)__popcll(m));
}

#endif

// ------------------------------------------------------------- split kernel
// A lattice too large for one CU's LDS held on chip by P work-groups for a
// whole PT interval (instantiated for L = 32 and 48, P = 4).  Work-group w
// of a lattice owns the slab pair z in [wD, wD + D) and [wD + L/2, wD + D +
// L/2), D = L / 2P, so each Philox quad it owns has all four of its sites
// there.  Thread t owns QPT = D L^2 / 2 NT quads q = t + m NT (m = 0 ..
// QPT-1), numbered through the planes in the order 0, D - 1, 1, 2, ..., D -
// 2, so that quads m < NB = L^2 / NT lie on the two boundary planes of the
// halves.  Each quad's word offset, site parity and row-wrap flags are packed
// in one VGPR, from which a phase derives the quad's neighbour offsets.
//
// A phase updates one colour c.  Its words stay in VGPRs; LDS holds a copy
// of the other colour: the slab pair's own planes plus a halo plane below
// and above each half, which is all the phase reads (six neighbours per
// site).  Each thread updates its boundary quads first and writes them into
// the exchange buffer, and the work-group publishes the phase in its flag;
// the inner quads are updated while the planes travel.  Once every wave has
// finished reading, the new colour-c words replace the copy (whose colour
// 1 - c words go back to the VGPRs for the next phase), the work-group waits
// for its two neighbours' flags and copies their planes into the halo rows:
// with HALO_FIRST every wave polls and requests its halo words before the
// swap, so they are in flight during it; otherwise thread 0 polls during the
// swap and the copies follow a barrier.
// Exchange data and flags go through device-coherent accesses (relaxed
// agent-scope atomics, which pass the per-XCD L2s to the memory side); each
// thread waits for its plane stores to complete before the barrier that
// precedes the flag store, so the planes are visible wherever the flag is.
// No L2 write-back or invalidate is needed (the L2s keep the coupling
// records).  The flag holds the absolute phase 2 (sweep + it) + c + 1, so
// flags only grow while the sweep counter does (the host clears them
// whenever the counter is set back); the buffer is double-buffered by phase
// parity, and a work-group is never more than one phase ahead of a
// neighbour.
//
// Work-group b holds slab pair w = b mod P, so the work-groups of one XCD
// (b mod 8, the dispatcher's round robin) all hold the same slab pair and
// its site records fit in that XCD's L2 (1.33 MB per slab pair at L = 48,
// against 5.3 MB for a whole lattice).  Its lattice comes from the ticket
// counter of its slab pair, so the k-th work-group of each slab pair to
// start takes lattice k: the slab pairs of a lattice go to work-groups
// dispatched together and a lattice never waits on a work-group that cannot
// start.  The last work-group to finish resets the counters for the next
// launch.  A wait gives
// up after kSplitSpin polls (far beyond any normal wait) and sets xerr, which
// the host turns into an error.
//
// PT runs before the launch (k_pt_chunk); the launch sweeps nsweeps >= 1
// times and, with want_energy, counts U (colour-0 ends) and K per
// work-group; the slab pair w = 0 zeroes the lattice's energy slots before
// its first flag and every work-group adds its counts at the end, after its
// last exchange, which orders the adds after the zeroing.
// Reference counterpart: perform_MC_steps, Program/src/eamsim.cu:112-208,
// whose lattice stays in one block's shared memory (L <= 18 there).
constexpr uint32_t kSplitSpin = 1u << 19;

// A value the compiler must recompute where it is used: keeps per-quad
// addresses out of the registers held across the sweep loop.
__device__ __forceinline__ int pinned(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
// word idx of a uniform base, addressed as base + 32-bit byte offset
__device__ __forceinline__ uint64_t ld_word(const SG_GLOBAL uint64_t *base, int idx) {
  return *(const SG_GLOBAL uint64_t *)((const SG_GLOBAL char *)base + (uint32_t)pinned(idx * 8));
}
__device__ __forceinline__ void st_word(SG_GLOBAL uint64_t *base, int idx, uint64_t v) {
  *(SG_GLOBAL uint64_t *)((SG_GLOBAL char *)base + (uint32_t)pinned(idx * 8)) = v;
}
// device-coherent word access (relaxed agent-scope atomics: sc1, past the
// per-XCD caches) for the exchange buffer
__device__ __forceinline__ void st_coherent(uint64_t *base, int idx, uint64_t v) {
  __hip_atomic_store((uint64_t *)((char *)base + (uint32_t)pinned(idx * 8)), v, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_coherent(const uint64_t *base, int idx) {
  return __hip_atomic_load((const uint64_t *)((const char *)base + (uint32_t)pinned(idx * 8)),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int L, int P, int NT>
struct SplitGeo {
  static constexpr int N = L * L * L;
  static constexpr int NH = N / 2;
  static constexpr int LH = L / 2;
  static constexpr int PL = LH * L;             // columns: words per plane of one colour
  static constexpr int D = L / (2 * P);         // planes per half slab
  static constexpr int QW = D * PL;             // quads per slab pair
  static constexpr int QPT = QW / NT;           // quads per thread
  static constexpr int NB = 2 * PL / NT;        // of them on the boundary planes: m < NB
  static constexpr int HPT = 4 * PL / NT;       // halo words per thread and phase
  static_assert(PL % 64 == 0, "a wave's halo words in one row");
  static constexpr int ROWS = D + 2;            // halo, D own planes, halo
  static constexpr int CW = 2 * ROWS * PL;      // the colour copy [half][row][column]
  static constexpr int PU = ceil_log2(12 * QPT + 1);
  static constexpr int PK = ceil_log2(4 * QPT + 1);
  static_assert(P >= 2 && L % (2 * P) == 0 && D >= 2, "slab pairs with two boundary planes");
  static_assert(P < kSplitTicketWords, "a ticket per slab pair and a done count");
  static_assert(NT % 64 == 0 && NT <= 1024 && QW % NT == 0 && (2 * PL) % NT == 0,
                "every thread owns the same number of boundary and inner quads");
  static_assert(CW * 8 + 4 * 64 * 4 + 4 <= 160 * 1024, "the colour copy in LDS");
  static_assert(PU * NT <= CW, "fold scratch in the copy");
};

template <int L, int P, int NT, bool HALO_FIRST, bool FIELD>
__global__ __launch_bounds__(NT) void k_split(const JobPtrs *__restrict__ jt,
                                             const TaskDesc *__restrict__ tasks,
                                             uint32_t *__restrict__ ticket, int nsweeps,
                                             int want_energy) {
  using G = SplitGeo<L, P, NT>;
  __shared__ uint64_t cb[G::CW];            // [half][row][column], one colour
  __shared__ int32_t redi[4 * 64];
  __shared__ uint32_t s_ticket;
  const int tid = threadIdx.x;
  const int w = (int)(blockIdx.x % P);
  if (tid == 0) s_ticket = atomicAdd(ticket + w, 1u);
  __syncthreads();
  const int ti = (int)(__builtin_amdgcn_readfirstlane(s_ticket) % (gridDim.x / P));
  const TaskDesc tk = ld_uniform(tasks + ti);
  const JobPtrs jp = ld_uniform(jt + tk.job);
  const LatDesc d = ld_uniform(jp.descs + tk.lat[0]);
  SG_GLOBAL uint64_t *__restrict__ g = to_global(jp.spins) + d.spin_off;
  SG_GLOBAL int32_t *__restrict__ energy = to_global(jp.energy) + (size_t)d.e_idx * 128;
  const SG_GLOBAL char *__restrict__ Js = (const SG_GLOBAL char *)(to_global(jp.Js) + 2 * d.j_off);
  uint64_t *xb = jp.xbuf + (size_t)tk.lat[0] * P * 8 * G::PL;
  uint32_t *xf = jp.xflag + (size_t)tk.lat[0] * P;
  const int wb = (w + P - 1) % P, wa = (w + 1) % P;   // neighbours below, above
  const int zw = w * G::D;                             // first plane of the lower half

  // quad m: q = tid + m NT; plane index pi = q / PL in the order 0, D-1, 1,
  // ...; column col = q mod PL = xh + (L/2) y.  Descriptor: bits 0-16 the
  // byte offset 8 (zl PL + col) of its word among a half's own rows; bit 17
  // (y + zl) & 1; bits 18, 19: xh = 0, xh = L/2 - 1; bits 20, 21: y = 0,
  // y = L - 1
  static_assert(8 * G::QW < (1 << 17), "descriptor byte offsets");
  const int wave_i = __builtin_amdgcn_readfirstlane(tid) & ~63;   // the wave's first thread
  uint32_t qd[G::QPT];
#pragma unroll
  for (int m = 0; m < G::QPT; ++m) {
    const int q = tid + m * NT, pi = q / G::PL, col = q % G::PL;
    const int zl = pi == 0 ? 0 : pi == 1 ? G::D - 1 : pi - 1;
    const int xh = col % G::LH, y = col / G::LH;
    qd[m] = (uint32_t)(8 * (zl * G::PL + col)) | ((uint32_t)((y + zl) & 1) << 17) |
            ((uint32_t)(xh == 0) << 18) | ((uint32_t)(xh == G::LH - 1) << 19) |
            ((uint32_t)(y == 0) << 20) | ((uint32_t)(y == L - 1) << 21);
  }
  // read through pinned(): each use recomputes its addresses, so none is
  // held across the sweep loop
  auto off_of = [&](int m) { return pinned((int)qd[m]) & 0x1FFF8; };
  // the copy's word at a byte offset; own site (m, hs) sits in row zl + 1 of
  // half hs
  auto cbw = [&](int off) -> uint64_t & { return *(uint64_t *)((char *)cb + off); };
  auto cb_off = [&](int m, int hs) { return off_of(m) + (hs * G::ROWS + 1) * G::PL * 8; };
  auto gidx = [&](int m, int hs, int c) {
    return c * G::NH + (off_of(m) >> 3) + G::PL * (zw + hs * G::LH);
  };

  // the registers hold the colour the phase updates (colour 0 first), the
  // copy the other one: own planes and halos, from the lattice
  uint64_t sp[G::QPT][2];                              // [quad][half]
#pragma unroll
  for (int m = 0; m < G::QPT; ++m)
#pragma unroll
    for (int hs = 0; hs < 2; ++hs) {
      sp[m][hs] = ld_word(g, gidx(m, hs, 0));
      cbw(cb_off(m, hs)) = ld_word(g, gidx(m, hs, 1));
    }
  for (int i = tid; i < 4 * G::PL; i += NT) {
    const int hs = i / (2 * G::PL), top = (i / G::PL) & 1, cl = i % G::PL;
    const int z = (zw + hs * G::LH + (top ? G::D : -1) + L) % L;
    cb[(hs * G::ROWS + (top ? G::D + 1 : 0)) * G::PL + cl] = ld_word(g, G::NH + cl + G::PL * z);
  }
  ThrRec tr = ld_uniform(jp.thr + d.thr_idx);
  const uint32_t n0 = ld_uniform(jp.js).sweep;
  if (want_energy && w == 0 && tid < (want_energy > 1 ? 128 : 64))
    atomicExch((int32_t *)(energy + tid), 0);
  __syncthreads();

  // byte offsets in the copy (of colour 1 - c) of the neighbours of own site
  // (m, 0) at colour c: +x, -x, +y, -y; +z and -z are the words PL rows
  // above and below a.  Half 1 adds ROWS PL words to each.  x = 2 xh + p,
  // p = (y + z + c) & 1 with z = zw + zl + hs L/2 (L/2 even: p is the same in
  // both halves); one x-neighbour shares the word index, the other is one
  // word away or across the row's wrap.
  struct NbOff {
    int a, px, mx, py, my;
  };
  auto nb_off = [&](int m, int c) {
    const uint32_t v = (uint32_t)pinned((int)qd[m]);
    const int a = (int)(v & 0x1FFF8u) + G::PL * 8;
    const bool p = (((v >> 17) ^ (uint32_t)(zw + c)) & 1u) != 0;
    const int t = (int)((v >> (p ? 19 : 18)) & 1u);   // the far x-neighbour wraps
    const int u = t * (G::LH * 8) - 8;                 // p = 0: it is at a + u, p = 1: a - u
    NbOff r;
    r.a = a;
    r.px = p ? a - u : a;
    r.mx = p ? a : a + u;
    r.py = a + G::LH * 8 - (int)((v >> 21) & 1u) * (L * G::LH * 8);
    r.my = a - G::LH * 8 + (int)((v >> 20) & 1u) * (L * G::LH * 8);
    return r;
  };
  auto neighbours = [&](const NbOff &r, int hs, uint64_t (&nv)[6]) {
    const int h = hs * G::ROWS * G::PL * 8;
    nv[0] = cbw(r.px + h);
    nv[1] = cbw(r.mx + h);
    nv[2] = cbw(r.py + h);
    nv[3] = cbw(r.my + h);
    nv[4] = cbw(r.a + h + G::PL * 8);
    nv[5] = cbw(r.a + h - G::PL * 8);
  };
  // own site (m, hs)'s coupling record at colour c: record index
  // zl PL + col of the slab's half, 48 bytes each
  auto record = [&](int m, int hs, int c, uint64_t (&jw)[6]) {
    site_record(Js + (size_t)(c * G::NH + hs * G::LH * G::PL + G::PL * zw) * 48,
                off_of(m) >> 3, true, jw);
  };

  const uint32_t sw = (kTagMC << 28) | (d.rt & 0x0FFFFFFFu);
  uint32_t r23[G::QPT][2];
  for (int it = 0; it < nsweeps; ++it) {
    const uint32_t n = n0 + (uint32_t)it;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const uint32_t phase = 2u * n + (uint32_t)c;
      uint64_t *xme = xb + ((size_t)w * 2 + (phase & 1)) * 4 * G::PL;   // [hs][side][PL]
      auto quad = [&](int m) {
        uint32_t rw[2];
        if (c == 0) {
          const uint32_t j = (uint32_t)((off_of(m) >> 3) + G::PL * zw);
          const u32x4 r = philox_sweep(j, n, d.g, sw, d.key0, d.key1);
          rw[0] = r.v[0];
          rw[1] = r.v[1];
          r23[m][0] = r.v[2];
          r23[m][1] = r.v[3];
        } else {
          rw[0] = r23[m][0];
          rw[1] = r23[m][1];
        }
        uint64_t jc[6], jn[6];
        record(m, 0, c, jc);
        const NbOff nb = nb_off(m, c);
#pragma unroll
        for (int hs = 0; hs < 2; ++hs) {
          if (hs == 0) record(m, 1, c, jn);
          uint64_t nv[6];
          neighbours(nb, hs, nv);
          update_site<L, FIELD>(sp[m][hs], nv, jc, rw[hs], tr);
#pragma unroll
          for (int e = 0; e < 6; ++e) jc[e] = jn[e];
          if (m < G::NB)
            st_coherent(xme, hs * 2 * G::PL + pinned(tid) + m * NT, sp[m][hs]);
        }
      };
      // boundary quads, their planes out, the phase published
#pragma unroll
      for (int m = 0; m < G::NB; ++m) {
        quad(m);
      }
      __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this thread's planes are out
      __syncthreads();
      if (tid == 0)
        __hip_atomic_store(xf + w, phase + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // inner quads while the planes travel
#pragma unroll
      for (int m = G::NB; m < G::QPT; ++m) quad(m);
      __syncthreads();   // every read of the colour 1 - c copy is done
      // the neighbours' planes of this phase: halo rows of colour c, below <-
      // the top planes of slab pair wb, above <- the bottom planes of wa;
      // across the wrap the halves trade places.  poll: until both flags
      // show this phase
      const uint64_t *xbl = xb + ((size_t)wb * 2 + (phase & 1)) * 4 * G::PL;
      const uint64_t *xab = xb + ((size_t)wa * 2 + (phase & 1)) * 4 * G::PL;
      auto poll = [&](auto uniform) {
        uint32_t polls = 0;
        for (;;) {
          const uint32_t fb = uniform(__hip_atomic_load(xf + wb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          const uint32_t fa = uniform(__hip_atomic_load(xf + wa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          if ((int32_t)(fb - (phase + 1u)) >= 0 && (int32_t)(fa - (phase + 1u)) >= 0) break;
          if (++polls > kSplitSpin) {
            if ((tid & 63) == 0) atomicOr(jp.xerr, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      };
      // halo word k of this thread (fetched before the swap): i = tid + k NT;
      // the 64 words of a wave lie in one row (PL is a multiple of 64), so
      // the row arithmetic is the wave's (scalar) and a lane adds its index
      auto halo_word = [&](int k) {
        const int i0 = wave_i + k * NT, hs = i0 / (2 * G::PL), top = (i0 / G::PL) & 1;
        const uint64_t *src = top ? xab + ((hs ^ (w == P - 1 ? 1 : 0)) * 2 + 0) * G::PL
                                  : xbl + ((hs ^ (w == 0 ? 1 : 0)) * 2 + 1) * G::PL;
        return ld_coherent(src + i0 % G::PL, tid & 63);
      };
      auto halo_slot = [&](int k) -> uint64_t & {
        const int i0 = wave_i + k * NT, hs = i0 / (2 * G::PL), top = (i0 / G::PL) & 1;
        return cb[(hs * G::ROWS + (top ? G::D + 1 : 0)) * G::PL + i0 % G::PL + (tid & 63)];
      };
      // the new colour-c words into the copy, the colour 1 - c words (the
      // next phase's) out of it
      auto swap = [&] {
#pragma unroll
        for (int m = 0; m < G::QPT; ++m)
#pragma unroll
          for (int hs = 0; hs < 2; ++hs) {
            uint64_t &slot = cbw(cb_off(m, hs));
            const uint64_t o = slot;
            slot = sp[m][hs];
            sp[m][hs] = o;
          }
      };
      if constexpr (HALO_FIRST) {
        // every wave polls, then its halo words are in flight during the swap
        poll([](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); });
        uint64_t hv[G::HPT];
#pragma unroll
        for (int k = 0; k < G::HPT; ++k) hv[k] = halo_word(k);
        swap();
#pragma unroll
        for (int k = 0; k < G::HPT; ++k) halo_slot(k) = hv[k];
      } else {
        // thread 0 polls while the others swap
        swap();
        if (tid == 0) poll([](uint32_t v) { return v; });
        __syncthreads();
        for (int i = tid; i < 4 * G::PL; i += NT) {
          const int hs = i / (2 * G::PL), top = (i / G::PL) & 1, cl = i % G::PL;
          cb[(hs * G::ROWS + (top ? G::D + 1 : 0)) * G::PL + cl] =
              top ? ld_coherent(xab, ((hs ^ (w == P - 1 ? 1 : 0)) * 2 + 0) * G::PL + cl)
                  : ld_coherent(xbl, ((hs ^ (w == 0 ? 1 : 0)) * 2 + 1) * G::PL + cl);
        }
      }
      __syncthreads();
    }
  }

  // after an even number of phases the registers hold colour 0 and the copy
  // colour 1 with its halos: U over the colour-0 ends, K over both colours
  uint64_t AU[1][G::PU] = {}, AK[1][G::PK] = {};
  if (want_energy) {
#pragma unroll
    for (int m = 0; m < G::QPT; ++m)
#pragma unroll
      for (int hs = 0; hs < 2; ++hs) {
        uint64_t jc[6], nv[6];
        record(m, hs, 0, jc);
        neighbours(nb_off(m, 0), hs, nv);
        uint64_t u0, u1, u2;
        bond_count(sp[m][hs], nv, jc, u0, u1, u2);
        vc_add3(AU[0], u0, u1, u2);
        __builtin_amdgcn_sched_barrier(0);   // one site's record live at a time
      }
  }
  if (want_energy > 1) {
#pragma unroll
    for (int m = 0; m < G::QPT; ++m)
#pragma unroll
      for (int hs = 0; hs < 2; ++hs) {
        vc_add1(AK[0], sp[m][hs]);
        vc_add1(AK[0], cbw(cb_off(m, hs)));
      }
  }
#pragma unroll
  for (int m = 0; m < G::QPT; ++m)
#pragma unroll
    for (int hs = 0; hs < 2; ++hs) {
      st_word(g, gidx(m, hs, 0), sp[m][hs]);
      st_word(g, gidx(m, hs, 1), cbw(cb_off(m, hs)));
    }
  if (want_energy) {
    __syncthreads();   // the copy serves as the fold scratch
    int32_t U[1];
    wg_count_lanes<G::PU, NT, 1>(AU, cb, redi, U);
    if (tid < 64) atomicAdd((int32_t *)(energy + tid), U[0]);
    if (want_energy > 1) {
      __syncthreads();
      int32_t Kc[1];
      wg_count_lanes<G::PK, NT, 1>(AK, cb, redi, Kc);
      if (tid < 64) atomicAdd((int32_t *)(energy + 64 + tid), Kc[0]);
    }
  }
  // the last work-group out resets the ticket counters for the next launch
  if (tid == 0 && atomicAdd(ticket + P, 1u) == gridDim.x - 1) {
    for (int i = 0; i <= P; ++i) atomicExch(ticket + i, 0u);
  }
}

// ---------------------------------------------------------------- dispatch
#define SG_LDS_SIZES(X) X(2) X(4) X(6) X(8) X(10) X(12) X(14) X(16)
#define SG_ALL_SIZES(X) SG_LDS_SIZES(X) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) \
  X(34) X(36) X(38) X(40) X(42) X(44) X(46) X(48)

// The resident sweep launchers (and so the only instances of k_block, k_mid,
// k_big and k_one) are compiled in sg_kernels_resident.hip, which defines
// SG_RESIDENT_PART and includes this file; everything else is compiled here.
#ifdef SG_RESIDENT_PART
bool sweep_kernel_supported(int L) {
  switch (L) {
#define X(l) case l: return true;
    SG_ALL_SIZES(X)
#undef X
    default: return false;
  }
}

hipError_t launch_block(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks,
                        int ntasks, int do_pt, int nsweeps, int want_energy, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  switch (L) {
#define X(l)                                                                             \
  case l:                                                                                \
    if (field)                                                                           \
      hipLaunchKernelGGL((k_block<l, true>), dim3(ntasks), dim3(Geo<l>::NT), 0, st,      \
                         jt, tasks, do_pt, nsweeps, want_energy);                        \
    else                                                                                 \
      hipLaunchKernelGGL((k_block<l, false>), dim3(ntasks), dim3(Geo<l>::NT), 0, st,     \
                         jt, tasks, do_pt, nsweeps, want_energy);                        \
    return hipGetLastError();
    SG_LDS_SIZES(X)
#undef X
#define X(l)                                                                             \
  case l:                                                                                \
    if (field)                                                                           \
      hipLaunchKernelGGL((k_mid<l, true>), dim3(ntasks), dim3(MidGeo<l>::NT), 0, st,     \
                         jt, tasks, do_pt, nsweeps, want_energy);                        \
    else                                                                                 \
      hipLaunchKernelGGL((k_mid<l, false>), dim3(ntasks), dim3(MidGeo<l>::NT), 0, st,    \
                         jt, tasks, do_pt, nsweeps, want_energy);                        \
    return hipGetLastError();
    X(18) X(20)
#undef X
#define X(l)                                                                             \
  case l:                                                                                \
    if (field)                                                                           \
      hipLaunchKernelGGL((k_big<l, true>), dim3(ntasks), dim3(1024), 0, st,              \
                         jt, tasks, do_pt, nsweeps, want_energy);                        \
    else                                                                                 \
      hipLaunchKernelGGL((k_big<l, false>), dim3(ntasks), dim3(1024), 0, st,             \
                         jt, tasks, do_pt, nsweeps, want_energy);                        \
    return hipGetLastError();
    X(22) X(24) X(26) X(28) X(30) X(32) X(34) X(36) X(38) X(40) X(42) X(44) X(46) X(48)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

bool single_kernel_supported(int L) { return L == 22 || L == 24 || L == 26; }

hipError_t launch_single(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int ntasks,
                         int nsweeps, int want_energy, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  switch (L) {
#define X(l)                                                                             \
  case l:                                                                                \
    if (field)                                                                           \
      hipLaunchKernelGGL((k_one<l, true>), dim3(ntasks), dim3(OneGeo<l>::NT), 0, st,     \
                         jt, tasks, nsweeps, want_energy);                               \
    else                                                                                 \
      hipLaunchKernelGGL((k_one<l, false>), dim3(ntasks), dim3(OneGeo<l>::NT), 0, st,    \
                         jt, tasks, nsweeps, want_energy);                               \
    return hipGetLastError();
    X(22) X(24) X(26)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

#elif defined(SG_MAIN_PART)

hipError_t launch_measure(int L, int R, const uint64_t *spins, const int32_t *energy,
                          const MeasDesc *md, int nmd, int32_t *out, int nfield, ObsAcc *acc,
                          const double *cs, int write_records, hipStream_t st) {
  if (nmd <= 0) return hipSuccess;
  if (R < 1 || R > 8) return hipErrorInvalidValue;
  switch (L) {
#define X(l)                                                                            \
  case l:                                                                               \
    hipLaunchKernelGGL(k_measure<l>, dim3(nmd), dim3(256), 0, st, spins, energy, md, R, out, \
                       nfield, acc, cs, write_records);                                 \
    return hipGetLastError();
    SG_ALL_SIZES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_init_spins(int L, uint64_t *spins, const LatDesc *descs, int ndesc,
                             hipStream_t st) {
  if (ndesc <= 0) return hipSuccess;
  const int N = L * L * L;
  switch (L) {
#define X(l)                                                                           \
  case l:                                                                              \
    hipLaunchKernelGGL(k_init_spins<l>, dim3((N + 255) / 256, ndesc), dim3(256), 0, st, \
                       spins, descs);                                                  \
    return hipGetLastError();
    SG_ALL_SIZES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_init_couplings(int L, uint64_t *J, uint64_t j_off, uint32_t g,
                                 uint32_t key0, uint32_t key1, hipStream_t st) {
  const int N = L * L * L;
  hipLaunchKernelGGL(k_init_couplings, dim3((N + 255) / 256, 3), dim3(256), 0, st, J, j_off,
                     N, g, key0, key1);
  return hipGetLastError();
}

#define SG_PHASE_SIZES(X) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) \
  X(34) X(36) X(38) X(40) X(42) X(44) X(46) X(48)

bool phase_kernel_supported(int L) {
  switch (L) {
#define X(l) case l:
    SG_PHASE_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

hipError_t launch_phase(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int ntasks,
                        int c, int it, int mode, int tile, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (tile < 0 || tile >= kPhaseTiles) return hipErrorInvalidValue;
  // launches that accumulate energies take the instance with the counting
  // code; the update-only instance holds fewer registers
  const bool eng = (mode & (kPhEnergy | kPhEnergyK)) != 0;
#define SG_PHASE_LAUNCH(l, f, e, q)                                                        \
  hipLaunchKernelGGL((k_phase<l, f, e, q>), dim3(ntasks), dim3(kPhaseNT / q), 0, st, jt,   \
                     tasks, c, it, mode)
#define SG_PHASE_TILE(l, f, e)                                                             \
  if (tile == 0) SG_PHASE_LAUNCH(l, f, e, 1); else SG_PHASE_LAUNCH(l, f, e, 2)
  switch (L) {
#define X(l)                                                                               \
  case l:                                                                                  \
    if (eng) {                                                                             \
      if (field) SG_PHASE_TILE(l, true, true); else SG_PHASE_TILE(l, false, true);         \
    } else {                                                                               \
      if (field) SG_PHASE_TILE(l, true, false); else SG_PHASE_TILE(l, false, false);       \
    }                                                                                      \
    return hipGetLastError();
    SG_PHASE_SIZES(X)
#undef X
#undef SG_PHASE_TILE
#undef SG_PHASE_LAUNCH
    default:
      return hipErrorInvalidValue;
  }
}

#define SG_SLAB_SIZES(X) X(32)

bool slab_kernel_supported(int L) {
  switch (L) {
#define X(l) case l: return true;
    SG_SLAB_SIZES(X)
#undef X
    default: return false;
  }
}

hipError_t launch_slab(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int ntasks,
                       int dir, int it, int mode, int tile, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (!(mode & kPhUpdate)) return hipErrorInvalidValue;   // sweeps only
  if (tile < 0 || tile >= kPhaseTiles) return hipErrorInvalidValue;
  const bool eng = (mode & (kPhEnergy | kPhEnergyK)) != 0;
  switch (L) {
#define SG_SLAB_ONE(l, f, e, cpt)                                                          \
  hipLaunchKernelGGL((k_slab<l, f, e, cpt>), dim3(ntasks), dim3((SlabGeo<l, cpt>::NT)), 0, \
                     st, jt, tasks, dir, it, mode)
#define SG_SLAB_LAUNCH(l, f, e)                                                            \
  if (tile == 0) SG_SLAB_ONE(l, f, e, 1); else SG_SLAB_ONE(l, f, e, 2)
#define X(l)                                                                               \
  case l:                                                                                  \
    if (field) {                                                                           \
      if (eng) SG_SLAB_LAUNCH(l, true, true); else SG_SLAB_LAUNCH(l, true, false);         \
    } else {                                                                               \
      if (eng) SG_SLAB_LAUNCH(l, false, true); else SG_SLAB_LAUNCH(l, false, false);       \
    }                                                                                      \
    return hipGetLastError();
    SG_SLAB_SIZES(X)
#undef X
#undef SG_SLAB_LAUNCH
#undef SG_SLAB_ONE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_pt_chunks(const JobPtrs *jt, const TaskDesc *tasks, int ntasks, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pt_chunk, dim3(ntasks), dim3(256), 0, st, jt, tasks);
  return hipGetLastError();
}

hipError_t launch_build_jsite(int L, uint64_t *Js, const uint64_t *J, uint64_t j_off,
                              hipStream_t st) {
  const int N = L * L * L;
  switch (L) {
#define X(l)                                                                               \
  case l:                                                                                  \
    hipLaunchKernelGGL(k_build_jsite<l>, dim3((N + 255) / 256), dim3(256), 0, st, Js, J,   \
                       j_off);                                                             \
    return hipGetLastError();
    SG_ALL_SIZES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_advance(const JobPtrs *jt, int njobs, int nsweeps, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_advance, dim3((njobs + 255) / 256), dim3(256), 0, st, jt, njobs,
                     nsweeps);
  return hipGetLastError();
}

#endif  // SG_RESIDENT_PART / SG_MAIN_PART

#ifdef SG_SPLIT_PART
// The split kernels, in a translation unit of their own (sg_kernels_split.hip).
int split_slabs(int L, int n_lattices) {
  // four slab pairs per lattice (4 + 4 planes each at L = 32, 6 + 6 at
  // L = 48), one work-group per CU: 64 lattices per launch
  return (L == 32 || L == 48) && n_lattices >= 1 ? 4 : 0;
}

int split_tiles(int L) { return L == 32 ? 2 : L == 48 ? 1 : 0; }

hipError_t launch_split(int L, bool field, const JobPtrs *jt, const TaskDesc *tasks, int nlat,
                        int P, uint32_t *ticket, int nsweeps, int want_energy, int tile,
                        hipStream_t st) {
  if (nlat <= 0) return hipSuccess;
  if (nsweeps < 1) return hipErrorInvalidValue;
  const dim3 grid(nlat * P);
  // tiles: threads per work-group, 1024 then 512 at L = 32; 768 at L = 48.
  // The halo words are fetched before the swap (every wave polls) at L = 48,
  // after it (one wave polls) at L = 32: 409-413 vs 432-437 us per 16-sweep
  // launch at L = 48, 176-180 vs 172-174 us at L = 32 (64 lattices)
#define SG_SPLIT(l, p, t, nt, hf)                                                          \
  if (L == l && P == p && tile == t) {                                                    \
    if (field)                                                                             \
      hipLaunchKernelGGL((k_split<l, p, nt, hf, true>), grid, dim3(nt), 0, st, jt, tasks,  \
                         ticket, nsweeps, want_energy);                                    \
    else                                                                                   \
      hipLaunchKernelGGL((k_split<l, p, nt, hf, false>), grid, dim3(nt), 0, st, jt, tasks, \
                         ticket, nsweeps, want_energy);                                    \
    return hipGetLastError();                                                              \
  }
  SG_SPLIT(32, 4, 0, 1024, false)
  SG_SPLIT(32, 4, 1, 512, false)
  SG_SPLIT(48, 4, 0, 768, true)
#undef SG_SPLIT
  return hipErrorInvalidValue;
}
#endif  // SG_SPLIT_PART

}  // namespace sg
