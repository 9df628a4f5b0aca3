/*
 * sg_oracle.c — CPU oracle (test infrastructure only; see sg_oracle.h for the
 * canonical algorithm).  Built with -ffp-contract=off so that the FP64
 * threshold and PT-acceptance arithmetic is a fixed sequence of IEEE
 * operations.
 *
 * Two independent update paths are kept:
 *   scalar_path = 1: per-lane loops, reading one spin and one bond bit at a
 *                    time (the definition, slow);
 *   scalar_path = 0: bit-sliced full-adder form on whole 64-lane words (fast;
 *                    also the CPU baseline).
 * The tests require the two to agree bit for bit.
 */
#include "sg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- Philox */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void philox_tagged(uint32_t c0, uint32_t c1, uint32_t g,
                                 uint32_t tag, uint32_t rho, uint32_t t,
                                 uint64_t seed, uint32_t out[4]) {
  uint32_t ctr[4] = {c0, c1, g, (tag << 28) | ((rho & 0xFFFu) << 16) | (t & 0xFFFFu)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox(ctr, key, out);
}

/* ------------------------------------------------------- deterministic exp */
/* exp(x) for x <= 0 from basic IEEE operations only: Cody–Waite reduction
 * x = k ln2 + r with the fdlibm split of ln2 (hi part has trailing zero bits
 * so k*hi is exact), |r| <= ln2/2, then the Taylor polynomial of e^r to
 * degree 14 (truncation < 1e-18 relative) by Horner, then an exact scaling
 * by 2^k built from the exponent field.  Returns 0 for x < -40, where every
 * use (floor(2^32 e^x)) is 0 anyway; so no subnormal is ever formed. */
double orc_exp(double x) {
  if (!(x <= 0.0)) return 1.0;             /* only x <= 0 is used */
  if (x < -40.0) return 0.0;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double kd = floor(x * inv_ln2 + 0.5);
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  /* 1/i! for i = 14 .. 0 */
  static const double inv_fact[15] = {
      1.0 / 87178291200.0, 1.0 / 6227020800.0, 1.0 / 479001600.0,
      1.0 / 39916800.0,    1.0 / 3628800.0,    1.0 / 362880.0,
      1.0 / 40320.0,       1.0 / 5040.0,       1.0 / 720.0,
      1.0 / 120.0,         1.0 / 24.0,         1.0 / 6.0,
      0.5,                 1.0,                1.0};
  double p = inv_fact[0];
  for (int i = 1; i < 15; ++i) p = p * r + inv_fact[i];
  int k = (int)kd; /* in [-58, 0] */
  union { uint64_t u; double d; } two_k;
  two_k.u = (uint64_t)(k + 1023) << 52;
  return p * two_k.d;
}

/* T[sigma*7 + u]: acceptance threshold for a spin with bit sigma and u
 * unsatisfied bonds (reference eamsim.cu:24-40: prob = exp(2*beta*e_i) with
 * e_i = 6 - 2S - H*sigma; here dE = -2 e_i = 12 - 4u + 2 H sigma). */
void orc_thresholds(double beta, double H, uint64_t T[14]) {
  for (int sig = 0; sig < 2; ++sig) {
    double sv = sig ? 1.0 : -1.0;
    for (int u = 0; u < 7; ++u) {
      double dE = (double)(12 - 4 * u) + 2.0 * H * sv;
      uint64_t thr;
      if (dE <= 0.0) {
        thr = (uint64_t)1 << 32;
      } else {
        double p = orc_exp(-beta * dE);
        thr = (uint64_t)(p * 4294967296.0);
      }
      T[sig * 7 + u] = thr;
    }
  }
}

/* ------------------------------------------------------------- geometry */
static inline int sidx(int L, int x, int y, int z) { return x + L * (y + L * z); }

/* Metropolis random numbers of one lattice for sweep n, indexed by site. */
static void sweep_randoms(int L, uint64_t seed, uint32_t g, uint32_t rho,
                          uint32_t t, uint32_t n, uint32_t *rnd) {
  const int N = L * L * L, Q = N / 4, Lh = L / 2;
  for (int j = 0; j < Q; ++j) {
    uint32_t out[4];
    philox_tagged((uint32_t)j, n, g, ORC_TAG_MC, rho, t, seed, out);
    for (int k = 0; k < 4; ++k) {
      int c = k >> 1, hi = k & 1;
      int h = j + hi * Q;
      int xh = h % Lh, y = (h / Lh) % L, z = h / (Lh * L);
      int x = 2 * xh + ((y + z + c) & 1);
      rnd[sidx(L, x, y, z)] = out[k];
    }
  }
}

static inline uint64_t maj3(uint64_t a, uint64_t b, uint64_t c) {
  return (a & b) | (a & c) | (b & c);
}

/* lanes where the bit-sliced 3-bit count (u2 u1 u0) is >= k, 0 <= k <= 7 */
static inline uint64_t count_ge(uint64_t u2, uint64_t u1, uint64_t u0, int k) {
  uint64_t gt = 0, eq = ~(uint64_t)0;
  uint64_t ub[3] = {u0, u1, u2};
  for (int b = 2; b >= 0; --b) {
    if ((k >> b) & 1) {
      eq &= ub[b];
    } else {
      gt |= eq & ub[b];
      eq &= ~ub[b];
    }
  }
  return gt | eq;
}

/* one colour of a sweep, in place */
static void update_colour(int L, uint64_t *S, const uint64_t *J,
                          const uint64_t T[14], const uint32_t *rnd, int c,
                          int scalar_path) {
  const int N = L * L * L;
  const uint64_t *Jx = J, *Jy = J + N, *Jz = J + 2 * N;
  for (int z = 0; z < L; ++z) {
    int zp = (z + 1) % L, zm = (z + L - 1) % L;
    for (int y = 0; y < L; ++y) {
      int yp = (y + 1) % L, ym = (y + L - 1) % L;
      for (int x = (y + z + c) & 1; x < L; x += 2) {
        int xp = (x + 1) % L, xm = (x + L - 1) % L;
        int i = sidx(L, x, y, z);
        int nb[6] = {sidx(L, xp, y, z), sidx(L, xm, y, z), sidx(L, x, yp, z),
                     sidx(L, x, ym, z), sidx(L, x, y, zp), sidx(L, x, y, zm)};
        uint64_t jb[6] = {Jx[i], Jx[nb[1]], Jy[i], Jy[nb[3]], Jz[i], Jz[nb[5]]};
        uint64_t s = S[i];
        uint64_t r = rnd[i];
        if (scalar_path) {
          uint64_t flip = 0;
          for (int l = 0; l < 64; ++l) {
            int sig = (int)((s >> l) & 1);
            int u = 0;
            for (int k = 0; k < 6; ++k) u += (int)(((s ^ S[nb[k]] ^ jb[k]) >> l) & 1);
            if (r < T[sig * 7 + u]) flip |= (uint64_t)1 << l;
          }
          S[i] = s ^ flip;
        } else {
          uint64_t b[6];
          for (int k = 0; k < 6; ++k) b[k] = s ^ S[nb[k]] ^ jb[k];
          uint64_t s1 = b[0] ^ b[1] ^ b[2], c1 = maj3(b[0], b[1], b[2]);
          uint64_t s2 = b[3] ^ b[4] ^ b[5], c2 = maj3(b[3], b[4], b[5]);
          uint64_t u0 = s1 ^ s2, c3 = s1 & s2;
          uint64_t u1 = c1 ^ c2 ^ c3, u2 = maj3(c1, c2, c3);
          int k0 = 0, k1 = 0;
          for (int u = 0; u < 7; ++u) {
            k0 += r >= T[u];
            k1 += r >= T[7 + u];
          }
          uint64_t f0 = count_ge(u2, u1, u0, k0), f1 = count_ge(u2, u1, u0, k1);
          uint64_t flip = (s & f1) | (~s & f0);
          S[i] = s ^ flip;
        }
      }
    }
  }
}

void orc_sweep_lattice(int L, uint64_t *S, const uint64_t *J,
                       const uint64_t T[14], uint64_t spin_seed, uint32_t g,
                       uint32_t rho, uint32_t t, uint32_t n, int scalar_path) {
  const int N = L * L * L;
  uint32_t *rnd = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)N);
  sweep_randoms(L, spin_seed, g, rho, t, n, rnd);
  update_colour(L, S, J, T, rnd, 0, scalar_path);
  update_colour(L, S, J, T, rnd, 1, scalar_path);
  free(rnd);
}

/* vertical counter: planes[p] bit l = bit p of lane l's count */
static inline void vc_add(uint64_t *planes, int np, uint64_t w) {
  uint64_t carry = w;
  for (int p = 0; p < np && carry; ++p) {
    uint64_t t = planes[p] & carry;
    planes[p] ^= carry;
    carry = t;
  }
}

static inline void vc_extract(const uint64_t *planes, int np, int32_t out[64]) {
  for (int l = 0; l < 64; ++l) {
    int32_t v = 0;
    for (int p = 0; p < np; ++p) v |= (int32_t)((planes[p] >> l) & 1) << p;
    out[l] = v;
  }
}

void orc_lattice_energy(int L, const uint64_t *S, const uint64_t *J,
                        int32_t U[64], int32_t K[64]) {
  const int N = L * L * L;
  const uint64_t *Jx = J, *Jy = J + N, *Jz = J + 2 * N;
  uint64_t pu[32] = {0}, pk[32] = {0};
  for (int z = 0; z < L; ++z)
    for (int y = 0; y < L; ++y)
      for (int x = 0; x < L; ++x) {
        int i = sidx(L, x, y, z);
        uint64_t s = S[i];
        vc_add(pu, 32, s ^ S[sidx(L, (x + 1) % L, y, z)] ^ Jx[i]);
        vc_add(pu, 32, s ^ S[sidx(L, x, (y + 1) % L, z)] ^ Jy[i]);
        vc_add(pu, 32, s ^ S[sidx(L, x, y, (z + 1) % L)] ^ Jz[i]);
        vc_add(pk, 32, s);
      }
  vc_extract(pu, 31, U);
  vc_extract(pk, 31, K);
}

/* ------------------------------------------------------------ simulation */
struct orc_sim {
  int L, N, G, R, T, pt_every, meas_every, scalar;
  double H;
  double *beta;
  uint64_t cseed, sseed;
  uint64_t *S;       /* [g][rho][t][N] */
  uint64_t *J;       /* [g][3][N] */
  uint64_t *thr;     /* [t][14] */
  int64_t sweep;
  int64_t n_meas, cap_meas;
  int32_t *meas;     /* n_meas * G*T*64*nfield */
  uint64_t *acc, *att; /* [T-1] */
};

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

orc_sim *orc_create(int L, int n_groups, int n_replicas, int n_temps,
                    const double *betas, double H, uint64_t coupling_seed,
                    uint64_t spin_seed, int pt_every, int meas_every,
                    int scalar_path) {
  if (L < 2 || (L & 1) || n_groups < 1 || n_replicas < 1 || n_temps < 1) return NULL;
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->L = L; s->N = L * L * L; s->G = n_groups; s->R = n_replicas; s->T = n_temps;
  s->H = H; s->cseed = coupling_seed; s->sseed = spin_seed;
  s->pt_every = pt_every; s->meas_every = meas_every; s->scalar = scalar_path;
  s->beta = (double *)malloc(sizeof(double) * (size_t)n_temps);
  memcpy(s->beta, betas, sizeof(double) * (size_t)n_temps);
  size_t nl = (size_t)n_groups * n_replicas * n_temps;
  s->S = (uint64_t *)calloc(nl * s->N, sizeof(uint64_t));
  s->J = (uint64_t *)calloc((size_t)n_groups * 3 * s->N, sizeof(uint64_t));
  s->thr = (uint64_t *)malloc(sizeof(uint64_t) * 14 * (size_t)n_temps);
  for (int t = 0; t < n_temps; ++t) orc_thresholds(s->beta[t], H, s->thr + 14 * t);
  s->acc = (uint64_t *)calloc((size_t)(n_temps > 1 ? n_temps - 1 : 1), sizeof(uint64_t));
  s->att = (uint64_t *)calloc((size_t)(n_temps > 1 ? n_temps - 1 : 1), sizeof(uint64_t));
  return s;
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  free(s->beta); free(s->S); free(s->J); free(s->thr); free(s->meas);
  free(s->acc); free(s->att);
  free(s);
}

static inline uint64_t *lat(const orc_sim *s, int g, int rho, int t) {
  return s->S + (((size_t)g * s->R + rho) * s->T + t) * (size_t)s->N;
}

void orc_init(orc_sim *s) {
  const int N = s->N;
  for (int g = 0; g < s->G; ++g)
    for (int d = 0; d < 3; ++d)
      for (int i = 0; i < N; ++i) {
        uint32_t out[4];
        philox_tagged((uint32_t)i, (uint32_t)d, (uint32_t)g, ORC_TAG_COUP, 0, 0, s->cseed, out);
        s->J[((size_t)g * 3 + d) * N + i] = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
      }
  for (int g = 0; g < s->G; ++g)
    for (int rho = 0; rho < s->R; ++rho)
      for (int t = 0; t < s->T; ++t) {
        uint64_t *S = lat(s, g, rho, t);
        for (int i = 0; i < N; ++i) {
          uint32_t out[4];
          philox_tagged((uint32_t)i, 0, (uint32_t)g, ORC_TAG_SPIN, (uint32_t)rho, (uint32_t)t, s->sseed, out);
          S[i] = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
        }
      }
  s->sweep = 0;
  s->n_meas = 0;
}

int orc_nfield(const orc_sim *s) { return 2 * s->R + 1 + 3 * s->L; }

void orc_pt_step(orc_sim *s, int64_t b) {
  const int N = s->N, T = s->T;
  const int nl = s->G * s->R;
  int32_t *U = (int32_t *)malloc(sizeof(int32_t) * 64 * (size_t)T * nl);
  int32_t *K = (int32_t *)malloc(sizeof(int32_t) * 64 * (size_t)T * nl);
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
  for (int idx = 0; idx < nl * T; ++idx) {
    int gr = idx / T, t = idx % T;
    int g = gr / s->R, rho = gr % s->R;
    orc_lattice_energy(s->L, lat(s, g, rho, t), s->J + (size_t)g * 3 * N,
                       U + (size_t)idx * 64, K + (size_t)idx * 64);
  }
  int mode = (int)(b & 1);
  for (int gr = 0; gr < nl; ++gr) {
    int g = gr / s->R, rho = gr % s->R;
    for (int t = mode; t + 1 < T; t += 2) {
      const int32_t *Ua = U + ((size_t)gr * T + t) * 64, *Ub = Ua + 64;
      const int32_t *Ka = K + ((size_t)gr * T + t) * 64, *Kb = Ka + 64;
      uint64_t m = 0;
      for (int q = 0; q < 16; ++q) {
        uint32_t out[4];
        philox_tagged((uint32_t)q, (uint32_t)b, (uint32_t)g, ORC_TAG_PT, (uint32_t)rho, (uint32_t)t, s->sseed, out);
        for (int k = 0; k < 4; ++k) {
          int l = 4 * q + k;
          double dE = (double)(2 * (Ua[l] - Ub[l])) - s->H * (double)(2 * (Ka[l] - Kb[l]));
          double x = (s->beta[t] - s->beta[t + 1]) * dE;
          int acc = (x >= 0.0) || ((uint64_t)out[k] < (uint64_t)(orc_exp(x) * 4294967296.0));
          if (acc) m |= (uint64_t)1 << l;
        }
      }
      s->acc[t] += (uint64_t)__builtin_popcountll(m);
      s->att[t] += 64;
      if (m) {
        uint64_t *A = lat(s, g, rho, t), *B = lat(s, g, rho, t + 1);
        for (int i = 0; i < N; ++i) {
          uint64_t d = (A[i] ^ B[i]) & m;
          A[i] ^= d;
          B[i] ^= d;
        }
      }
    }
  }
  free(U);
  free(K);
}

static void measure(orc_sim *s) {
  const int N = s->N, L = s->L, T = s->T, R = s->R, nf = orc_nfield(s);
  size_t per = (size_t)s->G * T * 64 * nf;
  if (s->n_meas == s->cap_meas) {
    s->cap_meas = s->cap_meas ? 2 * s->cap_meas : 16;
    s->meas = (int32_t *)realloc(s->meas, sizeof(int32_t) * per * (size_t)s->cap_meas);
  }
  int32_t *rec = s->meas + per * (size_t)s->n_meas;
  memset(rec, 0, sizeof(int32_t) * per);
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
  for (int gt = 0; gt < s->G * T; ++gt) {
    int g = gt / T, t = gt % T;
    int32_t *base = rec + (size_t)gt * 64 * nf;
    for (int rho = 0; rho < R; ++rho) {
      int32_t U[64], K[64];
      orc_lattice_energy(L, lat(s, g, rho, t), s->J + (size_t)g * 3 * N, U, K);
      for (int l = 0; l < 64; ++l) {
        base[l * nf + rho] = U[l];
        base[l * nf + R + rho] = K[l];
      }
    }
    if (R >= 2) {
      const uint64_t *A = lat(s, g, 0, t), *B = lat(s, g, 1, t);
      for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
          for (int x = 0; x < L; ++x) {
            uint64_t w = A[sidx(L, x, y, z)] ^ B[sidx(L, x, y, z)];
            while (w) {
              int l = __builtin_ctzll(w);
              w &= w - 1;
              int32_t *f = base + l * nf + 2 * R;
              f[0] += 1;
              f[1 + x] += 1;
              f[1 + L + y] += 1;
              f[1 + 2 * L + z] += 1;
            }
          }
    }
  }
  s->n_meas++;
}

void orc_run(orc_sim *s, int64_t n_sweeps) {
  const int N = s->N;
  const int nl = s->G * s->R * s->T;
  for (int64_t it = 0; it < n_sweeps; ++it) {
    int64_t n = s->sweep;
    if (s->pt_every > 0 && s->T > 1 && n % s->pt_every == 0) orc_pt_step(s, n / s->pt_every);
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
    for (int idx = 0; idx < nl; ++idx) {
      int t = idx % s->T, gr = idx / s->T;
      int g = gr / s->R, rho = gr % s->R;
      orc_sweep_lattice(s->L, lat(s, g, rho, t), s->J + (size_t)g * 3 * N,
                        s->thr + 14 * t, s->sseed, (uint32_t)g, (uint32_t)rho,
                        (uint32_t)t, (uint32_t)n, s->scalar);
    }
    s->sweep = n + 1;
    if (s->meas_every > 0 && s->sweep % s->meas_every == 0) measure(s);
  }
}

int64_t orc_sweep_count(const orc_sim *s) { return s->sweep; }
int64_t orc_n_meas(const orc_sim *s) { return s->n_meas; }

void orc_get_meas(const orc_sim *s, int64_t i, int32_t *rec) {
  size_t per = (size_t)s->G * s->T * 64 * orc_nfield(s);
  memcpy(rec, s->meas + per * (size_t)i, sizeof(int32_t) * per);
}

void orc_get_spins(const orc_sim *s, uint64_t *out) {
  memcpy(out, s->S, sizeof(uint64_t) * (size_t)s->G * s->R * s->T * s->N);
}
void orc_set_spins(orc_sim *s, const uint64_t *in) {
  memcpy(s->S, in, sizeof(uint64_t) * (size_t)s->G * s->R * s->T * s->N);
}
void orc_get_couplings(const orc_sim *s, uint64_t *out) {
  memcpy(out, s->J, sizeof(uint64_t) * (size_t)s->G * 3 * s->N);
}
void orc_set_couplings(orc_sim *s, const uint64_t *in) {
  memcpy(s->J, in, sizeof(uint64_t) * (size_t)s->G * 3 * s->N);
}

void orc_energies(const orc_sim *s, int32_t *U, int32_t *K) {
  for (int g = 0; g < s->G; ++g)
    for (int rho = 0; rho < s->R; ++rho)
      for (int t = 0; t < s->T; ++t) {
        size_t idx = ((size_t)g * s->R + rho) * s->T + t;
        orc_lattice_energy(s->L, lat(s, g, rho, t), s->J + (size_t)g * 3 * s->N,
                           U + idx * 64, K + idx * 64);
      }
}

void orc_swap_counts(const orc_sim *s, uint64_t *accepted, uint64_t *attempted) {
  for (int t = 0; t + 1 < s->T; ++t) {
    accepted[t] = s->acc[t];
    attempted[t] = s->att[t];
  }
}
