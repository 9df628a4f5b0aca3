/*
 * sg_oracle.h — CPU oracle for the parallel-tempering 3D Edwards–Anderson
 * ±J engine.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP product is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * It restates, in plain C, the algorithm of the reference's hot path
 * (Program/src/eamsim.cu) in the multispin form that the product uses.  The
 * canonical algorithm is defined here, and the product follows the same
 * definition.
 *
 * Lattice and words
 *   L even, N = L^3, periodic.  Natural site index s = x + L*(y + L*z).
 *   One uint64 word per site holds 64 disorder samples ("lanes") of one
 *   (group g, replica rho, temperature slot t) lattice; bit l = sample l.
 *   Spin bit 1 <=> sigma = +1 (reference eamsim.cu:483 "2*sbit-1").
 *   Couplings J[d][s], d = 0:+x 1:+y 2:+z, bond from s to s+e_d.
 *   J bit 1 <=> J = -1, so a bond is unsatisfied iff s_i ^ s_j ^ Jbit = 1.
 *   (The reference stores J=+1 as bit 1 and XORs it in to get the
 *   *satisfied* bit, eamsim.cu:162-167; this is the complement.)
 *
 * Energy (reference eamsim.cu:320-325)
 *   U = number of unsatisfied bonds (3N bonds), K = number of up spins.
 *   E = 2U - 3N - H*(2K - N),  M = 2K - N.
 *
 * Checkerboard Metropolis (reference eamsim.cu:135-206, 24-40)
 *   colour c = (x+y+z)&1; a sweep updates all c=0 sites, then all c=1 sites.
 *   For a site with u unsatisfied bonds and spin sigma:
 *     dE = 12 - 4u + 2*H*sigma  (= -2 * e_i of eamsim.cu:35).
 *   Threshold  T[sigma][u] = 2^32                      if dE <= 0
 *                          = floor(2^32 * exp(-beta*dE)) otherwise,
 *   computed in FP64 with orc_exp (a fixed sequence of IEEE operations).
 *   The spin flips iff (uint64) r < T, where r is a 32-bit Philox output
 *   shared by all 64 lanes of the word.  Unlike eamsim.cu:187, which draws
 *   one number per spin, all 64 samples of a word share r.
 *
 * Random numbers: Philox4x32-10 (Salmon et al., SC'11, Random123).
 *   ctr[3] = tag<<28 | (rho & 0xFFF)<<16 | (t & 0xFFFF)
 *   tags: COUP=1 SPIN=2 MC=3 PT=4.
 *   couplings: out = Philox((s, d, g, COUP<<28), coupling_seed)
 *              J[d][s] = out0 | out1<<32
 *   initial spins (hot start, eamsim.cu:562-575):
 *              out = Philox((s, 0, g, SPIN|rho|t), spin_seed)
 *              S[s] = out0 | out1<<32
 *   Metropolis r of a site at sweep n:
 *              h = (x>>1) + (L/2)*(y + L*z), Q = N/4, j = h % Q, hi = h / Q
 *              r = Philox((j, n, g, MC|rho|t), spin_seed)[2c + hi]
 *   PT r of (step b, pair starting at slot t, lane l):
 *              r = Philox((l>>2, b, g, PT|rho|t), spin_seed)[l & 3]
 *   A 64-bit seed is the key (lo, hi).
 *
 * Parallel tempering (reference eamsim.cu:43-90, 385-392)
 *   Before sweep n, if pt_every > 0 and n % pt_every == 0: step b = n/pt_every,
 *   pairs (t, t+1) for t = (b&1), (b&1)+2, ... with t+1 < T.  Per lane:
 *     dE = (double)(2*(U_t - U_t1)) - H*(double)(2*(K_t - K_t1))
 *     x  = (beta_t - beta_t1) * dE
 *     accept iff x >= 0 or (uint64) r < (uint64)(orc_exp(x) * 2^32).
 *   An accepted lane exchanges its spin bits between the two lattices.
 *
 * Measurement: after sweep n, if meas_every > 0 and (n+1) % meas_every == 0,
 *   one record per (g, t, lane) of int32 fields
 *     U[rho] (R), K[rho] (R), W, Wx[L], Wy[L], Wz[L]
 *   where W = #sites with s^0 != s^1 (replicas 0 and 1) and Wx[x] counts those
 *   sites in the plane x (likewise y, z).  With R < 2 the overlap fields are 0.
 *   Then q0 = (N - 2W)/N, and Re/Im q(k = 2pi/L e_d) = -(2/N) sum_c
 *   cos/sin(2 pi c / L) W_d[c] (reference eamsim.cu:451-511).
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TAG_COUP = 1, ORC_TAG_SPIN = 2, ORC_TAG_MC = 3, ORC_TAG_PT = 4 };

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_exp(double x);
void orc_thresholds(double beta, double H, uint64_t T[14]); /* [sigma*7 + u] */

typedef struct orc_sim orc_sim;

orc_sim *orc_create(int L, int n_groups, int n_replicas, int n_temps,
                    const double *betas, double H, uint64_t coupling_seed,
                    uint64_t spin_seed, int pt_every, int meas_every,
                    int scalar_path);
void orc_destroy(orc_sim *sim);
void orc_set_threads(int n);
void orc_init(orc_sim *sim);
void orc_run(orc_sim *sim, int64_t n_sweeps);
void orc_pt_step(orc_sim *sim, int64_t b);   /* one PT step b on the current state */
int64_t orc_sweep_count(const orc_sim *sim);
int orc_nfield(const orc_sim *sim);
int64_t orc_n_meas(const orc_sim *sim);
void orc_get_meas(const orc_sim *sim, int64_t i, int32_t *rec);
void orc_get_spins(const orc_sim *sim, uint64_t *out);
void orc_set_spins(orc_sim *sim, const uint64_t *in);
void orc_get_couplings(const orc_sim *sim, uint64_t *out);
void orc_set_couplings(orc_sim *sim, const uint64_t *in);
void orc_energies(const orc_sim *sim, int32_t *U, int32_t *K);
void orc_swap_counts(const orc_sim *sim, uint64_t *accepted, uint64_t *attempted);

/* single-lattice building blocks (used by the oracle's own tests) */
void orc_sweep_lattice(int L, uint64_t *S, const uint64_t *J,
                       const uint64_t T[14], uint64_t spin_seed, uint32_t g,
                       uint32_t rho, uint32_t t, uint32_t n, int scalar_path);
void orc_lattice_energy(int L, const uint64_t *S, const uint64_t *J,
                        int32_t U[64], int32_t K[64]);

#ifdef __cplusplus
}
#endif
#endif
