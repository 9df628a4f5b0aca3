// sg_handle.h — one simulation (one L; its sample groups, replicas and
// temperatures) on the device: buffers, descriptors, the PT-interval launch
// sequence, measurement, results and checkpoints.  The C-ABI of
// include/sg.h wraps one of these per handle; the batch engine holds one per
// job, with its buffers in the engine's arena and its launches on the
// engine's stream for its L.
//
// Mirrors the reference's host object `eamsim` (Program/src/eamsim.cuh:66-110,
// eamsim.cu:513-744): construction validates H and allocates, init_lattice
// draws couplings and spins, run_simulation loops over measurement intervals,
// save/load_checkpoint serialise the state.  Differences: every size is a
// run-time parameter, the random numbers are counter-based (runs are
// reproducible and resumable bit for bit), and errors are status codes.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sg.h"
#include "sg_arena.h"
#include "sg_common.h"
#include "sg_kernels.h"

namespace sg {
namespace detail {

inline thread_local std::string g_create_error;

constexpr uint32_t kNoPair = 0xFFFFFFFFu;

// The dispatcher hands work-group i of a launch to XCD i % 8, and each XCD
// has its own L2.  Streamed launches therefore order their (lattice, chunk)
// tasks so that every chunk of a lattice lands on one XCD, the same one in
// every launch: a lattice written by one colour launch is read by the next
// from that XCD's L2.  Lattice k of the list (in order of appearance) takes
// slot k % 8; within a slot the tasks go chunk by chunk across its lattices
// (chunk 0 of each, then chunk 1, ...), so the lattices of one coupling
// group that share an XCD read each chunk's coupling records close together
// in time, from L2; the slots interleave, short ones padded with no-op tasks
// (chunk index past the last quad: no spins, no energy slot touched).
constexpr int kXcds = 8;

inline std::vector<TaskDesc> xcd_order(const std::vector<TaskDesc> &in, uint32_t pad_chunk) {
  std::vector<std::vector<std::vector<TaskDesc>>> runs(kXcds);   // [slot][lattice][chunk]
  int k = -1;
  uint32_t last_lat = 0xFFFFFFFFu, last_job = 0xFFFFFFFFu;
  for (const TaskDesc &t : in) {
    if (t.lat[0] != last_lat || t.job != last_job) {
      ++k;
      last_lat = t.lat[0];
      last_job = t.job;
      runs[k % kXcds].emplace_back();
    }
    runs[k % kXcds].back().push_back(t);
  }
  std::vector<std::vector<TaskDesc>> slot(kXcds);
  for (int x = 0; x < kXcds; ++x) {
    size_t len = 0;
    for (const auto &r : runs[x]) len = std::max(len, r.size());
    for (size_t i = 0; i < len; ++i)
      for (const auto &r : runs[x])
        if (i < r.size()) slot[x].push_back(r[i]);
  }
  size_t m = 0;
  for (auto &v : slot) m = std::max(m, v.size());
  std::vector<TaskDesc> out;
  if (in.empty()) return out;
  out.reserve(m * kXcds);
  for (size_t i = 0; i < m; ++i)
    for (int x = 0; x < kXcds; ++x) {
      if (i < slot[x].size()) {
        out.push_back(slot[x][i]);
      } else {
        TaskDesc t = in.front();
        t.lat[1] = pad_chunk;
        out.push_back(t);
      }
    }
  return out;
}

// Resident tasks of one launch reordered so that all tasks with the same
// coupling group (key) run on one XCD: queue x collects the tasks of keys
// with key % kXcds == x, in order, and the queues are interleaved (work-group
// i goes to XCD i % kXcds).  The group's couplings are then fetched into one
// XCD's L2 instead of every one.  The two-lattice tasks stay ahead of the
// single ones (they are twice the work).  Unequal queues leave the tail in
// plain order.
template <class KeyFn>
inline void xcd_interleave(std::vector<TaskDesc> &tv, KeyFn key) {
  std::vector<TaskDesc> out;
  out.reserve(tv.size());
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<std::vector<TaskDesc>> q(kXcds);
    for (const TaskDesc &t : tv)
      if ((t.lat[1] != 0xFFFFFFFFu) == (pass == 0)) q[key(t) % kXcds].push_back(t);
    std::vector<size_t> pos(kXcds, 0);
    for (bool any = true; any;) {
      any = false;
      for (int x = 0; x < kXcds; ++x)
        if (pos[x] < q[x].size()) {
          out.push_back(q[x][pos[x]++]);
          any = true;
        }
    }
  }
  tv.swap(out);
}

struct Fail {
  sg_status code;
  std::string msg;
};

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw Fail{SG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)};  \
  } while (0)

inline uint64_t fnv1a(const void *data, size_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char *p = (const unsigned char *)data;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

// Autotuner (sg_runtime.cpp): the work-group shape of the sweeps for
// n_lattices lattices of size L — SG_KERNEL_RESIDENT (a lattice pair in one
// work-group for a PT interval), SG_KERNEL_RESIDENT_SINGLE (one lattice per
// work-group, twice the work-groups, PT in its own launch) or
// SG_KERNEL_STREAMED (one launch per colour over chunks of every lattice) —
// decided by timing each shape the L has on the device; cached per (device,
// L, field, power-of-two lattice count).  Each shape is timed at each of
// its tiles (shape_tiles); ms[k] receives the best time of shape k (0 where
// the L has no such shape), tms[k * kMaxTiles + t] that of tile t, and *tile
// the winning shape's tile.
constexpr int kMaxTiles = 4;
int tune_kernel_choice(int device, int L, bool field, size_t n_lattices, double *ms = nullptr,
                       int *tile = nullptr, double *tms = nullptr);
// The same decision without measuring: the cached result when the tuner has
// measured that (device, L, field, bucket), else the rule its measurements
// support (tile 0).  The batch engine decides with this, so no sg_step ever
// runs the tuner.
int cached_kernel_choice(int device, int L, bool field, size_t n_lattices, int *tile = nullptr);

// Metropolis thresholds of one temperature (see ThrRec): FP64 on the host,
// floor(2^32 det_exp(-beta dE)) for every (spin bit, u) with dE > 0.
inline void build_threshold_rec(double b, double H, ThrRec &r) {
#pragma clang fp contract(off)
  std::memset(&r, 0, sizeof(r));
  for (int sig = 0; sig < 2; ++sig) {
    const double sv = sig ? 1.0 : -1.0;
    uint32_t nfin = 0;
    for (int u = 0; u < 7; ++u) {
      const double dE = (double)(12 - 4 * u) + 2.0 * H * sv;
      if (dE <= 0.0) break;
      const double p = det_exp(-b * dE);
      const uint64_t thr = (uint64_t)(p * 4294967296.0);
      if (thr >= (1ull << 32)) throw Fail{SG_ERR_INVALID, "threshold overflow"};
      r.t[sig][u] = (uint32_t)thr;
      nfin = (uint32_t)(u + 1);
    }
    r.nfin[sig] = nfin;
    // the sweep kernels take u = 5, 6 as always accepted (true for H <= 4)
    if (nfin > 5) throw Fail{SG_ERR_INVALID, "H out of range [0, 4]"};
  }
  if (H == 0.0 && (r.nfin[0] != 3 || r.nfin[1] != 3))
    throw Fail{SG_ERR_INVALID, "unexpected threshold table at H = 0"};
  // the sweep kernels count how many thresholds of a spin value the random
  // number is below, which needs them rising with u (exp(-beta dE) falls
  // with dE)
  for (int sig = 0; sig < 2; ++sig)
    for (uint32_t u = 1; u < r.nfin[sig]; ++u)
      if (r.t[sig][u - 1] > r.t[sig][u])
        throw Fail{SG_ERR_INVALID, "thresholds are not monotone in u"};
}

// Device bytes one engine job needs, block by block as sg_handle::create
// allocates them from an arena (each block rounded to the arena's
// alignment; no record buffer).
inline size_t job_footprint(int L, int n_samples, int R, int T) {
  const bool streamed = phase_kernel_supported(L);  // both task sets are kept
  const size_t N = (size_t)L * L * L, G = (size_t)(n_samples + 63) / 64;
  const size_t nlat = G * R * T;
  const size_t npairs = G * R * (size_t)(T > 1 ? T - 1 : 0);
  auto r = [](size_t b) { return DevArena::round_up(b ? b : 1); };
  size_t b = r(8 * nlat * N) + r(8 * G * 3 * N) + r(8 * G * 6 * N) + r(sizeof(ThrRec) * T) +
             r(4 * nlat * 128) + r(sizeof(LatDesc) * nlat) + r(sizeof(MeasDesc) * G * T) +
             r(sizeof(JobState)) + r(8 * (size_t)(T > 1 ? T - 1 : 1)) +
             r(sizeof(PairDesc) * npairs) + 2 * r(sizeof(TaskDesc) * nlat) +
             r(sizeof(ObsAcc) * G * T * 64) + r(16 * (size_t)L) + r(sizeof(JobPtrs));
  // PT launches of the one-lattice shapes: (pair, chunk) tasks per parity
  b += 2 * r(sizeof(TaskDesc) * (npairs / 2 + G * R) * pt_chunks(L));
  if (!single_kernel_supported(L)) b += r(sizeof(TaskDesc) * nlat);   // one-lattice tasks
  if (streamed)
    // chunk tasks: XCD padding adds at most 7 lattices' worth per group
    b += r(sizeof(TaskDesc) * (nlat + 7 * G) * phase_chunks(L));
  if (streamed && slab_kernel_supported(L))
    // slab sweeps: the second spin buffer and the (lattice, slab pair) tasks
    b += r(8 * nlat * N) + r(sizeof(TaskDesc) * (nlat + 7 * G) * kSlabParts);
  return b;
}

}  // namespace detail
}  // namespace sg

using namespace sg;
using namespace sg::detail;

struct sg_handle {
  std::string err;
  int L = 0, N = 0, G = 0, R = 0, T = 0, S = 0, pt_every = 0, meas_every = 0, form = 0;
  int device = 0;
  double H = 0;
  std::vector<double> beta;
  uint64_t cseed = 0, sseed = 0;
  bool field = false;
  hipStream_t st = nullptr;
  bool own_stream = true;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device memory comes from this arena when set (engine jobs), else from
  // hipMalloc
  DevArena *arena = nullptr;

  template <class T>
  void dev_alloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (arena) {
      *p = (T *)arena->alloc(count * sizeof(T));
      if (!*p) throw Fail{SG_ERR_NOMEM, "arena exhausted"};
      return;
    }
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
      *p = nullptr;
      throw Fail{SG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)};
    }
  }
  template <class T>
  void dev_free(T *&p) {
    if (p) {
      if (arena)
        arena->release((void *)p);
      else
        (void)hipFree((void *)p);
    }
    p = nullptr;
  }

  // this job's buffers as the kernels see them, and a one-entry device table
  // of them for the launches of this handle alone
  JobPtrs self{};
  JobPtrs *d_self = nullptr;
  std::vector<TaskDesc> h_tasks[2];

  // streamed sweeps: (lattice, chunk) tasks of the colour launches and
  // (pair, chunk) tasks of the PT launches of parity 0 and 1, ordered by
  // group (chunk_off / ptc_off).  Both kernels use the same data, so a
  // handle that can stream keeps both task sets and `streamed` may change
  // between PT intervals.
  bool can_stream = false, streamed = false;
  uint64_t *d_Js = nullptr;  // coupling site records [group][colour][h][6]
  TaskDesc *d_chunks = nullptr, *d_ptc[2] = {nullptr, nullptr};
  std::vector<TaskDesc> h_chunks, h_ptc[2];
  std::vector<int> chunk_off, ptc_off[2];
  // slab sweeps (L with a slab kernel): second spin buffer, (lattice, slab
  // pair) tasks per group like the chunks
  uint64_t *d_spins2 = nullptr;
  TaskDesc *d_slabs = nullptr;
  std::vector<TaskDesc> h_slabs;
  std::vector<int> slab_off;

  uint64_t *d_spins = nullptr, *d_J = nullptr;
  ThrRec *d_thr = nullptr;
  int32_t *d_energy = nullptr, *d_rec = nullptr;
  LatDesc *d_lat = nullptr;
  PairDesc *d_pairs = nullptr;
  MeasDesc *d_meas = nullptr;
  JobState *d_jobs = nullptr;
  unsigned long long *d_cnt = nullptr;
  ObsAcc *d_acc = nullptr;        // [group][temp][lane]
  double *d_cs = nullptr;         // cos, sin of 2 pi c / L
  int npairs = 0, nfield = 0;
  // work-group tasks of the sweep kernel for PT steps of parity 0 and 1:
  // pairs (t, t+1) with t of that parity, the other slots single; ordered by
  // group (task_off[p][g] .. task_off[p][g+1])
  TaskDesc *d_tasks[2] = {nullptr, nullptr};
  std::vector<int> task_off[2];
  // split sweeps (SG_KERNEL_SPLIT, L = 32, 40, 48): a lattice held by P
  // work-groups for a PT interval (P from the lattices per launch); one task
  // per lattice; boundary-plane exchange buffer, phase flags, error word and
  // the launch's ticket counters
  bool can_split = false, split_shape = false;
  int tile = 0;                  // the shape's tile (threads per work-group variant)
  int split_P = 0;
  TaskDesc *d_split = nullptr;
  uint64_t *d_xbuf = nullptr;
  uint32_t *d_xflag = nullptr, *d_xerr = nullptr, *d_ticket = nullptr;
  // one lattice per work-group with a two-lattice kernel
  // (SG_KERNEL_RESIDENT_SINGLE): every lattice its own task, the same list
  // for both parities, PT in its own launch; ordered by group (single_off)
  bool can_single = false, single_shape = false;
  TaskDesc *d_single = nullptr;
  std::vector<TaskDesc> h_single;
  std::vector<int> single_off;

  // records for the sink: the measurement kernel writes d_rec, an async copy
  // moves it to one of kRing pinned buffers, and the sink runs once the copy
  // has landed (while later intervals already run on the device)
  static constexpr int kRing = 3;
  int32_t *h_ring[kRing] = {nullptr, nullptr, nullptr};
  hipEvent_t ring_ev[kRing] = {nullptr, nullptr, nullptr};
  struct PendingRec { int slot; int64_t sweep; };
  std::vector<PendingRec> rec_queue;
  int ring_next = 0;
  size_t rec_count = 0;
  std::vector<uint64_t> swap_att; // [T-1]
  bool initialized = false, energy_valid = false;
  bool energy_k_valid = false;   // the energy slots hold U and K of the current states
  int64_t sweep = 0, n_meas = 0;
  double kernel_ms = 0, flips = 0;

  // per-kernel profiling (sg_set_profiling)
  struct KStat { int64_t launches = 0; double ms = 0, flips = 0; };
  struct Pending { hipEvent_t a, b; int kind; double flips; };
  bool profiling = false;
  KStat kstat[3];
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_pool;

  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  template <class F>
  void timed(int kind, double fl, F &&launch) {
    if (!profiling) {
      launch();
      return;
    }
    Pending p{take_event(), take_event(), kind, fl};
    HIPCHK(hipEventRecord(p.a, st));
    launch();
    HIPCHK(hipEventRecord(p.b, st));
    pending.push_back(p);
  }
  void collect_profile() {
    if (pending.empty()) return;
    HIPCHK(hipStreamSynchronize(st));
    for (auto &p : pending) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
      kstat[p.kind].launches += 1;
      kstat[p.kind].ms += ms;
      kstat[p.kind].flips += p.flips;
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }

  // batched form: one PT interval (sweep launch + counter advance) per
  // (parity, length, energy) captured once as a hipGraph and replayed
  struct GraphKey { int parity, seg, energy, count; hipGraphExec_t exec; };
  std::vector<GraphKey> graphs;

  ~sg_handle() { release(); }

  void release() {
    dev_free(d_spins); dev_free(d_J); dev_free(d_thr); dev_free(d_energy); dev_free(d_rec);
    dev_free(d_lat); dev_free(d_pairs); dev_free(d_meas); dev_free(d_jobs); dev_free(d_cnt);
    dev_free(d_acc); dev_free(d_cs); dev_free(d_tasks[0]); dev_free(d_tasks[1]); dev_free(d_single);
    dev_free(d_self); dev_free(d_Js); dev_free(d_chunks); dev_free(d_ptc[0]); dev_free(d_ptc[1]);
    dev_free(d_spins2); dev_free(d_slabs);
    dev_free(d_split); dev_free(d_xbuf); dev_free(d_xflag); dev_free(d_xerr); dev_free(d_ticket);
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
    for (int i = 0; i < kRing; ++i) {
      if (h_ring[i]) (void)hipHostFree(h_ring[i]);
      if (ring_ev[i]) (void)hipEventDestroy(ring_ev[i]);
      h_ring[i] = nullptr;
      ring_ev[i] = nullptr;
    }
    for (auto &p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    pending.clear();
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    ev_pool.clear();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (st && own_stream) (void)hipStreamDestroy(st);
    ev0 = ev1 = nullptr; st = nullptr;
  }

  size_t lat_index(int g, int rho, int t) const { return ((size_t)g * R + rho) * T + t; }
  size_t n_lat() const { return (size_t)G * R * T; }

  // ------------------------------------------------------------ set-up
  // stream: the engine's stream for this job's L (not owned), or nullptr for
  // a stream of its own
  void create(const sg_params &p, DevArena *from = nullptr, hipStream_t stream = nullptr) {
    arena = from;
    L = p.L; G = (p.n_samples + 63) / 64; S = p.n_samples; R = p.n_replicas; T = p.n_temps;
    N = L * L * L; H = p.H; cseed = p.coupling_seed; sseed = p.spin_seed;
    pt_every = p.pt_every; meas_every = p.meas_every; form = p.form; device = p.device;
    field = H != 0.0;
    beta.resize(T);
    if (p.betas) {
      for (int t = 0; t < T; ++t) beta[t] = p.betas[t];
    } else {
      for (int t = 0; t < T; ++t) {
        const double f = T > 1 ? (double)t / (double)(T - 1) : 0.0;
        beta[t] = 1.0 / (p.T_min * std::pow(p.T_max / p.T_min, f));
      }
    }
    for (int t = 0; t < T; ++t)
      if (!(beta[t] > 0.0) || !std::isfinite(beta[t]))
        throw Fail{SG_ERR_INVALID, "every beta must be positive and finite"};

    HIPCHK(hipSetDevice(device));
    if (stream) {
      st = stream;
      own_stream = false;
    } else {
      HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    }
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    nfield = 2 * R + 1 + 3 * L;

    can_stream = phase_kernel_supported(L);
    can_single = !single_kernel_supported(L);
    // split sweeps need the streamed kernels too (energy-only passes)
    split_P = split_slabs(L, form == SG_FORM_PLAIN ? R * T : (int)((size_t)G * R * T));
    can_split = split_P > 0 && can_stream && !arena;
    if (p.kernel == SG_KERNEL_STREAMED && !can_stream)
      throw Fail{SG_ERR_INVALID, "no streamed kernel for this L"};
    if (p.kernel == SG_KERNEL_SPLIT && !can_split)
      throw Fail{SG_ERR_INVALID, "no split kernel for this L and size"};
    dev_alloc(&d_spins, n_lat() * N);
    dev_alloc(&d_J, (size_t)G * 3 * N);
    dev_alloc(&d_Js, (size_t)G * 6 * N);
    dev_alloc(&d_thr, (size_t)T);
    dev_alloc(&d_energy, n_lat() * 128);
    // engine jobs (arena) never deliver records: no buffer for them
    if (!arena) dev_alloc(&d_rec, (size_t)G * T * 64 * nfield);
    dev_alloc(&d_lat, n_lat());
    dev_alloc(&d_meas, (size_t)G * T);
    dev_alloc(&d_jobs, 1);
    dev_alloc(&d_cnt, (size_t)std::max(T - 1, 1));

    // thresholds
    std::vector<ThrRec> thr(T);
    for (int t = 0; t < T; ++t) build_thresholds(beta[t], thr[t]);
    HIPCHK(hipMemcpy(d_thr, thr.data(), sizeof(ThrRec) * T, hipMemcpyHostToDevice));

    // lattice descriptors [g][rho][t]
    std::vector<LatDesc> ld(n_lat());
    for (int g = 0; g < G; ++g)
      for (int rho = 0; rho < R; ++rho)
        for (int t = 0; t < T; ++t) {
          const size_t i = lat_index(g, rho, t);
          LatDesc &d = ld[i];
          std::memset(&d, 0, sizeof(d));
          d.spin_off = i * (uint64_t)N;
          d.j_off = (uint64_t)g * 3 * N;
          d.key0 = (uint32_t)sseed; d.key1 = (uint32_t)(sseed >> 32);
          d.g = (uint32_t)g;
          d.rt = ((uint32_t)rho << 16) | (uint32_t)t;
          d.thr_idx = (uint32_t)t;
          d.e_idx = (uint32_t)i;
        }
    HIPCHK(hipMemcpy(d_lat, ld.data(), sizeof(LatDesc) * ld.size(), hipMemcpyHostToDevice));

    // PT pairs [g][rho][t], t = 0..T-2
    npairs = G * R * std::max(T - 1, 0);
    dev_alloc(&d_pairs, (size_t)npairs);
    std::vector<PairDesc> pd;
    pd.reserve(npairs);
    for (int g = 0; g < G; ++g)
      for (int rho = 0; rho < R; ++rho)
        for (int t = 0; t + 1 < T; ++t) {
          PairDesc d;
          std::memset(&d, 0, sizeof(d));
          d.a_off = lat_index(g, rho, t) * (uint64_t)N;
          d.b_off = lat_index(g, rho, t + 1) * (uint64_t)N;
          d.beta_a = beta[t]; d.beta_b = beta[t + 1]; d.H = H;
          d.e_a = (uint32_t)lat_index(g, rho, t); d.e_b = (uint32_t)lat_index(g, rho, t + 1);
          d.key0 = (uint32_t)sseed; d.key1 = (uint32_t)(sseed >> 32);
          d.g = (uint32_t)g; d.rt = ((uint32_t)rho << 16) | (uint32_t)t;
          d.cnt_idx = (uint32_t)t; d.n_words = (uint32_t)N;
          d.parity = (uint32_t)(t & 1);
          pd.push_back(d);
        }
    if (npairs)
      HIPCHK(hipMemcpy(d_pairs, pd.data(), sizeof(PairDesc) * npairs, hipMemcpyHostToDevice));

    // resident tasks: PT pairs of the parity and the lattices left over, or
    // (single-lattice kernel) every lattice on its own, PT in its own launch
    const bool single = single_kernel_supported(L);
    for (int par = 0; par < 2; ++par) {
      std::vector<TaskDesc> tv;
      task_off[par].assign(G + 1, 0);
      for (int g = 0; g < G; ++g) {
        task_off[par][g] = (int)tv.size();
        for (int rho = 0; rho < R; ++rho)
          for (int t = 0; t < T;) {
            TaskDesc k;
            k.job = 0;
            k.lat[0] = (uint32_t)lat_index(g, rho, t);
            if (!single && (t & 1) == par && t + 1 < T) {
              k.lat[1] = (uint32_t)lat_index(g, rho, t + 1);
              k.pair = (uint32_t)(((size_t)g * R + rho) * (T - 1) + t);
              t += 2;
            } else {
              k.lat[1] = 0xFFFFFFFFu;
              k.pair = 0xFFFFFFFFu;
              t += 1;
            }
            tv.push_back(k);
          }
        // within a group, pairs (twice the work) first, then the single
        // lattices, so that the dispatcher's greedy fill ends evenly
        auto first_single = std::stable_partition(
            tv.begin() + task_off[par][g], tv.end(),
            [](const TaskDesc &k) { return k.lat[1] != 0xFFFFFFFFu; });
        // two-lattice kernels: the lattices left without a partner at this
        // parity (slots 0 and T-1 at odd parity) share a work-group, two by
        // two, with no PT between them (they share the group's couplings)
        if (!single) {
          std::vector<TaskDesc> singles(first_single, tv.end());
          tv.erase(first_single, tv.end());
          std::vector<TaskDesc> rest;
          for (size_t i = 0; i < singles.size(); i += 2) {
            TaskDesc k = singles[i];
            if (i + 1 < singles.size()) {
              k.lat[1] = singles[i + 1].lat[0];
              tv.push_back(k);
            } else {
              rest.push_back(k);
            }
          }
          tv.insert(tv.end(), rest.begin(), rest.end());
        }
      }
      if (form == SG_FORM_BATCHED) {  // one launch: all pairs of all groups first
        const size_t per_group = (size_t)R * T;
        xcd_interleave(tv, [per_group](const TaskDesc &k) {
          return (uint32_t)(k.lat[0] / per_group);
        });
      }
      task_off[par][G] = (int)tv.size();
      h_tasks[par] = tv;
      dev_alloc(&d_tasks[par], tv.size());
      HIPCHK(hipMemcpy(d_tasks[par], tv.data(), sizeof(TaskDesc) * tv.size(),
                       hipMemcpyHostToDevice));
    }

    std::vector<MeasDesc> md((size_t)G * T);
    for (int g = 0; g < G; ++g)
      for (int t = 0; t < T; ++t) {
        MeasDesc &d = md[(size_t)g * T + t];
        d.spin_off = lat_index(g, 0, t) * (uint64_t)N;
        d.rho_stride = (uint64_t)T * N;
        d.e_idx = lat_index(g, 0, t);
        d.e_stride = (uint64_t)T;
        d.out_off = ((uint64_t)g * T + t) * 64 * nfield;
        d.acc_off = ((uint64_t)g * T + t) * 64;
        d.cs_off = 0;
      }
    HIPCHK(hipMemcpy(d_meas, md.data(), sizeof(MeasDesc) * md.size(), hipMemcpyHostToDevice));

    dev_alloc(&d_acc, (size_t)G * T * 64);
    HIPCHK(hipMemset(d_acc, 0, sizeof(ObsAcc) * (size_t)G * T * 64));
    std::vector<double> cs(2 * L);
    for (int c = 0; c < L; ++c) {
      cs[c] = std::cos(2.0 * M_PI * c / L);
      cs[L + c] = std::sin(2.0 * M_PI * c / L);
    }
    dev_alloc(&d_cs, (size_t)2 * L);
    HIPCHK(hipMemcpy(d_cs, cs.data(), sizeof(double) * 2 * L, hipMemcpyHostToDevice));
    rec_count = (size_t)G * T * 64 * nfield;
    swap_att.assign((size_t)std::max(T - 1, 1), 0);
    reset_counters();

    build_ptc_tasks();
    if (can_single) build_single_tasks();
    if (can_stream) {
      build_stream_tasks();
      if (slab_kernel_supported(L)) {
        build_slab_tasks();
        dev_alloc(&d_spins2, n_lat() * N);
      }
    }
    if (can_split) build_split();
    self.spins = d_spins; self.J = d_J; self.Js = d_Js;
    self.spins2 = d_spins2;
    self.xbuf = d_xbuf; self.xflag = d_xflag; self.xerr = d_xerr;
    self.descs = d_lat; self.pairs = d_pairs;
    self.thr = d_thr; self.js = d_jobs; self.energy = d_energy; self.counters = d_cnt;
    dev_alloc(&d_self, 1);
    HIPCHK(hipMemcpy(d_self, &self, sizeof(JobPtrs), hipMemcpyHostToDevice));
    if (p.kernel == SG_KERNEL_AUTO && !arena) {
      int t = 0;
      const int k = tune_kernel_choice(device, L, field, n_lat(), nullptr, &t);
      set_mode(k, t);
    } else if (p.kernel != SG_KERNEL_AUTO) {
      set_mode(p.kernel, p.tile > 0 ? p.tile - 1 : 0);
    }
  }

  // the tiles (threads per work-group) shape k has at this L: the streamed
  // kernels kPhaseTiles, the split kernel split_tiles(L), the resident ones 1
  int shape_tiles(int k) const {
    if (k == SG_KERNEL_STREAMED) return can_stream ? kPhaseTiles : 0;
    if (k == SG_KERNEL_SPLIT) return can_split ? split_tiles(L) : 0;
    return 1;
  }

  // switches between the work-group shapes (between PT intervals): k =
  // SG_KERNEL_RESIDENT, _RESIDENT_SINGLE (the same as _RESIDENT where the
  // resident kernel holds one lattice), _STREAMED or _SPLIT, at tile t;
  // captured graphs hold the old launch sequence
  void set_mode(int k, int t = 0) {
    const bool s = k == SG_KERNEL_STREAMED, one = k == SG_KERNEL_RESIDENT_SINGLE && can_single;
    const bool sp = k == SG_KERNEL_SPLIT;
    if (s && !can_stream) throw Fail{SG_ERR_INVALID, "no streamed kernel for this L"};
    if (sp && !can_split) throw Fail{SG_ERR_INVALID, "no split kernel for this L and size"};
    if (t < 0 || t >= shape_tiles(k)) throw Fail{SG_ERR_INVALID, "no such tile for this shape"};
    if (s == streamed && one == single_shape && sp == split_shape && t == tile) return;
    streamed = s;
    single_shape = one;
    split_shape = sp;
    tile = t;
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  void build_single_tasks() {
    h_single.clear();
    single_off.assign(G + 1, 0);
    for (int g = 0; g < G; ++g) {
      single_off[g] = (int)h_single.size();
      for (int rho = 0; rho < R; ++rho)
        for (int t = 0; t < T; ++t)
          h_single.push_back(TaskDesc{{(uint32_t)lat_index(g, rho, t), 0xFFFFFFFFu}, kNoPair, 0});
    }
    single_off[G] = (int)h_single.size();
    if (form == SG_FORM_BATCHED) {   // each coupling group on one XCD
      const size_t per_group = (size_t)R * T;
      xcd_interleave(h_single, [per_group](const TaskDesc &k) {
        return (uint32_t)(k.lat[0] / per_group);
      });
    }
    dev_alloc(&d_single, h_single.size());
    HIPCHK(hipMemcpy(d_single, h_single.data(), sizeof(TaskDesc) * h_single.size(),
                     hipMemcpyHostToDevice));
  }

  void build_split() {
    std::vector<TaskDesc> tv(n_lat());
    for (size_t i = 0; i < n_lat(); ++i) tv[i] = TaskDesc{{(uint32_t)i, 0xFFFFFFFFu}, kNoPair, 0};
    dev_alloc(&d_split, tv.size());
    HIPCHK(hipMemcpy(d_split, tv.data(), sizeof(TaskDesc) * tv.size(), hipMemcpyHostToDevice));
    const size_t PL = (size_t)L * L / 2;
    dev_alloc(&d_xbuf, n_lat() * split_P * 8 * PL);
    dev_alloc(&d_xflag, n_lat() * split_P);
    dev_alloc(&d_xerr, 1);
    dev_alloc(&d_ticket, kSplitTicketWords);
    HIPCHK(hipMemset(d_xflag, 0, sizeof(uint32_t) * n_lat() * split_P));
    HIPCHK(hipMemset(d_xerr, 0, sizeof(uint32_t)));
    HIPCHK(hipMemset(d_ticket, 0, sizeof(uint32_t) * kSplitTicketWords));
  }

  // the split launches' phase flags hold absolute phases 2 sweep + c + 1;
  // wherever the sweep counter may move back (init, load, an aborted launch)
  // they are cleared at once, in stream order before any later launch (a
  // captured interval then finds them cleared too)
  void clear_split_flags() {
    if (!can_split) return;
    HIPCHK(launch_fill_u32(d_xflag, 0u, (int)(n_lat() * split_P), st));
  }

  // the error word of the split launches: a halo wait that timed out
  void check_split_error() {
    if (!can_split) return;
    uint32_t e = 0;
    HIPCHK(hipMemcpy(&e, d_xerr, sizeof(e), hipMemcpyDeviceToHost));
    if (e) {
      HIPCHK(hipMemset(d_xerr, 0, sizeof(uint32_t)));
      HIPCHK(hipMemset(d_ticket, 0, sizeof(uint32_t) * kSplitTicketWords));
      clear_split_flags();
      throw Fail{SG_ERR_DEVICE, "split sweeps: a boundary-plane wait timed out (results invalid)"};
    }
  }

  void build_stream_tasks() {
    const int nc = phase_chunks(L);
    h_chunks.clear();
    chunk_off.assign(G + 1, 0);
    const bool per_group = form == SG_FORM_PLAIN;
    std::vector<TaskDesc> run;
    for (int g = 0; g < G; ++g) {
      if (per_group) chunk_off[g] = (int)h_chunks.size();
      for (int rho = 0; rho < R; ++rho)
        for (int t = 0; t < T; ++t)
          for (int k = 0; k < nc; ++k)
            run.push_back(TaskDesc{{(uint32_t)lat_index(g, rho, t), (uint32_t)k}, kNoPair, 0});
      if (per_group || g == G - 1) {
        const std::vector<TaskDesc> o = xcd_order(run, (uint32_t)nc);
        h_chunks.insert(h_chunks.end(), o.begin(), o.end());
        run.clear();
      }
    }
    chunk_off[G] = (int)h_chunks.size();
    dev_alloc(&d_chunks, h_chunks.size());
    HIPCHK(hipMemcpy(d_chunks, h_chunks.data(), sizeof(TaskDesc) * h_chunks.size(),
                     hipMemcpyHostToDevice));
  }

  // (pair, chunk) tasks of the PT launches of parity 0 and 1 (streamed and
  // one-lattice shapes), ordered by group (ptc_off)
  void build_ptc_tasks() {
    const int np = pt_chunks(L);
    for (int par = 0; par < 2; ++par) {
      h_ptc[par].clear();
      ptc_off[par].assign(G + 1, 0);
      for (int g = 0; g < G; ++g) {
        ptc_off[par][g] = (int)h_ptc[par].size();
        for (int rho = 0; rho < R; ++rho)
          for (int t = par; t + 1 < T; t += 2)
            for (int k = 0; k < np; ++k)
              h_ptc[par].push_back(TaskDesc{
                  {(uint32_t)(((size_t)g * R + rho) * (T - 1) + t), (uint32_t)k}, kNoPair, 0});
      }
      ptc_off[par][G] = (int)h_ptc[par].size();
      dev_alloc(&d_ptc[par], h_ptc[par].size());
      if (!h_ptc[par].empty())
        HIPCHK(hipMemcpy(d_ptc[par], h_ptc[par].data(), sizeof(TaskDesc) * h_ptc[par].size(),
                         hipMemcpyHostToDevice));
    }
  }

  void build_slab_tasks() {
    h_slabs.clear();
    slab_off.assign(G + 1, 0);
    const bool per_group = form == SG_FORM_PLAIN;
    std::vector<TaskDesc> run;
    for (int g = 0; g < G; ++g) {
      if (per_group) slab_off[g] = (int)h_slabs.size();
      for (int rho = 0; rho < R; ++rho)
        for (int t = 0; t < T; ++t)
          for (int k = 0; k < kSlabParts; ++k)
            run.push_back(TaskDesc{{(uint32_t)lat_index(g, rho, t), (uint32_t)k}, kNoPair, 0});
      if (per_group || g == G - 1) {
        const std::vector<TaskDesc> o = xcd_order(run, (uint32_t)kSlabParts);
        h_slabs.insert(h_slabs.end(), o.begin(), o.end());
        run.clear();
      }
    }
    slab_off[G] = (int)h_slabs.size();
    dev_alloc(&d_slabs, h_slabs.size());
    HIPCHK(hipMemcpy(d_slabs, h_slabs.data(), sizeof(TaskDesc) * h_slabs.size(),
                     hipMemcpyHostToDevice));
  }

  // coupling site records from the [3][N] couplings (after every change of J)
  void build_jsite() {
    for (int g = 0; g < G; ++g)
      HIPCHK(launch_build_jsite(L, d_Js, d_J, (uint64_t)g * 3 * N, st));
  }

  void ensure_ring() {
    if (h_ring[0]) return;
    for (int i = 0; i < kRing; ++i) {
      hipError_t e = hipHostMalloc((void **)&h_ring[i], sizeof(int32_t) * rec_count);
      if (e != hipSuccess) {
        h_ring[i] = nullptr;
        throw Fail{SG_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e)};
      }
      HIPCHK(hipEventCreateWithFlags(&ring_ev[i], hipEventDisableTiming));
    }
  }

  void build_thresholds(double b, ThrRec &r) const { build_threshold_rec(b, H, r); }

  void reset_counters() {
    JobState js;
    std::memset(&js, 0, sizeof(js));
    js.sweep = (uint32_t)sweep;
    js.pt_every = (uint32_t)(T > 1 ? pt_every : 0);
    HIPCHK(hipMemcpy(d_jobs, &js, sizeof(js), hipMemcpyHostToDevice));
  }

  // ------------------------------------------------------------ lattice
  void init_lattice() {
    for (int g = 0; g < G; ++g)
      HIPCHK(launch_init_couplings(L, d_J, (uint64_t)g * 3 * N, (uint32_t)g, (uint32_t)cseed,
                                   (uint32_t)(cseed >> 32), st));
    build_jsite();
    HIPCHK(launch_init_spins(L, d_spins, d_lat, (int)n_lat(), st));
    sweep = 0;
    n_meas = 0;
    HIPCHK(hipMemsetAsync(d_acc, 0, sizeof(ObsAcc) * (size_t)G * T * 64, st));
    HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * std::max(T - 1, 1), st));
    std::fill(swap_att.begin(), swap_att.end(), 0);
    reset_counters();
    energy_valid = energy_k_valid = false;
    initialized = true;
    clear_split_flags();
  }

  // natural <-> checkerboard-split order
  int cb(int s) const {
    const int x = s % L, y = (s / L) % L, z = s / (L * L);
    return ((x + y + z) & 1) * (N / 2) + (x >> 1) + (L / 2) * (y + L * z);
  }

  void get_spins(uint64_t *out) {
    std::vector<uint64_t> raw(n_lat() * N);
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(raw.data(), d_spins, sizeof(uint64_t) * raw.size(), hipMemcpyDeviceToHost));
    std::vector<int> perm(N);
    for (int s = 0; s < N; ++s) perm[s] = cb(s);
    for (size_t l = 0; l < n_lat(); ++l)
      for (int s = 0; s < N; ++s) out[l * N + s] = raw[l * N + perm[s]];
  }

  void set_spins(const uint64_t *in) {
    std::vector<uint64_t> raw(n_lat() * N);
    std::vector<int> perm(N);
    for (int s = 0; s < N; ++s) perm[s] = cb(s);
    for (size_t l = 0; l < n_lat(); ++l)
      for (int s = 0; s < N; ++s) raw[l * N + perm[s]] = in[l * N + s];
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(d_spins, raw.data(), sizeof(uint64_t) * raw.size(), hipMemcpyHostToDevice));
    energy_valid = energy_k_valid = false;
  }

  // ------------------------------------------------------------ run loop
  int energy_mode() const { return field ? 2 : 1; }   // K enters PT only when H != 0

  // One sweep-kernel launch over the tasks of parity `par` (all groups in the
  // batched form, one launch per group in the plain form), then the device
  // sweep counter advance.  do_pt: the launch starts at a PT step.
  void launch_interval(int par, int do_pt, int seg, int want_energy) {
    if (split_shape && seg > 0) {
      launch_interval_split(par, do_pt, seg, want_energy);
      return;
    }
    if (streamed || split_shape) {
      launch_interval_streamed(par, do_pt, seg, want_energy);
      return;
    }
    const bool per_group = form == SG_FORM_PLAIN;
    const bool single = single_kernel_supported(L);
    // one lattice per work-group: PT in its own launch first
    const bool own_pt = single || single_shape;
    const TaskDesc *tasks = single_shape ? d_single : d_tasks[par];
    const std::vector<int> &off = single_shape ? single_off : task_off[par];
    for (int g = 0; g < (per_group ? G : 1); ++g) {
      const int a = per_group ? off[g] : 0;
      const int b = per_group ? off[g + 1] : off[G];
      const double fl = (double)seg * N * 64.0 * R * T * (per_group ? 1 : G);
      if (own_pt && do_pt) {
        const int pa = per_group ? ptc_off[par][g] : 0;
        const int pb = per_group ? ptc_off[par][g + 1] : ptc_off[par][G];
        HIPCHK(launch_pt_chunks(d_self, d_ptc[par] + pa, pb - pa, st));
      }
      timed(0, fl, [&] {
        if (single)
          HIPCHK(launch_single(L, field, d_self, tasks + a, b - a, seg, want_energy, st));
        else
          HIPCHK(launch_block(L, field, d_self, tasks + a, b - a, own_pt ? 0 : do_pt, seg,
                              want_energy, st));
      });
    }
    if (seg > 0) HIPCHK(launch_advance(d_self, 1, seg, st));
  }

  void launch_energy_all(int mode) {
    if (streamed || split_shape) {
      launch_interval_streamed(0, 0, 0, mode);
      return;
    }
    if (single_kernel_supported(L))
      HIPCHK(launch_single(L, field, d_self, d_tasks[0], task_off[0][G], 0, mode, st));
    else
      HIPCHK(launch_block(L, field, d_self, d_tasks[0], task_off[0][G], 0, 0, mode, st));
  }

  // Split interval: the PT launch over the (pair, chunk) tasks, then one split
  // launch (all lattices of the launch, P work-groups each) for the seg sweeps
  // and the energies.
  void launch_interval_split(int par, int do_pt, int seg, int want_energy) {
    const bool per_group = form == SG_FORM_PLAIN;
    const int per = R * T;
    for (int g = 0; g < (per_group ? G : 1); ++g) {
      const int pa = per_group ? ptc_off[par][g] : 0;
      const int pb = per_group ? ptc_off[par][g + 1] : ptc_off[par][G];
      if (do_pt) HIPCHK(launch_pt_chunks(d_self, d_ptc[par] + pa, pb - pa, st));
      const int a = per_group ? g * per : 0, n = per_group ? per : (int)n_lat();
      // at most split_batch(P) lattices a launch (all their work-groups on
      // the CUs at once); more go out as consecutive launches
      for (int b = 0; b < n; b += split_batch(split_P)) {
        const int nb = std::min(n - b, split_batch(split_P));
        const double fl = (double)seg * N * 64.0 * nb;
        timed(0, fl, [&] {
          HIPCHK(launch_split(L, field, d_self, d_split + a + b, nb, split_P, d_ticket, seg,
                              want_energy, tile, st));
        });
      }
    }
    HIPCHK(launch_advance(d_self, 1, seg, st));
  }

  // Streamed interval: the PT launch, then one launch per colour and sweep;
  // the first zeroes the energy slots, the last accumulates the energies of
  // the final states.  seg = 0 with want_energy: energies only.
  void launch_interval_streamed(int par, int do_pt, int seg, int want_energy) {
    const bool per_group = form == SG_FORM_PLAIN;
    const int ptile = streamed ? tile : 0;   // the split shape's energy passes: tile 0
    const int emode = want_energy ? (kPhEnergy | (want_energy > 1 ? kPhEnergyK : 0)) : 0;
    for (int g = 0; g < (per_group ? G : 1); ++g) {
      const int ca = per_group ? chunk_off[g] : 0, cb = per_group ? chunk_off[g + 1] : chunk_off[G];
      const int pa = per_group ? ptc_off[par][g] : 0;
      const int pb = per_group ? ptc_off[par][g + 1] : ptc_off[par][G];
      if (do_pt) HIPCHK(launch_pt_chunks(d_self, d_ptc[par] + pa, pb - pa, st));
      if (seg == 0) {
        if (want_energy) {
          HIPCHK(launch_phase(L, field, d_self, d_chunks + ca, cb - ca, 0, 0, kPhZero, ptile, st));
          HIPCHK(launch_phase(L, field, d_self, d_chunks + ca, cb - ca, 1, 0, emode, ptile, st));
        }
        continue;
      }
      const double fl = (double)seg * N * 64.0 * R * T * (per_group ? 1 : G);
      // pairs of sweeps by the slab kernel (buffer 1 and back), a last odd
      // sweep by the colour launches
      const int ns = d_slabs ? seg & ~1 : 0;
      const int sa = per_group && d_slabs ? slab_off[g] : 0;
      const int sb = d_slabs ? (per_group ? slab_off[g + 1] : slab_off[G]) : 0;
      timed(0, fl, [&] {
        for (int it = 0; it < ns; ++it) {
          int mode = kPhUpdate;
          if (it == 0 && want_energy && ns == seg) mode |= kPhZero;
          if (it == seg - 1) mode |= emode;
          HIPCHK(launch_slab(L, field, d_self, d_slabs + sa, sb - sa, it & 1, it, mode, ptile, st));
        }
        for (int it = ns; it < seg; ++it)
          for (int c = 0; c < 2; ++c) {
            int mode = kPhUpdate;
            if (it == ns && c == 0 && want_energy) mode |= kPhZero;
            if (it == seg - 1 && c == 1) mode |= emode;
            HIPCHK(launch_phase(L, field, d_self, d_chunks + ca, cb - ca, c, it, mode, ptile, st));
          }
      });
    }
    if (seg > 0) HIPCHK(launch_advance(d_self, 1, seg, st));
  }

  // Batched form: `count` consecutive full PT intervals (parities par,
  // 1 - par, ...) are one hipGraph per (parity, length, energy, count),
  // replayed while the device sweep counter advances; fewer graph launches,
  // fewer gaps between them.  graph_for captures and instantiates one on
  // first use (nothing runs during a capture).
  hipGraphExec_t graph_for(int par, int seg, int want_energy, int count) {
    for (auto &k : graphs)
      if (k.parity == par && k.seg == seg && k.energy == want_energy && k.count == count)
        return k.exec;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      for (int i = 0; i < count; ++i) launch_interval((par + i) & 1, 1, seg, want_energy);
    } catch (...) {
      // leave the stream out of capture mode before reporting
      if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st, &g));
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    graphs.push_back(GraphKey{par, seg, want_energy, count, exec});
    return exec;
  }

  // full PT intervals from sweep n that one graph may replay: up to 4, all
  // before the next measurement and within the remaining sweeps
  int graph_run_length(int64_t n, int64_t remaining) const {
    int64_t room = remaining / pt_every;
    if (meas_every > 0) {
      const int64_t m = n - n % meas_every + meas_every;   // next measurement sweep
      room = std::min<int64_t>(room, (m - n) / pt_every);
    }
    return room >= 4 ? 4 : room >= 2 ? 2 : 1;
  }

  // One step of the run loop from sweep n with `remaining` sweeps to go:
  // seg sweeps per interval (up to the next PT step or measurement), count
  // intervals (several only when replayed from one graph), PT at the start
  // of each, and the energy mode of the last: U (and K at H != 0) for the
  // next PT step, U and K when a measurement follows.
  struct StepPlan {
    int64_t seg;
    int par, count, emode;
    bool pt_now, graphed;
  };
  StepPlan plan_step(int64_t n, int64_t remaining) const {
    const bool pt = pt_every > 0 && T > 1;
    StepPlan s{};
    s.seg = remaining;
    if (pt) s.seg = std::min<int64_t>(s.seg, pt_every - n % pt_every);
    if (meas_every > 0) s.seg = std::min<int64_t>(s.seg, meas_every - n % meas_every);
    const bool want_energy = pt && (n + s.seg) % pt_every == 0;
    s.pt_now = pt && n % pt_every == 0;
    s.par = s.pt_now ? (int)((n / pt_every) & 1) : 0;
    s.count = 1;
    s.graphed = form == SG_FORM_BATCHED && s.pt_now && s.seg == pt_every && !profiling;
    if (s.graphed) s.count = graph_run_length(n, remaining);
    const bool meas_end = meas_every > 0 && (n + s.count * s.seg) % meas_every == 0;
    s.emode = meas_end ? 2 : want_energy ? energy_mode() : 0;
    return s;
  }

  // Instantiates every graph that a call run(n_sweeps) starting at sweep
  // `from` (the current sweep when negative) would replay, and runs nothing:
  // a caller that times that call afterwards times no capture or
  // instantiation.
  void prepare(int64_t from, int64_t n_sweeps) {
    if (!initialized) throw Fail{SG_ERR_STATE, "lattice not initialised"};
    if (n_sweeps < 0) throw Fail{SG_ERR_INVALID, "negative sweep count"};
    int64_t n = from < 0 ? sweep : from, remaining = n_sweeps;
    while (remaining > 0) {
      const StepPlan s = plan_step(n, remaining);
      if (s.graphed) graph_for(s.par, (int)s.seg, s.emode, s.count);
      n += s.count * s.seg;
      remaining -= s.count * s.seg;
    }
  }

  void run(int64_t n_sweeps, sg_obs_sink sink, void *user) {
    if (!initialized) throw Fail{SG_ERR_STATE, "lattice not initialised"};
    if (n_sweeps < 0) throw Fail{SG_ERR_INVALID, "negative sweep count"};
    if (sweep + n_sweeps > 0xFFFFFFFFll) throw Fail{SG_ERR_INVALID, "sweep counter overflow"};
    HIPCHK(hipEventRecord(ev0, st));
    int64_t remaining = n_sweeps;
    while (remaining > 0) {
      const int64_t n = sweep;
      const StepPlan s = plan_step(n, remaining);
      if (s.pt_now && !energy_valid) launch_energy_all(energy_mode());
      if (s.graphed)
        HIPCHK(hipGraphLaunch(graph_for(s.par, (int)s.seg, s.emode, s.count), st));
      else
        launch_interval(s.par, s.pt_now ? 1 : 0, (int)s.seg, s.emode);
      for (int i = 0; i < s.count; ++i) {
        if (s.pt_now)
          for (int t = (s.par + i) & 1; t + 1 < T; t += 2) swap_att[t] += (uint64_t)G * R * 64;
        flips += (double)s.seg * N * 64.0 * G * R * T;
      }
      sweep = n + s.count * s.seg;
      energy_valid = s.emode > 0;
      energy_k_valid = s.emode >= 2;
      remaining -= s.count * s.seg;
      // records of earlier measurements whose copies have landed
      if (!rec_queue.empty()) drain_records(sink, user, 0);
      if (meas_every > 0 && sweep % meas_every == 0) measure(sink, user);
    }
    drain_records(sink, user, -1);
    HIPCHK(hipEventRecord(ev1, st));
    HIPCHK(hipEventSynchronize(ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    kernel_ms += ms;
    collect_profile();
    if (split_shape) check_split_error();
  }

  // The measurement folds itself into the device sums (ObsAcc); records leave
  // the device only when a sink wants them.
  void measure(sg_obs_sink sink, void *user) {
    const bool want = sink != nullptr;
    if (want && !d_rec) throw Fail{SG_ERR_STATE, "this handle delivers no records"};
    // U and K of the current states come from the energy slots (the sweep
    // launch ending here fills them when it knows a measurement follows)
    if (!energy_k_valid) {
      launch_energy_all(2);
      energy_valid = energy_k_valid = true;
    }
    timed(2, 0.0, [&] {
      HIPCHK(launch_measure(L, R, d_spins, d_energy, d_meas, G * T, d_rec, nfield, d_acc, d_cs,
                            want ? 1 : 0, st));
    });
    ++n_meas;
    if (!want) return;
    ensure_ring();
    // a ring slot is reused only after its sink call has run
    if ((int)rec_queue.size() == kRing) drain_records(sink, user, 1);
    const int slot = ring_next;
    ring_next = (ring_next + 1) % kRing;
    HIPCHK(hipMemcpyAsync(h_ring[slot], d_rec, sizeof(int32_t) * rec_count,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ring_ev[slot], st));
    rec_queue.push_back(PendingRec{slot, sweep});
  }

  // Runs the sink on copied records in measurement order: all of them with
  // at_least < 0, otherwise at least that many plus any already landed.
  void drain_records(sg_obs_sink sink, void *user, int at_least) {
    size_t i = 0;
    for (; i < rec_queue.size(); ++i) {
      const PendingRec &p = rec_queue[i];
      const bool must = at_least < 0 || (int)i < at_least;
      if (!must) {
        const hipError_t q = hipEventQuery(ring_ev[p.slot]);
        if (q == hipErrorNotReady) break;
        HIPCHK(q);
      } else {
        HIPCHK(hipEventSynchronize(ring_ev[p.slot]));
      }
      if (sink) sink(user, p.sweep, h_ring[p.slot], (int64_t)G * T * 64, nfield);
    }
    rec_queue.erase(rec_queue.begin(), rec_queue.begin() + (long)i);
  }

  void read_acc(std::vector<ObsAcc> &a) {
    a.resize((size_t)G * T * 64);
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(a.data(), d_acc, sizeof(ObsAcc) * a.size(), hipMemcpyDeviceToHost));
  }

  // Per temperature: exact integer totals over all measurements, samples and
  // replicas (128-bit; Q^4 sums need them at L = 48), each turned into a mean
  // by one division, so the result does not depend on the order in which
  // samples are reduced; the one FP64 observable (the k_min overlap, summed
  // per lane on the device in measurement order) is added over samples in
  // ascending order.  analyze-all.py:38-71 averages per sample first; the two
  // agree up to rounding because every sample has the same measurement count.
  void results(double *res) {
    std::vector<ObsAcc> acc;
    read_acc(acc);
    results_from(acc.data(), res);
  }

  void results_from(const ObsAcc *acc, double *res) const {
#pragma clang fp contract(off)
    const double Nd = N;
    const double nS = (double)n_meas * S, nRS = (double)n_meas * R * S;
    for (int t = 0; t < T; ++t) {
      __int128 sU = 0, sK = 0, sAbsQ = 0;
      unsigned __int128 sQ2 = 0, sQ4 = 0;
      double sq1 = 0.0;
      for (int s = 0; s < S; ++s) {
        const ObsAcc &a = acc[((size_t)(s / 64) * T + t) * 64 + (s % 64)];
        sU += a.sU;
        sK += a.sK;
        sAbsQ += a.sAbsQ;
        sQ2 += a.sQ2;
        sQ4 += ((unsigned __int128)a.sQ4hi << 64) | a.sQ4lo;
        sq1 += a.sq1;
      }
      double *r = res + (size_t)t * SG_NRES;
      for (int k = 0; k < SG_NRES; ++k) r[k] = 0.0;
      r[SG_RES_BETA] = beta[t];
      r[SG_RES_NMEAS] = (double)n_meas;
      r[SG_RES_BINDER] = r[SG_RES_XI_L] = NAN;
      if (n_meas == 0) continue;
      const double U = (double)sU / nRS, K = (double)sK / nRS;
      const double M = 2.0 * K - Nd;
      const double q2 = (double)sQ2 / nS / (Nd * Nd);
      const double q4 = (double)sQ4 / nS / (Nd * Nd * Nd * Nd);
      const double c1 = sq1 / nS;
      r[SG_RES_E] = ((2.0 * U - 3.0 * Nd) - H * M) / Nd;
      r[SG_RES_M] = M / Nd;
      r[SG_RES_Q] = (double)sAbsQ / nS / Nd;
      r[SG_RES_Q2] = q2;
      r[SG_RES_Q4] = q4;
      r[SG_RES_BINDER] = q2 > 0 ? 0.5 * (3.0 - q4 / (q2 * q2)) : NAN;
      r[SG_RES_CHI0] = q2;
      r[SG_RES_CHI1] = c1;
      r[SG_RES_XI_L] = (c1 > 0 && q2 / c1 > 1.0)
                           ? std::sqrt(q2 / c1 - 1.0) / (2.0 * std::sin(M_PI / L)) / L
                           : NAN;
    }
  }

  // ------------------------------------------------------------ checkpoint
  struct CkHeader {
    char magic[8];
    uint32_t version, L, G, R, T, S, pt_every, meas_every;
    double H;
    uint64_t cseed, sseed;
    int64_t sweep, n_meas;
    uint64_t payload_bytes;
  };

  size_t ck_size() const {
    return sizeof(CkHeader) + sizeof(double) * T + sizeof(uint64_t) * n_lat() * N +
           sizeof(uint64_t) * (size_t)G * 3 * N + sizeof(uint64_t) * 2 * std::max(T - 1, 1) +
           sizeof(ObsAcc) * (size_t)G * T * 64 + sizeof(uint64_t);
  }

  void save(void *buf, size_t *size) {
    const size_t need = ck_size();
    if (!buf) { *size = need; return; }
    if (*size < need) throw Fail{SG_ERR_INVALID, "checkpoint buffer too small"};
    if (!initialized) throw Fail{SG_ERR_STATE, "lattice not initialised"};
    char *p = (char *)buf;
    CkHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "SGCHKPT1", 8);
    h.version = 2; h.L = L; h.G = G; h.R = R; h.T = T; h.S = S;
    h.pt_every = pt_every; h.meas_every = meas_every; h.H = H; h.cseed = cseed; h.sseed = sseed;
    h.sweep = sweep; h.n_meas = n_meas;
    h.payload_bytes = need - sizeof(CkHeader) - sizeof(uint64_t);
    std::memcpy(p, &h, sizeof(h)); p += sizeof(h);
    std::memcpy(p, beta.data(), sizeof(double) * T); p += sizeof(double) * T;
    get_spins((uint64_t *)p); p += sizeof(uint64_t) * n_lat() * N;
    HIPCHK(hipMemcpy(p, d_J, sizeof(uint64_t) * (size_t)G * 3 * N, hipMemcpyDeviceToHost));
    p += sizeof(uint64_t) * (size_t)G * 3 * N;
    std::vector<uint64_t> cnt(std::max(T - 1, 1));
    HIPCHK(hipMemcpy(cnt.data(), d_cnt, sizeof(uint64_t) * cnt.size(), hipMemcpyDeviceToHost));
    std::memcpy(p, cnt.data(), sizeof(uint64_t) * cnt.size()); p += sizeof(uint64_t) * cnt.size();
    std::memcpy(p, swap_att.data(), sizeof(uint64_t) * cnt.size()); p += sizeof(uint64_t) * cnt.size();
    std::vector<ObsAcc> acc;
    read_acc(acc);
    std::memcpy(p, acc.data(), sizeof(ObsAcc) * acc.size()); p += sizeof(ObsAcc) * acc.size();
    const uint64_t ck = fnv1a(buf, need - sizeof(uint64_t));
    std::memcpy(p, &ck, sizeof(ck));
    *size = need;
  }

  void load(const void *buf, size_t size) {
    const size_t need = ck_size();
    if (size != need) throw Fail{SG_ERR_FORMAT, "checkpoint size mismatch"};
    const char *p = (const char *)buf;
    CkHeader h;
    std::memcpy(&h, p, sizeof(h));
    if (std::memcmp(h.magic, "SGCHKPT1", 8) != 0 || h.version != 2)
      throw Fail{SG_ERR_FORMAT, "checkpoint magic/version mismatch"};
    uint64_t ck;
    std::memcpy(&ck, p + need - sizeof(uint64_t), sizeof(ck));
    if (ck != fnv1a(buf, need - sizeof(uint64_t)))
      throw Fail{SG_ERR_FORMAT, "checkpoint checksum mismatch"};
    if ((int)h.L != L || (int)h.G != G || (int)h.R != R || (int)h.T != T || (int)h.S != S ||
        (int)h.pt_every != pt_every || (int)h.meas_every != meas_every || h.H != H ||
        h.cseed != cseed || h.sseed != sseed)
      throw Fail{SG_ERR_FORMAT, "checkpoint parameters differ from this handle"};
    p += sizeof(h);
    if (std::memcmp(p, beta.data(), sizeof(double) * T) != 0)
      throw Fail{SG_ERR_FORMAT, "checkpoint ladder differs"};
    p += sizeof(double) * T;
    initialized = true;
    set_spins((const uint64_t *)p); p += sizeof(uint64_t) * n_lat() * N;
    HIPCHK(hipMemcpy(d_J, p, sizeof(uint64_t) * (size_t)G * 3 * N, hipMemcpyHostToDevice));
    build_jsite();
    p += sizeof(uint64_t) * (size_t)G * 3 * N;
    const size_t nc = std::max(T - 1, 1);
    HIPCHK(hipMemcpy(d_cnt, p, sizeof(uint64_t) * nc, hipMemcpyHostToDevice)); p += sizeof(uint64_t) * nc;
    std::memcpy(swap_att.data(), p, sizeof(uint64_t) * nc); p += sizeof(uint64_t) * nc;
    HIPCHK(hipMemcpy(d_acc, p, sizeof(ObsAcc) * (size_t)G * T * 64, hipMemcpyHostToDevice));
    sweep = h.sweep;
    n_meas = h.n_meas;
    reset_counters();
    energy_valid = energy_k_valid = false;
    clear_split_flags();
  }
};

