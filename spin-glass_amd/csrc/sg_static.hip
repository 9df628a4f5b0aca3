// sg_static.hip — sg_run_static (include/sg.h): the PT sweeps of sg_run over
// lattices the caller holds in device memory, allocating nothing.
//
// sg_run keeps its working buffers in the handle that sg_create allocated:
// coupling site records, lattice / pair / task descriptors, threshold tables,
// the sweep counter and the job table.  sg_run_static builds the same
// buffers, for the same sizes, in one file-scope device array (`scratch`, a
// fixed size per device, held until the process ends) and runs the same
// lattice-pair kernels (launch_block) on them, so that its spins, energies
// and swap counts equal sg_run's bit for bit.  Inputs and outputs are the
// caller's: spins (updated in place), couplings, per-lane energies, swap
// counts.  The descriptors are written by kernels from values passed as
// kernel arguments, so no host-to-device copy is needed either.
//
// Every call on a device uses that device's one array: calls that overlap on
// one device race.  Nothing in the array is assumed from an earlier call;
// everything the kernels read is rebuilt, and the swap counters are zeroed
// by a kernel, by each call.
#include "sg_handle.h"

#include <mutex>

namespace {

using namespace sg;
using namespace sg::detail;

// the array (a host-side name only: host code uses the address that
// hipGetSymbolAddress returns for the current device)
__device__ __align__(256) unsigned char scratch[SG_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
unsigned char *g_static_ptr[kMaxDevices] = {};

// device address of `scratch` on `device` (current), looked up once per device
hipError_t static_scratch(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_static_ptr[device] = (unsigned char *)p;
  }
  *out = g_static_ptr[device];
  return hipSuccess;
}

// Byte offsets of the buffers inside the array.
struct Layout {
  size_t Js, lat, pairs, thr, js, jt, tasks[2], end;
  int ntasks[2];
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Pair tasks (t, t + 1) of parity par per (group, replica); the lattices left
// without a partner at that parity, two by two within a group (k_block holds
// the couplings of one group per work-group); one last lone lattice per group
// when their count is odd.  The same work as the handle's task lists, in
// another order (tasks are independent, so the order changes no bit).
struct TaskShape {
  int np[2];          // pair tasks per (group, replica) and parity
  int ns[2];          // lattices without a partner per (group, replica)
  __host__ __device__ int single_tasks(int par, int R) const { return (R * ns[par] + 1) / 2; }
};

TaskShape task_shape(int T) {
  TaskShape s;
  for (int par = 0; par < 2; ++par) {
    s.np[par] = 0;
    for (int t = par; t + 1 < T; t += 2) ++s.np[par];
    s.ns[par] = T - 2 * s.np[par];
  }
  return s;
}

Layout layout(int L, int G, int R, int T) {
  const size_t N = (size_t)L * L * L, nlat = (size_t)G * R * T;
  const size_t npairs = (size_t)G * R * (T > 1 ? T - 1 : 0);
  const TaskShape ts = task_shape(T);
  Layout l;
  size_t o = 0;
  l.Js = o;    o = align256(o + sizeof(uint64_t) * (size_t)G * 6 * N);
  l.lat = o;   o = align256(o + sizeof(LatDesc) * nlat);
  l.pairs = o; o = align256(o + sizeof(PairDesc) * (npairs ? npairs : 1));
  l.thr = o;   o = align256(o + sizeof(ThrRec) * (size_t)T);
  l.js = o;    o = align256(o + sizeof(JobState));
  l.jt = o;    o = align256(o + sizeof(JobPtrs));
  for (int par = 0; par < 2; ++par) {
    l.ntasks[par] = G * (R * ts.np[par] + ts.single_tasks(par, R));
    l.tasks[par] = o;
    o = align256(o + sizeof(TaskDesc) * (size_t)l.ntasks[par]);
  }
  l.end = o;
  return l;
}

// descriptors of the temperatures [t0, t0 + nt), passed by value
constexpr int kChunkT = 16;
struct TempChunk {
  ThrRec thr[kChunkT];
  double beta[kChunkT + 1];   // beta[i] of slot t0 + i, and of t0 + nt when it exists
  int32_t t0, nt;
};

struct Sizes {
  int32_t L, G, R, T, N;
  uint32_t key0, key1;
  double H;
};

// LatDesc of every (group, replica, t in the chunk), PairDesc of every
// (group, replica, t in the chunk with t + 1 < T), ThrRec of the chunk's
// temperatures; the same fields sg_handle::create writes.
__global__ void k_static_descs(LatDesc *lat, PairDesc *pairs, ThrRec *thr, Sizes z,
                               TempChunk c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;   // (g, rho, tt)
  if (i >= z.G * z.R * c.nt) return;
  const int tt = i % c.nt, gr = i / c.nt, rho = gr % z.R, g = gr / z.R;
  const int t = c.t0 + tt;
  const uint32_t li = (uint32_t)(gr * z.T + t);
  LatDesc d = {};
  d.spin_off = (uint64_t)li * z.N;
  d.j_off = (uint64_t)g * 3 * z.N;
  d.key0 = z.key0; d.key1 = z.key1;
  d.g = (uint32_t)g;
  d.rt = ((uint32_t)rho << 16) | (uint32_t)t;
  d.thr_idx = (uint32_t)t;
  d.e_idx = li;
  lat[li] = d;
  if (t + 1 < z.T) {
    PairDesc p = {};
    p.a_off = (uint64_t)li * z.N;
    p.b_off = (uint64_t)(li + 1) * z.N;
    p.beta_a = c.beta[tt];
    p.beta_b = c.beta[tt + 1];
    p.H = z.H;
    p.e_a = li; p.e_b = li + 1;
    p.key0 = z.key0; p.key1 = z.key1;
    p.g = (uint32_t)g;
    p.rt = d.rt;
    p.cnt_idx = (uint32_t)t;
    p.n_words = (uint32_t)z.N;
    p.parity = (uint32_t)(t & 1);
    pairs[(size_t)gr * (z.T - 1) + t] = p;
  }
  if (gr == 0) thr[t] = c.thr[tt];
}

// task lists of both parities, the sweep counter and the job table
__global__ void k_static_tasks(TaskDesc *tasks0, TaskDesc *tasks1, JobState *js, JobPtrs *jt,
                               Sizes z, TaskShape ts, JobPtrs self, uint32_t sweep0,
                               uint32_t pt_every) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0 && blockIdx.y == 0) {
    JobState s = {};
    s.sweep = sweep0;
    s.pt_every = pt_every;
    *js = s;
    *jt = self;
  }
  const int par = blockIdx.y;
  const int np = ts.np[par], ns = ts.ns[par];
  const int per_gr = z.R * np;
  const int nsingle = ts.single_tasks(par, z.R);
  const int npair_tasks = z.G * per_gr;
  if (k >= npair_tasks + z.G * nsingle) return;
  TaskDesc d;
  d.job = 0;
  if (k < npair_tasks) {
    const int g = k / per_gr, rho = (k % per_gr) / np, j = k % np;
    const int t = 2 * j + par;
    const int gr = g * z.R + rho;
    d.lat[0] = (uint32_t)(gr * z.T + t);
    d.lat[1] = d.lat[0] + 1;
    d.pair = (uint32_t)(gr * (z.T - 1) + t);
  } else {
    // lone lattices of group g in (replica, slot) order: slot 0 at odd
    // parity, slot T - 1 when it has no partner
    const int q = k - npair_tasks, g = q / nsingle, m = q % nsingle;
    auto lone = [&](int a) {
      const int rho = a / ns, w = a % ns;
      const int t = (par == 1 && w == 0) ? 0 : z.T - 1;
      return (uint32_t)((g * z.R + rho) * z.T + t);
    };
    d.lat[0] = lone(2 * m);
    d.lat[1] = 2 * m + 1 < z.R * ns ? lone(2 * m + 1) : 0xFFFFFFFFu;
    d.pair = 0xFFFFFFFFu;
  }
  (par ? tasks1 : tasks0)[k] = d;
}

__global__ void k_static_zero(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

extern "C" sg_status sg_run_static(const sg_params *p, uint64_t *spins, const uint64_t *couplings,
                                   int64_t sweep0, int64_t n_sweeps, int32_t *energies,
                                   uint64_t *swaps_accepted, void *stream) {
  g_create_error.clear();
  auto bad = [](const char *m) { g_create_error = m; return SG_ERR_INVALID; };
  if (!p || !spins || !couplings || !energies) return bad("NULL argument");
  if (p->L < 2 || p->L > kMaxL || (p->L & 1)) return bad("L must be even and in [2, 48]");
  if (p->n_samples < 1) return bad("n_samples must be >= 1");
  if (p->n_temps < 1 || p->n_temps > 4096) return bad("n_temps must be in [1, 4096]");
  if (p->n_replicas < 1 || p->n_replicas > 8) return bad("n_replicas must be in [1, 8]");
  if (!(p->H >= 0.0 && p->H <= 4.0)) return bad("H out of range [0, 4]");
  if (p->pt_every < 0) return bad("negative cadence");
  if (p->meas_every != 0) return bad("sg_run_static takes no measurements (meas_every must be 0)");
  if (p->n_temps > 1 && !swaps_accepted) return bad("swaps_accepted is NULL");
  if (sweep0 < 0 || n_sweeps < 0 || sweep0 + n_sweeps > 0xFFFFFFFFll)
    return bad("sweep0 and sweep0 + n_sweeps must be in [0, 2^32)");
  if (!p->betas && !(p->T_min > 0.0 && p->T_max > 0.0)) return bad("need betas or T_min/T_max > 0");
  if (p->device < 0 || p->device >= kMaxDevices) return bad("device ordinal out of range");
  const int L = p->L, N = L * L * L, G = (p->n_samples + 63) / 64, R = p->n_replicas,
            T = p->n_temps;
  const Layout lay = layout(L, G, R, T);
  if (lay.end > (size_t)SG_STATIC_SCRATCH_BYTES) {
    static thread_local char msg[160];
    std::snprintf(msg, sizeof(msg),
                  "sg_run_static needs %zu bytes of scratch for this size; the static array "
                  "holds %u (SG_STATIC_SCRATCH_BYTES)",
                  lay.end, (unsigned)SG_STATIC_SCRATCH_BYTES);
    return bad(msg);
  }
  const bool field = p->H != 0.0;
  const hipStream_t st = (hipStream_t)stream;
  try {
    HIPCHK(hipSetDevice(p->device));
    unsigned char *base = nullptr;
    HIPCHK(static_scratch(p->device, &base));
    uint64_t *Js = (uint64_t *)(base + lay.Js);
    LatDesc *lat = (LatDesc *)(base + lay.lat);
    PairDesc *pairs = (PairDesc *)(base + lay.pairs);
    ThrRec *thr = (ThrRec *)(base + lay.thr);
    JobState *js = (JobState *)(base + lay.js);
    JobPtrs *jt = (JobPtrs *)(base + lay.jt);
    TaskDesc *tasks[2] = {(TaskDesc *)(base + lay.tasks[0]), (TaskDesc *)(base + lay.tasks[1])};

    Sizes z;
    z.L = L; z.G = G; z.R = R; z.T = T; z.N = N;
    z.key0 = (uint32_t)p->spin_seed; z.key1 = (uint32_t)(p->spin_seed >> 32);
    z.H = p->H;
    // the ladder as sg_handle::create computes it, in chunks of temperatures
    auto beta_of = [&](int t) {
      if (p->betas) return p->betas[t];
      const double f = T > 1 ? (double)t / (double)(T - 1) : 0.0;
      return 1.0 / (p->T_min * std::pow(p->T_max / p->T_min, f));
    };
    for (int t0 = 0; t0 < T; t0 += kChunkT) {
      TempChunk c;
      std::memset(&c, 0, sizeof(c));
      c.t0 = t0;
      c.nt = std::min(kChunkT, T - t0);
      for (int i = 0; i <= c.nt && t0 + i < T; ++i) {
        c.beta[i] = beta_of(t0 + i);
        if (!(c.beta[i] > 0.0) || !std::isfinite(c.beta[i]))
          throw Fail{SG_ERR_INVALID, "every beta must be positive and finite"};
        if (i < c.nt) build_threshold_rec(c.beta[i], p->H, c.thr[i]);
      }
      const int n = G * R * c.nt;
      hipLaunchKernelGGL(k_static_descs, dim3((n + 255) / 256), dim3(256), 0, st, lat, pairs,
                         thr, z, c);
      HIPCHK(hipGetLastError());
    }
    JobPtrs self = {};
    self.spins = spins;
    self.J = couplings;
    self.Js = Js;
    self.descs = lat;
    self.pairs = pairs;
    self.thr = thr;
    self.js = js;
    self.energy = energies;
    self.counters = (unsigned long long *)swaps_accepted;
    self.spins2 = nullptr;
    const TaskShape ts = task_shape(T);
    const int nt_max = std::max(lay.ntasks[0], lay.ntasks[1]);
    hipLaunchKernelGGL(k_static_tasks, dim3((std::max(nt_max, 1) + 255) / 256, 2), dim3(256), 0,
                       st, tasks[0], tasks[1], js, jt, z, ts, self, (uint32_t)sweep0,
                       (uint32_t)(T > 1 ? p->pt_every : 0));
    HIPCHK(hipGetLastError());
    for (int g = 0; g < G; ++g)
      HIPCHK(launch_build_jsite(L, Js, couplings, (uint64_t)g * 3 * N, st));
    if (T > 1) {
      hipLaunchKernelGGL(k_static_zero, dim3((T - 1 + 255) / 256), dim3(256), 0, st,
                         (unsigned long long *)swaps_accepted, T - 1);
      HIPCHK(hipGetLastError());
    }

    // U and K of the current states (the first PT step reads them), then
    // intervals as sg_run runs them, then U and K of the final states
    HIPCHK(launch_block(L, field, jt, tasks[0], lay.ntasks[0], 0, 0, 2, st));
    const bool pt = p->pt_every > 0 && T > 1;
    const int emode_pt = field ? 2 : 1;
    int64_t n = sweep0, remaining = n_sweeps;
    while (remaining > 0) {
      int64_t seg = remaining;
      if (pt) seg = std::min<int64_t>(seg, p->pt_every - n % p->pt_every);
      const bool pt_now = pt && n % p->pt_every == 0;
      const int par = pt_now ? (int)((n / p->pt_every) & 1) : 0;
      const bool want_energy = pt && (n + seg) % p->pt_every == 0;
      HIPCHK(launch_block(L, field, jt, tasks[par], lay.ntasks[par], pt_now ? 1 : 0, (int)seg,
                          want_energy ? emode_pt : 0, st));
      HIPCHK(launch_advance(jt, 1, (int)seg, st));
      n += seg;
      remaining -= seg;
    }
    HIPCHK(launch_block(L, field, jt, tasks[0], lay.ntasks[0], 0, 0, 2, st));
  } catch (const Fail &f) {
    g_create_error = f.msg;
    return f.code;
  }
  return SG_OK;
}
