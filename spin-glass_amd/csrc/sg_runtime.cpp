// sg_runtime.cpp — C-ABI of libsg (include/sg.h) over sg_handle.
#include "sg_handle.h"

#include <map>
#include <mutex>
#include <tuple>

using namespace sg;
using namespace sg::detail;

// ---------------------------------------------------------------- autotuner
// Work-group shape and tile of the sweeps for (device, L, field, lattice
// count rounded up to a power of two): lattice pairs, single lattices (where
// the resident kernel holds two), streamed chunks (where the L has a
// streamed kernel; tiles: 1 or 2 quads or columns per thread) and split
// lattices (where the L has a split kernel and the lattices fit on the chip;
// tiles: threads per work-group) run the same PT intervals on a scratch
// handle of that shape (random couplings and spins, 24 temperatures or
// fewer, 2 replicas, PT every 16, the batched form's captured intervals) and
// the fastest (shape, tile) is kept for the life of the process.
namespace {
struct TuneRes {
  int kernel, tile;
  double ms[5];                // by SG_KERNEL_* value: the shape's best tile; 0 where not measured
  double tms[5][kMaxTiles];    // every (shape, tile)
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, bool, size_t>, TuneRes> g_tune;
int64_t g_tune_runs = 0;   // tune_measure calls in this process

TuneRes tune_measure(int device, int L, bool field, size_t nlat) {
  ++g_tune_runs;
  sg_params p;
  std::memset(&p, 0, sizeof(p));
  const int T = (int)std::max<size_t>(1, std::min<size_t>(24, nlat / 2));
  const int G = (int)std::max<size_t>(1, nlat / (2 * (size_t)T));
  p.L = L; p.n_samples = 64 * G; p.n_temps = T; p.n_replicas = 2;
  p.T_min = 0.5; p.T_max = 2.0; p.H = field ? 0.5 : 0.0;
  p.coupling_seed = 0x70E0001; p.spin_seed = 0x70E0002;
  p.pt_every = 16; p.meas_every = 0; p.device = device; p.form = SG_FORM_BATCHED;
  p.kernel = SG_KERNEL_RESIDENT;
  sg_handle h;
  h.create(p);
  h.init_lattice();
  std::vector<int> modes = {SG_KERNEL_RESIDENT};
  if (h.can_single) modes.push_back(SG_KERNEL_RESIDENT_SINGLE);
  if (h.can_stream) modes.push_back(SG_KERNEL_STREAMED);
  if (h.can_split) modes.push_back(SG_KERNEL_SPLIT);
  TuneRes res{SG_KERNEL_RESIDENT, 0, {0, 0, 0, 0, 0}, {}};
  double best = 0;
  for (int m : modes)
    for (int t = 0; t < std::min(h.shape_tiles(m), kMaxTiles); ++t) {
      h.set_mode(m, t);
      h.run(64, nullptr, nullptr);  // warm-up: code, caches, the same captured graphs
      const double before = h.kernel_ms;
      h.run(64, nullptr, nullptr);
      const double ms = h.kernel_ms - before;
      res.tms[m][t] = ms;
      if (res.ms[m] <= 0 || ms < res.ms[m]) res.ms[m] = ms;
      if (best <= 0 || ms < best) {
        best = ms;
        res.kernel = m;
        res.tile = t;
      }
    }
  return res;
}

size_t tune_bucket(size_t n_lattices) {
  size_t bucket = 1;
  while (bucket < n_lattices) bucket <<= 1;
  return bucket;
}

const TuneRes &tune_lookup(int device, int L, bool field, size_t n_lattices) {
  const size_t bucket = tune_bucket(n_lattices);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const auto key = std::make_tuple(device, L, field, bucket);
  auto it = g_tune.find(key);
  if (it == g_tune.end()) it = g_tune.emplace(key, tune_measure(device, L, field, bucket)).first;
  return it->second;   // entries are never erased
}

// The shape the tuner's measurements on the MI355X favour, for a (device, L,
// field, lattice count) it has not measured: two lattices per work-group
// once there are enough of them for about two work-groups per CU (from 512
// lattices at L <= 16, 256 at L = 18, 20), one lattice per work-group below;
// the one-lattice resident kernel at L = 22..26; streamed chunks above, where
// a lattice no longer fits in LDS.
int rule_shape(int L, size_t n_lattices) {
  if (single_kernel_supported(L)) return SG_KERNEL_RESIDENT;
  if (L >= 28 && phase_kernel_supported(L)) return SG_KERNEL_STREAMED;
  const size_t pairs_from = L <= 16 ? 512 : 256;
  return n_lattices >= pairs_from ? SG_KERNEL_RESIDENT : SG_KERNEL_RESIDENT_SINGLE;
}
}  // namespace

namespace sg {
namespace detail {
int tune_kernel_choice(int device, int L, bool field, size_t n_lattices, double *ms, int *tile,
                       double *tms) {
  const TuneRes &r = tune_lookup(device, L, field, n_lattices);
  if (ms) std::copy(r.ms, r.ms + 5, ms);
  if (tms) std::copy(&r.tms[0][0], &r.tms[0][0] + 5 * kMaxTiles, tms);
  if (tile) *tile = r.tile;
  return r.kernel;
}

int cached_kernel_choice(int device, int L, bool field, size_t n_lattices, int *tile) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (tile) *tile = 0;
  auto it = g_tune.find(std::make_tuple(device, L, field, tune_bucket(n_lattices)));
  if (it == g_tune.end()) return rule_shape(L, n_lattices);
  // the engine's classes run concurrently, so they never take the split
  // shape (its work-groups must run together): the best other shape
  const TuneRes &r = it->second;
  int best = r.kernel;
  if (best == SG_KERNEL_SPLIT) {
    best = SG_KERNEL_RESIDENT;
    for (int k = 1; k < 4; ++k)
      if (r.ms[k] > 0 && (r.ms[best] <= 0 || r.ms[k] < r.ms[best])) best = k;
  }
  if (tile) {
    for (int t = 0; t < kMaxTiles; ++t)
      if (r.tms[best][t] > 0 && r.tms[best][t] <= r.tms[best][*tile]) *tile = t;
  }
  return best;
}
}  // namespace detail
}  // namespace sg

// ---------------------------------------------------------------- C-ABI
template <class F>
static sg_status guarded(sg_handle *h, F &&f) {
  if (!h) return SG_ERR_INVALID;
  try {
    f();
    h->err.clear();
    return SG_OK;
  } catch (const Fail &e) {
    h->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    h->err = "host allocation failed";
    return SG_ERR_NOMEM;
  }
}

extern "C" {

void sg_params_default(sg_params *p) {
  static double ref_betas[24];
  for (int i = 0; i < 24; ++i) ref_betas[i] = 2.0 * std::pow(2.0, -4.0 * i / 23.0);
  std::memset(p, 0, sizeof(*p));
  p->L = 16; p->n_samples = 2048; p->n_temps = 24; p->n_replicas = 2;
  p->betas = ref_betas; p->H = 0.0;
  p->coupling_seed = 0x5EED0001ull; p->spin_seed = 0x5EED0002ull;
  p->pt_every = 16; p->meas_every = 1024; p->device = 0; p->form = SG_FORM_BATCHED;
}

sg_status sg_create(const sg_params *p, sg_handle **out) {
  if (!out) return SG_ERR_INVALID;
  *out = nullptr;
  g_create_error.clear();
  auto bad = [](const char *m) { g_create_error = m; return SG_ERR_INVALID; };
  if (!p) return bad("params is NULL");
  if (p->L < 2 || p->L > kMaxL || (p->L & 1)) return bad("L must be even and in [2, 48]");
  if (!sweep_kernel_supported(p->L)) return bad("no sweep kernel for this L");
  if (p->n_samples < 1) return bad("n_samples must be >= 1");
  if (p->n_temps < 1 || p->n_temps > 4096) return bad("n_temps must be in [1, 4096]");
  if (p->n_replicas < 1 || p->n_replicas > 8) return bad("n_replicas must be in [1, 8]");
  if (!(p->H >= 0.0 && p->H <= 4.0)) return bad("H out of range [0, 4]");
  if (p->pt_every < 0 || p->meas_every < 0) return bad("negative cadence");
  if (!p->betas && !(p->T_min > 0.0 && p->T_max > 0.0)) return bad("need betas or T_min/T_max > 0");
  if (p->form != SG_FORM_PLAIN && p->form != SG_FORM_BATCHED) return bad("unknown form");
  if (p->kernel < SG_KERNEL_AUTO || p->kernel > SG_KERNEL_SPLIT) return bad("unknown kernel");
  if (p->tile < 0) return bad("tile must be >= 0");
  if (p->kernel == SG_KERNEL_STREAMED && !phase_kernel_supported(p->L))
    return bad("no streamed kernel for this L");
  if (p->kernel == SG_KERNEL_SPLIT) {
    const int64_t per = (int64_t)p->n_replicas * p->n_temps;
    const int64_t n = p->form == SG_FORM_PLAIN ? per : per * ((p->n_samples + 63) / 64);
    if (n > 4096 || split_slabs(p->L, (int)n) == 0)
      return bad("no split kernel for this L and size (L = 32 or 48)");
  }
  sg_handle *h = new (std::nothrow) sg_handle();
  if (!h) return bad("host allocation failed");
  try {
    h->create(*p);
  } catch (const Fail &e) {
    g_create_error = e.msg;
    delete h;
    return e.code;
  }
  *out = h;
  return SG_OK;
}

sg_status sg_destroy(sg_handle *h) {
  if (!h) return SG_ERR_INVALID;
  delete h;
  return SG_OK;
}

const char *sg_last_error(const sg_handle *h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

sg_status sg_init_lattice(sg_handle *h) {
  return guarded(h, [&] { h->init_lattice(); });
}

sg_status sg_run(sg_handle *h, int64_t n_sweeps, sg_obs_sink sink, void *user) {
  return guarded(h, [&] { h->run(n_sweeps, sink, user); });
}

sg_status sg_prepare(sg_handle *h, int64_t from_sweep, int64_t n_sweeps) {
  return guarded(h, [&] { h->prepare(from_sweep, n_sweeps); });
}

int32_t sg_graph_count(const sg_handle *h) { return h ? (int32_t)h->graphs.size() : -1; }

sg_status sg_save_checkpoint(sg_handle *h, void *buf, size_t *size) {
  if (!size) return SG_ERR_INVALID;
  return guarded(h, [&] { h->save(buf, size); });
}

sg_status sg_load_checkpoint(sg_handle *h, const void *buf, size_t size) {
  if (!buf) return SG_ERR_INVALID;
  return guarded(h, [&] { h->load(buf, size); });
}

int32_t sg_nfield(const sg_handle *h) { return h ? h->nfield : 0; }
int32_t sg_n_groups(const sg_handle *h) { return h ? h->G : 0; }
int64_t sg_sweep_count(const sg_handle *h) { return h ? h->sweep : -1; }

sg_status sg_get_betas(const sg_handle *h, double *betas) {
  if (!h || !betas) return SG_ERR_INVALID;
  std::memcpy(betas, h->beta.data(), sizeof(double) * h->T);
  return SG_OK;
}

sg_status sg_get_spins(sg_handle *h, uint64_t *out) {
  if (!out) return SG_ERR_INVALID;
  return guarded(h, [&] { h->get_spins(out); });
}

sg_status sg_set_spins(sg_handle *h, const uint64_t *in) {
  if (!in) return SG_ERR_INVALID;
  return guarded(h, [&] {
    h->set_spins(in);
    h->initialized = true;
  });
}

sg_status sg_get_couplings(sg_handle *h, uint64_t *out) {
  if (!out) return SG_ERR_INVALID;
  return guarded(h, [&] {
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(out, h->d_J, sizeof(uint64_t) * (size_t)h->G * 3 * h->N,
                     hipMemcpyDeviceToHost));
  });
}

sg_status sg_set_couplings(sg_handle *h, const uint64_t *in) {
  if (!in) return SG_ERR_INVALID;
  return guarded(h, [&] {
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(h->d_J, in, sizeof(uint64_t) * (size_t)h->G * 3 * h->N,
                     hipMemcpyHostToDevice));
    h->build_jsite();
    h->energy_valid = h->energy_k_valid = false;
  });
}

sg_status sg_get_energies(sg_handle *h, int32_t *U, int32_t *K) {
  if (!U || !K) return SG_ERR_INVALID;
  return guarded(h, [&] {
    // always a fresh pass: the slots of the last PT step hold K only at H != 0
    h->launch_energy_all(2);
    h->energy_valid = h->energy_k_valid = true;
    std::vector<int32_t> e(h->n_lat() * 128);
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(e.data(), h->d_energy, sizeof(int32_t) * e.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h->n_lat(); ++i) {
      std::memcpy(U + i * 64, e.data() + i * 128, sizeof(int32_t) * 64);
      std::memcpy(K + i * 64, e.data() + i * 128 + 64, sizeof(int32_t) * 64);
    }
  });
}

sg_status sg_get_swap_counts(sg_handle *h, uint64_t *accepted, uint64_t *attempted) {
  if (!accepted || !attempted) return SG_ERR_INVALID;
  return guarded(h, [&] {
    if (h->T < 2) return;
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(accepted, h->d_cnt, sizeof(uint64_t) * (h->T - 1), hipMemcpyDeviceToHost));
    std::memcpy(attempted, h->swap_att.data(), sizeof(uint64_t) * (h->T - 1));
  });
}

sg_status sg_get_results(sg_handle *h, double *res) {
  if (!res) return SG_ERR_INVALID;
  return guarded(h, [&] { h->results(res); });
}

sg_status sg_get_timing(sg_handle *h, double *kernel_ms, double *flips) {
  if (!kernel_ms || !flips) return SG_ERR_INVALID;
  return guarded(h, [&] {
    *kernel_ms = h->kernel_ms;
    *flips = h->flips;
  });
}

sg_status sg_set_profiling(sg_handle *h, int32_t on) {
  return guarded(h, [&] { h->profiling = on != 0; });
}

sg_status sg_tune_kernel(int32_t device, int32_t L, int32_t n_lattices, int32_t field,
                         int32_t *kernel, double *resident_ms, double *streamed_ms) {
  g_create_error.clear();
  if (!kernel || n_lattices < 1 || L < 2 || L > kMaxL || (L & 1)) return SG_ERR_INVALID;
  try {
    double ms[5];
    *kernel = tune_kernel_choice(device, L, field != 0, (size_t)n_lattices, ms);
    if (resident_ms) *resident_ms = ms[SG_KERNEL_RESIDENT];
    if (streamed_ms) *streamed_ms = ms[SG_KERNEL_STREAMED];
    return SG_OK;
  } catch (const Fail &e) {
    g_create_error = e.msg;
    return e.code;
  }
}

sg_status sg_tune_shapes(int32_t device, int32_t L, int32_t n_lattices, int32_t field,
                         int32_t *kernel, double ms[5]) {
  g_create_error.clear();
  if (!kernel || !ms || n_lattices < 1 || L < 2 || L > kMaxL || (L & 1)) return SG_ERR_INVALID;
  try {
    *kernel = tune_kernel_choice(device, L, field != 0, (size_t)n_lattices, ms);
    return SG_OK;
  } catch (const Fail &e) {
    g_create_error = e.msg;
    return e.code;
  }
}

sg_status sg_tune_tiles(int32_t device, int32_t L, int32_t n_lattices, int32_t field,
                        int32_t *kernel, int32_t *tile, double ms[20]) {
  g_create_error.clear();
  if (!kernel || !tile || !ms || n_lattices < 1 || L < 2 || L > kMaxL || (L & 1))
    return SG_ERR_INVALID;
  try {
    int t = 0;
    *kernel = tune_kernel_choice(device, L, field != 0, (size_t)n_lattices, nullptr, &t, ms);
    *tile = t;
    return SG_OK;
  } catch (const Fail &e) {
    g_create_error = e.msg;
    return e.code;
  }
}

sg_status sg_get_shape(const sg_handle *h, int32_t *kernel, int32_t *tile) {
  if (!h || !kernel || !tile) return SG_ERR_INVALID;
  *kernel = h->split_shape ? SG_KERNEL_SPLIT
            : h->streamed  ? SG_KERNEL_STREAMED
            : h->single_shape ? SG_KERNEL_RESIDENT_SINGLE
                              : SG_KERNEL_RESIDENT;
  *tile = h->tile;
  return SG_OK;
}

int64_t sg_tune_count(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return g_tune_runs;
}

sg_status sg_get_kernel_stats(sg_handle *h, int32_t kind, int64_t *launches, double *total_ms,
                              double *flips) {
  if (!launches || !total_ms || !flips || kind < 0 || kind > 2) return SG_ERR_INVALID;
  return guarded(h, [&] {
    h->collect_profile();
    *launches = h->kstat[kind].launches;
    *total_ms = h->kstat[kind].ms;
    *flips = h->kstat[kind].flips;
  });
}

}  // extern "C"
