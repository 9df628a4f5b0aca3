// sg_engine.cpp — batch and serving engine of libsg (sg_batch_* in
// include/sg.h).
//
// Many simulations ("jobs") share one device.  Their state lives in one
// device arena (sg_arena.h) from admission to retirement; a job is admitted,
// in submission order, once its footprint fits.  Jobs of one (L, field)
// class share a HIP stream, and in the batched form one sweep-kernel launch
// per PT interval covers every task of every running job of the class
// (TaskDesc.job indexes the class's JobPtrs table).  Classes run on separate
// streams, so the device overlaps them, and each is refilled as its own
// queued slices finish, so a class of short intervals is not held to the
// pace of a slower one.  A class's work-group shape comes
// from the tuner's cache (or the rule its measurements support): no step runs
// the tuner.  Because each job's random numbers
// are keyed by its own seeds and counters, the mix never changes a result.
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "sg_handle.h"

namespace {

using Clock = std::chrono::steady_clock;

// Pinned host blocks for the jobs' final copies, kept for reuse: a retired
// job gives its blocks back here rather than to hipHostFree (which
// synchronises the device); sizes are powers of two from 4 KiB.  Freed with
// the batch.
class HostPool {
 public:
  HostPool() = default;
  HostPool(const HostPool &) = delete;
  HostPool &operator=(const HostPool &) = delete;
  ~HostPool() {
    for (auto &b : free_) (void)hipHostFree(b.second);
  }
  static size_t size_class(size_t n) {
    size_t c = 4096;
    while (c < n) c <<= 1;
    return c;
  }
  void *take(size_t n) {
    const size_t c = size_class(n);
    auto it = free_.find(c);
    if (it != free_.end()) {
      void *p = it->second;
      free_.erase(it);
      return p;
    }
    void *p = nullptr;
    HIPCHK(hipHostMalloc(&p, c));
    ++fresh_;
    return p;
  }
  void give(void *p, size_t n) {
    if (p) free_.emplace(size_class(n), p);
  }
  int64_t fresh() const { return fresh_; }   // blocks allocated so far

 private:
  std::multimap<size_t, void *> free_;
  int64_t fresh_ = 0;
};

struct Job {
  int64_t id = -1;
  sg_params p{};
  std::vector<double> betas;
  int64_t n_sweeps = 0;
  int state = SG_JOB_QUEUED;
  bool finishing = false;   // final copies queued, not yet seen complete
  size_t footprint = 0;
  int cls = -1;
  std::unique_ptr<sg_handle> h;
  ObsAcc *h_acc = nullptr;  // pinned copies of the final sums and counters
  uint64_t *h_cnt = nullptr;
  size_t acc_bytes = 0, cnt_bytes = 0;
  HostPool *pool = nullptr;  // where h_acc and h_cnt came from
  hipEvent_t done_ev = nullptr;
  std::vector<double> res;
  std::vector<uint64_t> swaps;
  double t_submit = 0, t_admit = 0, t_done = 0;

  void free_host() {
    if (pool) {
      pool->give(h_acc, acc_bytes);
      pool->give(h_cnt, cnt_bytes);
    }
    if (done_ev) (void)hipEventDestroy(done_ev);
    h_acc = nullptr;
    h_cnt = nullptr;
    done_ev = nullptr;
  }
  ~Job() { free_host(); }
};

// A device array with pinned staging, grown on demand by doubling.  The
// arrays it outgrows go to a graveyard (launches still queued may read them)
// and are freed with their class, so growing never waits for a stream and
// never calls hipFree (which synchronises the device) while jobs run.
struct Grave {
  std::vector<std::pair<void *, void *>> arrays;  // (device, pinned host)
  void bury(void *d, void *h) {
    if (d || h) arrays.emplace_back(d, h);
  }
  void free_all() {
    for (auto &a : arrays) {
      if (a.first) (void)hipFree(a.first);
      if (a.second) (void)hipHostFree(a.second);
    }
    arrays.clear();
  }
};

template <class T>
struct DevList {
  T *d = nullptr, *h = nullptr;
  size_t cap = 0;
  int n = 0;
  void free_all() {
    (void)hipFree(d);
    if (h) (void)hipHostFree(h);
    d = h = nullptr;
    cap = 0;
  }
};

// Jobs of one (L, field, requested kernel) class: their stream, the JobPtrs
// table and the task lists of the class's batched launches.  Every list
// exists twice: [0] every job at its current PT parity, [1] every parity
// flipped (the lists alternate from one PT interval to the next).  With
// SG_KERNEL_AUTO the class picks its work-group shape (lattice pairs, single
// lattices or streamed chunks) whenever its membership changes, from the
// autotuner's cached measurement for its lattice count (sg_tune_shapes before
// serving fills the cache) or, for a count it has not measured, the rule its
// measurements support.
struct Class {
  int L = 0;
  int kernel = SG_KERNEL_AUTO;
  bool field = false, streamed = false;
  bool single = false;          // one lattice per work-group with a two-lattice kernel
  int tile = 0;                 // the streamed kernels' tile (the tuner's cached pick)
  hipStream_t st = nullptr;
  std::vector<Job *> jobs;
  bool dirty = true;
  int64_t steps_since_build = 0;
  DevList<JobPtrs> jt;
  DevList<TaskDesc> tasks[2];   // resident: pair (or one-lattice) tasks
  DevList<TaskDesc> chunks;     // streamed: (lattice, chunk) tasks
  DevList<TaskDesc> slabs;      // streamed with a slab kernel: (lattice, slab pair) tasks
  DevList<TaskDesc> ptc[2];     // streamed: (pair, chunk) tasks of the PT launch
  hipEvent_t up_ev = nullptr;   // the last upload from the staging
  bool uploaded = false;
  std::deque<hipEvent_t> inflight;   // one event after each slice queued on st
  Grave grave;                  // outgrown table arrays
  int64_t grown = 0;            // table growths so far
  // Batched intervals captured as hipGraphs: one per (first parity list,
  // intervals, energy mode of the last interval), valid while the tables
  // above are.  Graphs of earlier tables wait in `old` until the event
  // recorded after their last launch has completed.
  struct Graph { int k, count, emode; hipGraphExec_t exec; };
  std::vector<Graph> graphs;
  std::vector<std::pair<hipEvent_t, std::vector<Graph>>> old;

  void retire_graphs() {
    if (graphs.empty()) return;
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventRecord(e, st));
    old.emplace_back(e, std::move(graphs));
    graphs.clear();
  }
  // frees the retired graphs whose launches have finished (all, with wait)
  void reap_graphs(bool wait) {
    for (size_t i = 0; i < old.size();) {
      if (!wait && hipEventQuery(old[i].first) == hipErrorNotReady) {
        ++i;
        continue;
      }
      (void)hipEventSynchronize(old[i].first);
      (void)hipEventDestroy(old[i].first);
      for (auto &g : old[i].second) (void)hipGraphExecDestroy(g.exec);
      old.erase(old.begin() + (ptrdiff_t)i);
    }
  }

  void free_all() {
    if (st) (void)hipStreamSynchronize(st);
    retire_graphs();
    reap_graphs(true);
    grave.free_all();
    jt.free_all();
    chunks.free_all();
    slabs.free_all();
    for (int k = 0; k < 2; ++k) {
      tasks[k].free_all();
      ptc[k].free_all();
    }
    if (up_ev) (void)hipEventDestroy(up_ev);
    if (st) (void)hipStreamDestroy(st);
  }
};

}  // namespace

struct sg_batch {
  std::string err;
  sg_batch_params bp{};
  DevArena arena;
  HostPool host_pool;   // declared before jobs: they give their blocks back to it
  Clock::time_point t0;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::deque<Job *> queue;            // submitted, not admitted
  std::vector<Class> classes;
  std::vector<hipEvent_t> ev_pool;
  int64_t next_id = 0, steps = 0;
  int64_t iv_direct = 0, iv_graphed = 0, graphs_built = 0;   // batched class intervals
  double flips = 0;

  ~sg_batch() {
    for (auto &c : classes)
      if (c.st) (void)hipStreamSynchronize(c.st);
    jobs.clear();  // job handles free into the arena before it goes
    for (auto &c : classes)
      for (auto e : c.inflight) (void)hipEventDestroy(e);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (auto &c : classes) c.free_all();
  }

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
  }

  void create(const sg_batch_params &p) {
    bp = p;
    t0 = Clock::now();
    HIPCHK(hipSetDevice(bp.device));
    HIPCHK(arena.init((size_t)bp.arena_bytes));
  }

  Class &class_for(int L, bool field, int kernel) {
    for (auto &c : classes)
      if (c.L == L && c.field == field && c.kernel == kernel) return c;
    classes.emplace_back();
    Class &c = classes.back();
    c.L = L;
    c.field = field;
    c.kernel = kernel;
    HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c.up_ev, hipEventDisableTiming));
    return c;
  }

  int class_index(const Class &c) const { return (int)(&c - classes.data()); }

  // ------------------------------------------------------------ submission
  int64_t submit(const sg_params &in, int64_t n_sweeps, double t_submit) {
    const int pe = bp.pt_every;
    if (in.pt_every != 0 && in.pt_every != pe)
      throw Fail{SG_ERR_INVALID, "job pt_every differs from the engine's"};
    if (n_sweeps <= 0 || n_sweeps % pe != 0)
      throw Fail{SG_ERR_INVALID, "n_sweeps must be a positive multiple of pt_every"};
    if (n_sweeps > 0xFFFFFFFFll) throw Fail{SG_ERR_INVALID, "n_sweeps too large"};
    if (in.meas_every < 0 || in.meas_every % pe != 0)
      throw Fail{SG_ERR_INVALID, "meas_every must be a multiple of pt_every"};
    if (in.L < 2 || in.L > kMaxL || (in.L & 1) || !sweep_kernel_supported(in.L))
      throw Fail{SG_ERR_INVALID, "L must be even and in [2, 48]"};
    if (in.n_samples < 1 || in.n_temps < 1 || in.n_temps > 4096 || in.n_replicas < 1 ||
        in.n_replicas > 8)
      throw Fail{SG_ERR_INVALID, "bad sizes"};
    if (!(in.H >= 0.0 && in.H <= 4.0)) throw Fail{SG_ERR_INVALID, "H out of range [0, 4]"};
    if (!in.betas && !(in.T_min > 0.0 && in.T_max > 0.0))
      throw Fail{SG_ERR_INVALID, "need betas or T_min/T_max > 0"};
    auto j = std::make_unique<Job>();
    j->id = next_id++;
    j->p = in;
    j->p.pt_every = pe;
    j->p.device = bp.device;
    j->p.form = SG_FORM_BATCHED;
    if (in.kernel < SG_KERNEL_AUTO || in.kernel > SG_KERNEL_RESIDENT_SINGLE)
      throw Fail{SG_ERR_INVALID, "unknown kernel (the engine takes no split shape)"};
    if (in.kernel == SG_KERNEL_STREAMED && !phase_kernel_supported(in.L))
      throw Fail{SG_ERR_INVALID, "no streamed kernel for this L"};
    if (in.betas) {
      j->betas.assign(in.betas, in.betas + in.n_temps);
      j->p.betas = j->betas.data();
    }
    j->n_sweeps = n_sweeps;
    j->footprint = job_footprint(in.L, in.n_samples, in.n_replicas, in.n_temps);
    if (j->footprint > arena.capacity())
      throw Fail{SG_ERR_NOMEM, "job footprint exceeds the arena"};
    j->t_submit = t_submit;
    Job *raw = j.get();
    jobs[raw->id] = std::move(j);
    queue.push_back(raw);
    return raw->id;
  }

  // ------------------------------------------------------------ admission
  bool try_admit(Job &j) {
    if (arena.capacity() - arena.in_use() < j.footprint) return false;
    Class &c = class_for(j.p.L, j.p.H != 0.0, j.p.kernel);
    auto h = std::make_unique<sg_handle>();
    try {
      h->create(j.p, &arena, c.st);
    } catch (const Fail &f) {
      if (f.code == SG_ERR_NOMEM) return false;  // fragmented: wait for a release
      throw;
    }
    h->init_lattice();
    h->launch_energy_all(h->energy_mode());
    const size_t nc = (size_t)std::max(j.p.n_temps - 1, 1);
    j.acc_bytes = sizeof(ObsAcc) * (size_t)h->G * h->T * 64;
    j.cnt_bytes = sizeof(uint64_t) * nc;
    j.pool = &host_pool;
    j.h_acc = (ObsAcc *)host_pool.take(j.acc_bytes);
    j.h_cnt = (uint64_t *)host_pool.take(j.cnt_bytes);
    HIPCHK(hipEventCreateWithFlags(&j.done_ev, hipEventDisableTiming));
    j.h = std::move(h);
    j.cls = class_index(c);
    j.state = SG_JOB_RUNNING;
    j.t_admit = now();
    c.jobs.push_back(&j);
    c.dirty = true;
    return true;
  }

  void admit() {
    while (!queue.empty()) {
      if (!try_admit(*queue.front())) break;  // strict submission order
      queue.pop_front();
    }
  }

  // ------------------------------------------------------------ retirement
  // A job whose final copies have landed gets its results and gives its
  // device memory back.
  void retire(bool wait) {
    for (auto &kv : jobs) {
      Job &j = *kv.second;
      if (!j.finishing) continue;
      if (wait) {
        HIPCHK(hipEventSynchronize(j.done_ev));
      } else {
        const hipError_t q = hipEventQuery(j.done_ev);
        if (q == hipErrorNotReady) continue;
        HIPCHK(q);
      }
      sg_handle &h = *j.h;
      j.res.assign((size_t)h.T * SG_NRES, 0.0);
      h.results_from(j.h_acc, j.res.data());
      j.swaps.assign(j.h_cnt, j.h_cnt + std::max(h.T - 1, 0));
      j.h.reset();
      j.free_host();
      j.finishing = false;
      j.state = SG_JOB_DONE;
      j.t_done = now();
    }
  }

  void finish(Job &j) {
    sg_handle &h = *j.h;
    HIPCHK(hipMemcpyAsync(j.h_acc, h.d_acc, sizeof(ObsAcc) * (size_t)h.G * h.T * 64,
                          hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipMemcpyAsync(j.h_cnt, h.d_cnt, sizeof(uint64_t) * std::max(h.T - 1, 1),
                          hipMemcpyDeviceToHost, h.st));
    HIPCHK(hipEventRecord(j.done_ev, h.st));
    j.finishing = true;
  }

  // ------------------------------------------------------------ a class's tables
  // Grows a list so it holds n entries.  Queued launches may still read the
  // old device array and a queued upload the old staging, so both go to the
  // class's graveyard rather than being freed.
  template <class T>
  void reserve(Class &c, DevList<T> &l, size_t n) {
    if (n <= l.cap && l.d) return;
    c.grave.bury(l.d, l.h);
    l.d = l.h = nullptr;
    l.cap = std::max<size_t>(2 * n, 64);
    HIPCHK(hipMalloc((void **)&l.d, sizeof(T) * l.cap));
    HIPCHK(hipHostMalloc((void **)&l.h, sizeof(T) * l.cap));
    ++c.grown;
  }

  template <class T>
  void upload(Class &c, DevList<T> &l) {
    if (l.n > 0)
      HIPCHK(hipMemcpyAsync(l.d, l.h, sizeof(T) * l.n, hipMemcpyHostToDevice, c.st));
  }

  // One job's tasks appended with their job index; with pairs_first, the
  // pair tasks of all jobs go first, so the dispatcher's greedy fill ends
  // evenly.
  void gather(Class &c, DevList<TaskDesc> &out, bool pairs_first,
              const std::vector<TaskDesc> &(*pick)(sg_handle &, int), int k) {
    size_t total = 0;
    for (Job *j : c.jobs) total += pick(*j->h, (int)(((j->h->sweep / bp.pt_every) + k) & 1)).size();
    reserve(c, out, total);
    out.n = 0;
    for (int pass = 0; pass < (pairs_first ? 2 : 1); ++pass)
      for (size_t ji = 0; ji < c.jobs.size(); ++ji) {
        sg_handle &h = *c.jobs[ji]->h;
        const int par = (int)(((h.sweep / bp.pt_every) + k) & 1);
        for (TaskDesc t : pick(h, par)) {
          if (pairs_first && ((t.lat[1] != 0xFFFFFFFFu) != (pass == 0))) continue;
          t.job = (uint32_t)ji;
          out.h[out.n++] = t;
        }
      }
    if (pairs_first) {
      // resident tasks: each coupling group (job, group) on one XCD
      std::vector<uint32_t> first(c.jobs.size() + 1, 0);   // class-wide group numbering
      for (size_t ji = 0; ji < c.jobs.size(); ++ji) first[ji + 1] = first[ji] + c.jobs[ji]->h->G;
      std::vector<TaskDesc> tv(out.h, out.h + out.n);
      xcd_interleave(tv, [&](const TaskDesc &t) {
        const sg_handle &h = *c.jobs[t.job]->h;
        return first[t.job] + (uint32_t)(t.lat[0] / ((size_t)h.R * h.T));
      });
      std::copy(tv.begin(), tv.end(), out.h);
    }
  }

  static const std::vector<TaskDesc> &pick_tasks(sg_handle &h, int par) { return h.h_tasks[par]; }
  static const std::vector<TaskDesc> &pick_single(sg_handle &h, int) { return h.h_single; }
  static const std::vector<TaskDesc> &pick_chunks(sg_handle &h, int) { return h.h_chunks; }
  static const std::vector<TaskDesc> &pick_slabs(sg_handle &h, int) { return h.h_slabs; }
  static const std::vector<TaskDesc> &pick_ptc(sg_handle &h, int par) { return h.h_ptc[par]; }

  void build(Class &c) {
    c.retire_graphs();   // they hold the old tables' pointers and counts
    c.reap_graphs(false);
    if (c.uploaded) HIPCHK(hipEventSynchronize(c.up_ev));  // staging free again
    int mode = c.kernel, tile = 0;
    if (mode == SG_KERNEL_AUTO) {
      size_t nlat = 0;
      for (Job *j : c.jobs) nlat += j->h->n_lat();
      mode = cached_kernel_choice(bp.device, c.L, c.field, nlat, &tile);
    }
    c.streamed = mode == SG_KERNEL_STREAMED;
    c.single = mode == SG_KERNEL_RESIDENT_SINGLE && !single_kernel_supported(c.L);
    c.tile = tile;
    for (Job *j : c.jobs) j->h->set_mode(mode, tile);
    reserve(c, c.jt, c.jobs.size());
    c.jt.n = (int)c.jobs.size();
    for (size_t ji = 0; ji < c.jobs.size(); ++ji) c.jt.h[ji] = c.jobs[ji]->h->self;
    upload(c, c.jt);
    if (c.streamed) {
      gather(c, c.chunks, false, pick_chunks, 0);
      upload(c, c.chunks);
      if (slab_kernel_supported(c.L)) {
        gather(c, c.slabs, false, pick_slabs, 0);
        upload(c, c.slabs);
      }
      for (int k = 0; k < 2; ++k) {
        gather(c, c.ptc[k], false, pick_ptc, k);
        upload(c, c.ptc[k]);
      }
    } else {
      for (int k = 0; k < 2; ++k) {
        gather(c, c.tasks[k], true, c.single ? pick_single : pick_tasks, k);
        upload(c, c.tasks[k]);
        if (single_kernel_supported(c.L) || c.single) {   // PT in its own launch
          gather(c, c.ptc[k], false, pick_ptc, k);
          upload(c, c.ptc[k]);
        }
      }
    }
    HIPCHK(hipEventRecord(c.up_ev, c.st));
    c.uploaded = true;
    c.dirty = false;
    c.steps_since_build = 0;
  }

  // ------------------------------------------------------------ one step
  // After each PT interval: host counters, measurements, and the final
  // copies of jobs that reached their sweep count (which leave the class).
  void after_interval(Class &c, int emode) {
    const int pe = bp.pt_every;
    bool left = false;
    for (Job *jp : c.jobs) {
      sg_handle &h = *jp->h;
      h.energy_valid = true;
      h.energy_k_valid = emode >= 2;
      const int par = (int)((h.sweep / pe) & 1);
      for (int t = par; t + 1 < h.T; t += 2) h.swap_att[t] += (uint64_t)h.G * h.R * 64;
      h.sweep += pe;
      const double fl = (double)pe * h.N * 64.0 * h.G * h.R * h.T;
      h.flips += fl;
      flips += fl;
      if (h.meas_every > 0 && h.sweep % h.meas_every == 0) h.measure(nullptr, nullptr);
      if (h.sweep >= jp->n_sweeps) {
        finish(*jp);
        left = true;
      }
    }
    if (left) {
      c.jobs.erase(std::remove_if(c.jobs.begin(), c.jobs.end(),
                                  [](Job *j) { return j->finishing; }),
                   c.jobs.end());
      c.dirty = true;
    }
  }

  // the launches of one batched PT interval of the class: PT, pt_every
  // sweeps, the energies of the final states (emode as in launch_block), the
  // sweep counters; k selects the parity list
  void launch_class_interval(Class &c, int k, int emode) {
    const int pe = bp.pt_every;
    if (c.streamed) {
      const int pm = kPhEnergy | (emode > 1 ? kPhEnergyK : 0);
      HIPCHK(launch_pt_chunks(c.jt.d, c.ptc[k].d, c.ptc[k].n, c.st));
      // pairs of sweeps by the slab kernel, a last odd sweep by the colour
      // launches
      const int ns = slab_kernel_supported(c.L) ? pe & ~1 : 0;
      for (int it = 0; it < ns; ++it) {
        int mode = kPhUpdate;
        if (it == 0 && ns == pe) mode |= kPhZero;
        if (it == pe - 1) mode |= pm;
        HIPCHK(launch_slab(c.L, c.field, c.jt.d, c.slabs.d, c.slabs.n, it & 1, it, mode, c.tile,
                           c.st));
      }
      for (int it = ns; it < pe; ++it)
        for (int col = 0; col < 2; ++col) {
          int mode = kPhUpdate;
          if (it == ns && col == 0) mode |= kPhZero;
          if (it == pe - 1 && col == 1) mode |= pm;
          HIPCHK(launch_phase(c.L, c.field, c.jt.d, c.chunks.d, c.chunks.n, col, it, mode, c.tile,
                              c.st));
        }
    } else if (single_kernel_supported(c.L)) {
      HIPCHK(launch_pt_chunks(c.jt.d, c.ptc[k].d, c.ptc[k].n, c.st));
      HIPCHK(launch_single(c.L, c.field, c.jt.d, c.tasks[k].d, c.tasks[k].n, pe, emode, c.st));
    } else if (c.single) {
      HIPCHK(launch_pt_chunks(c.jt.d, c.ptc[k].d, c.ptc[k].n, c.st));
      HIPCHK(launch_block(c.L, c.field, c.jt.d, c.tasks[k].d, c.tasks[k].n, 0, pe, emode, c.st));
    } else {
      HIPCHK(launch_block(c.L, c.field, c.jt.d, c.tasks[k].d, c.tasks[k].n, 1, pe, emode, c.st));
    }
    HIPCHK(launch_advance(c.jt.d, c.jt.n, pe, c.st));
  }

  // Batched PT intervals of the class that may run back to back from here,
  // at most `room`: every one but the last ends with no job measuring or
  // finishing (so the host bookkeeping of the earlier ones launches
  // nothing).  1, 2 or 4, to keep the number of captured shapes small.
  int run_length(const Class &c, int room) const {
    const int64_t pe = bp.pt_every;
    int64_t n = room;
    for (const Job *jp : c.jobs) {
      const sg_handle &h = *jp->h;
      n = std::min<int64_t>(n, (jp->n_sweeps - h.sweep) / pe);
      if (h.meas_every > 0) n = std::min<int64_t>(n, (h.meas_every - h.sweep % h.meas_every) / pe);
    }
    return n >= 4 ? 4 : n >= 2 ? 2 : 1;
  }

  // `count` batched intervals from parity list k, the last with energy mode
  // e_last and the others with e_mid.  Once the class's tables have served
  // kGraphAfter intervals unchanged, each shape is captured as a hipGraph
  // and replayed; until then (jobs coming and going) the launches go out
  // directly, which costs no instantiation.
  static constexpr int kGraphAfter = 8;
  void run_batched(Class &c, int k, int count, int e_mid, int e_last) {
    if (c.steps_since_build < kGraphAfter) {
      for (int i = 0; i < count; ++i) launch_class_interval(c, k ^ (i & 1), i + 1 < count ? e_mid : e_last);
      iv_direct += count;
      return;
    }
    hipGraphExec_t exec = nullptr;
    for (const auto &g : c.graphs)
      if (g.k == k && g.count == count && g.emode == e_last) exec = g.exec;
    if (!exec) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
      try {
        for (int i = 0; i < count; ++i)
          launch_class_interval(c, k ^ (i & 1), i + 1 < count ? e_mid : e_last);
      } catch (...) {
        // leave the stream out of capture mode before reporting
        if (hipStreamEndCapture(c.st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHK(hipStreamEndCapture(c.st, &g));
      const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(e);
      c.graphs.push_back(Class::Graph{k, count, e_last, exec});
      ++graphs_built;
    }
    HIPCHK(hipGraphLaunch(exec, c.st));
    iv_graphed += count;
  }

  // slice_pt_steps PT intervals of every job of the class
  void run_class(Class &c) {
    const int pe = bp.pt_every;
    for (int i = 0; i < bp.slice_pt_steps && !c.jobs.empty();) {
      if (bp.form == SG_FORM_BATCHED && c.dirty) build(c);
      const int count = bp.form == SG_FORM_BATCHED ? run_length(c, bp.slice_pt_steps - i) : 1;
      // U (and K at H != 0) for the next PT step; U and K when a job of the
      // class measures after the last interval
      bool meas = false;
      for (Job *jp : c.jobs)
        meas = meas || (jp->h->meas_every > 0 &&
                        (jp->h->sweep + (int64_t)count * pe) % jp->h->meas_every == 0);
      const int e_mid = c.field ? 2 : 1, e_last = c.field || meas ? 2 : 1;
      if (bp.form == SG_FORM_BATCHED) {
        run_batched(c, (int)(c.steps_since_build & 1), count, e_mid, e_last);
        c.steps_since_build += count;
      } else {
        for (Job *jp : c.jobs) {
          sg_handle &h = *jp->h;
          h.launch_interval((int)((h.sweep / pe) & 1), 1, pe, e_last);
        }
      }
      for (int r = 0; r < count; ++r) after_interval(c, r + 1 < count ? e_mid : e_last);
      i += count;
    }
  }

  // One step: retire finished jobs, admit queued ones, and queue one slice
  // (slice_pt_steps PT intervals) of every class whose stream holds fewer
  // than max_inflight unfinished slices.  Classes are paced by their own
  // queues, so a class of short intervals runs ahead of a slower one instead
  // of waiting for it; when every busy class is full the host waits for the
  // first free slot.
  void step() {
    retire(false);
    admit();
    const int cap = std::max(bp.max_inflight, 1);
    bool ran = false;
    for (auto &c : classes) {
      reap_slices(c);
      if (c.jobs.empty() || (int)c.inflight.size() >= cap) continue;
      run_class(c);
      hipEvent_t e = take_event();
      HIPCHK(hipEventRecord(e, c.st));
      c.inflight.push_back(e);
      ran = true;
    }
    ++steps;
    if (!ran) wait_for_slot(cap);
    retire(false);
  }

  // the finished slices at the head of the class's queue
  void reap_slices(Class &c) {
    while (!c.inflight.empty()) {
      const hipError_t q = hipEventQuery(c.inflight.front());
      if (q == hipErrorNotReady) break;
      HIPCHK(q);
      ev_pool.push_back(c.inflight.front());
      c.inflight.pop_front();
    }
  }

  // until a class with jobs has room for another slice (none: at once)
  void wait_for_slot(int cap) {
    for (;;) {
      bool busy_class = false;
      for (auto &c : classes) {
        reap_slices(c);
        if (c.jobs.empty()) continue;
        busy_class = true;
        if ((int)c.inflight.size() < cap) return;
      }
      if (!busy_class) return;
      std::this_thread::yield();
    }
  }

  int n_running() const {
    int n = 0;
    for (auto &kv : jobs) n += kv.second->state == SG_JOB_RUNNING;
    return n;
  }

  bool busy() const {
    for (auto &kv : jobs)
      if (kv.second->state != SG_JOB_DONE) return true;
    return false;
  }

  void drain() {
    while (busy()) {
      bool any_active = !queue.empty();
      for (auto &c : classes) any_active = any_active || !c.jobs.empty();
      if (any_active)
        step();
      else
        retire(true);
    }
  }

  Job &job(int64_t id) {
    auto it = jobs.find(id);
    if (it == jobs.end()) throw Fail{SG_ERR_INVALID, "unknown job id"};
    return *it->second;
  }
};

// ---------------------------------------------------------------- C-ABI
namespace {
template <class F>
sg_status bguarded(sg_batch *b, F &&f) {
  if (!b) return SG_ERR_INVALID;
  try {
    f();
    b->err.clear();
    return SG_OK;
  } catch (const Fail &e) {
    b->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    b->err = "host allocation failed";
    return SG_ERR_NOMEM;
  }
}
thread_local std::string g_batch_create_error;
}  // namespace

extern "C" {

void sg_batch_params_default(sg_batch_params *bp) {
  std::memset(bp, 0, sizeof(*bp));
  bp->device = 0;
  bp->pt_every = 16;
  // slices of four intervals with four queued per class: the best of (2, 2),
  // (2, 8), (4, 4) and (4, 16) on the serving workload
  bp->slice_pt_steps = 4;
  bp->form = SG_FORM_BATCHED;
  bp->arena_bytes = (int64_t)4 << 30;
  bp->max_inflight = 4;
}

sg_status sg_batch_create(const sg_batch_params *bp, sg_batch **out) {
  if (!out) return SG_ERR_INVALID;
  *out = nullptr;
  g_batch_create_error.clear();
  auto bad = [](const char *m) { g_batch_create_error = m; return SG_ERR_INVALID; };
  if (!bp) return bad("params is NULL");
  if (bp->pt_every < 1) return bad("pt_every must be >= 1");
  if (bp->slice_pt_steps < 1) return bad("slice_pt_steps must be >= 1");
  if (bp->form != SG_FORM_PLAIN && bp->form != SG_FORM_BATCHED) return bad("unknown form");
  if (bp->arena_bytes < (int64_t)DevArena::kAlign) return bad("arena too small");
  sg_batch *b = new (std::nothrow) sg_batch();
  if (!b) return bad("host allocation failed");
  try {
    b->create(*bp);
  } catch (const Fail &e) {
    g_batch_create_error = e.msg;
    delete b;
    return e.code;
  }
  *out = b;
  return SG_OK;
}

sg_status sg_batch_destroy(sg_batch *b) {
  if (!b) return SG_ERR_INVALID;
  delete b;
  return SG_OK;
}

const char *sg_batch_last_error(const sg_batch *b) {
  return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int64_t sg_job_footprint(const sg_params *p) {
  if (!p || p->L < 2 || p->n_samples < 1 || p->n_temps < 1 || p->n_replicas < 1) return -1;
  return (int64_t)job_footprint(p->L, p->n_samples, p->n_replicas, p->n_temps);
}

sg_status sg_submit(sg_batch *b, const sg_params *p, int64_t n_sweeps, int64_t *id) {
  if (!p || !id) return SG_ERR_INVALID;
  return bguarded(b, [&] { *id = b->submit(*p, n_sweeps, b->now()); });
}

sg_status sg_step(sg_batch *b, int32_t *n_running, int32_t *n_queued) {
  return bguarded(b, [&] {
    b->step();
    if (n_running) *n_running = b->n_running();
    if (n_queued) *n_queued = (int32_t)b->queue.size();
  });
}

sg_status sg_drain(sg_batch *b) {
  return bguarded(b, [&] { b->drain(); });
}

sg_status sg_poll(sg_batch *b, int64_t id, int32_t *state) {
  if (!state) return SG_ERR_INVALID;
  return bguarded(b, [&] {
    Job &j = b->job(id);
    if (j.finishing) b->retire(false);
    *state = j.state;
  });
}

sg_status sg_result(sg_batch *b, int64_t id, double *res, uint64_t *swaps_accepted) {
  if (!res) return SG_ERR_INVALID;
  return bguarded(b, [&] {
    Job &j = b->job(id);
    if (j.finishing) b->retire(false);
    if (j.state != SG_JOB_DONE) throw Fail{SG_ERR_STATE, "job not finished"};
    std::memcpy(res, j.res.data(), sizeof(double) * j.res.size());
    if (swaps_accepted && !j.swaps.empty())
      std::memcpy(swaps_accepted, j.swaps.data(), sizeof(uint64_t) * j.swaps.size());
    b->jobs.erase(id);
  });
}

sg_status sg_job_times(sg_batch *b, int64_t id, double *t_submit, double *t_admit,
                       double *t_done) {
  if (!t_submit || !t_admit || !t_done) return SG_ERR_INVALID;
  return bguarded(b, [&] {
    Job &j = b->job(id);
    *t_submit = j.t_submit;
    *t_admit = j.t_admit;
    *t_done = j.t_done;
  });
}

sg_status sg_batch_stats(sg_batch *b, double *flips, int64_t *steps, int64_t *arena_in_use,
                         int64_t *arena_capacity) {
  return bguarded(b, [&] {
    if (flips) *flips = b->flips;
    if (steps) *steps = b->steps;
    if (arena_in_use) *arena_in_use = (int64_t)b->arena.in_use();
    if (arena_capacity) *arena_capacity = (int64_t)b->arena.capacity();
  });
}

sg_status sg_batch_launch_stats(sg_batch *b, int64_t *intervals_direct,
                                int64_t *intervals_graphed, int64_t *graphs_built) {
  return bguarded(b, [&] {
    if (intervals_direct) *intervals_direct = b->iv_direct;
    if (intervals_graphed) *intervals_graphed = b->iv_graphed;
    if (graphs_built) *graphs_built = b->graphs_built;
  });
}

sg_status sg_batch_alloc_stats(sg_batch *b, int64_t *table_growths, int64_t *pinned_blocks) {
  return bguarded(b, [&] {
    int64_t g = 0;
    for (auto &c : b->classes) g += c.grown;
    if (table_growths) *table_growths = g;
    if (pinned_blocks) *pinned_blocks = b->host_pool.fresh();
  });
}

sg_status sg_serve(sg_batch *b, const sg_params *jp, const int64_t *n_sweeps,
                   const double *arrival_s, int32_t n_jobs, int64_t *ids, double *wall_s) {
  if (!jp || !n_sweeps || !arrival_s || !ids || n_jobs < 0) return SG_ERR_INVALID;
  return bguarded(b, [&] {
    for (int32_t i = 1; i < n_jobs; ++i)
      if (arrival_s[i] < arrival_s[i - 1])
        throw Fail{SG_ERR_INVALID, "arrival times must be non-decreasing"};
    const double start = b->now();
    int32_t next = 0;
    while (next < n_jobs || b->busy()) {
      const double t = b->now() - start;
      while (next < n_jobs && arrival_s[next] <= t) {
        ids[next] = b->submit(jp[next], n_sweeps[next], start + arrival_s[next]);
        ++next;
      }
      bool active = !b->queue.empty();
      for (auto &c : b->classes) active = active || !c.jobs.empty();
      if (active) {
        b->step();
      } else if (b->busy()) {
        b->retire(false);
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      } else if (next < n_jobs) {
        const double wait = arrival_s[next] - (b->now() - start);
        if (wait > 0)
          std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
      }
    }
    if (wall_s) *wall_s = b->now() - start;
  });
}

}  // extern "C"
