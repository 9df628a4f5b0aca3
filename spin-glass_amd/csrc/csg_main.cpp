// csg_main.cpp — command-line driver over the libsg C-ABI, a drop-in for the
// reference's `csg <sim_dir> <H>` (Program/src/main.cu:8-84).
//
// Same positional interface, file names and exit codes: H is parsed with
// std::stof (main.cu:18); files are <sim_dir>/<H %06.4f>-obs.bin (float32
// obs_s records [sample][temperature] appended per measurement, main.cu:57-61
// and eamsim.cuh:55-62) and <sim_dir>/<H %06.4f>-chkpt.bin (resumed when
// present, main.cu:40-54); a <unix_time>.log in the working directory is kept
// only on failure (main.cu:30-32, 77-80).  Unlike the reference, the
// checkpoint is versioned and checksummed, is written to a temporary file,
// synced and renamed, and a checkpoint that does not match the run is
// rejected.
//
// Crash consistency.  The reference appends a record at every measurement
// but writes its checkpoint only at the end of a run (main.cu:40-68,
// eamsim.cu:655-676): after a crash, the next run appends the lost records
// again from the old state.  Here the checkpoint carries the byte length of
// obs.bin it corresponds to and the sweep count the invocation is to reach,
// and is written (with obs.bin synced first) every --checkpoint-every
// measurements as well as at the end.  On resume obs.bin is cut back to that
// length, and a checkpoint short of its invocation's target finishes that
// invocation first, so an interrupted run followed by a resume leaves exactly
// the obs.bin of an uninterrupted one.
//
// Optional flags after the two positional arguments set what the reference
// fixes at compile time (eamsim.cuh:13-25):
//   --L 16 --samples 2048 --replicas 2 --sweeps 262144 --pt-every 16
//   --meas-every 1024 --temps 24 --tmin T --tmax T (geometric in T; default:
//   the reference ladder beta_i = 2*2^(-4i/23)) --coupling-seed S --spin-seed S
//   --form plain|batched --device D --checkpoint-every M (measurements, 64;
//   0 = only at the end) --stop-after-records K (stop after appending K
//   records, leaving the state a crash there would leave: for tests)
#include <fcntl.h>
#include <sg.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::ofstream g_log;

std::string stamp() {
  std::time_t t = std::time(nullptr);
  char buf[32];
  std::strftime(buf, sizeof(buf), "[%d/%m/%y %H:%M:%S] ", std::localtime(&t));
  return buf;
}

void record(const std::string &msg) {
  std::cout << stamp() << msg << std::endl;
  if (g_log.is_open()) g_log << stamp() << msg << std::endl;
}

struct Fail {
  std::string msg;
};

void check(sg_status st, sg_handle *h) {
  if (st != SG_OK) throw Fail{std::string("libsg: ") + sg_last_error(h)};
}

std::string fixed(double v, int width, int prec) {
  std::ostringstream o;
  o << std::setw(width) << std::setfill('0') << std::setprecision(prec) << std::fixed << v;
  return o.str();
}

struct ObsWriter {
  FILE *out;
  int L, R, T, S;
  double H;
  std::vector<double> cw, sw;
  std::vector<float> buf;   // [sample][temp][11]
  int64_t records = 0;
  bool failed = false;
};

// The checkpoint file: this header, then libsg's checkpoint image.
struct RunHeader {
  char magic[8];          // "CSGRUN01"
  uint64_t obs_bytes;     // obs.bin length the state corresponds to
  uint64_t target_sweep;  // sweep count the writing invocation is to reach
  uint64_t image_bytes;
};

// writes data to path through a synced temporary file and a rename
void write_atomic(const std::string &path, const std::vector<char> &data) {
  const std::string tmp = path + ".tmp";
  const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) throw Fail{"unable to open " + tmp};
  size_t off = 0;
  while (off < data.size()) {
    const ssize_t n = ::write(fd, data.data() + off, data.size() - off);
    if (n <= 0) { ::close(fd); throw Fail{"unable to write " + tmp}; }
    off += (size_t)n;
  }
  if (::fsync(fd) != 0) { ::close(fd); throw Fail{"unable to sync " + tmp}; }
  ::close(fd);
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw Fail{"unable to rename " + tmp};
}

void save_run(sg_handle *h, const std::string &path, FILE *obs, int64_t target) {
  if (std::fflush(obs) != 0 || ::fsync(fileno(obs)) != 0) throw Fail{"unable to sync obs.bin"};
  struct stat st;
  if (::fstat(fileno(obs), &st) != 0) throw Fail{"unable to stat obs.bin"};
  size_t size = 0;
  check(sg_save_checkpoint(h, nullptr, &size), h);
  std::vector<char> blob(sizeof(RunHeader) + size);
  check(sg_save_checkpoint(h, blob.data() + sizeof(RunHeader), &size), h);
  RunHeader rh;
  std::memcpy(rh.magic, "CSGRUN01", 8);
  rh.obs_bytes = (uint64_t)st.st_size;
  rh.target_sweep = (uint64_t)target;
  rh.image_bytes = size;
  std::memcpy(blob.data(), &rh, sizeof(rh));
  write_atomic(path, blob);
}

// integer measurement fields -> the reference's float32 obs_s records
void obs_sink(void *user, int64_t, const int32_t *rec, int64_t, int32_t nfield) {
  ObsWriter &w = *static_cast<ObsWriter *>(user);
  const int L = w.L, R = w.R, T = w.T;
  const double N = (double)L * L * L;
  w.buf.assign((size_t)w.S * T * 11, 0.0f);
  for (int s = 0; s < w.S; ++s) {
    const int g = s / 64, l = s % 64;
    for (int t = 0; t < T; ++t) {
      const int32_t *f = rec + (((size_t)g * T + t) * 64 + l) * nfield;
      float *o = &w.buf[((size_t)s * T + t) * 11];
      for (int rho = 0; rho < R && rho < 2; ++rho) {
        const double M = 2.0 * f[R + rho] - N;
        o[rho] = (float)(((2.0 * f[rho] - 3.0 * N) - w.H * M) / N);
        o[2 + rho] = (float)(M / N);
      }
      if (R >= 2) {
        o[4] = (float)((N - 2.0 * f[2 * R]) / N);
        for (int d = 0; d < 3; ++d) {
          const int32_t *Wd = f + 2 * R + 1 + d * L;
          double re = 0, im = 0;
          for (int c = 0; c < L; ++c) {
            re += w.cw[c] * Wd[c];
            im += w.sw[c] * Wd[c];
          }
          o[5 + d] = (float)(-2.0 / N * re);
          o[8 + d] = (float)(-2.0 / N * im);
        }
      }
    }
  }
  if (std::fwrite(w.buf.data(), sizeof(float), w.buf.size(), w.out) != w.buf.size() ||
      std::fflush(w.out) != 0)
    w.failed = true;
  ++w.records;
}

bool file_exists(const std::string &p) {
  std::ifstream f(p, std::ios::binary);
  return f.good();
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) { std::cout << "missing arguments\n"; return EXIT_FAILURE; }
  const std::string sim_dir = argv[1];
  float Hf;
  try {
    Hf = std::stof(argv[2]);
  } catch (...) {
    std::cout << "invalid H\n";
    return EXIT_FAILURE;
  }
  sg_params p;
  sg_params_default(&p);
  p.H = Hf;
  p.form = SG_FORM_BATCHED;
  int64_t sweeps = 262144;
  int64_t ck_every = 64, stop_after = -1;
  int ntemps = 24;
  double tmin = 0, tmax = 0;
  for (int i = 3; i < argc; ++i) {
    const std::string a = argv[i];
    if (i + 1 >= argc) { std::cout << "extra arguments\n"; return EXIT_FAILURE; }
    const char *v = argv[++i];
    if (a == "--L") p.L = std::atoi(v);
    else if (a == "--samples") p.n_samples = std::atoi(v);
    else if (a == "--replicas") p.n_replicas = std::atoi(v);
    else if (a == "--sweeps") sweeps = std::atoll(v);
    else if (a == "--pt-every") p.pt_every = std::atoi(v);
    else if (a == "--meas-every") p.meas_every = std::atoi(v);
    else if (a == "--temps") ntemps = std::atoi(v);
    else if (a == "--tmin") tmin = std::atof(v);
    else if (a == "--tmax") tmax = std::atof(v);
    else if (a == "--coupling-seed") p.coupling_seed = std::strtoull(v, nullptr, 0);
    else if (a == "--spin-seed") p.spin_seed = std::strtoull(v, nullptr, 0);
    else if (a == "--device") p.device = std::atoi(v);
    else if (a == "--checkpoint-every") ck_every = std::atoll(v);
    else if (a == "--stop-after-records") stop_after = std::atoll(v);
    else if (a == "--form") p.form = std::strcmp(v, "plain") == 0 ? SG_FORM_PLAIN : SG_FORM_BATCHED;
    else { std::cout << "extra arguments\n"; return EXIT_FAILURE; }
  }
  if (tmin > 0 && tmax > 0) {
    p.betas = nullptr;
    p.T_min = tmin;
    p.T_max = tmax;
    p.n_temps = ntemps;
  } else if (ntemps != 24) {
    std::cout << "--temps needs --tmin and --tmax\n";
    return EXIT_FAILURE;
  }

  const std::string prefix = sim_dir + "/" + fixed(Hf, 6, 4) + "-";
  const std::time_t t_s = std::time(nullptr);
  const std::string log_path = std::to_string(t_s) + ".log";
  g_log.open(log_path);

  sg_handle *h = nullptr;
  try {
    check(sg_create(&p, &h), nullptr);
    record("H = " + fixed(Hf, 6, 4));
    record("eamsim initialized");
    const std::string ck = prefix + "chkpt.bin", obs_path = prefix + "obs.bin";
    int64_t target = -1;                // sweep count this invocation runs to
    int64_t obs_bytes = -1;             // obs.bin length of the loaded state
    if (!file_exists(ck)) {
      check(sg_init_lattice(h), h);
      record("lattice array initialized");
    } else {
      std::ifstream in(ck, std::ios::binary);
      std::vector<char> blob((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      RunHeader rh;
      if (blob.size() >= sizeof(rh) && std::memcmp(blob.data(), "CSGRUN01", 8) == 0) {
        std::memcpy(&rh, blob.data(), sizeof(rh));
        if (rh.image_bytes != blob.size() - sizeof(rh)) throw Fail{"checkpoint size mismatch"};
        check(sg_load_checkpoint(h, blob.data() + sizeof(rh), (size_t)rh.image_bytes), h);
        obs_bytes = (int64_t)rh.obs_bytes;
        if (sg_sweep_count(h) < (int64_t)rh.target_sweep) target = (int64_t)rh.target_sweep;
      } else {   // a bare libsg image: no obs.bin length, no interrupted run
        check(sg_load_checkpoint(h, blob.data(), blob.size()), h);
      }
      record("checkpoint loaded");
    }
    if (target < 0) target = sg_sweep_count(h) + sweeps;
    else record("resuming an interrupted run at sweep " + std::to_string(sg_sweep_count(h)) +
                " of " + std::to_string(target));
    if (obs_bytes >= 0) {
      // records appended after the checkpoint belong to the lost part of the
      // run, which is about to be redone bit for bit
      struct stat st;
      const int64_t have = ::stat(obs_path.c_str(), &st) == 0 ? (int64_t)st.st_size : 0;
      if (have < obs_bytes) throw Fail{"obs.bin is shorter than the checkpoint's records"};
      if (have > obs_bytes) {
        if (::truncate(obs_path.c_str(), (off_t)obs_bytes) != 0)
          throw Fail{"unable to truncate " + obs_path};
        record("obs.bin cut back to the checkpoint (" + std::to_string(have - obs_bytes) +
               " bytes of an interrupted run dropped)");
      }
    }
    FILE *obs = std::fopen(obs_path.c_str(), "ab");
    if (!obs) throw Fail{"unable to open " + obs_path};
    ObsWriter w;
    w.out = obs;
    w.L = p.L; w.R = p.n_replicas; w.T = p.n_temps; w.S = p.n_samples; w.H = p.H;
    for (int c = 0; c < p.L; ++c) {
      w.cw.push_back(std::cos(2.0 * M_PI * c / p.L));
      w.sw.push_back(std::sin(2.0 * M_PI * c / p.L));
    }
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t start = sg_sweep_count(h), total = target - start;
    // run in slices that end at measurements, so progress can be shown (the
    // reference prints a percentage per measurement section, eamsim.cu:692)
    // and checkpoints fall between records
    int64_t since_ck = 0;
    for (int64_t sw = start; sw < target;) {
      int64_t n = target - sw;
      if (p.meas_every > 0) n = std::min<int64_t>(n, p.meas_every - sw % p.meas_every);
      const int64_t before = w.records;
      check(sg_run(h, n, obs_sink, &w), h);
      if (w.failed) throw Fail{"unable to write " + obs_path};
      sw += n;
      since_ck += w.records - before;
      std::cout << "\rprogress: " << fixed(100.0 * (sw - start) / std::max<int64_t>(total, 1), 5, 1)
                << "%" << std::flush;
      if (stop_after >= 0 && w.records >= stop_after) {
        std::cout << "\n";
        std::fclose(obs);
        throw Fail{"stopped after " + std::to_string(w.records) +
                   " records (--stop-after-records)"};
      }
      if (ck_every > 0 && since_ck >= ck_every && sw < target) {
        save_run(h, ck, obs, target);
        since_ck = 0;
      }
    }
    std::cout << "\n";
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double attempts = (double)total * p.L * p.L * p.L * p.n_samples * p.n_replicas * p.n_temps;
    record("t_e = " + fixed(secs * 1e12 / std::max(attempts, 1.0), 5, 2) + " ps (" +
           fixed(attempts / std::max(secs, 1e-9) * 1e-9, 1, 1) + " spin flips/ns)");
    record("simulation ended");
    save_run(h, ck, obs, target);
    std::fclose(obs);
    record("checkpoint saved");
  } catch (const Fail &e) {
    record(e.msg);
    if (h) sg_destroy(h);
    return EXIT_FAILURE;
  }
  sg_destroy(h);
  g_log.close();
  std::remove(log_path.c_str());
  return EXIT_SUCCESS;
}
