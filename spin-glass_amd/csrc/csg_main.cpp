// csg_main.cpp — command-line driver over the libsg C-ABI, a drop-in for the
// reference's `csg <sim_dir> <H>` (Program/src/main.cu:8-84).
//
// Same positional interface, file names and exit codes: H is parsed with
// std::stof (main.cu:18); files are <sim_dir>/<H %06.4f>-obs.bin (float32
// obs_s records [sample][temperature] appended per measurement, main.cu:57-61
// and eamsim.cuh:55-62) and <sim_dir>/<H %06.4f>-chkpt.bin (resumed when
// present, main.cu:40-54); a <unix_time>.log in the working directory is kept
// only on failure (main.cu:30-32, 77-80).  Unlike the reference, the
// checkpoint is versioned and checksummed, is written to a temporary file and
// renamed, and a checkpoint that does not match the run is rejected.
//
// Optional flags after the two positional arguments set what the reference
// fixes at compile time (eamsim.cuh:13-25):
//   --L 16 --samples 2048 --replicas 2 --sweeps 262144 --pt-every 16
//   --meas-every 1024 --temps 24 --tmin T --tmax T (geometric in T; default:
//   the reference ladder beta_i = 2*2^(-4i/23)) --coupling-seed S --spin-seed S
//   --form plain|batched --device D
#include <sg.h>

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
  std::ofstream *out;
  int L, R, T, S;
  double H;
  std::vector<double> cw, sw;
  std::vector<float> buf;   // [sample][temp][11]
  int64_t records = 0;
};

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
  w.out->write(reinterpret_cast<const char *>(w.buf.data()),
               (std::streamsize)(w.buf.size() * sizeof(float)));
  w.out->flush();
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
    const std::string ck = prefix + "chkpt.bin";
    if (!file_exists(ck)) {
      check(sg_init_lattice(h), h);
      record("lattice array initialized");
    } else {
      std::ifstream in(ck, std::ios::binary);
      std::vector<char> blob((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
      check(sg_load_checkpoint(h, blob.data(), blob.size()), h);
      record("checkpoint loaded");
    }
    std::ofstream obs(prefix + "obs.bin", std::ios::binary | std::ios::app);
    if (!obs.is_open()) throw Fail{"unable to open " + prefix + "obs.bin"};
    ObsWriter w;
    w.out = &obs;
    w.L = p.L; w.R = p.n_replicas; w.T = p.n_temps; w.S = p.n_samples; w.H = p.H;
    for (int c = 0; c < p.L; ++c) {
      w.cw.push_back(std::cos(2.0 * M_PI * c / p.L));
      w.sw.push_back(std::sin(2.0 * M_PI * c / p.L));
    }
    const auto t0 = std::chrono::steady_clock::now();
    // run in slices so progress can be shown (the reference prints a
    // percentage per measurement section, eamsim.cu:692)
    const int64_t slice = p.meas_every > 0 ? p.meas_every : sweeps;
    for (int64_t done = 0; done < sweeps;) {
      const int64_t n = std::min<int64_t>(slice, sweeps - done);
      check(sg_run(h, n, obs_sink, &w), h);
      done += n;
      std::cout << "\rprogress: " << fixed(100.0 * done / sweeps, 5, 1) << "%" << std::flush;
    }
    std::cout << "\n";
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double attempts = (double)sweeps * p.L * p.L * p.L * p.n_samples * p.n_replicas * p.n_temps;
    record("t_e = " + fixed(secs * 1e12 / attempts, 5, 2) + " ps (" +
           fixed(attempts / secs * 1e-9, 1, 1) + " spin flips/ns)");
    record("simulation ended");
    size_t size = 0;
    check(sg_save_checkpoint(h, nullptr, &size), h);
    std::vector<char> blob(size);
    check(sg_save_checkpoint(h, blob.data(), &size), h);
    const std::string tmp = ck + ".tmp";
    {
      std::ofstream out(tmp, std::ios::binary);
      if (!out.is_open()) throw Fail{"unable to open " + tmp};
      out.write(blob.data(), (std::streamsize)size);
      if (!out) throw Fail{"unable to write " + tmp};
    }
    if (std::rename(tmp.c_str(), ck.c_str()) != 0) throw Fail{"unable to rename " + tmp};
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
