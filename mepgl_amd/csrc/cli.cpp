// The mepgl command-line program: the process boundary of the reference
// (mepgl.cu:394-418 parseArgs, mepgl.cu:421-820 main), driving libmepgl.
//
//   mepgl --simname S --mode {S,M,L,C} --Fin F --Fout F' --iterations I
//
// reads ./simulations/S/input_data/ (io.cu:10-126), runs I outer iterations
// of the relaxation loop (mepgl.cu:148-360) with the stdin keys q + - C
// (mepgl.cu:635-671), resamples to F' images (mepgl.cu:699-766) and writes
// ./simulations/S/output_data/ (io.cu:128-291) and ./energy.csv.
//
// What the reference fixes at compile time (config.cuh:14-34) is read from
// the input or given as options: N from the arrays, the component count
// from the presence of 0/u_2.npy, dx from --dx, input_data/dx.npy or the
// x mesh.  Arrays are converted from their stored dtype, never
// reinterpreted.
#include <hip/hip_runtime.h>
#include <poll.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "mepgl.h"
#include "npyio.hpp"

namespace {

using mepgl::npy_load;
using mepgl::npy_save;
using mepgl::NpyArray;

struct Options {
  std::string simname;
  char mode = 0;
  int F = -1;
  int F_out = -1;
  int iterations = -1;
  int device = 0;
  int ncomp = 0;  // 0 = detect
  int no_self_field = 0;
  int precision = MEPGL_FP64;
  int relaxation_steps = 25;  // default_relaxation_step_number
  unsigned flags = 0;
  int segment_rows = 0;
  double dx = 0.0;  // 0 = from input
  bool stdin_keys = true;
  bool dry_run = false;
  // The FP64 statement of a mixed-precision run's error: 'c' continues the
  // end state in FP64 for fp64_check_iterations outer iterations, 's'
  // repeats the whole run in FP64 from the input, 0 skips it.
  char fp64_check = 'c';
  int fp64_check_iterations = 1;
  std::string root = ".";
};

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "mepgl: %s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
}

void usage() {
  std::printf(
      "Options:\n"
      "  -h [ --help ]              Help screen\n"
      "  --simname arg              Simulation name\n"
      "  --mode arg                 Mode selected (M, S, L, C)\n"
      "  --Fin arg                  Number of frames of the input guess\n"
      "  --Fout arg                 Number of frames of the output (default Fin)\n"
      "  --iterations arg           Number of iterations\n"
      "Run-time configuration (compile-time constants of the original):\n"
      "  --device arg               HIP device (default 0)\n"
      "  --dx arg                   lattice spacing (default: input_data/dx.npy, else x mesh)\n"
      "  --ncomp arg                1 or 2 components (default: 2 iff 0/u_2.npy exists)\n"
      "  --no-self-field            thin film: A fixed, no magnetic energy\n"
      "  --relaxation-steps arg     initial relaxation steps per iteration (default 25)\n"
      "  --precision arg            fp64 (default) or mixed\n"
      "  --poly-exact               exact line-search quartic\n"
      "  --cubic-textbook           textbook spline weight in mode C\n"
      "  --gauge-align              align neighbouring image phases in modes L, C\n"
      "  --no-graph                 launch eagerly instead of replaying a hipGraph\n"
      "  --segment-rows arg         rows per wavefront (default: chosen by the library)\n"
      "  --fp64-check arg           with --precision mixed: error against the FP64 path,\n"
      "                             continue (default), shadow or off\n"
      "  --fp64-check-iterations arg  outer iterations of the FP64 continuation (default 1)\n"
      "  --no-stdin                 do not read the q + - C keys\n"
      "  --dry-run                  load and validate the input, print its summary, exit\n"
      "  --root arg                 directory holding simulations/ (default .)\n");
}

int to_int(const std::string& k, const std::string& v) {
  char* end = nullptr;
  errno = 0;
  const long r = std::strtol(v.c_str(), &end, 10);
  if (errno || end == v.c_str() || *end) die("option --" + k + " expects an integer, got '" + v + "'");
  return (int)r;
}

Options parse(int argc, char** argv) {
  Options o;
  static const char* with_value[] = {"simname", "mode", "Fin", "Fout", "iterations", "device",
                                     "dx", "ncomp", "relaxation-steps", "precision",
                                     "segment-rows", "root", "fp64-check",
                                     "fp64-check-iterations"};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage();
      std::exit(0);
    }
    if (a.rfind("--", 0) != 0) die("unexpected argument '" + a + "'");
    std::string key = a.substr(2), val;
    const size_t eq = key.find('=');
    bool has_val = false;
    if (eq != std::string::npos) {
      val = key.substr(eq + 1);
      key = key.substr(0, eq);
      has_val = true;
    }
    bool needs = false;
    for (const char* w : with_value) needs |= key == w;
    if (needs && !has_val) {
      if (i + 1 >= argc) die("option --" + key + " requires a value");
      val = argv[++i];
    }
    if (key == "simname") o.simname = val;
    else if (key == "mode") {
      if (val.size() != 1) die("--mode expects one of S, M, L, C");
      o.mode = val[0];
    } else if (key == "Fin") o.F = to_int(key, val);
    else if (key == "Fout") o.F_out = to_int(key, val);
    else if (key == "iterations") o.iterations = to_int(key, val);
    else if (key == "device") o.device = to_int(key, val);
    else if (key == "ncomp") o.ncomp = to_int(key, val);
    else if (key == "relaxation-steps") o.relaxation_steps = to_int(key, val);
    else if (key == "segment-rows") o.segment_rows = to_int(key, val);
    else if (key == "root") o.root = val;
    else if (key == "fp64-check") {
      if (val == "continue") o.fp64_check = 'c';
      else if (val == "shadow") o.fp64_check = 's';
      else if (val == "off") o.fp64_check = 0;
      else die("--fp64-check expects continue, shadow or off");
    } else if (key == "fp64-check-iterations") o.fp64_check_iterations = to_int(key, val);
    else if (key == "dx") {
      char* end = nullptr;
      o.dx = std::strtod(val.c_str(), &end);
      if (end == val.c_str() || *end || !(o.dx > 0)) die("--dx expects a positive number");
    } else if (key == "precision") {
      if (val == "fp64") o.precision = MEPGL_FP64;
      else if (val == "mixed") o.precision = MEPGL_MIXED;
      else die("--precision expects fp64 or mixed");
    } else if (key == "no-self-field") o.no_self_field = 1;
    else if (key == "poly-exact") o.flags |= MEPGL_POLY_EXACT;
    else if (key == "cubic-textbook") o.flags |= MEPGL_CUBIC_TEXTBOOK;
    else if (key == "gauge-align") o.flags |= MEPGL_GAUGE_ALIGN;
    else if (key == "no-graph") o.flags |= MEPGL_NO_GRAPH;
    else if (key == "no-stdin") o.stdin_keys = false;
    else if (key == "dry-run") o.dry_run = true;
    else die("unrecognised option '--" + key + "'");
  }
  if (o.simname.empty()) die("--simname is required");
  if (o.mode != 'S' && o.mode != 'M' && o.mode != 'L' && o.mode != 'C')
    die("--mode must be one of S, M, L, C");
  if (o.F < 1) die("--Fin must be a positive number of frames");
  if (o.F_out < 0) o.F_out = o.F;
  if (o.F_out < 1) die("--Fout must be positive");
  if (o.iterations < 0) die("--iterations is required and must be >= 0");
  if (o.relaxation_steps < 1) die("--relaxation-steps must be >= 1");
  if (o.ncomp != 0 && o.ncomp != 1 && o.ncomp != 2) die("--ncomp must be 1 or 2");
  if (o.fp64_check_iterations < 1) die("--fp64-check-iterations must be >= 1");
  return o;
}

bool exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

void make_dirs(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty() && !exists(cur) && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
        throw std::runtime_error("cannot create directory " + cur);
    }
    if (i < path.size()) cur.push_back(path[i]);
  }
}

// A square (N, N) array of the file, [i][j] with i along x, returned in the
// library's layout (index i + N*j), as io.cu:23-30 transposes.
std::vector<double> load_plane(const std::string& path, int& N) {
  NpyArray a = npy_load(path);
  if (a.shape.size() != 2 || a.shape[0] != a.shape[1])
    throw std::runtime_error(path + ": expected a square 2-d array");
  if (N < 0) N = (int)a.shape[0];
  if ((int)a.shape[0] != N)
    throw std::runtime_error(path + ": size " + std::to_string(a.shape[0]) + " differs from N = " +
                             std::to_string(N));
  const std::vector<double> v = a.as_double();
  std::vector<double> out((size_t)N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) out[(size_t)i + (size_t)N * j] = v[(size_t)j + (size_t)N * i];
  return out;
}

std::vector<int32_t> load_mask(const std::string& path, int N) {
  NpyArray a = npy_load(path);
  if (a.shape.size() != 2 || (int)a.shape[0] != N || (int)a.shape[1] != N)
    throw std::runtime_error(path + ": expected an (N, N) array");
  if (a.descr[1] == 'f') throw std::runtime_error(path + ": domain masks must be integer or bool");
  const std::vector<int32_t> v = a.as_int32();
  std::vector<int32_t> out((size_t)N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) out[(size_t)i + (size_t)N * j] = v[(size_t)j + (size_t)N * i];
  return out;
}

// Library layout (index i + N*j) -> file layout [i][j], as io.cu:163-173.
template <class T>
std::vector<T> to_file(const double* lib, int N, size_t stride = 1) {
  std::vector<T> out((size_t)N * N);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      out[(size_t)j + (size_t)N * i] = (T)lib[((size_t)i + (size_t)N * j) * stride];
  return out;
}

struct Problem {
  int N = -1, F = 0, ncomp = 1;
  double dx = 0;
  std::vector<double> x, y;
  std::vector<int32_t> comp, sc;
  std::map<std::string, std::vector<double>> planes;  // a_1 ... h
  double q[5] = {0, 0, 0, 0, 0};
  std::vector<double> a, psi_1, psi_2;  // (F, N, N, 2)
};

const char* kPlaneNames[] = {"a_1", "b_1", "m_xx_1", "m_yy_1", "a_2",
                             "b_2", "m_xx_2", "m_yy_2", "h"};

Problem load_input(const Options& o) {
  const std::string dir = o.root + "/simulations/" + o.simname + "/input_data/";
  if (!exists(dir)) throw std::runtime_error("input directory " + dir + " not found");
  Problem p;
  p.F = o.F;
  p.x = load_plane(dir + "x.npy", p.N);
  p.y = load_plane(dir + "y.npy", p.N);
  const int N = p.N;
  if (N < 2) throw std::runtime_error("N must be at least 2");
  p.comp = load_mask(dir + "comp_domain.npy", N);
  p.sc = load_mask(dir + "sc_domain.npy", N);
  for (const char* name : kPlaneNames) p.planes[name] = load_plane(dir + name + ".npy", p.N);
  {
    NpyArray q = npy_load(dir + "q.npy");
    const std::vector<double> v = q.as_double();
    if (v.size() < 5) throw std::runtime_error("q.npy must hold [q_1, q_2, eta, gamma, delta]");
    for (int k = 0; k < 5; ++k) p.q[k] = v[k];
  }
  if (o.dx > 0) p.dx = o.dx;
  else if (exists(dir + "dx.npy")) {
    const std::vector<double> v = npy_load(dir + "dx.npy").as_double();
    if (v.empty()) throw std::runtime_error("dx.npy is empty");
    p.dx = v[0];
  } else {
    // x is meshgrid(x_ax, y_ax, indexing="ij") (builder.py:32-34): the file's
    // x[1][0] - x[0][0] is x_ax[1] - x_ax[0] (config.py:56).
    p.dx = p.x[1] - p.x[0];
  }
  if (!(p.dx > 0) || !std::isfinite(p.dx)) throw std::runtime_error("lattice spacing dx must be positive");
  p.ncomp = o.ncomp ? o.ncomp : (exists(dir + "0/u_2.npy") ? 2 : 1);
  const size_t NN = (size_t)N * N;
  p.a.assign((size_t)p.F * NN * 2, 0.0);
  p.psi_1.assign((size_t)p.F * NN * 2, 0.0);
  if (p.ncomp == 2) p.psi_2.assign((size_t)p.F * NN * 2, 0.0);
  auto put = [&](std::vector<double>& dst, int n, int part, const std::vector<double>& src) {
    for (size_t c = 0; c < NN; ++c) dst[((size_t)n * NN + c) * 2 + part] = src[c];
  };
  for (int n = 0; n < p.F; ++n) {
    const std::string fd = dir + std::to_string(n) + "/";
    put(p.a, n, 0, load_plane(fd + "ax.npy", p.N));
    put(p.a, n, 1, load_plane(fd + "ay.npy", p.N));
    put(p.psi_1, n, 0, load_plane(fd + "u_1.npy", p.N));
    put(p.psi_1, n, 1, load_plane(fd + "v_1.npy", p.N));
    if (p.ncomp == 2) {
      put(p.psi_2, n, 0, load_plane(fd + "u_2.npy", p.N));
      put(p.psi_2, n, 1, load_plane(fd + "v_2.npy", p.N));
    }
  }
  return p;
}

// A plane whose values are all equal goes to the library as a uniform value.
mepgl_param as_param(const std::vector<double>& plane) {
  mepgl_param r{plane[0], nullptr};
  for (double v : plane)
    if (v != plane[0] || std::signbit(v) != std::signbit(plane[0])) {
      r.plane = plane.data();
      return r;
    }
  return r;
}

// The table of the reference's loop (mepgl.cu:325-335) and of main's start
// and end (mepgl.cu:620-626, 780-786).
void print_energy_table(const std::vector<double>& e, int F) {
  std::printf("┌──────┬────────────────┐ \n");
  std::printf("| Sys# |  Free Energy   | \n");
  std::printf("├──────┼────────────────┤ \n");
  for (int n = 0; n < F; n += std::max((F - 1) / 10, 1)) std::printf("|  %3.3d | %.8e | \n", n, e[n]);
  std::printf("└──────┴────────────────┘ \n");
}

struct Lib {
  mepgl_ctx* ctx = nullptr;
  void check(int rc, const char* what) {
    if (rc != MEPGL_OK) {
      const char* msg = ctx ? mepgl_last_error(ctx) : "";
      throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + msg);
    }
  }
  std::vector<double> scalars(int which, int F) {
    std::vector<double> v((size_t)F);
    check(mepgl_get_scalars(ctx, 0, which, v.data()), "mepgl_get_scalars");
    return v;
  }
  ~Lib() {
    if (ctx) mepgl_destroy(ctx);
  }
};

// The stdin thread of mepgl.cu:637-671: q stops after the current
// iteration, + and - change the relaxation steps, C toggles NLCG.
void stdin_keys(mepgl_ctx* ctx, std::atomic<bool>* done) {
  while (!done->load()) {
    struct pollfd pf = {0, POLLIN, 0};
    const int r = ::poll(&pf, 1, 50);
    if (r <= 0) continue;
    if (pf.revents & (POLLERR | POLLNVAL)) break;
    char buffer[4];
    const ssize_t n = ::read(0, buffer, sizeof buffer);
    if (n <= 0) break;  // end of input: keys no longer arrive
    int v = 0;
    switch (buffer[0]) {
      case 'q':
        std::printf("I am quitting! Please wait.\n");
        mepgl_set_control(ctx, MEPGL_CONTROL_STOP, 1);
        return;
      case '+':
        mepgl_get_control(ctx, MEPGL_CONTROL_RELAXATION_STEPS, &v);
        mepgl_set_control(ctx, MEPGL_CONTROL_RELAXATION_STEPS, v + 1);
        std::printf("relaxation_steps = %d\n", v + 1);
        break;
      case '-':
        mepgl_get_control(ctx, MEPGL_CONTROL_RELAXATION_STEPS, &v);
        if (v > 1) --v;
        mepgl_set_control(ctx, MEPGL_CONTROL_RELAXATION_STEPS, v);
        std::printf("relaxation_steps = %d\n", v);
        break;
      case 'C':
        mepgl_get_control(ctx, MEPGL_CONTROL_NLCG, &v);
        mepgl_set_control(ctx, MEPGL_CONTROL_NLCG, !v);
        std::printf("NLCG set %s", !v ? "ON" : "OFF");
        break;
      default:
        break;
    }
  }
}

void save_output(const Options& o, const Problem& p, int F, const std::vector<double>& a,
                 const std::vector<double>& psi_1, const std::vector<double>& psi_2,
                 const std::vector<double>& b, const std::vector<double>& j_1,
                 const std::vector<double>& j_2, const std::vector<double>& fenergy,
                 const std::vector<double>& s, const std::vector<double>& s_jb) {
  const std::string dir = o.root + "/simulations/" + o.simname + "/output_data/";
  make_dirs(dir);
  const int N = p.N;
  const size_t NN = (size_t)N * N;
  const bool f64 = o.precision == MEPGL_FP64;
  const std::vector<size_t> sq = {(size_t)N, (size_t)N};
  // Arrays are written as `real`: float64 for FP64, float32 for MIXED, as the
  // reference's DOUBLE_PRECISION switch does (io.cu:139-290, real.cuh:7-14).
  auto save_vec = [&](const std::string& name, const std::vector<double>& v) {
    if (f64) npy_save(dir + name, v.data(), {v.size()});
    else {
      std::vector<float> f(v.begin(), v.end());
      npy_save(dir + name, f.data(), {f.size()});
    }
  };
  auto save_plane = [&](const std::string& path, const double* lib, size_t stride) {
    if (f64) npy_save(path, to_file<double>(lib, N, stride).data(), sq);
    else npy_save(path, to_file<float>(lib, N, stride).data(), sq);
  };
  save_vec("fenergy.npy", fenergy);
  save_vec("s.npy", s);
  save_vec("s_jb.npy", s_jb);
  save_vec("q.npy", std::vector<double>(p.q, p.q + 5));
  {
    std::vector<int32_t> c(NN), m(NN);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        c[(size_t)j + (size_t)N * i] = p.comp[(size_t)i + (size_t)N * j];
        m[(size_t)j + (size_t)N * i] = p.sc[(size_t)i + (size_t)N * j];
      }
    npy_save(dir + "comp_domain.npy", c.data(), sq);
    npy_save(dir + "sc_domain.npy", m.data(), sq);
  }
  save_plane(dir + "x.npy", p.x.data(), 1);
  save_plane(dir + "y.npy", p.y.data(), 1);
  for (const char* name : kPlaneNames) save_plane(dir + name + ".npy", p.planes.at(name).data(), 1);
  {
    const std::vector<double> dxv = {p.dx};
    npy_save(dir + "dx.npy", dxv.data(), {1});
  }
  for (int n = 0; n < F; ++n) {
    const std::string fd = dir + std::to_string(n) + "/";
    make_dirs(fd);
    const size_t off2 = (size_t)n * NN * 2, off1 = (size_t)n * NN;
    save_plane(fd + "u_1.npy", psi_1.data() + off2, 2);
    save_plane(fd + "v_1.npy", psi_1.data() + off2 + 1, 2);
    save_plane(fd + "ax.npy", a.data() + off2, 2);
    save_plane(fd + "ay.npy", a.data() + off2 + 1, 2);
    if (p.ncomp == 2) {
      save_plane(fd + "u_2.npy", psi_2.data() + off2, 2);
      save_plane(fd + "v_2.npy", psi_2.data() + off2 + 1, 2);
    }
    save_plane(fd + "jx_1.npy", j_1.data() + off2, 2);
    save_plane(fd + "jy_1.npy", j_1.data() + off2 + 1, 2);
    save_plane(fd + "b.npy", b.data() + off1, 1);
    if (p.ncomp == 2) {
      save_plane(fd + "jx_2.npy", j_2.data() + off2, 2);
      save_plane(fd + "jy_2.npy", j_2.data() + off2 + 1, 2);
    }
  }
}

// Relative difference, absolute where the reference value is zero.
double rel_diff(double x, double ref) {
  return ref != 0.0 ? std::fabs(x - ref) / std::fabs(ref) : std::fabs(x - ref);
}

// Entry and exit barriers of a string (post.py:188-189): max(F - F_0) and
// max(F - F_last).
std::pair<double, double> barriers(const std::vector<double>& e) {
  double in = -INFINITY, out = -INFINITY;
  for (double v : e) {
    in = std::max(in, v - e.front());
    out = std::max(out, v - e.back());
  }
  return {in, out};
}

// The error of a mixed-precision run against the FP64 path, per image and
// for the barriers, printed and written to output_data/fp64_check.npy as
// (F_out, 2) float64: the mixed run's energies, then the FP64 path's.
//  'c'  relaxes the mixed end state further in FP64, for
//       o.fp64_check_iterations outer iterations with the run's last
//       relaxation-step count and NLCG setting: on a relaxed string the
//       difference is what float arithmetic left in it; on a string still
//       relaxing it also holds that further relaxation, so it bounds the
//       precision error from above.
//  's'  repeats the run's iterations in FP64 from the input (the same
//       relaxation-step counts and NLCG settings, and the same resampling to
//       F_out), which also counts where the two trajectories part.
void fp64_check(const Options& o, mepgl_string_desc d, const Problem& p, int F_out,
                const std::vector<double>& e_mixed, const std::vector<double>& a,
                const std::vector<double>& psi_1, const std::vector<double>& psi_2,
                const std::vector<std::pair<int, int>>& schedule, int relax, int nlcg) {
  Lib chk;
  const bool two = p.ncomp == 2;
  char what[160];
  if (o.fp64_check == 'c') {
    d.F = F_out;
    chk.check(mepgl_create(&d, 1, MEPGL_FP64, o.flags, o.device, &chk.ctx), "mepgl_create");
    chk.check(mepgl_upload(chk.ctx, 0, a.data(), psi_1.data(), two ? psi_2.data() : nullptr),
              "mepgl_upload");
    chk.check(mepgl_loop(chk.ctx, o.fp64_check_iterations, relax, o.mode, nlcg), "mepgl_loop");
    std::snprintf(what, sizeof what, "the end state relaxed further in FP64, %d iteration(s) of %d steps",
                  o.fp64_check_iterations, relax);
  } else {
    d.F = p.F;
    chk.check(mepgl_create(&d, 1, MEPGL_FP64, o.flags, o.device, &chk.ctx), "mepgl_create");
    chk.check(mepgl_upload(chk.ctx, 0, p.a.data(), p.psi_1.data(), two ? p.psi_2.data() : nullptr),
              "mepgl_upload");
    for (const auto& it : schedule) chk.check(mepgl_loop(chk.ctx, 1, it.first, o.mode, it.second), "mepgl_loop");
    if (F_out != p.F) chk.check(mepgl_resample(chk.ctx, 0, F_out), "mepgl_resample");
    std::snprintf(what, sizeof what, "the same %zu iteration(s) run in FP64 from the input",
                  schedule.size());
  }
  chk.check(mepgl_free_energy(chk.ctx), "mepgl_free_energy");
  const std::vector<double> e64 = chk.scalars(MEPGL_SCALAR_ENERGY, F_out);

  std::printf("\nMixed precision against the FP64 path (%s):\n", what);
  std::printf("┌──────┬─────────────────┬─────────────────┬────────────┐\n");
  std::printf("│ Sys# │  Mixed energy   │   FP64 energy   │ rel. error │\n");
  std::printf("├──────┼─────────────────┼─────────────────┼────────────┤\n");
  for (int n = 0; n < F_out; n += std::max((F_out - 1) / 10, 1))
    std::printf("│  %3.3d │ %+.8e │ %+.8e │ %.4e │\n", n, e_mixed[n], e64[n], rel_diff(e_mixed[n], e64[n]));
  std::printf("└──────┴─────────────────┴─────────────────┴────────────┘\n");
  double worst = 0.0;
  for (int n = 0; n < F_out; ++n) worst = std::max(worst, rel_diff(e_mixed[n], e64[n]));
  const auto bm = barriers(e_mixed), bd = barriers(e64);
  std::printf("energy, largest relative error over images   %.4e\n", worst);
  std::printf("entry barrier  mixed %+.10e  FP64 %+.10e  relative error %.4e\n", bm.first, bd.first,
              rel_diff(bm.first, bd.first));
  std::printf("exit barrier   mixed %+.10e  FP64 %+.10e  relative error %.4e\n", bm.second, bd.second,
              rel_diff(bm.second, bd.second));
  std::vector<double> both((size_t)F_out * 2);
  for (int n = 0; n < F_out; ++n) {
    both[2 * (size_t)n] = e_mixed[n];
    both[2 * (size_t)n + 1] = e64[n];
  }
  npy_save(o.root + "/simulations/" + o.simname + "/output_data/fp64_check.npy", both.data(),
           {(size_t)F_out, 2});
}

int run(const Options& o) {
  setvbuf(stdout, nullptr, _IONBF, 0);
  if (o.dry_run) {
    Problem p = load_input(o);
    int uniform = 0;
    for (const char* name : kPlaneNames) uniform += as_param(p.planes[name]).plane == nullptr;
    std::printf("N %d\nF %d\nncomp %d\ndx %.17g\nq %.17g %.17g %.17g %.17g %.17g\n"
                "uniform_planes %d\n",
                p.N, p.F, p.ncomp, p.dx, p.q[0], p.q[1], p.q[2], p.q[3], p.q[4], uniform);
    return 0;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
  std::printf("Detected %d HIP device(s)\n", count);
  if (o.device < 0 || o.device >= count) die("device " + std::to_string(o.device) + " is not available");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, o.device) == hipSuccess)
    std::printf("\nUsing device %d: \"%s\"\n", o.device, prop.name);
  std::printf("\nSimulation name: %s\n", o.simname.c_str());

  std::printf("Loading input data...  ");
  Problem p = load_input(o);
  std::printf("done.\n\n");

  mepgl_string_desc d;
  std::memset(&d, 0, sizeof d);
  d.N = p.N;
  d.F = p.F;
  d.ncomp = p.ncomp;
  d.no_self_field = o.no_self_field;
  d.dx = p.dx;
  d.a_1 = as_param(p.planes["a_1"]);
  d.b_1 = as_param(p.planes["b_1"]);
  d.m_xx_1 = as_param(p.planes["m_xx_1"]);
  d.m_yy_1 = as_param(p.planes["m_yy_1"]);
  d.a_2 = as_param(p.planes["a_2"]);
  d.b_2 = as_param(p.planes["b_2"]);
  d.m_xx_2 = as_param(p.planes["m_xx_2"]);
  d.m_yy_2 = as_param(p.planes["m_yy_2"]);
  d.h = as_param(p.planes["h"]);
  d.q_1 = p.q[0];
  d.q_2 = p.q[1];
  d.eta = p.q[2];
  d.gamma = p.q[3];
  d.delta = p.q[4];
  d.comp_domain = p.comp.data();
  d.sc_domain = p.sc.data();
  d.segment_rows = o.segment_rows;

  Lib lib;
  lib.check(mepgl_create(&d, 1, o.precision, o.flags, o.device, &lib.ctx), "mepgl_create");
  lib.check(mepgl_upload(lib.ctx, 0, p.a.data(), p.psi_1.data(),
                         p.ncomp == 2 ? p.psi_2.data() : nullptr),
            "mepgl_upload");
  lib.check(mepgl_free_energy(lib.ctx), "mepgl_free_energy");
  print_energy_table(lib.scalars(MEPGL_SCALAR_ENERGY, p.F), p.F);
  std::printf("Starting MEP-MGL.\n\n");

  mepgl_set_control(lib.ctx, MEPGL_CONTROL_RELAXATION_STEPS, o.relaxation_steps);
  mepgl_set_control(lib.ctx, MEPGL_CONTROL_NLCG, 1);
  std::atomic<bool> done{false};
  std::thread keys;
  if (o.stdin_keys) keys = std::thread(stdin_keys, lib.ctx, &done);

  const int F = p.F;
  const int N = p.N;
  std::vector<double> fenergy_old((size_t)F, INFINITY);
  std::vector<std::pair<int, int>> schedule;  // (relaxation steps, NLCG) per iteration
  int relax_last = o.relaxation_steps, nlcg_last = 1;
  int rc_loop = MEPGL_OK;
  for (int counter = 0; counter < o.iterations; ++counter) {
    int relax = 0, nlcg = 1, stop = 0;
    mepgl_get_control(lib.ctx, MEPGL_CONTROL_RELAXATION_STEPS, &relax);
    mepgl_get_control(lib.ctx, MEPGL_CONTROL_NLCG, &nlcg);
    std::printf("\n");
    std::printf("Iteration:   %d\n", counter + 1);
    if (o.mode == 'L') std::printf("=> Linear interpolation step \n");
    if (o.mode == 'C') std::printf("=> Cubic interpolation step \n");
    std::printf("=> Relaxation steps: ");
    rc_loop = mepgl_loop(lib.ctx, 1, relax, o.mode, nlcg);
    schedule.emplace_back(relax, nlcg);
    relax_last = relax;
    nlcg_last = nlcg;
    for (int k = 0; k < relax; ++k) std::printf(" %d ", k + 1);
    std::printf("\n");
    if (rc_loop != MEPGL_OK) break;

    const std::vector<double> fenergy = lib.scalars(MEPGL_SCALAR_FENERGY, F);
    const std::vector<double> s_jb = lib.scalars(MEPGL_SCALAR_S_JB, F);
    const std::vector<double> g2 = lib.scalars(MEPGL_SCALAR_G_NORM2, F);
    const std::vector<double> alpha = lib.scalars(MEPGL_SCALAR_ALPHA, F);
    double g_norm_mean = 0;
    for (int n = 0; n < F; ++n) g_norm_mean += std::sqrt(g2[n]);
    g_norm_mean = g_norm_mean / ((double)F * N * N);
    std::printf("\n");
    std::printf("┌──────┬──────┬─────────────────┬─────────────────┬────────────────┬────────────────┐\n");
    std::printf("│ Sys# │   s  │  Free Energy    │  Delta Energy   │      ||g||     │      alpha     │\n");
    std::printf("├──────┼──────┼─────────────────┼─────────────────┼────────────────┼────────────────┤\n");
    for (int n = 0; n < F; n += std::max((F - 1) / 10, 1))
      std::printf("│  %3.3d │ %4.2f │ %+.8e │ %+.8e │ %.8e │ %.8e │\n", n, s_jb[n], fenergy[n],
                  fenergy[n] - fenergy_old[n], std::sqrt(g2[n]) / ((double)N * N), alpha[n]);
    std::printf("└──────┴──────┴─────────────────┴─────────────────┴────────────────┴────────────────┘\n");
    std::printf("\n");
    std::printf("||g||_mean      %e       \n", g_norm_mean);
    std::printf("\n");
    fenergy_old = fenergy;
    {
      // energy.csv in the working directory, rewritten every iteration
      // (mepgl.cu:348-355).
      std::FILE* f = std::fopen("energy.csv", "w");
      if (f) {
        for (int n = 0; n < F; ++n) std::fprintf(f, n + 1 < F ? "%e," : "%e", fenergy[n]);
        std::fclose(f);
      }
    }
    mepgl_get_control(lib.ctx, MEPGL_CONTROL_STOP, &stop);
    if (stop) break;
  }
  done.store(true);
  if (keys.joinable()) keys.join();
  lib.check(rc_loop, "mepgl_loop");

  // End of run (mepgl.cu:699-796): resample to F_out images, final energy,
  // supercurrents and field, copy back.
  const int F_out = o.F_out;
  if (F_out != F) lib.check(mepgl_resample(lib.ctx, 0, F_out), "mepgl_resample");
  lib.check(mepgl_free_energy(lib.ctx), "mepgl_free_energy");
  const std::vector<double> fenergy = lib.scalars(MEPGL_SCALAR_ENERGY, F_out);
  print_energy_table(fenergy, F_out);

  const size_t NN = (size_t)N * N;
  std::vector<double> a((size_t)F_out * NN * 2), psi_1(a.size()), psi_2;
  if (p.ncomp == 2) psi_2.resize(a.size());
  std::vector<double> b((size_t)F_out * NN), j_1(a.size()), j_2(p.ncomp == 2 ? a.size() : 0);
  lib.check(mepgl_download(lib.ctx, 0, a.data(), psi_1.data(), p.ncomp == 2 ? psi_2.data() : nullptr),
            "mepgl_download");
  lib.check(mepgl_compute_jb(lib.ctx, 0, b.data(), j_1.data(), p.ncomp == 2 ? j_2.data() : nullptr),
            "mepgl_compute_jb");
  // s and s_jb as the loop left them: the last reparametrisation's values,
  // or those of the resampled string when F_out != F_in.
  const std::vector<double> s = lib.scalars(MEPGL_SCALAR_S, F_out);
  const std::vector<double> s_jb = lib.scalars(MEPGL_SCALAR_S_JB, F_out);

  std::printf("Saving results...  ");
  save_output(o, p, F_out, a, psi_1, psi_2, b, j_1, j_2, fenergy, s, s_jb);
  std::printf("done.\n");

  if (o.precision == MEPGL_MIXED && o.fp64_check) {
    // The mixed context goes first: the FP64 one needs twice its field memory.
    mepgl_destroy(lib.ctx);
    lib.ctx = nullptr;
    try {
      fp64_check(o, d, p, F_out, fenergy, a, psi_1, psi_2, schedule, relax_last, nlcg_last);
    } catch (const std::exception& e) {
      // The run's outputs are saved; the check's failure is reported, not fatal.
      std::fflush(stdout);
      std::fprintf(stderr, "\nmepgl: FP64 check not completed: %s\n", e.what());
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const Options o = parse(argc, argv);
  try {
    return run(o);
  } catch (const std::exception& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "\nmepgl: %s\n", e.what());
    return EXIT_FAILURE;
  }
}
