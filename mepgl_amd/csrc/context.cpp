// Host side of libmepgl: context, memory, work lists, the NLCG step
// sequence and the relaxation loop (loop(), mepgl.cu:38-391), behind the C
// ABI declared in include/mepgl.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mepgl.h"
#include "common.hpp"
#include "gauge_local.hpp"
#include "kernels.hpp"

using namespace mepgl;

// Most parts of each class's images whose NLCG steps run side by side on
// their own streams (MEPGL_CONTROL_SPLIT_IMAGES).  The best count depends on
// the workload, so the autotuner picks 2 or 4 with the segment heights.  On
// one device, vortex_entry_256 in mixed precision (64 images of 256^2) ran a
// step in 0.1109 ms with four parts at 16-row segments, against 0.1117 ms
// with two at their best height (22 rows).  Four parts made the same
// workload's FP64 steps 0.8 % slower, and thin_film_disk_1024's mixed steps
// 1.4 % slower.
constexpr int kMaxParts = 4;

namespace {

struct Alloc {
  void* p = nullptr;
  size_t bytes = 0;
};

struct StringHost {
  mepgl_string_desc desc;
  DevString dev;            // host copy of the device descriptor
  std::vector<Alloc> owned; // device allocations of this string
  int ntiles = 0;
  std::vector<uint16_t> nn_host;  // neighbour map, [row][column] at the row pitch
  double* jb = nullptr;     // JB planes (lazy)
  double* dct = nullptr;    // DCT-II matrix and Laplacian eigenvalues (lazy)
  size_t elem = 8;          // bytes per stored real
};

struct Timing {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, std::pair<size_t, size_t>>> marks;  // kernel, (start, stop)
  double total_ms[kKernelCount] = {0};
  int count[kKernelCount] = {0};
};

}  // namespace

struct mepgl_ctx {
  int precision = MEPGL_FP64;
  unsigned flags = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  // Streams of the image parts (split_images): between reparametrisations
  // every image's NLCG state is independent of the others', so the NLCG
  // steps of `parts` parts of each class's images run as launch sequences
  // side by side, one part's P (mostly arithmetic) beside another part's G
  // (mostly memory traffic).  Part 0 runs on `stream`, part p on xs[p].
  hipStream_t xs[kMaxParts] = {};
  hipEvent_t fork_ev = nullptr, join_evs[kMaxParts] = {};
  bool split_images = true;  // MEPGL_CONTROL_SPLIT_IMAGES
  int parts = 2;             // parts of a split sequence, 2..kMaxParts
  bool edge_exchange = true; // MEPGL_CONTROL_EDGE_EXCHANGE
  // Global ids of each part's images, for the finishing launches of a split
  // sequence.
  int* d_part_imgs[kMaxParts] = {};
  int part_images[kMaxParts] = {};
  std::vector<StringHost> strings;
  int total_images = 0;
  DevString* d_strings = nullptr;
  ImageScalars* d_sc = nullptr;
  int2* d_img = nullptr;     // global image id -> (string, image)
  LineSearchTable* d_ls = nullptr;
  WorkItem* d_items = nullptr;
  int nitems = 0;
  // Items sorted by class (components, fixed A, parameter planes); G and P
  // launch once per non-empty class so each launch runs one specialised
  // kernel.
  // `cut`: item offsets of the parts of the class's images, each at an
  // image boundary; part p is items [cut[p], cut[p + 1]).
  struct Group {
    int begin, count, nc, nsf, pl;
    int cut[kMaxParts + 1];
  };
  std::vector<Group> groups;
  double** d_jb = nullptr;
  double* d_out = nullptr;
  std::string err;
  int xp = 0, dp = 0;        // current field / direction buffers
  bool pending = false;      // an update x += alpha d is owed
  // The last P's line search is left to the next G (its partials hold the
  // quartic; alpha and the line-search scalars are not yet stored).
  bool ls_deferred = false;
  bool fold_ls = true;       // MEPGL_CONTROL_FOLD_LINE_SEARCH
  std::atomic<int> stop{0}, relax_steps{25}, nlcg{1};
  Timing timing;
  bool dry = false;          // advance the parity state without launching
  std::map<std::tuple<int, int, int, int, int, char>, hipGraphExec_t> graphs;
  long long bytes = 0;
  // Non-finite flag: set to 1 by the step kernels (a gradient total) or the
  // finishing launch (a line-search coefficient) when a per-image value is
  // NaN or +-Inf.  Pinned host memory, written by the device, read by the
  // host after an event that follows the writing launch.
  int* health = nullptr;
  hipEvent_t iter_done[2] = {nullptr, nullptr};

  bool fp32() const { return precision == MEPGL_MIXED; }
};

namespace {

int fail(mepgl_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, MEPGL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int dev_alloc(mepgl_ctx* c, StringHost* s, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return MEPGL_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, MEPGL_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                        hipGetErrorString(e));
  }
  if (s) s->owned.push_back({*p, bytes});
  c->bytes += (long long)bytes;
  return MEPGL_OK;
}

void free_string(mepgl_ctx* c, StringHost& s) {
  for (auto& a : s.owned) {
    (void)hipFree(a.p);
    c->bytes -= (long long)a.bytes;
  }
  s.owned.clear();
}

// Shortest segment a schedule uses; the per-tile partial buffers are sized
// for it, so a schedule change never reallocates.
int min_rows(int N) { return std::min(N, 4); }

int clamp_rows(int rows, int N) { return std::max(std::min(rows, N), min_rows(N)); }

size_t max_tiles(const DevString& d) {
  return (size_t)d.tiles_x * (size_t)((d.N + min_rows(d.N) - 1) / min_rows(d.N));
}

// Segment heights the library picks from.  The G and P kernels march rows
// in rotations of three after the first, so heights of 3k + 1 rows waste no
// work (any other height pads its last rotation).
constexpr int kRowChoices[] = {4, 7, 10, 13, 16, 19, 22, 25, 31, 64, 127};

// Rows marched by one wavefront.  Taller segments re-read fewer halo rows;
// shorter ones give more wavefronts per launch.  Pick the tallest segment
// that still yields about eight wavefronts per SIMD (1024 SIMDs).
int choose_rows(const mepgl_string_desc& d, long long total_images) {
  if (d.segment_rows > 0) return clamp_rows(d.segment_rows, d.N);
  const long long cols = (d.N + kTileX - 1) / kTileX;
  for (int h : {64, 31, 16, 10, 7}) {
    const long long items = cols * ((d.N + h - 1) / h) * total_images;
    if (h <= d.N && items >= 8192) return h;
  }
  return std::min(d.N, 4);
}

bool param_ok(const mepgl_param& p) { return p.plane != nullptr || std::isfinite(p.value); }

void build_line_search_table(LineSearchTable* t) {
  // lineSearch, nlcg.cu:1414-1436: x_0 = 1e-10, x_{k+1} = x_k * 1.025f.
  const double m = (double)1.025f;
  double x = 1e-10;
  for (int k = 0; k <= kLsSteps; ++k) {
    t->x[k] = x;
    t->x4[k] = std::pow(x, 4);
    if (k < kLsSteps) {
      const double xn = x * m;
      t->alpha[k] = xn / m;
      x = xn;
    }
  }
}

StepArgs step_args(mepgl_ctx* c) {
  StepArgs a{};
  a.strings = c->d_strings;
  a.items = c->d_items;
  a.nitems = c->nitems;
  a.sc = c->d_sc;
  a.ls = c->d_ls;
  a.nlcg = 1;
  a.edges = c->edge_exchange ? 1 : 0;
  a.health = c->health;
  return a;
}

AuxArgs aux_args(mepgl_ctx* c) {
  AuxArgs a{};
  a.strings = c->d_strings;
  a.items = c->d_items;
  a.nitems = c->nitems;
  a.nstrings = (int)c->strings.size();
  a.sc = c->d_sc;
  a.xin = c->xp;
  a.xout = 1 - c->xp;
  a.dcur = c->dp;
  a.out = c->d_out;
  a.jb = c->d_jb;
  a.only_string = -1;
  return a;
}

// Records the next event of the timing pool on the stream, creating it on
// first use.
hipError_t record_timing_event(mepgl_ctx* c) {
  Timing& t = c->timing;
  if (t.used == t.pool.size()) {
    hipEvent_t e;
    const hipError_t err = hipEventCreate(&e);
    if (err != hipSuccess) return err;
    t.pool.push_back(e);
  }
  return hipEventRecord(t.pool[t.used++], c->stream);
}

int run_step_group(mepgl_ctx* c, const StepArgs& a, int which, hipStream_t st);

hipStream_t part_stream(const mepgl_ctx* c, int part);

// Launches `which` once per item class: on all of its images (part -1) on
// the context's stream, or on part `part` of its images on that part's
// stream.
int run_step(mepgl_ctx* c, const StepArgs& a0, int which, int part = -1) {
  if (c->dry) return MEPGL_OK;
  for (const auto& g : c->groups) {
    StepArgs a = a0;
    const int b = part < 0 ? 0 : g.cut[part];
    const int e = part < 0 ? g.count : g.cut[part + 1];
    if (e <= b) continue;
    a.items = c->d_items + g.begin + b;
    a.nitems = e - b;
    a.nc = g.nc;
    a.nsf = g.nsf;
    a.pl = g.pl;
    int rc = run_step_group(c, a, which, part_stream(c, part));
    if (rc) return rc;
  }
  return MEPGL_OK;
}

int run_step_group(mepgl_ctx* c, const StepArgs& a, int which, hipStream_t st) {
  if (c->timing.on && which != kKernelUpdate) {
    const size_t s0 = c->timing.used;
    hipError_t e = record_timing_event(c);
    if (e == hipSuccess) e = launch_step(a, which, c->fp32(), c->stream);
    if (e == hipSuccess) e = record_timing_event(c);
    if (e != hipSuccess) return fail(c, MEPGL_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    // Timing ids: 0 = G with the pending update, 1 = P, 2 = G without it.
    const int tid = (which == kKernelGrad && !a.upd) ? (int)kKernelUpdate : which;
    c->timing.marks.push_back({tid, {s0, s0 + 1}});
    return MEPGL_OK;
  }
  hipError_t e = launch_step(a, which, c->fp32(), st);
  if (e != hipSuccess) return fail(c, MEPGL_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  return MEPGL_OK;
}

// The second stream starts after the work queued so far on the first
// (fork) and the first continues after the second's work (join).
hipStream_t part_stream(const mepgl_ctx* c, int part) {
  return part >= 1 ? c->xs[part] : c->stream;
}

int fork_parts(mepgl_ctx* c) {
  if (c->dry) return MEPGL_OK;
  HIP_TRY(c, hipEventRecord(c->fork_ev, c->stream));
  for (int p = 1; p < c->parts; ++p) HIP_TRY(c, hipStreamWaitEvent(c->xs[p], c->fork_ev, 0));
  return MEPGL_OK;
}

int join_parts(mepgl_ctx* c) {
  if (c->dry) return MEPGL_OK;
  for (int p = 1; p < c->parts; ++p) {
    HIP_TRY(c, hipEventRecord(c->join_evs[p], c->xs[p]));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->join_evs[p], 0));
  }
  return MEPGL_OK;
}

int settle_ls(mepgl_ctx* c);

int run_aux(mepgl_ctx* c, const AuxArgs& a, int which, int arg) {
  if (c->ls_deferred) {
    if (int rc = settle_ls(c)) return rc;
  }
  if (c->dry) return MEPGL_OK;
  if (which == kKernelArclen) {
    // Launched once per item class, like the step kernels.
    for (const auto& g : c->groups) {
      AuxArgs ag = a;
      ag.items = c->d_items + g.begin;
      ag.nitems = g.count;
      ag.nc = g.nc;
      ag.pl = g.pl;
      hipError_t e = launch_aux(ag, which, arg, c->fp32(), c->stream);
      if (e != hipSuccess)
        return fail(c, MEPGL_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    }
    return MEPGL_OK;
  }
  hipError_t e = launch_aux(a, which, arg, c->fp32(), c->stream);
  if (e != hipSuccess) return fail(c, MEPGL_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  return MEPGL_OK;
}

// `part` in [0, parts): only that part's images (a split sequence).
FinishArgs finish_args(mepgl_ctx* c, int part = -1) {
  FinishArgs f{};
  f.strings = c->d_strings;
  f.img = c->d_img;
  f.nimages = c->total_images;
  if (part >= 0) {
    f.gids = c->d_part_imgs[part];
    f.nimages = c->part_images[part];
  }
  f.sc = c->d_sc;
  f.ls = c->d_ls;
  f.out = c->d_out;
  f.xin = c->xp;
  f.health = c->health;
  return f;
}

// Timing id of the finishing launch after P (id 3, a finishing launch after
// G, is no longer made: P reduces beta itself).
constexpr int kTimeFinishPoly = 4;

int run_finish(mepgl_ctx* c, const FinishArgs& f, int which, hipStream_t st = nullptr) {
  if (c->dry) return MEPGL_OK;
  const bool timed = c->timing.on && which == kKernelPoly;
  const size_t s0 = c->timing.used;
  hipError_t e = timed ? record_timing_event(c) : hipSuccess;
  if (e == hipSuccess) e = launch_finish(f, which, c->fp32(), st ? st : c->stream);
  if (timed && e == hipSuccess) e = record_timing_event(c);
  if (e != hipSuccess) return fail(c, MEPGL_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  if (timed)
    c->timing.marks.push_back({kTimeFinishPoly, {s0, s0 + 1}});
  return MEPGL_OK;
}

// Runs a line search left to the next G as a finishing launch, for anything
// other than that G (nothing is left pending across calls; this covers a
// sequence of steps cut short by an error).
int settle_ls(mepgl_ctx* c) {
  if (!c->ls_deferred) return MEPGL_OK;
  c->ls_deferred = false;
  return run_finish(c, finish_args(c), kKernelPoly);
}

// Applies an owed update in place.
int flush(mepgl_ctx* c) {
  if (int rc = settle_ls(c)) return rc;
  if (!c->pending) return MEPGL_OK;
  StepArgs a = step_args(c);
  a.xin = c->xp;
  a.din = c->dp;
  int rc = run_step(c, a, kKernelUpdate);
  if (rc) return rc;
  c->pending = false;
  return MEPGL_OK;
}

// Whether G takes over P's line search: only when every string has at most
// one tile per lane of the partials' reduction.  With more tiles per image
// every G wavefront waits for several rounds of partial loads before its
// first row, which cost more than the finishing launch saves (two_band_512,
// 256 tiles per image: 1 % slower in both precisions; vortex_entry_256, 64
// tiles: 1.2-1.8 % faster).
bool fold_line_search(const mepgl_ctx* c) {
  if (!c->fold_ls) return false;
  for (const auto& s : c->strings)
    if (s.ntiles > kWave) return false;
  return true;
}

// One NLCG step k of a call (nlcg.cu:1478-1583).  The reference swaps local
// copies of the gradient pointers (nlcg.cu:1484-1498), so step k writes the
// caller's g_old buffer when k is even and its g buffer when k is odd.
// `more`: another step follows at once, so P's line search may be left to
// that step's G (every G wavefront scans the table from P's partials), which
// saves the finishing launch and one kernel boundary per step.
// `split`: the image parts run on their streams (the caller forked them and
// joins them); each part's finishing launch covers its own images.
int nlcg_step(mepgl_ctx* c, int k, int nlcg, int maxwell, bool more = false, bool split = false) {
  const int gcur = (k % 2 == 0) ? 1 : 0;
  StepArgs a = step_args(c);
  a.gcur = gcur;
  a.gold = 1 - gcur;
  a.scur = gcur;
  a.sold = 1 - gcur;
  a.nlcg = nlcg;
  a.maxwell = maxwell;
  a.xin = c->xp;
  a.din = c->dp;
  a.upd = c->pending ? 1 : 0;
  a.ls_fold = c->pending && c->ls_deferred ? 1 : 0;
  a.xout = 1 - c->xp;
  if (!a.ls_fold) {
    if (int rc = settle_ls(c)) return rc;
  }
  int rc = MEPGL_OK;
  for (int p = split ? 0 : -1; p < (split ? c->parts : 0) && !rc; ++p) rc = run_step(c, a, kKernelGrad, p);
  if (rc) return rc;
  c->ls_deferred = false;
  if (c->pending) {
    c->xp = 1 - c->xp;
    c->pending = false;
  }
  a.upd = 0;
  a.ls_fold = 0;
  a.xin = c->xp;
  a.din = c->dp;
  a.dout = 1 - c->dp;
  // P totals the gradient partials into beta itself (there is no finishing
  // launch after G); the quartic totals and the line search follow P.
  for (int p = split ? 0 : -1; p < (split ? c->parts : 0) && !rc; ++p) rc = run_step(c, a, kKernelPoly, p);
  if (rc) return rc;
  if (more && fold_line_search(c)) {
    c->ls_deferred = true;
  } else if (split) {
    for (int p = 0; p < c->parts; ++p)
      if ((rc = run_finish(c, finish_args(c, p), kKernelPoly, part_stream(c, p)))) return rc;
  } else if ((rc = run_finish(c, finish_args(c), kKernelPoly))) {
    return rc;
  }
  c->dp = 1 - c->dp;
  c->pending = true;
  return MEPGL_OK;
}

// Names the first image whose energy, line-search coefficients or gradient
// norms are not finite (a gradient norm may be +Inf only in the slot of the
// previous step, where a solver reset puts it), after the stream is idle.
std::string first_nonfinite_image(mepgl_ctx* c) {
  std::vector<ImageScalars> h(c->total_images);
  if (hipMemcpy(h.data(), c->d_sc, sizeof(ImageScalars) * h.size(), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return "";
  for (size_t si = 0; si < c->strings.size(); ++si) {
    const DevString& s = c->strings[si].dev;
    for (int n = 0; n < s.F; ++n) {
      const ImageScalars& x = h[s.img0 + n];
      bool bad = !std::isfinite(x.fenergy) || std::isnan(x.g_norm2[0]) ||
                 std::isnan(x.g_norm2[1]) || !std::isfinite(x.g_delta_g);
      for (double v : x.poly) bad = bad || !std::isfinite(v);
      if (bad) return " (string " + std::to_string(si) + " image " + std::to_string(n) + ")";
    }
  }
  return "";
}

// Reads and clears the non-finite flag once the stream is idle.  Every
// value first_nonfinite_image checks is also checked by the kernel that
// writes it (the gradient totals by P, the energy and line-search
// coefficients by the line search), so the scalars are read back only to
// name the image once the flag is set.
int check_finite(mepgl_ctx* c, const std::string& where = "") {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (*(volatile int*)c->health) {
    *(volatile int*)c->health = 0;
    return fail(c, MEPGL_ERR_NONFINITE,
                "non-finite energy or gradient norm" + where + first_nonfinite_image(c));
  }
  return MEPGL_OK;
}

// Per-tile item flags of a string at its current segment height.
std::vector<int> tile_flags(const StringHost& sh) {
  const DevString& s = sh.dev;
  std::vector<int> f((size_t)s.tiles_x * s.tiles_y, 0);
  if (sh.nn_host.empty()) return f;
  for (int t = 0; t < s.tiles_x * s.tiles_y; ++t) {
    const int x0 = (t % s.tiles_x) * kTileX, x1 = std::min(x0 + kTileX, s.N);
    const int j0 = (t / s.tiles_x) * s.seg_rows, j1 = std::min(j0 + s.seg_rows, s.N);
    bool bonds = false;
    for (int j = j0; j < j1 && !bonds; ++j)
      for (int i = x0; i < x1 && !bonds; ++i) bonds = sh.nn_host[(size_t)j * s.pitch + i] != 0;
    f[t] = bonds ? 0 : kItemNoBonds;
  }
  return f;
}

int build_items(mepgl_ctx* c) {
  std::vector<WorkItem> items;
  std::vector<int> part_gids[kMaxParts];
  c->groups.clear();
  for (int nc = 1; nc <= 2; ++nc)
    for (int nsf = 0; nsf <= 1; ++nsf)
      for (int pl = 0; pl <= 1; ++pl) {
        const int begin = (int)items.size();
        std::vector<int> image_ends;  // item index after each image of the class
        std::vector<int> image_gids;  // its global image id
        for (size_t si = 0; si < c->strings.size(); ++si) {
          const DevString& s = c->strings[si].dev;
          const int spl = (s.planes || s.h.p) ? 1 : 0;
          if (s.nc != nc || s.nsf != nsf || spl != pl) continue;
          const std::vector<int> flags = tile_flags(c->strings[si]);
          for (int n = 0; n < s.F; ++n) {
            for (int t = 0; t < s.tiles_x * s.tiles_y; ++t)
              items.push_back({(int)si, n, t, flags[t]});
            image_ends.push_back((int)items.size() - begin);
            image_gids.push_back(s.img0 + n);
          }
        }
        const int count = (int)items.size() - begin;
        // Cut k is the image boundary nearest to k / parts of the class's
        // items, at or after cut k - 1 (a part may be empty).
        const int np = c->parts;
        mepgl_ctx::Group g{begin, count, nc, nsf, pl, {}};
        g.cut[0] = 0;
        for (int k = np; k <= kMaxParts; ++k) g.cut[k] = count;
        for (int k = 1; k < np; ++k) {
          int best = g.cut[k - 1];
          for (int e : image_ends)
            if (e >= g.cut[k - 1] && std::abs(np * e - k * count) < std::abs(np * best - k * count))
              best = e;
          g.cut[k] = best;
        }
        if (count > 0) c->groups.push_back(g);
        for (size_t q = 0; q < image_ends.size(); ++q) {
          int p = 0;
          while (p + 1 < np && image_ends[q] > g.cut[p + 1]) ++p;
          part_gids[p].push_back(image_gids[q]);
        }
      }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->d_items) (void)hipFree(c->d_items);
  c->d_items = nullptr;
  c->nitems = (int)items.size();
  HIP_TRY(c, hipMalloc(&c->d_items, sizeof(WorkItem) * std::max<size_t>(items.size(), 1)));
  HIP_TRY(c, hipMemcpyAsync(c->d_items, items.data(), sizeof(WorkItem) * items.size(),
                            hipMemcpyHostToDevice, c->stream));
  for (int h = 0; h < kMaxParts; ++h) {
    if (c->d_part_imgs[h]) (void)hipFree(c->d_part_imgs[h]);
    c->d_part_imgs[h] = nullptr;
    c->part_images[h] = (int)part_gids[h].size();
    HIP_TRY(c, hipMalloc(&c->d_part_imgs[h], sizeof(int) * std::max<size_t>(part_gids[h].size(), 1)));
    if (!part_gids[h].empty())
      HIP_TRY(c, hipMemcpyAsync(c->d_part_imgs[h], part_gids[h].data(),
                                sizeof(int) * part_gids[h].size(), hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MEPGL_OK;
}

int upload_descriptors(mepgl_ctx* c) {
  // The descriptor arrays keep their addresses for the life of the context,
  // so captured graphs stay valid when a descriptor changes.
  std::vector<DevString> ds;
  for (auto& s : c->strings) ds.push_back(s.dev);
  if (!c->d_strings) HIP_TRY(c, hipMalloc(&c->d_strings, sizeof(DevString) * ds.size()));
  HIP_TRY(c, hipMemcpyAsync(c->d_strings, ds.data(), sizeof(DevString) * ds.size(),
                            hipMemcpyHostToDevice, c->stream));
  std::vector<double*> jb;
  for (auto& s : c->strings) jb.push_back(s.jb);
  if (!c->d_jb) HIP_TRY(c, hipMalloc(&c->d_jb, sizeof(double*) * jb.size()));
  HIP_TRY(c, hipMemcpyAsync(c->d_jb, jb.data(), sizeof(double*) * jb.size(),
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MEPGL_OK;
}

int upload_param(mepgl_ctx* c, StringHost& s, const mepgl_param& p, DevParam& out) {
  out.v = p.value;
  out.p = nullptr;
  if (!p.plane) return MEPGL_OK;
  const int N = s.desc.N, P = s.dev.pitch;
  std::vector<double> h((size_t)N * P, 0.0);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) h[(size_t)j * P + i] = p.plane[(size_t)j * N + i];
  void* d;
  int rc = dev_alloc(c, &s, &d, h.size() * sizeof(double));
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  out.p = (const double*)d;
  return MEPGL_OK;
}

// Allocates the per-image buffers of a string for F images.
int alloc_images(mepgl_ctx* c, StringHost& s, int F) {
  DevString& d = s.dev;
  d.F = F;
  const size_t fs = (size_t)F * d.image_stride * s.elem;
  for (int b = 0; b < 2; ++b) {
    int rc;
    if ((rc = dev_alloc(c, &s, &d.x[b], fs))) return rc;
    if ((rc = dev_alloc(c, &s, &d.d[b], fs))) return rc;
    if ((rc = dev_alloc(c, &s, &d.g[b], fs))) return rc;
    HIP_TRY(c, hipMemsetAsync(d.x[b], 0, fs, c->stream));
    HIP_TRY(c, hipMemsetAsync(d.d[b], 0, fs, c->stream));
    HIP_TRY(c, hipMemsetAsync(d.g[b], 0, fs, c->stream));
  }
  d.gs = nullptr;
  d.cubic_rhs = nullptr;
  s.jb = nullptr;
  const size_t nt = (size_t)F * max_tiles(d);
  void* p;
  int rc;
  if ((rc = dev_alloc(c, &s, &p, nt * kPartG * sizeof(double)))) return rc;
  d.part_g = (double*)p;
  if ((rc = dev_alloc(c, &s, &p, nt * kPartP * sizeof(double)))) return rc;
  d.part_p = (double*)p;
  if ((rc = dev_alloc(c, &s, &p, nt * kPartA * sizeof(double)))) return rc;
  d.part_a = (double*)p;
  if ((rc = dev_alloc(c, &s, &p, (size_t)F * 4 * sizeof(double)))) return rc;
  HIP_TRY(c, hipMemsetAsync(p, 0, (size_t)F * 4 * sizeof(double), c->stream));
  d.cubic = (double*)p;
  if ((rc = dev_alloc(c, &s, &p, (size_t)F * (2 * sizeof(int) + 2 * sizeof(double))))) return rc;
  HIP_TRY(c, hipMemsetAsync(p, 0, (size_t)F * (2 * sizeof(int) + 2 * sizeof(double)), c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  d.plan_w = (double*)p;
  d.cplan_w = d.plan_w + F;
  d.plan_k = (int*)(d.cplan_w + F);
  d.cplan_k = d.plan_k + F;
  return MEPGL_OK;
}

int alloc_scalars(mepgl_ctx* c) {
  int img = 0;
  for (auto& s : c->strings) {
    s.dev.img0 = img;
    img += s.dev.F;
  }
  c->total_images = img;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->d_sc) (void)hipFree(c->d_sc);
  if (c->d_out) (void)hipFree(c->d_out);
  c->d_sc = nullptr;
  c->d_out = nullptr;
  HIP_TRY(c, hipMalloc(&c->d_sc, sizeof(ImageScalars) * img));
  HIP_TRY(c, hipMemsetAsync(c->d_sc, 0, sizeof(ImageScalars) * img, c->stream));
  HIP_TRY(c, hipMalloc(&c->d_out, sizeof(double) * img));
  HIP_TRY(c, hipMemsetAsync(c->d_out, 0, sizeof(double) * img, c->stream));
  std::vector<int2> map;
  for (size_t si = 0; si < c->strings.size(); ++si)
    for (int n = 0; n < c->strings[si].dev.F; ++n) map.push_back(make_int2((int)si, n));
  if (c->d_img) (void)hipFree(c->d_img);
  c->d_img = nullptr;
  HIP_TRY(c, hipMalloc(&c->d_img, sizeof(int2) * map.size()));
  HIP_TRY(c, hipMemcpyAsync(c->d_img, map.data(), sizeof(int2) * map.size(),
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MEPGL_OK;
}

void destroy_graphs(mepgl_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
}

int reset_solver(mepgl_ctx* c) {
  for (auto& s : c->strings) {
    const long long n = (long long)s.dev.F * s.dev.image_stride;
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(c, launch_fill(s.dev.g[b], n, c->fp32(), 0.0, c->stream));
      HIP_TRY(c, launch_fill(s.dev.d[b], n, c->fp32(), 0.0, c->stream));
    }
  }
  AuxArgs a = aux_args(c);
  int rc = run_aux(c, a, kKernelResetNorms, 0);
  if (rc) return rc;
  c->pending = false;
  return MEPGL_OK;
}

bool valid_string(mepgl_ctx* c, int s) { return c && s >= 0 && s < (int)c->strings.size(); }

// Field conversion between the host layout (F, N, N, 2) and the device planes.
template <class T>
void pack_fields(const StringHost& s, const double* a, const double* p1, const double* p2,
                 std::vector<T>& out) {
  const DevString& d = s.dev;
  const int N = d.N, P = d.pitch, F = d.F;
  out.assign((size_t)F * d.image_stride, (T)0);
  const double* src[3] = {a, p1, p2};
  for (int n = 0; n < F; ++n)
    for (int f = 0; f < 1 + d.nc; ++f) {
      if (!src[f]) continue;
      for (int comp = 0; comp < 2; ++comp) {
        T* dst = out.data() + (size_t)n * d.image_stride + (size_t)(2 * f + comp) * d.plane_stride;
        const double* sp = src[f] + (size_t)n * N * N * 2;
        for (int j = 0; j < N; ++j)
          for (int i = 0; i < N; ++i) dst[(size_t)j * P + i] = (T)sp[((size_t)j * N + i) * 2 + comp];
      }
    }
}

template <class T>
void unpack_fields(const StringHost& s, const std::vector<T>& in, double* a, double* p1,
                   double* p2) {
  const DevString& d = s.dev;
  const int N = d.N, P = d.pitch, F = d.F;
  double* dst[3] = {a, p1, p2};
  for (int n = 0; n < F; ++n)
    for (int f = 0; f < 3; ++f) {
      if (!dst[f]) continue;
      for (int comp = 0; comp < 2; ++comp) {
        double* dp = dst[f] + (size_t)n * N * N * 2;
        if (f > d.nc) {
          for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) dp[((size_t)j * N + i) * 2 + comp] = 0.0;
          continue;
        }
        const T* sp = in.data() + (size_t)n * d.image_stride + (size_t)(2 * f + comp) * d.plane_stride;
        for (int j = 0; j < N; ++j)
          for (int i = 0; i < N; ++i) dp[((size_t)j * N + i) * 2 + comp] = (double)sp[(size_t)j * P + i];
      }
    }
}

int copy_out_fields(mepgl_ctx* c, int si, const void* dev, double* a, double* p1, double* p2) {
  StringHost& s = c->strings[si];
  const size_t n = (size_t)s.dev.F * s.dev.image_stride;
  if (c->fp32()) {
    std::vector<float> h(n);
    HIP_TRY(c, hipMemcpyAsync(h.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    unpack_fields(s, h, a, p1, p2);
  } else {
    std::vector<double> h(n);
    HIP_TRY(c, hipMemcpyAsync(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    unpack_fields(s, h, a, p1, p2);
  }
  return MEPGL_OK;
}

// Host fields (F, N, N, 2) into a device field set of one string.
int copy_in_fields(mepgl_ctx* c, int si, void* dev, const double* a, const double* p1,
                   const double* p2) {
  StringHost& s = c->strings[si];
  const size_t n = (size_t)s.dev.F * s.dev.image_stride;
  if (c->fp32()) {
    std::vector<float> h;
    pack_fields(s, a, p1, s.dev.nc == 2 ? p2 : nullptr, h);
    HIP_TRY(c, hipMemcpyAsync(dev, h.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<double> h;
    pack_fields(s, a, p1, s.dev.nc == 2 ? p2 : nullptr, h);
    HIP_TRY(c, hipMemcpyAsync(dev, h.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  return MEPGL_OK;
}

// reset_interior: the finalize launch also resets the interior images'
// gradient norms, as an outer iteration does after its reparametrisation.
int arc_length_all(mepgl_ctx* c, int reset_interior = 0) {
  for (auto& s : c->strings)
    if (s.dev.F < 2)
      return fail(c, MEPGL_ERR_ARG, "arclength and reparametrisation need at least 2 images");
  // k_arc_finalize reads the arclength partial totals of every image.
  int rc;
  AuxArgs a = aux_args(c);
  if (c->pending && c->fp32()) {
    // Mixed precision: the arclength pass applies the pending update itself
    // and writes the new fields to the other buffer.
    a.dcur = c->dp;
    a.upd = 1;
    if ((rc = run_aux(c, a, kKernelArclen, 0))) return rc;
    c->xp = 1 - c->xp;
    c->pending = false;
    a = aux_args(c);
  } else {
    if ((rc = flush(c))) return rc;
    a = aux_args(c);
    if ((rc = run_aux(c, a, kKernelArclen, 0))) return rc;
  }
  if ((rc = run_finish(c, finish_args(c), kKernelArclen))) return rc;
  a.reset_interior = reset_interior;
  return run_aux(c, a, kKernelArcFinalize, (int)c->strings.size());
}

int ensure_cubic_rhs(mepgl_ctx* c) {
  bool changed = false;
  for (auto& s : c->strings) {
    if (s.dev.cubic_rhs) continue;
    int rc = dev_alloc(c, &s, &s.dev.cubic_rhs,
                       (size_t)s.dev.F * s.dev.image_stride * sizeof(double));
    if (rc) return rc;
    changed = true;
  }
  return changed ? upload_descriptors(c) : MEPGL_OK;
}

// Redistribution of all strings onto the other field buffer (uses the plan
// computed by the last arclength pass).
int redistribute_all(mepgl_ctx* c, int cubic) {
  int rc = flush(c);
  if (rc) return rc;
  AuxArgs a = aux_args(c);
  if (cubic) {
    for (auto& s : c->strings)
      if (s.dev.F < 4)
        return fail(c, MEPGL_ERR_ARG, "cubic reparametrisation needs at least 4 images");
    if ((rc = run_aux(c, a, kKernelCubicPlan, (int)c->strings.size())) ||
        (rc = run_aux(c, a, kKernelCubic, (c->flags & MEPGL_CUBIC_TEXTBOOK) ? 1 : 0)))
      return rc;
  } else {
    if ((rc = run_aux(c, a, kKernelLinear, 0))) return rc;
  }
  // NO_SELF_FIELD strings keep the same fixed A in both field buffers: copy
  // the new A (buffer xout) over the old one (buffer xin).  A shared by every
  // image is not interpolated, so both buffers still hold it.
  bool copy_a = false;
  for (auto& s : c->strings) copy_a = copy_a || (s.dev.nsf && !s.dev.a_shared);
  if (copy_a && (rc = run_aux(c, a, kKernelCopyA, 0))) return rc;
  c->xp = 1 - c->xp;
  return MEPGL_OK;
}

// The orthonormal DCT-II matrix C[i][k] = w_k cos(pi k (i + 1/2) / N) and the
// eigenvalues lam_k = 2 - 2 cos(pi k / N) of the Neumann graph Laplacian of a
// path of N sites, per string (N * N + N doubles on the device).
int ensure_dct(mepgl_ctx* c) {
  for (auto& s : c->strings) {
    if (s.dct || s.dev.nsf) continue;
    const int N = s.dev.N;
    const double pi = 3.14159265358979323846;
    std::vector<double> h((size_t)N * N + N);
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < N; ++k) {
        const double w = k == 0 ? std::sqrt(1.0 / N) : std::sqrt(2.0 / N);
        h[(size_t)i * N + k] = w * std::cos(pi * k * (i + 0.5) / N);
      }
    for (int k = 0; k < N; ++k) h[(size_t)N * N + k] = 2.0 - 2.0 * std::cos(pi * k / N);
    void* p;
    int rc = dev_alloc(c, &s, &p, h.size() * sizeof(double));
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    s.dct = (double*)p;
  }
  return MEPGL_OK;
}

// Local gauge alignment of neighbouring images (optional; no reference
// function): chi_n from the least-squares problem of every neighbouring pair
// (rhs, four DCT products, prefix along the string), then applied to the
// fields, the direction and both gradient buffers (gauge_local.hip).
int gauge_align_local_all(mepgl_ctx* c) {
  for (auto& s : c->strings)
    if (s.dev.F < 2) return fail(c, MEPGL_ERR_ARG, "alignment needs at least 2 images");
  int rc = flush(c);
  if (rc) return rc;
  if ((rc = ensure_cubic_rhs(c)) || (rc = ensure_dct(c))) return rc;
  if (c->dry) return MEPGL_OK;
  AuxArgs a = aux_args(c);
  hipError_t e = launch_glocal(a, 0, c->fp32(), c->stream);
  for (auto& s : c->strings) {
    const DevString& d = s.dev;
    if (d.nsf || e != hipSuccess) continue;
    const int N = d.N, F = d.F;
    const long long IS = d.image_stride, PS = d.plane_stride, P = d.pitch;
    double* w0 = (double*)d.cubic_rhs;
    double* w1 = w0 + PS;
    const double* C = s.dct;
    const double* lam = s.dct + (size_t)N * N;
    // b C, C^T (b C) / (lam_l + lam_k), C (.), (.) C^T
    const GemmArgs st[4] = {
        {w0, P, 1, IS, C, N, 1, 0, w1, P, IS, N, nullptr},
        {C, 1, N, 0, w1, P, 1, IS, w0, P, IS, N, lam},
        {C, N, 1, 0, w0, P, 1, IS, w1, P, IS, N, nullptr},
        {w1, P, 1, IS, C, 1, N, 0, w0, P, IS, N, nullptr}};
    for (const auto& g : st)
      if (e == hipSuccess) e = launch_dct_gemm(g, F, c->stream);
    if (e == hipSuccess) e = launch_glocal_prefix(w0, N, F, P, IS, c->stream);
  }
  if (e == hipSuccess) e = launch_glocal(a, 1, c->fp32(), c->stream);
  if (e != hipSuccess) return fail(c, MEPGL_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  return MEPGL_OK;
}

// Constant-phase alignment of neighbouring images (optional; the reference
// has no alignment step and relies on the gauge-invariant metric).
int gauge_align_all(mepgl_ctx* c) {
  for (auto& s : c->strings)
    if (s.dev.F < 2) return fail(c, MEPGL_ERR_ARG, "alignment needs at least 2 images");
  int rc = flush(c);
  if (rc) return rc;
  AuxArgs a = aux_args(c);
  if ((rc = run_aux(c, a, kKernelAlign, 0))) return rc;
  if ((rc = run_finish(c, finish_args(c), kKernelAlign))) return rc;
  if ((rc = run_aux(c, a, kKernelAlignFinalize, (int)c->strings.size()))) return rc;
  return run_aux(c, a, kKernelAlignApply, 0);
}

// One outer iteration of loop() (mepgl.cu:153-303).
int outer_iteration(mepgl_ctx* c, int relax, char mode, int nlcg) {
  int rc;
  if (mode == 'L' || mode == 'C') {
    if ((c->flags & MEPGL_GAUGE_ALIGN_LOCAL) && (rc = gauge_align_local_all(c))) return rc;
    if ((c->flags & MEPGL_GAUGE_ALIGN) && (rc = gauge_align_all(c))) return rc;
    // The interior norms are reset (beta = 0 on the next step) by the
    // arclength finalize; the redistribution does not read them.
    if ((rc = arc_length_all(c, 1))) return rc;
    if ((rc = redistribute_all(c, mode == 'C'))) return rc;
  }
  const bool maxwell = mode == 'M';
  // The steps of the image parts run side by side (per-kernel timing
  // keeps one stream).
  const bool split = c->split_images && relax > 1 && !c->timing.on;
  if (split && (rc = fork_parts(c))) return rc;
  for (int k = 0; k < relax; ++k)
    if ((rc = nlcg_step(c, k, nlcg, maxwell ? 1 : 0, k + 1 < relax, split))) {
      if (split) (void)join_parts(c);
      return rc;
    }
  if (split && (rc = join_parts(c))) return rc;
  return MEPGL_OK;
}

// ---------------------------------------------------------------------------
// Autotuning of the schedule: the segment height of each class of strings
// (same N, components and fixed-A setting), timed on the live problem.

// Makes the descriptors, work list and graphs follow the strings' heights.
int apply_schedule(mepgl_ctx* c) {
  for (auto& s : c->strings) {
    DevString& d = s.dev;
    d.tiles_y = (d.N + d.seg_rows - 1) / d.seg_rows;
    s.ntiles = d.tiles_x * d.tiles_y;
  }
  destroy_graphs(c);
  int rc;
  if ((rc = upload_descriptors(c)) || (rc = build_items(c))) return rc;
  return MEPGL_OK;
}

// Device copy of everything a step reads or writes, so that timed trial
// steps leave no trace.
struct Snapshot {
  std::vector<std::pair<void*, void*>> buf;  // (copy, live)
  std::vector<size_t> bytes;
  void* sc = nullptr;
  int xp = 0, dp = 0;
  bool pending = false, ls_deferred = false;
};

void drop_snapshot(mepgl_ctx* c, Snapshot& sn) {
  (void)hipStreamSynchronize(c->stream);
  for (auto& b : sn.buf) (void)hipFree(b.first);
  if (sn.sc) (void)hipFree(sn.sc);
  sn.buf.clear();
  sn.sc = nullptr;
}

int take_snapshot(mepgl_ctx* c, Snapshot& sn) {
  auto save = [&](void* live, size_t bytes) -> int {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess)
      return fail(c, MEPGL_ERR_NOMEM, "autotune needs a copy of the solver state (" +
                                          std::to_string(bytes) + " bytes)");
    sn.buf.push_back({p, live});
    sn.bytes.push_back(bytes);
    HIP_TRY(c, hipMemcpyAsync(p, live, bytes, hipMemcpyDeviceToDevice, c->stream));
    return MEPGL_OK;
  };
  int rc;
  for (auto& s : c->strings) {
    const size_t fs = (size_t)s.dev.F * s.dev.image_stride * s.elem;
    for (int b = 0; b < 2; ++b)
      if ((rc = save(s.dev.x[b], fs)) || (rc = save(s.dev.d[b], fs)) || (rc = save(s.dev.g[b], fs)))
        return rc;
    // The reparametrisation plans and spline factors: timed outer iterations
    // recompute them, and a later redistribution must use the caller's.
    const size_t F = (size_t)s.dev.F;
    if ((rc = save(s.dev.plan_w, F * (2 * sizeof(int) + 2 * sizeof(double)))) ||
        (rc = save(s.dev.cubic, F * 4 * sizeof(double))))
      return rc;
  }
  const size_t scb = sizeof(ImageScalars) * c->total_images;
  if (hipMalloc(&sn.sc, scb) != hipSuccess) return fail(c, MEPGL_ERR_NOMEM, "autotune scalars");
  HIP_TRY(c, hipMemcpyAsync(sn.sc, c->d_sc, scb, hipMemcpyDeviceToDevice, c->stream));
  sn.xp = c->xp;
  sn.dp = c->dp;
  sn.pending = c->pending;
  sn.ls_deferred = c->ls_deferred;
  return MEPGL_OK;
}

int restore_snapshot(mepgl_ctx* c, const Snapshot& sn) {
  for (size_t k = 0; k < sn.buf.size(); ++k)
    HIP_TRY(c, hipMemcpyAsync(sn.buf[k].second, sn.buf[k].first, sn.bytes[k],
                              hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(c->d_sc, sn.sc, sizeof(ImageScalars) * c->total_images,
                            hipMemcpyDeviceToDevice, c->stream));
  c->xp = sn.xp;
  c->dp = sn.dp;
  c->pending = sn.pending;
  c->ls_deferred = sn.ls_deferred;
  return MEPGL_OK;
}

// Copies the live solver state into an existing snapshot (same buffers).
int refresh_snapshot(mepgl_ctx* c, Snapshot& sn) {
  for (size_t k = 0; k < sn.buf.size(); ++k)
    HIP_TRY(c, hipMemcpyAsync(sn.buf[k].first, sn.buf[k].second, sn.bytes[k],
                              hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(sn.sc, c->d_sc, sizeof(ImageScalars) * c->total_images,
                            hipMemcpyDeviceToDevice, c->stream));
  sn.xp = c->xp;
  sn.dp = c->dp;
  sn.pending = c->pending;
  sn.ls_deferred = c->ls_deferred;
  return MEPGL_OK;
}

// Timing rounds per candidate height.
constexpr int kTuneRounds = 5;

// Milliseconds per NLCG step of the current schedule, starting from the
// snapshot: one untimed step, then an outer iteration of mode L with `steps`
// steps (arclength, linear redistribution, the steps) when every string has
// two images or more, else `steps` steps.  The reparametrisation kernels run
// on the same work items as the steps, so a height is timed on what the
// relaxation loop does: with very tall segments the steps alone can time
// well while the arclength and redistribution launches have too few work
// items.
int time_schedule(mepgl_ctx* c, const Snapshot& sn, int steps, double* ms) {
  int rc;
  if ((rc = apply_schedule(c)) || (rc = restore_snapshot(c, sn))) return rc;
  if ((rc = nlcg_step(c, 0, 1, 0))) return rc;
  bool reparam = true;
  for (auto& s : c->strings) reparam = reparam && s.dev.F >= 2;
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(c, MEPGL_ERR_HIP, "autotune timing: event creation failed");
  }
  hipError_t e = hipEventRecord(e0, c->stream);
  if (e == hipSuccess && reparam) {
    rc = outer_iteration(c, steps, 'L', 1);
  } else {
    for (int k = 0; k < steps && !rc && e == hipSuccess; ++k)
      rc = nlcg_step(c, k + 1, 1, 0, k + 1 < steps);
  }
  if (!rc && e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (e != hipSuccess) return fail(c, MEPGL_ERR_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
  *ms = (double)t / std::max(steps, 1);
  return MEPGL_OK;
}

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

int mepgl_create(const mepgl_string_desc* strings, int n_strings, int precision, unsigned flags,
                 int device, mepgl_ctx** out) {
  if (!out) return MEPGL_ERR_ARG;
  *out = nullptr;
  if (!strings || n_strings <= 0) return MEPGL_ERR_ARG;
  if (precision != MEPGL_FP64 && precision != MEPGL_MIXED) return MEPGL_ERR_ARG;
  // The step kernels address one image's fields through a buffer resource
  // with 32-bit byte offsets: every image's field set must stay below 2 GiB.
  // Checked before any HIP call.
  for (int i = 0; i < n_strings; ++i) {
    const long long N = strings[i].N;
    if (N < 2 || N > (1 << 20) || (strings[i].ncomp != 1 && strings[i].ncomp != 2))
      continue;  // rejected with a message below
    const long long elem = precision == MEPGL_MIXED ? 4 : 8;
    const long long image_bytes =
        (N + kPlanePadRows) * (2 + 2 * strings[i].ncomp) * ((N + 63) / 64 * 64) * elem;
    if (image_bytes > 0x7fffffffLL) return MEPGL_ERR_ARG;
  }
  std::unique_ptr<mepgl_ctx> c(new mepgl_ctx());
  c->precision = precision;
  c->flags = flags;
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return MEPGL_ERR_HIP;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return MEPGL_ERR_HIP;
  bool made = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) == hipSuccess;
  for (int p = 1; p < kMaxParts && made; ++p)
    made = hipStreamCreateWithFlags(&c->xs[p], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&c->join_evs[p], hipEventDisableTiming) == hipSuccess;
  if (!made) {
    mepgl_destroy(c.release());
    return MEPGL_ERR_HIP;
  }
  if (hipHostMalloc((void**)&c->health, 64, hipHostMallocDefault) != hipSuccess) {
    c->health = nullptr;
    mepgl_destroy(c.release());
    return MEPGL_ERR_NOMEM;
  }
  *(volatile int*)c->health = 0;
  for (auto& ev : c->iter_done)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
      mepgl_destroy(c.release());
      return MEPGL_ERR_HIP;
    }

  long long total_img = 0;
  for (int i = 0; i < n_strings; ++i) total_img += std::max(strings[i].F, 0);

  c->strings.resize(n_strings);
  for (int si = 0; si < n_strings; ++si) {
    const mepgl_string_desc& in = strings[si];
    const mepgl_param* ps[] = {&in.a_1, &in.b_1, &in.m_xx_1, &in.m_yy_1, &in.a_2,
                               &in.b_2, &in.m_xx_2, &in.m_yy_2, &in.h};
    bool ok = in.N >= 2 && in.F >= 1 && in.F <= kMaxImages && (in.ncomp == 1 || in.ncomp == 2) && in.dx > 0 &&
              std::isfinite(in.dx) && in.comp_domain && in.sc_domain;
    for (auto* p : ps) ok = ok && param_ok(*p);
    if (!ok) {
      mepgl_ctx* raw = c.release();
      *out = raw;
      return fail(raw, MEPGL_ERR_ARG, "invalid string descriptor " + std::to_string(si));
    }
    StringHost& s = c->strings[si];
    s.desc = in;
    s.elem = c->fp32() ? 4 : 8;
    DevString& d = s.dev;
    std::memset(&d, 0, sizeof(d));
    d.N = in.N;
    d.pitch = (in.N + 63) / 64 * 64;
    d.nc = in.ncomp;
    d.nplanes = 2 + 2 * in.ncomp;
    d.nsf = in.no_self_field ? 1 : 0;
    d.a_shared = d.nsf;  // the fields start at zero in every image
    d.poly_exact = (flags & MEPGL_POLY_EXACT) ? 1 : 0;
    d.seg_rows = choose_rows(in, total_img);
    d.tiles_x = (in.N + kTileX - 1) / kTileX;
    d.tiles_y = (in.N + d.seg_rows - 1) / d.seg_rows;
    s.ntiles = d.tiles_x * d.tiles_y;
    // Planes interleaved by row: row j of an image holds row j of every
    // plane, one after another (plane stride = one plane row), so the planes
    // a wavefront reads at one row lie in one contiguous span.  One padding
    // row per image keeps the image stride off powers of two, so the streams
    // of neighbouring images do not fall on the same HBM channels.
    const int plane_row = d.pitch;
    d.pitch = d.nplanes * plane_row;
    d.plane_stride = plane_row;
    d.image_stride = (long long)(in.N + kPlanePadRows) * d.pitch;
    d.dx = in.dx;
    d.inv_dx = 1.0 / in.dx;
    d.dx2 = in.dx * in.dx;
    d.q[0] = in.q_1;
    d.q[1] = in.q_2;
    d.eta = in.eta;
    d.gamma = in.gamma;
    d.delta = in.delta;
    const mepgl_param* comp_p[2][4] = {{&in.a_1, &in.b_1, &in.m_xx_1, &in.m_yy_1},
                                       {&in.a_2, &in.b_2, &in.m_xx_2, &in.m_yy_2}};
    d.planes = 0;
    for (int q = 0; q < 2; ++q) {
      for (int k = 0; k < 4; ++k) d.planes |= comp_p[q][k]->plane != nullptr;
      const double a = comp_p[q][0]->value, b = comp_p[q][1]->value;
      const double mx = comp_p[q][2]->value, my = comp_p[q][3]->value;
      d.kc[q][0] = q == 0 ? in.q_1 : in.q_2;
      d.kc[q][1] = 2.0 * a;
      d.kc[q][2] = 2.0 * b;
      d.kc[q][3] = b;
      d.kc[q][4] = a / b;
      d.kc[q][5] = 0.5 / mx;
      d.kc[q][6] = 0.5 / my;
    }
    int rc;
    mepgl_ctx* cp = c.get();
#define UP(field, dst)                                          \
  if ((rc = upload_param(cp, s, in.field, dst))) {              \
    *out = c.release();                                         \
    return rc;                                                  \
  }
    UP(a_1, d.a[0]) UP(b_1, d.b[0]) UP(m_xx_1, d.mxx[0]) UP(m_yy_1, d.myy[0])
    UP(a_2, d.a[1]) UP(b_2, d.b[1]) UP(m_xx_2, d.mxx[1]) UP(m_yy_2, d.myy[1])
    UP(h, d.h)
#undef UP
    // Neighbour maps.
    void *dsc = nullptr, *dcm = nullptr, *nn = nullptr;
    const size_t db = (size_t)in.N * in.N * sizeof(int32_t);
    if ((rc = dev_alloc(cp, &s, &nn, (size_t)in.N * d.pitch * sizeof(uint16_t))) ||
        (rc = dev_alloc(cp, nullptr, &dsc, db)) || (rc = dev_alloc(cp, nullptr, &dcm, db))) {
      *out = c.release();
      return rc;
    }
    hipError_t he = hipMemsetAsync(nn, 0, (size_t)in.N * d.pitch * sizeof(uint16_t), c->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(dsc, in.sc_domain, db, hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess)
      he = hipMemcpyAsync(dcm, in.comp_domain, db, hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess)
      he = launch_nn_map((const int*)dsc, (const int*)dcm, in.N, d.pitch, (uint16_t*)nn, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he == hipSuccess) {
      s.nn_host.resize((size_t)in.N * d.pitch);
      he = hipMemcpy(s.nn_host.data(), nn, s.nn_host.size() * sizeof(uint16_t),
                     hipMemcpyDeviceToHost);
    }
    (void)hipFree(dsc);
    (void)hipFree(dcm);
    c->bytes -= 2 * (long long)db;
    d.nn = (const uint16_t*)nn;
    if (he != hipSuccess) {
      *out = c.release();
      return fail(*out, MEPGL_ERR_HIP, std::string("neighbour map: ") + hipGetErrorString(he));
    }
    if ((rc = alloc_images(cp, s, in.F))) {
      *out = c.release();
      return rc;
    }
  }
  mepgl_ctx* cp = c.get();
  int rc;
  if ((rc = alloc_scalars(cp)) || (rc = upload_descriptors(cp)) || (rc = build_items(cp))) {
    *out = c.release();
    return rc;
  }
  LineSearchTable* t = new LineSearchTable;
  build_line_search_table(t);
  if (hipMalloc(&c->d_ls, sizeof(LineSearchTable)) != hipSuccess) {
    delete t;
    *out = c.release();
    return fail(*out, MEPGL_ERR_NOMEM, "line-search table");
  }
  hipError_t he = hipMemcpyAsync(c->d_ls, t, sizeof(LineSearchTable), hipMemcpyHostToDevice,
                                 c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  delete t;
  if (he != hipSuccess) {
    *out = c.release();
    return fail(*out, MEPGL_ERR_HIP, std::string("line-search table: ") + hipGetErrorString(he));
  }
  if ((rc = reset_solver(cp))) {
    *out = c.release();
    return rc;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    *out = c.release();
    return fail(*out, MEPGL_ERR_HIP, "initialisation");
  }
  *out = c.release();
  return MEPGL_OK;
}

void mepgl_destroy(mepgl_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);  // best effort
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_graphs(c);
  for (auto& s : c->strings) free_string(c, s);
  (void)hipFree(c->d_strings);
  (void)hipFree(c->d_sc);
  (void)hipFree(c->d_img);
  (void)hipFree(c->d_ls);
  (void)hipFree(c->d_items);
  (void)hipFree(c->d_jb);
  (void)hipFree(c->d_out);
  for (auto e : c->timing.pool) (void)hipEventDestroy(e);
  for (auto e : c->iter_done)
    if (e) (void)hipEventDestroy(e);
  if (c->health) (void)hipHostFree(c->health);
  for (int* p : c->d_part_imgs)
    if (p) (void)hipFree(p);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  for (int p = 1; p < kMaxParts; ++p) {
    if (c->join_evs[p]) (void)hipEventDestroy(c->join_evs[p]);
    if (c->xs[p]) (void)hipStreamDestroy(c->xs[p]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* mepgl_last_error(const mepgl_ctx* c) { return c ? c->err.c_str() : "no context"; }

int mepgl_upload(mepgl_ctx* c, int si, const double* a, const double* p1, const double* p2) {
  if (!valid_string(c, si) || !a || !p1) return fail(c, MEPGL_ERR_ARG, "upload: bad arguments");
  StringHost& s = c->strings[si];
  if (s.dev.nc == 2 && !p2) return fail(c, MEPGL_ERR_ARG, "upload: psi_2 required");
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);  // other strings keep their owed update
  if (rc) return rc;
  // A NO_SELF_FIELD string whose A is the same in every image keeps one A
  // for all of them (image 0's planes); any other keeps A per image.
  if (s.dev.nsf) {
    const size_t img = (size_t)s.dev.N * s.dev.N * 2;
    bool same = true;
    for (int k = 1; k < s.dev.F && same; ++k)
      same = std::memcmp(a, a + (size_t)k * img, img * sizeof(double)) == 0;
    if ((int)same != s.dev.a_shared) {
      s.dev.a_shared = same ? 1 : 0;
      destroy_graphs(c);
      if ((rc = upload_descriptors(c))) return rc;
    }
  }
  const size_t n = (size_t)s.dev.F * s.dev.image_stride;
  for (int b = 0; b < 2; ++b) {
    if (c->fp32()) {
      std::vector<float> h;
      pack_fields(s, a, p1, s.dev.nc == 2 ? p2 : nullptr, h);
      HIP_TRY(c, hipMemcpyAsync(s.dev.x[b], h.data(), n * sizeof(float), hipMemcpyHostToDevice,
                                c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
      std::vector<double> h;
      pack_fields(s, a, p1, s.dev.nc == 2 ? p2 : nullptr, h);
      HIP_TRY(c, hipMemcpyAsync(s.dev.x[b], h.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
  }
  return MEPGL_OK;
}

int mepgl_download(mepgl_ctx* c, int si, double* a, double* p1, double* p2) {
  if (!valid_string(c, si) || !a || !p1) return fail(c, MEPGL_ERR_ARG, "download: bad arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  return copy_out_fields(c, si, c->strings[si].dev.x[c->xp], a, p1, p2);
}

int mepgl_reset_solver(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  if ((rc = reset_solver(c))) return rc;
  // The current fields and direction become buffer 0 (pointer swaps in the
  // descriptors; captured graphs reach the buffers through them), so that
  // every reset leaves the same buffer parity and the loop after it replays
  // the same graphs, whatever ran before.
  if (c->xp || c->dp) {
    for (auto& s : c->strings) {
      if (c->xp) std::swap(s.dev.x[0], s.dev.x[1]);
      if (c->dp) std::swap(s.dev.d[0], s.dev.d[1]);
    }
    c->xp = c->dp = 0;
    if ((rc = upload_descriptors(c))) return rc;
  }
  return MEPGL_OK;
}

int mepgl_winding(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  for (auto& s : c->strings)
    if (s.dev.N < 2) return fail(c, MEPGL_ERR_ARG, "winding needs N >= 2");
  int rc = flush(c);
  if (rc) return rc;
  return run_finish(c, finish_args(c), kKernelWinding);
}

int mepgl_free_energy(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  AuxArgs a = aux_args(c);
  if ((rc = run_aux(c, a, kKernelEnergy, 0))) return rc;
  return run_finish(c, finish_args(c), kKernelEnergy);
}

int mepgl_gradient(mepgl_ctx* c, int si, double* a_g, double* p1_g, double* p2_g) {
  if (!valid_string(c, si) || !a_g || !p1_g) return fail(c, MEPGL_ERR_ARG, "gradient: bad arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  bool changed = false;
  for (auto& s : c->strings) {
    if (s.dev.gs) continue;
    const size_t fs = (size_t)s.dev.F * s.dev.image_stride * s.elem;
    if ((rc = dev_alloc(c, &s, &s.dev.gs, fs))) return rc;
    HIP_TRY(c, hipMemsetAsync(s.dev.gs, 0, fs, c->stream));
    changed = true;
  }
  if (changed && (rc = upload_descriptors(c))) return rc;
  StepArgs a = step_args(c);
  a.xin = a.xout = c->xp;
  a.din = c->dp;
  a.scratch = 1;
  if ((rc = run_step(c, a, kKernelGrad))) return rc;
  return copy_out_fields(c, si, c->strings[si].dev.gs, a_g, p1_g, p2_g);
}

int mepgl_nlcg_steps(mepgl_ctx* c, int iter_n, int nlcg, int maxwell) {
  if (!c || iter_n < 0) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  for (int k = 0; k < iter_n; ++k) {
    int rc = nlcg_step(c, k, nlcg ? 1 : 0, maxwell ? 1 : 0, k + 1 < iter_n);
    if (rc) return rc;
  }
  return check_finite(c, " after nlcg_steps");
}

int mepgl_compute_jb(mepgl_ctx* c, int si, double* b, double* j1, double* j2) {
  if (!valid_string(c, si) || !b || !j1) return fail(c, MEPGL_ERR_ARG, "compute_jb: bad arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  bool changed = false;
  for (auto& s : c->strings) {
    if (s.jb) continue;
    void* p;
    // b and j in the fields' layout (planes 0, 1 + 2c, 2 + 2c), in double.
    if ((rc = dev_alloc(c, &s, &p, (size_t)s.dev.F * s.dev.image_stride * sizeof(double))))
      return rc;
    s.jb = (double*)p;
    changed = true;
  }
  if (changed && (rc = upload_descriptors(c))) return rc;
  AuxArgs a = aux_args(c);
  if ((rc = run_aux(c, a, kKernelJB, 0))) return rc;
  StringHost& s = c->strings[si];
  const DevString& d = s.dev;
  std::vector<double> h((size_t)d.F * d.image_stride);
  HIP_TRY(c, hipMemcpyAsync(h.data(), s.jb, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int N = d.N, P = d.pitch;
  for (int n = 0; n < d.F; ++n) {
    const double* base = h.data() + (size_t)n * d.image_stride;
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        const size_t o = (size_t)j * P + i, q = (size_t)n * N * N + (size_t)j * N + i;
        b[q] = base[o];
        j1[2 * q] = base[d.plane_stride + o];
        j1[2 * q + 1] = base[2 * d.plane_stride + o];
        if (j2) {
          j2[2 * q] = d.nc == 2 ? base[3 * d.plane_stride + o] : 0.0;
          j2[2 * q + 1] = d.nc == 2 ? base[4 * d.plane_stride + o] : 0.0;
        }
      }
  }
  return MEPGL_OK;
}

int mepgl_arc_length(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  return arc_length_all(c);
}

int mepgl_redistribute(mepgl_ctx* c, int cubic) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc;
  if (cubic && (rc = ensure_cubic_rhs(c))) return rc;
  return redistribute_all(c, cubic);
}

int mepgl_gauge_align(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  return gauge_align_all(c);
}

int mepgl_gauge_align_local(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  return gauge_align_local_all(c);
}

int mepgl_reset_interior_beta(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  AuxArgs a = aux_args(c);
  return run_aux(c, a, kKernelResetNorms, 1);
}

int mepgl_resample(mepgl_ctx* c, int si, int F_new) {
  if (!valid_string(c, si) || F_new < 2 || F_new > kMaxImages)
    return fail(c, MEPGL_ERR_ARG, "resample: bad arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = arc_length_all(c);
  if (rc) return rc;
  StringHost& s = c->strings[si];
  const int F = s.dev.F;
  std::vector<ImageScalars> h(c->total_images);
  HIP_TRY(c, hipMemcpyAsync(h.data(), c->d_sc, sizeof(ImageScalars) * h.size(),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  std::vector<double> S(F), sp(F), sj(F);
  for (int n = 0; n < F; ++n) {
    S[n] = h[s.dev.img0 + n].s_jb;
    sp[n] = h[s.dev.img0 + n].s;
  }
  // generateNewFramesLinear's segment search (interpolation.cu:200-219).
  std::vector<int> pk(F_new);
  std::vector<double> pw(F_new);
  int k = 1;
  double seg_start = 0, seg_end = S[1];
  for (int n = 0; n < F_new; ++n) {
    const double s_new = (double)n / ((double)F_new - 1);
    while (seg_end < s_new && k + 1 < F) {
      ++k;
      seg_start = S[k - 1];
      seg_end = S[k];
    }
    pk[n] = k;
    pw[n] = (s_new - seg_start) / (S[k] - S[k - 1]);
  }
  // Old fields to host, resample there in double with the kernel's formula,
  // then rebuild the string's buffers for F_new images.
  std::vector<double> fa((size_t)F * s.dev.N * s.dev.N * 2), f1(fa.size()), f2(fa.size());
  if ((rc = copy_out_fields(c, si, s.dev.x[c->xp], fa.data(), f1.data(), f2.data()))) return rc;
  const size_t img = (size_t)s.dev.N * s.dev.N * 2;
  std::vector<double> na(img * F_new), n1(img * F_new), n2(img * F_new);
  for (int n = 0; n < F_new; ++n) {
    const double w = pw[n];
    for (size_t e = 0; e < img; ++e) {
      const size_t o1 = (size_t)pk[n] * img + e, o0 = o1 - img;
      na[n * img + e] = std::fma(w, fa[o1], (1 - w) * fa[o0]);
      n1[n * img + e] = std::fma(w, f1[o1], (1 - w) * f1[o0]);
      n2[n * img + e] = std::fma(w, f2[o1], (1 - w) * f2[o0]);
    }
  }
  // Free per-image buffers of this string (keep nn and parameter planes).
  std::vector<Alloc> keep;
  for (auto& a : s.owned) {
    const void* p = a.p;
    bool shared = p == (const void*)s.dev.nn;
    const DevParam* prm[] = {&s.dev.a[0], &s.dev.b[0], &s.dev.mxx[0], &s.dev.myy[0], &s.dev.a[1],
                             &s.dev.b[1], &s.dev.mxx[1], &s.dev.myy[1], &s.dev.h};
    for (auto* q : prm) shared = shared || p == (const void*)q->p;
    if (shared) {
      keep.push_back(a);
    } else {
      (void)hipFree(a.p);
      c->bytes -= (long long)a.bytes;
    }
  }
  s.owned = keep;
  if ((rc = alloc_images(c, s, F_new))) return rc;
  s.desc.F = F_new;
  destroy_graphs(c);
  // The other strings keep their solver state: their per-image scalars move
  // to their new global image ids, and their gradients and directions are
  // untouched.  String si starts from loop()'s initial state (zero gradients
  // and direction from alloc_images, infinite norms).
  std::vector<int> old_img0;
  for (auto& t : c->strings) old_img0.push_back(t.dev.img0);
  if ((rc = alloc_scalars(c))) return rc;
  std::vector<ImageScalars> hn(c->total_images);
  for (size_t k = 0; k < c->strings.size(); ++k) {
    if ((int)k == si) continue;
    const DevString& d = c->strings[k].dev;
    for (int n = 0; n < d.F; ++n) hn[d.img0 + n] = h[old_img0[k] + n];
  }
  HIP_TRY(c, hipMemcpyAsync(c->d_sc, hn.data(), sizeof(ImageScalars) * hn.size(),
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if ((rc = upload_descriptors(c)) || (rc = build_items(c))) return rc;
  // arc_length_all applied any owed update, so nothing is pending here.
  if ((rc = mepgl_upload(c, si, na.data(), n1.data(), n2.data()))) return rc;
  AuxArgs ra = aux_args(c);
  ra.only_string = si;
  if ((rc = run_aux(c, ra, kKernelResetNorms, 0))) return rc;
  return arc_length_all(c);
}

int mepgl_loop(mepgl_ctx* c, int iterations, int relax, char mode, int nlcg) {
  if (!c || iterations < 0 || relax < 0) return MEPGL_ERR_ARG;
  if (mode != 'S' && mode != 'M' && mode != 'L' && mode != 'C')
    return fail(c, MEPGL_ERR_ARG, std::string("unknown mode ") + mode);
  HIP_TRY(c, hipSetDevice(c->device));
  int rc;
  // Buffers an outer iteration needs are allocated here, outside any capture.
  const bool local = (c->flags & MEPGL_GAUGE_ALIGN_LOCAL) && (mode == 'L' || mode == 'C');
  if ((mode == 'C' || local) && (rc = ensure_cubic_rhs(c))) return rc;
  if (local && (rc = ensure_dct(c))) return rc;
  const bool use_graph = !(c->flags & MEPGL_NO_GRAPH) && !c->timing.on;
  const bool rollback = (c->flags & MEPGL_ROLLBACK) != 0;
  // Non-finite guard, once per outer iteration.  Without rollback the check
  // of iteration it runs while iteration it + 1 is already queued, so the
  // device never waits for the host; with rollback the solver state is
  // copied before every iteration and checked right after it.
  // A flag left by calls that do not report (the autotuner's trial steps)
  // is stale here.
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *(volatile int*)c->health = 0;
  Snapshot sn;
  if (rollback && (rc = take_snapshot(c, sn))) {
    drop_snapshot(c, sn);
    return rc;
  }
  auto bad_iteration = [&](int it) {
    std::string msg = " in outer iteration " + std::to_string(it);
    if (rollback) {
      if (restore_snapshot(c, sn) == MEPGL_OK && hipStreamSynchronize(c->stream) == hipSuccess)
        msg += "; the solver state was rolled back to the start of that iteration";
      drop_snapshot(c, sn);
    }
    *(volatile int*)c->health = 0;
    return fail(c, MEPGL_ERR_NONFINITE,
                "non-finite energy or gradient norm" + msg + first_nonfinite_image(c));
  };
  int it = 0;
  for (; it < iterations; ++it) {
    if (it > 0 && c->stop.load()) break;
    if (rollback && it > 0 && (rc = refresh_snapshot(c, sn))) break;
    if (!use_graph) {
      if ((rc = outer_iteration(c, relax, mode, nlcg ? 1 : 0))) break;
    } else {
      // Replay a captured outer iteration.  The capture depends on the parity
      // state it starts from, which repeats after one or two iterations.
      auto key = std::make_tuple(c->xp, c->dp, (int)c->pending, relax, nlcg ? 1 : 0, mode);
      auto itg = c->graphs.find(key);
      if (itg == c->graphs.end()) {
        const int xp0 = c->xp, dp0 = c->dp;
        const bool pend0 = c->pending;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
          rc = fail(c, MEPGL_ERR_HIP, "graph capture failed to start");
          break;
        }
        rc = outer_iteration(c, relax, mode, nlcg ? 1 : 0);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(c->stream, &g);
        c->xp = xp0;
        c->dp = dp0;
        c->pending = pend0;
        c->ls_deferred = false;
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          break;
        }
        if (e != hipSuccess) {
          rc = fail(c, MEPGL_ERR_HIP, "graph capture failed");
          break;
        }
        hipGraphExec_t ex;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          rc = fail(c, MEPGL_ERR_HIP, "graph instantiate failed");
          break;
        }
        itg = c->graphs.emplace(key, ex).first;
      }
      hipError_t e = hipGraphLaunch(itg->second, c->stream);
      if (e != hipSuccess) {
        rc = fail(c, MEPGL_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e));
        break;
      }
      // Advance the parity state exactly as the captured sequence does.
      c->dry = true;
      rc = outer_iteration(c, relax, mode, nlcg ? 1 : 0);
      c->dry = false;
      if (rc) break;
    }
    hipEvent_t done = c->iter_done[it & 1];
    hipError_t e = hipEventRecord(done, c->stream);
    if (e == hipSuccess && rollback) e = hipEventSynchronize(done);
    if (e != hipSuccess) {
      rc = fail(c, MEPGL_ERR_HIP, std::string("iteration event: ") + hipGetErrorString(e));
      break;
    }
    if (rollback) {
      if (*(volatile int*)c->health) return bad_iteration(it);
    } else if (it > 0) {
      // Iteration it - 1 is done once its event is; iteration it runs.
      if ((e = hipEventSynchronize(c->iter_done[(it - 1) & 1])) != hipSuccess) {
        rc = fail(c, MEPGL_ERR_HIP, std::string("iteration event: ") + hipGetErrorString(e));
        break;
      }
      if (*(volatile int*)c->health) {
        (void)hipStreamSynchronize(c->stream);
        return bad_iteration(it - 1);
      }
    }
  }
  if (rollback) drop_snapshot(c, sn);
  if (rc) return rc;
  // The last iteration (checked right after it with rollback).
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (!rollback && it > 0 && *(volatile int*)c->health) return bad_iteration(it - 1);
  return check_finite(c, " after the loop");
}

int mepgl_get_scalars(mepgl_ctx* c, int si, int which, double* out) {
  if (!valid_string(c, si) || !out) return fail(c, MEPGL_ERR_ARG, "get_scalars: bad arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = settle_ls(c)) return rc;
  const DevString& d = c->strings[si].dev;
  std::vector<ImageScalars> h(d.F);
  HIP_TRY(c, hipMemcpyAsync(h.data(), c->d_sc + d.img0, sizeof(ImageScalars) * d.F,
                            hipMemcpyDeviceToHost, c->stream));
  std::vector<double> e(d.F);
  if (which == MEPGL_SCALAR_ENERGY)
    HIP_TRY(c, hipMemcpyAsync(e.data(), c->d_out + d.img0, sizeof(double) * d.F,
                              hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int n = 0; n < d.F; ++n) {
    const ImageScalars& x = h[n];
    switch (which) {
      case MEPGL_SCALAR_FENERGY: out[n] = x.fenergy; break;
      case MEPGL_SCALAR_G_NORM2: out[n] = x.g_norm2[0]; break;
      case MEPGL_SCALAR_G_NORM2_OLD: out[n] = x.g_norm2[1]; break;
      case MEPGL_SCALAR_ALPHA: out[n] = x.alpha; break;
      case MEPGL_SCALAR_S: out[n] = x.s; break;
      case MEPGL_SCALAR_S_JB: out[n] = x.s_jb; break;
      case MEPGL_SCALAR_BETA: out[n] = x.beta; break;
      case MEPGL_SCALAR_G_DELTA_G: out[n] = x.g_delta_g; break;
      case MEPGL_SCALAR_LS_INDEX: out[n] = x.ls_index; break;
      case MEPGL_SCALAR_THETA: out[n] = x.theta; break;
      case MEPGL_SCALAR_WINDING_1: out[n] = x.winding[0]; break;
      case MEPGL_SCALAR_WINDING_2: out[n] = x.winding[1]; break;
      case MEPGL_SCALAR_ENERGY: out[n] = e[n]; break;
      case MEPGL_SCALAR_POLY:
        for (int q = 0; q < 4; ++q) out[4 * n + q] = x.poly[q];
        break;
      default:
        return fail(c, MEPGL_ERR_ARG, "get_scalars: unknown scalar");
    }
  }
  return MEPGL_OK;
}

int mepgl_download_state(mepgl_ctx* c, int si, int which, double* a, double* p1, double* p2) {
  if (!valid_string(c, si) || !a || !p1 || which < 0 || which > 2)
    return fail(c, MEPGL_ERR_ARG, "download_state: bad arguments");
  HIP_TRY(c, hipSetDevice(c->device));
  const DevString& d = c->strings[si].dev;
  const void* src = which == 0 ? d.g[0] : which == 1 ? d.g[1] : d.d[c->dp];
  return copy_out_fields(c, si, src, a, p1, p2);
}

int mepgl_upload_state(mepgl_ctx* c, int si, int which, const double* a, const double* p1,
                       const double* p2) {
  if (!valid_string(c, si) || !a || !p1 || which < 0 || which > 2)
    return fail(c, MEPGL_ERR_ARG, "upload_state: bad arguments");
  if (c->strings[si].dev.nc == 2 && !p2) return fail(c, MEPGL_ERR_ARG, "upload_state: psi_2 required");
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);  // an owed update uses the direction being replaced
  if (rc) return rc;
  const DevString& d = c->strings[si].dev;
  void* dst = which == 0 ? d.g[0] : which == 1 ? d.g[1] : d.d[c->dp];
  return copy_in_fields(c, si, dst, a, p1, p2);
}

int mepgl_set_scalars(mepgl_ctx* c, int si, int which, const double* in) {
  if (!valid_string(c, si) || !in) return fail(c, MEPGL_ERR_ARG, "set_scalars: bad arguments");
  if (which != MEPGL_SCALAR_G_NORM2 && which != MEPGL_SCALAR_G_NORM2_OLD &&
      which != MEPGL_SCALAR_S && which != MEPGL_SCALAR_S_JB)
    return fail(c, MEPGL_ERR_ARG, "set_scalars: only g_norm2, g_norm2_old, s and s_jb are settable");
  HIP_TRY(c, hipSetDevice(c->device));
  const DevString& d = c->strings[si].dev;
  std::vector<ImageScalars> h(d.F);
  HIP_TRY(c, hipMemcpyAsync(h.data(), c->d_sc + d.img0, sizeof(ImageScalars) * d.F,
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int n = 0; n < d.F; ++n) {
    switch (which) {
      case MEPGL_SCALAR_G_NORM2: h[n].g_norm2[0] = in[n]; break;
      case MEPGL_SCALAR_G_NORM2_OLD: h[n].g_norm2[1] = in[n]; break;
      case MEPGL_SCALAR_S: h[n].s = in[n]; break;
      default: h[n].s_jb = in[n]; break;
    }
  }
  HIP_TRY(c, hipMemcpyAsync(c->d_sc + d.img0, h.data(), sizeof(ImageScalars) * d.F,
                            hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MEPGL_OK;
}

int mepgl_set_control(mepgl_ctx* c, int key, int value) {
  if (!c) return MEPGL_ERR_ARG;
  switch (key) {
    case MEPGL_CONTROL_STOP: c->stop.store(value); break;
    case MEPGL_CONTROL_RELAXATION_STEPS: c->relax_steps.store(value); break;
    case MEPGL_CONTROL_NLCG: c->nlcg.store(value); break;
    case MEPGL_CONTROL_FOLD_LINE_SEARCH:
      if (c->fold_ls != (value != 0)) {
        c->fold_ls = value != 0;
        destroy_graphs(c);
      }
      break;
    case MEPGL_CONTROL_SPLIT_IMAGES:
      if (value < 0 || value > kMaxParts) return MEPGL_ERR_ARG;
      if (value >= 2 && value != c->parts) {
        HIP_TRY(c, hipSetDevice(c->device));
        c->parts = value;
        c->split_images = true;
        return apply_schedule(c);
      }
      if (c->split_images != (value != 0)) {
        c->split_images = value != 0;
        destroy_graphs(c);
      }
      break;
    case MEPGL_CONTROL_EDGE_EXCHANGE:
      if (c->edge_exchange != (value != 0)) {
        c->edge_exchange = value != 0;
        destroy_graphs(c);
      }
      break;
    default: return MEPGL_ERR_ARG;
  }
  return MEPGL_OK;
}

int mepgl_get_control(mepgl_ctx* c, int key, int* value) {
  if (!c || !value) return MEPGL_ERR_ARG;
  switch (key) {
    case MEPGL_CONTROL_STOP: *value = c->stop.load(); break;
    case MEPGL_CONTROL_RELAXATION_STEPS: *value = c->relax_steps.load(); break;
    case MEPGL_CONTROL_NLCG: *value = c->nlcg.load(); break;
    case MEPGL_CONTROL_FOLD_LINE_SEARCH: *value = c->fold_ls ? 1 : 0; break;
    case MEPGL_CONTROL_SPLIT_IMAGES: *value = c->split_images ? c->parts : 0; break;
    case MEPGL_CONTROL_EDGE_EXCHANGE: *value = c->edge_exchange ? 1 : 0; break;
    default: return MEPGL_ERR_ARG;
  }
  return MEPGL_OK;
}

int mepgl_profile(mepgl_ctx* c, int enable) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->timing.on = enable != 0;
  c->timing.used = 0;
  c->timing.marks.clear();
  for (int k = 0; k < kKernelCount; ++k) {
    c->timing.total_ms[k] = 0;
    c->timing.count[k] = 0;
  }
  return MEPGL_OK;
}

int mepgl_profile_read(mepgl_ctx* c, int kernel, double* avg_ms, int* count) {
  if (!c || kernel < 0 || kernel >= kKernelCount || !avg_ms || !count) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  Timing& t = c->timing;
  for (auto& m : t.marks) {
    float ms = 0;
    HIP_TRY(c, hipEventElapsedTime(&ms, t.pool[m.second.first], t.pool[m.second.second]));
    t.total_ms[m.first] += ms;
    t.count[m.first] += 1;
  }
  t.marks.clear();
  t.used = 0;
  *count = t.count[kernel];
  *avg_ms = t.count[kernel] ? t.total_ms[kernel] / t.count[kernel] : 0.0;
  return MEPGL_OK;
}

int mepgl_synchronize(mepgl_ctx* c) {
  if (!c) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MEPGL_OK;
}

int mepgl_set_segment_rows(mepgl_ctx* c, int si, int rows) {
  if (!valid_string(c, si) || rows < 1) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  StringHost& s = c->strings[si];
  s.dev.seg_rows = clamp_rows(rows, s.dev.N);
  return apply_schedule(c);
}

int mepgl_autotune(mepgl_ctx* c, int steps, int* rows_out) {
  if (!c || steps < 1) return MEPGL_ERR_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = flush(c);
  if (rc) return rc;
  // Classes of strings that share a kernel specialisation and a grid size,
  // tuned one after another, largest total work first.
  std::map<std::tuple<int, int, int>, std::vector<int>> cls;
  for (size_t si = 0; si < c->strings.size(); ++si) {
    const DevString& d = c->strings[si].dev;
    cls[std::make_tuple(d.N, d.nc, d.nsf)].push_back((int)si);
  }
  std::vector<std::pair<double, std::vector<int>>> order;
  for (auto& kv : cls) {
    double work = 0;
    for (int si : kv.second) {
      const DevString& d = c->strings[si].dev;
      work += (double)d.F * d.N * d.N * d.nplanes;
    }
    order.push_back({work, kv.second});
  }
  std::sort(order.begin(), order.end(),
            [](const auto& a, const auto& b) { return a.first > b.first; });
  Snapshot sn;
  if ((rc = take_snapshot(c, sn))) {
    drop_snapshot(c, sn);
    return rc;
  }
  const bool timing = c->timing.on;
  c->timing.on = false;
  // The heights are tuned for each part count of a split sequence (2 and
  // kMaxParts), and the count whose tuned schedule runs fastest is kept with
  // its heights.
  const int parts0 = c->parts;
  std::vector<int> part_cand{c->parts};
  if (c->split_images)
    for (int p : {2, kMaxParts})
      if (p != c->parts) part_cand.push_back(p);
  double best_score = INFINITY;
  int best_parts = c->parts;
  std::vector<int> best_rows(c->strings.size());
  for (int parts : part_cand) {
    c->parts = parts;
    // Candidate heights are timed in rounds that visit every candidate once,
    // and each keeps its fastest round: neighbouring heights often differ by a
    // few percent, about as much as one timing varies, and a drift of the
    // clock over the tuning run then favours no candidate.
    for (auto& cl : order) {
      const int N = c->strings[cl.second[0]].dev.N;
      std::vector<int> cand{c->strings[cl.second[0]].dev.seg_rows};
      for (int h : kRowChoices) {
        const int r = clamp_rows(h, N);
        if (h > N && h != kRowChoices[0]) continue;
        if (std::find(cand.begin(), cand.end(), r) == cand.end()) cand.push_back(r);
      }
      std::vector<double> best(cand.size(), INFINITY);
      for (int round = 0; round < kTuneRounds && !rc; ++round)
        for (size_t k = 0; k < cand.size() && !rc; ++k) {
          for (int si : cl.second) c->strings[si].dev.seg_rows = cand[k];
          double ms = 0;
          if ((rc = time_schedule(c, sn, steps, &ms))) break;
          best[k] = std::min(best[k], ms);
        }
      if (rc) break;
      const size_t pick = (size_t)(std::min_element(best.begin(), best.end()) - best.begin());
      for (int si : cl.second) c->strings[si].dev.seg_rows = cand[pick];
    }
    if (rc) break;
    // This part count's score: the fastest round of its tuned schedule.
    double score = INFINITY;
    for (int round = 0; round < kTuneRounds && !rc; ++round) {
      double ms = 0;
      if ((rc = time_schedule(c, sn, steps, &ms))) break;
      score = std::min(score, ms);
    }
    if (rc) break;
    if (score < best_score) {
      best_score = score;
      best_parts = parts;
      for (size_t si = 0; si < c->strings.size(); ++si) best_rows[si] = c->strings[si].dev.seg_rows;
    }
  }
  if (rc) {
    c->parts = parts0;
  } else {
    c->parts = best_parts;
    for (size_t si = 0; si < c->strings.size(); ++si) c->strings[si].dev.seg_rows = best_rows[si];
  }
  c->timing.on = timing;
  int rc2 = apply_schedule(c);
  if (rc2 == MEPGL_OK) rc2 = restore_snapshot(c, sn);
  drop_snapshot(c, sn);
  if (rc) return rc;
  if (rc2) return rc2;
  if (rows_out)
    for (size_t si = 0; si < c->strings.size(); ++si) rows_out[si] = c->strings[si].dev.seg_rows;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return MEPGL_OK;
}

int mepgl_info(mepgl_ctx* c, int si, int* tile_rows, int* work_items, long long* bytes) {
  if (!valid_string(c, si)) return MEPGL_ERR_ARG;
  if (tile_rows) *tile_rows = c->strings[si].dev.seg_rows;
  if (work_items) *work_items = c->nitems;
  if (bytes) *bytes = c->bytes;
  return MEPGL_OK;
}

}  // extern "C"
