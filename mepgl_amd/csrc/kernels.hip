// HIP kernels of the MI355X gauged-string-method relaxation loop (gfx950),
// other than the step kernels (step_kernels.hpp): the per-image finishing
// launches of an NLCG step, the energy and JB passes, arclength and
// reparametrisation, gauge alignment, scalar maintenance, and the host
// launchers of all kernels.
#include "row_window.hpp"

namespace mepgl {

// ---------------------------------------------------------------------------
// Finishing launches: one wavefront per image (global image id = block).
__device__ __forceinline__ const DevString& string_of_image(const FinishArgs& f, int& n) {
  const int2 si = f.img[blockIdx.x];
  n = si.y;
  return f.strings[si.x];
}

// After P: the quartic A..E and the line search (lineSearch).  One
// 256-thread workgroup per image.  Thread t holds the scan candidates
// 8t .. 8t + 7 in registers (x and x^4 at p and p + 1) while wavefront 0
// reduces the partials.  The first rise is the lowest p with
// f(x_{p+1}) > f(x_p): each lane finds its own, a ballot gives each
// wavefront's lowest, and thread 0 takes the minimum over the wavefronts.
constexpr int kFinishPolyBlock = 256;
constexpr int kLsPerLane = 8;
static_assert(kFinishPolyBlock * kLsPerLane >= kLsSteps, "the scan must cover every candidate");

template <bool COMP>
__global__ __launch_bounds__(kFinishPolyBlock) void k_finish_poly(FinishArgs f) {
  __shared__ double s_q[5];
  __shared__ int s_first[kFinishPolyBlock / kWave];
  const int gid = f.gids ? f.gids[blockIdx.x] : (int)blockIdx.x;
  const int2 si = f.img[gid];
  const DevString& s = f.strings[si.x];
  const LineSearchTable& t = *f.ls;
  const int tid = threadIdx.x;
  const int p0 = tid * kLsPerLane;
  double x[kLsPerLane + 1], x4[kLsPerLane + 1];
#pragma unroll
  for (int k = 0; k <= kLsPerLane; ++k) {
    const int p = min(p0 + k, kLsSteps);
    x[k] = t.x[p];
    x4[k] = t.x4[p];
  }
  if (tid < kWave) {
    double fin[5];
    reduce_partials<5, COMP>(s.part_p, si.y, s.tiles_x * s.tiles_y, fin);
    if (tid == 0)
#pragma unroll
      for (int k = 0; k < 5; ++k) s_q[k] = fin[k];
  }
  __syncthreads();
  const double q[5] = {s_q[0], s_q[1], s_q[2], s_q[3], s_q[4]};
  const double A = q[0], B = q[1], C = q[2], D = q[3], E = q[4];
  double fv[kLsPerLane + 1];
#pragma unroll
  for (int k = 0; k <= kLsPerLane; ++k) fv[k] = ls_quartic(q, x[k], x4[k]);
  int first = kLsSteps;
#pragma unroll
  for (int k = kLsPerLane - 1; k >= 0; --k)
    if (p0 + k < kLsSteps && fv[k + 1] > fv[k]) first = p0 + k;
  // Lane ranges increase with the lane, so the lowest lane with a rise
  // holds the wavefront's first rise.
  const unsigned long long rise = __ballot(first < kLsSteps);
  const int wfirst = __shfl(first, rise ? __builtin_ctzll(rise) : 0);
  if (tid % kWave == 0) s_first[tid / kWave] = rise ? wfirst : kLsSteps;
  __syncthreads();
  if (tid == 0) {
    int fr = kLsSteps;
#pragma unroll
    for (int w = 0; w < kFinishPolyBlock / kWave; ++w) fr = min(fr, s_first[w]);
    const int first_rise = fr < kLsSteps ? fr : kLsSteps - 1;
    ImageScalars& sc = f.sc[gid];
    sc.fenergy = A;
    sc.poly[0] = B;
    sc.poly[1] = C;
    sc.poly[2] = D;
    sc.poly[3] = E;
    sc.alpha = t.alpha[first_rise];
    sc.ls_index = first_rise;
    if (!__builtin_isfinite(A) || !__builtin_isfinite(B) || !__builtin_isfinite(C) ||
        !__builtin_isfinite(D) || !__builtin_isfinite(E))
      *f.health = 1;
  }
}

// After the energy pass (computeFreeEnergy).
__global__ __launch_bounds__(kBlock) void k_finish_energy(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[1];
  reduce_partials<1, true>(s.part_p, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) f.out[blockIdx.x] = fin[0];
}

// After the arclength pass: squared distances to the previous image.
__global__ __launch_bounds__(kBlock) void k_finish_arclen(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[2];
  reduce_partials<2, true>(s.part_a, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) {
    f.sc[blockIdx.x].ds2 = fin[0];
    f.sc[blockIdx.x].ds2_jb = fin[1];
  }
}

// ---------------------------------------------------------------------------
// Shared gather of the 5-point neighbourhood in double (energy, JB).
template <class T, int NC>
__device__ __forceinline__ void gather5(const DevString& s, const T* X, int i, int j,
                                        PolyNbhd<double, NC>& nb, double (&um)[NC],
                                        double (&vm)[NC], double (&uym)[NC], double (&vym)[NC]) {
  const int N = s.N, P = s.pitch;
  const long long ps = s.plane_stride;
  const int ip = min(i + 1, N - 1), im = max(i - 1, 0);
  const int jp = min(j + 1, N - 1), jm = max(j - 1, 0);
  const long long off = (long long)j * P + i;
  const long long offs[5] = {off, (long long)j * P + ip, (long long)j * P + im,
                             (long long)jp * P + i, (long long)jm * P + i};
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    nb.ax[t] = (double)X[PAX * ps + offs[t]];
    nb.ay[t] = (double)X[PAY * ps + offs[t]];
    nb.dax[t] = nb.day[t] = 0;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    nb.u[c][0] = (double)X[(PU1 + 2 * c) * ps + offs[0]];
    nb.v[c][0] = (double)X[(PV1 + 2 * c) * ps + offs[0]];
    nb.u[c][1] = (double)X[(PU1 + 2 * c) * ps + offs[1]];
    nb.v[c][1] = (double)X[(PV1 + 2 * c) * ps + offs[1]];
    nb.u[c][2] = (double)X[(PU1 + 2 * c) * ps + offs[3]];
    nb.v[c][2] = (double)X[(PV1 + 2 * c) * ps + offs[3]];
    um[c] = (double)X[(PU1 + 2 * c) * ps + offs[2]];
    vm[c] = (double)X[(PV1 + 2 * c) * ps + offs[2]];
    uym[c] = (double)X[(PU1 + 2 * c) * ps + offs[4]];
    vym[c] = (double)X[(PV1 + 2 * c) * ps + offs[4]];
#pragma unroll
    for (int t = 0; t < 3; ++t) nb.du[c][t] = nb.dv[c][t] = 0;
  }
  nb.h0 = hval(s, off);
}

// computeFreeEnergy (nlcg.cu:105-240), always evaluated in double.
template <class T, int NC>
__device__ void energy_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  const T* X = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
  const double idx = s.inv_dx, dx2 = s.dx2;
  double acc_s[1] = {0}, acc_c[1] = {0};
  if (t.i < s.N) {
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      PolyNbhd<double, NC> nb;
      double um[NC], vm[NC], uym[NC], vym[NC];
      gather5<T, NC>(s, X, t.i, j, nb, um, vm, uym, vym);
      CompConst<double> k[NC];
      load_comp_const<double, NC>(s, off, k);
      const uint16_t nnv = s.nn[off];
      const double e = site_energy<NC>(nb, k, s.eta, s.gamma, s.delta, idx, dx2, nnv & 0xff,
                                       nnv >> 8, s.nsf);
      two_sum_acc(acc_s[0], acc_c[0], e);
    }
  }
  publish_partial<1, true>(acc_s, acc_c, s.part_p, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_energy(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    energy_body<T, 2>(a, s, it);
  else
    energy_body<T, 1>(a, s, it);
}

// computeJB (nlcg.cu:243-380) into planes b, j_1, j_2 of a double buffer
// laid out [image][plane][row][pitch] with 1 + 2*nc planes.
template <class T, int NC>
__device__ void jb_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const T* X = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
  double* O = a.jb[it.string] + (long long)t.n * s.image_stride;  // the fields' layout
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
    PolyNbhd<double, NC> nb;
    double um[NC], vm[NC], uym[NC], vym[NC];
    gather5<T, NC>(s, X, t.i, j, nb, um, vm, uym, vym);
    CompConst<double> k[NC];
    load_comp_const<double, NC>(s, off, k);
    const uint16_t nnv = s.nn[off];
    double b, jj[NC][2];
    site_jb<double, NC>(nb, um, vm, uym, vym, k, s.inv_dx, nnv & 0xff, nnv >> 8, b, jj);
    O[off] = b;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      O[(1 + 2 * c) * s.plane_stride + off] = jj[c][0];
      O[(2 + 2 * c) * s.plane_stride + off] = jj[c][1];
    }
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_jb(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    jb_body<T, 2>(a, s, it);
  else
    jb_body<T, 1>(a, s, it);
}

// Arclength partials (arcLengthParametrization + arcLengthParametrizationJB,
// interpolation.cu:39-92): squared distance of image n to image n-1 in the
// plain metric and in the gauge-invariant (b, j) metric, with b and j
// recomputed on the fly for both images.  Image 0 contributes nothing.  The
// wavefront keeps a three-row window of both images, as the G kernel does:
// buffer loads of row j + 2 issued while row j is computed, x-neighbours
// through DPP, and neighbours outside the grid clamped to its edge, so each
// site's values are those of the reference's clamped 5-point gather.
// Buffers of images n - 1 and n of an arclength work item (image 0 uses
// itself; it contributes nothing), with the directions and step lengths of
// a pending update.  They are built by the kernel before it branches on the
// component count, so that the buffer descriptors stay wave-uniform scalars.
struct ArcBufs {
  GBufs b0, b1;       // b1.xo: the new fields of image n (pending update)
  double al0, al1;    // alpha of images n - 1 and n
};

template <class T>
__device__ __forceinline__ ArcBufs arc_bufs(const AuxArgs& a, const DevString& s, const Strip& t) {
  const long long ibytes = s.image_stride * (long long)sizeof(T);
  const T* X = (const T*)s.x[a.xin];
  const T* D = (const T*)s.d[a.dcur];
  const int n0 = t.n > 0 ? t.n - 1 : 0;
  ArcBufs ab;
  ab.b0.x = rsrc_of(X + (long long)n0 * s.image_stride, ibytes);
  ab.b1.x = rsrc_of(X + (long long)t.n * s.image_stride, ibytes);
  ab.b0.xa = s.a_shared ? rsrc_of(X, ibytes) : ab.b0.x;
  ab.b1.xa = s.a_shared ? rsrc_of(X, ibytes) : ab.b1.x;
  ab.b0.d = rsrc_of(D + (long long)n0 * s.image_stride, ibytes);
  ab.b1.d = rsrc_of(D + (long long)t.n * s.image_stride, ibytes);
  ab.b1.xo = rsrc_of((const T*)s.x[a.xout] + (long long)t.n * s.image_stride, ibytes);
  ab.b0.nn = ab.b1.nn = rsrc_of(s.nn, (long long)s.N * s.pitch * (long long)sizeof(uint16_t));
  ab.b0.ps = ab.b1.ps = (unsigned)(s.plane_stride * (long long)sizeof(T));
  ab.al0 = a.sc[s.img0 + n0].alpha;
  ab.al1 = a.sc[s.img0 + t.n].alpha;
  return ab;
}

// PL: some parameter is a per-site plane (read per row); otherwise the
// parameters are read once, so no load outside the row batches drains the
// loads in flight.
//
// UPD: the fields carry a pending update (updateFields, nlcg.cu:1372-1396).
// Both images are read with their directions and updated in the window, in
// the expression of the G kernel and of k_update, the distances are taken
// between the updated images, and the wavefront stores its rows of the
// updated image n to the other field buffer: the update and the arclength
// pass share one sweep over the fields.  As in the G kernel, the fixed A of
// a NO_SELF_FIELD string takes no update and is not stored.
//
// b and j are computed in double from the stored fields in both precisions:
// on a nearly degenerate string, neighbouring images differ by less than
// float rounding of b and j (two_band_512 in mixed precision moved energies
// by 2.5e-5 relative after three outer iterations with float b and j).
template <class T, int NC, bool PL, bool UPD>
__device__ __forceinline__ void arclen_body(const DevString& s, const Strip& t, const ArcBufs& ab) {
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  const int N = s.N, P_ = s.pitch;
  const bool nsf = s.nsf;
  const T al0 = UPD ? (T)ab.al0 : (T)0, al1 = UPD ? (T)ab.al1 : (T)0;
  if (t.n > 0) {
    const GBufs& b0 = ab.b0;  // image n - 1
    const GBufs& b1 = ab.b1;  // image n
    const bool lane_in = t.i < N;
    const double idx = s.inv_dx;
    CompConst<double> kconst[NC];
    if constexpr (!PL) load_comp_const<double, NC>(s, 0, kconst);
    const ParamView<NC> pv = param_view<NC>(s);
    auto nnoff = [&](int row) { return (unsigned)((row * P_ + t.ic) * (int)sizeof(uint16_t)); };
    // A row's raw loads into window values; the fixed A takes no update.
    auto process = [&](GRow<T, NC>& w, GRaw<T, NC>& r, T al) {
      if constexpr (UPD) {
        r.d[PAX] = nsf ? (T)0 : r.d[PAX];
        r.d[PAY] = nsf ? (T)0 : r.d[PAY];
        r.hd[PAX] = nsf ? (T)0 : r.hd[PAX];
        r.hd[PAY] = nsf ? (T)0 : r.hd[PAY];
      }
      g_process<T, NC>(w, r, al, UPD);
    };
    GRaw<T, NC> r0, r1;
    GRow<T, NC> w0[3], w1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int row = clampr(t.j0 - 1 + k, N);
      g_load<T, NC, UPD, false>(r0, b0, row, P_, t, nnoff(row));
      g_load<T, NC, UPD, false>(r1, b1, row, P_, t, nnoff(row));
      process(w0[k], r0, al0);
      process(w1[k], r1, al1);
    }
    // b and j of one image at row j from its window rows (previous,
    // current, next).
    auto jb = [&](const GRow<T, NC>& wp, const GRow<T, NC>& wc, const GRow<T, NC>& wn,
                  const CompConst<double> (&k)[NC], double& b, double (&jj)[NC][2]) {
      PolyNbhd<double, NC> nb;
      double um[NC], vm[NC], uym[NC], vym[NC];
      nb.ax[0] = (double)wc.ax;  nb.ay[0] = (double)wc.ay;
      nb.ax[1] = (double)wc.axr; nb.ay[1] = (double)wc.ayr;  // +x
      nb.ax[2] = (double)wc.axl; nb.ay[2] = (double)wc.ayl;  // -x
      nb.ax[3] = (double)wn.ax;  nb.ay[3] = (double)wn.ay;   // +y
      nb.ax[4] = (double)wp.ax;  nb.ay[4] = (double)wp.ay;   // -y
#pragma unroll
      for (int q = 0; q < 5; ++q) nb.dax[q] = nb.day[q] = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        nb.u[c][0] = (double)wc.u[c];
        nb.v[c][0] = (double)wc.v[c];
        nb.u[c][1] = (double)from_right(wc.u[c], wc.eu[c]);
        nb.v[c][1] = (double)from_right(wc.v[c], wc.ev[c]);
        nb.u[c][2] = (double)wn.u[c];
        nb.v[c][2] = (double)wn.v[c];
        um[c] = (double)from_left(wc.u[c], wc.eu[c]);
        vm[c] = (double)from_left(wc.v[c], wc.ev[c]);
        uym[c] = (double)wp.u[c];
        vym[c] = (double)wp.v[c];
#pragma unroll
        for (int q = 0; q < 3; ++q) nb.du[c][q] = nb.dv[c][q] = 0;
      }
      nb.h0 = 0;
      site_jb<double, NC>(nb, um, vm, uym, vym, k, idx, wc.nn & 0xff, wc.nn >> 8, b, jj);
    };
    // Row j: loads of row j + 2 for both images, the row's distances, the
    // loads into the previous row's slots (see the G kernel's loop).
    auto row = [&](int j, auto ip_tag, auto ic_tag, auto in_tag, auto batch_tag) {
      constexpr int IP = decltype(ip_tag)::value, IC = decltype(ic_tag)::value;
      constexpr int IN = decltype(in_tag)::value;
      constexpr bool BATCH = decltype(batch_tag)::value;
      if constexpr (BATCH) {
        const int r2 = clampr(j + 2, N);
        g_load<T, NC, UPD, false>(r0, b0, r2, P_, t, nnoff(r2));
        g_load<T, NC, UPD, false>(r1, b1, r2, P_, t, nnoff(r2));
      }
      __builtin_amdgcn_sched_barrier(0);
      if (lane_in && j < t.j1) {
        const GRow<T, NC>& c0 = w0[IC];
        const GRow<T, NC>& c1 = w1[IC];
        // Plain metric, planes in storage order.
        T p0[2 + 2 * NC], p1[2 + 2 * NC];
        p0[PAX] = c0.ax, p0[PAY] = c0.ay, p1[PAX] = c1.ax, p1[PAY] = c1.ay;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          p0[PU1 + 2 * c] = c0.u[c], p0[PV1 + 2 * c] = c0.v[c];
          p1[PU1 + 2 * c] = c1.u[c], p1[PV1 + 2 * c] = c1.v[c];
        }
        double d2 = 0;
#pragma unroll
        for (int p = 0; p < 2 + 2 * NC; ++p) {
          const double dv = (double)p1[p] - (double)p0[p];
          d2 += dv * dv;
        }
        CompConst<double> k[NC];
        if constexpr (PL) {
          load_comp_planes<double, NC>(pv, (long long)j * P_ + t.i, k);
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) k[c] = kconst[c];
        }
        double b[2], jj[2][NC][2];
        jb(w0[IP], c0, w0[IN], k, b[0], jj[0]);
        jb(w1[IP], c1, w1[IN], k, b[1], jj[1]);
        double djb = sq(b[1] - b[0]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          djb += sq(jj[1][c][0] - jj[0][c][0]) + sq(jj[1][c][1] - jj[0][c][1]);
        two_sum_acc(acc_s[0], acc_c[0], d2);
        two_sum_acc(acc_s[1], acc_c[1], djb);
        if constexpr (UPD) {
          const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
          if (!nsf) {
            bstore_keep(c1.ax, b1.xo, o, PAX * b1.ps);
            bstore_keep(c1.ay, b1.xo, o, PAY * b1.ps);
          }
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            bstore_keep(c1.u[c], b1.xo, o, (PU1 + 2 * c) * b1.ps);
            bstore_keep(c1.v[c], b1.xo, o, (PV1 + 2 * c) * b1.ps);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (BATCH) {
        process(w0[IP], r0, al0);
        process(w1[IP], r1, al1);
        pin_row<T, NC>(w0[IP]);
        pin_row<T, NC>(w1[IP]);
      }
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    const int groups = (t.j1 - 1 - t.j0 + 2) / 3;
    int j = t.j0;
#pragma clang loop unroll(disable)
    for (int g = 0; g < groups; ++g, j += 3) {
      row(j, S0(), S1(), S2(), std::true_type());
      row(j + 1, S1(), S2(), S0(), std::true_type());
      row(j + 2, S2(), S0(), S1(), std::true_type());
    }
    row(j, S0(), S1(), S2(), std::false_type());
  } else if constexpr (UPD) {
    // Image 0 has no distance; its rows only take the update.
    const GBufs& b1 = ab.b1;
    if (t.i < N) {
      const int p0 = nsf ? 2 : 0;
      for (int j = t.j0; j < t.j1; ++j) {
        const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
#pragma unroll
        for (int p = 0; p < 2 + 2 * NC; ++p)
          if (p >= p0)
            bstore_keep(fma(al1, bload<T>(b1.d, o, p * b1.ps), bload<T>(b1.x, o, p * b1.ps)),
                        b1.xo, o, p * b1.ps);
      }
    }
  }
  publish_partial<2, true>(acc_s, acc_c, s.part_a, t.n, t.tile, s.tiles_x * s.tiles_y);
}

// The fused update is built for the mixed-precision fields only: in FP64
// the update's direction loads would push the kernel past 256 VGPRs (one
// wavefront per SIMD), so FP64 applies the update with k_update first.  One
// launch per item class (components, parameter planes), specialised like
// the step kernels, so that each class runs at its own register count.
template <class T, int NC, bool PL, bool UPD>
__global__ __launch_bounds__(kBlock) void k_arclen(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  const ArcBufs ab = arc_bufs<T>(a, s, t);
  arclen_body<T, NC, PL, UPD>(s, t, ab);
}

template <class T, bool UPD>
static void launch_arclen_upd(const AuxArgs& a, hipStream_t st) {
  const dim3 g((unsigned)a.nitems), b(kBlock);
  if (a.nc == 2) {
    if (a.pl)
      hipLaunchKernelGGL((k_arclen<T, 2, true, UPD>), g, b, 0, st, a);
    else
      hipLaunchKernelGGL((k_arclen<T, 2, false, UPD>), g, b, 0, st, a);
  } else {
    if (a.pl)
      hipLaunchKernelGGL((k_arclen<T, 1, true, UPD>), g, b, 0, st, a);
    else
      hipLaunchKernelGGL((k_arclen<T, 1, false, UPD>), g, b, 0, st, a);
  }
}

template <class T>
static void launch_arclen(const AuxArgs& a, hipStream_t st) {
  if constexpr (sizeof(T) == 4) {
    if (a.upd) {
      launch_arclen_upd<T, true>(a, st);
      return;
    }
  }
  launch_arclen_upd<T, false>(a, st);
}

// computeArcLength (interpolation.cu:126-140) for both metrics, then the
// linear redistribution plan of every new image on s_jb: the segment search
// of generateNewFramesLinear (interpolation.cu:200-219).  One workgroup per
// string.  The square roots run one image per thread; the prefix sums run in
// one thread in the reference's order (so s and s_jb are the reference's
// sums bit for bit); the segment searches run one new image per thread: for
// a non-decreasing s the reference's incremental search stops at the first
// node k >= 1 with s[k] >= s_new, which is what each thread finds directly.
__global__ __launch_bounds__(kArcBlock) void k_arc_finalize(AuxArgs a, int nstrings) {
  __shared__ double ls[kMaxImages], lsj[kMaxImages];
  const int si = blockIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  ImageScalars* sc = a.sc + s.img0;
  const int F = s.F;
  const int tid = threadIdx.x;
  if (F < 2 || F > kMaxImages) return;
  for (int n = tid; n < F; n += kArcBlock) {
    ls[n] = n == 0 ? 0.0 : sqrt(sc[n].ds2);
    lsj[n] = n == 0 ? 0.0 : sqrt(sc[n].ds2_jb);
  }
  __syncthreads();
  if (tid == 0) {
    double run = 0, runj = 0;
#pragma unroll 8
    for (int n = 1; n < F; ++n) {
      run += ls[n];
      runj += lsj[n];
      ls[n] = run;
      lsj[n] = runj;
    }
  }
  __syncthreads();
  const double tot = ls[F - 1], totj = lsj[F - 1];
  __syncthreads();
  for (int n = tid; n < F; n += kArcBlock) {
    ls[n] /= tot;
    lsj[n] /= totj;
    sc[n].s = ls[n];
    sc[n].s_jb = lsj[n];
  }
  __syncthreads();
  // Linear plan (F_new = F).
  for (int n = tid; n < F; n += kArcBlock) {
    const double s_new = (double)n / ((double)F - 1);
    int k = 1;
    while (k + 1 < F && lsj[k] < s_new) ++k;
    const double seg_start = k == 1 ? 0.0 : lsj[k - 1];
    s.plan_k[n] = k;
    s.plan_w[n] = (s_new - seg_start) / (lsj[k] - lsj[k - 1]);
    // The reset of beta after a reparametrisation (mepgl.cu:231-234), done
    // here rather than by a launch of its own: nothing reads the norms
    // before the next NLCG step.
    if (a.reset_interior && n != 0 && n != F - 1) {
      sc[n].g_norm2[0] = INFINITY;
      sc[n].g_norm2[1] = INFINITY;
    }
  }
}

// Cubic redistribution plan on the s_jb of the last arclength pass: the
// running sums of the spacings h, the Thomas factors of the spline system
// built by buildHMatrix (interpolation.cu:252-290) and the segment search of
// generateNewFramesCubic (interpolation.cu:373-390).  One workgroup per
// string, launched only before a cubic redistribution; the recurrences run
// in one thread in the reference's order.
__global__ __launch_bounds__(kArcBlock) void k_cubic_plan(AuxArgs a, int nstrings) {
  __shared__ double lsj[kMaxImages], cum[kMaxImages];
  const int si = blockIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  const ImageScalars* sc = a.sc + s.img0;
  const int F = s.F;
  const int tid = threadIdx.x;
  if (F < 4 || F > kMaxImages) return;
  double* cb = s.cubic;  // [F][4]: c', m, sub-diagonal, h
  for (int n = tid; n < F; n += kArcBlock) lsj[n] = sc[n].s_jb;
  __syncthreads();
  for (int n = tid; n < F; n += kArcBlock) cb[4 * n + 3] = n == 0 ? 0.0 : lsj[n] - lsj[n - 1];
  if (tid == 0) {
    auto h = [&](int n) { return n == 0 ? 0.0 : lsj[n] - lsj[n - 1]; };
    cum[0] = 0;
    for (int n = 1; n < F; ++n) cum[n] = cum[n - 1] + h(n);
    double cprev = 0;
    for (int n = 0; n < F; ++n) {
      double dl, d, du;
      if (n == 0) {
        dl = 0; d = -h(1); du = h(1);
      } else if (n == F - 1) {
        dl = h(F - 1); d = -h(F - 1); du = 0;
      } else {
        dl = h(n); d = 2 * (h(n) + h(n + 1)); du = h(n + 1);
      }
      const double m = n == 0 ? d : d - dl * cprev;
      cprev = du / m;
      cb[4 * n + 0] = cprev;
      cb[4 * n + 1] = m;
      cb[4 * n + 2] = dl;
    }
  }
  __syncthreads();
  for (int n = tid; n < F; n += kArcBlock) {
    const double s_new = (double)n / ((double)F - 1);
    int k = 1;
    while (k + 1 < F && cum[k] < s_new) ++k;
    s.cplan_k[n] = k;
    s.cplan_w[n] = (s_new - cum[k - 1]) / (lsj[k] - lsj[k - 1]);
  }
}

// generateNewFramesLinear (interpolation.cu:190-227) over all planes of a
// string: image n of the output is w x[k] + (1 - w) x[k-1].
template <class T>
__global__ __launch_bounds__(kBlock) void k_redistribute_linear(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const int k = s.plan_k[t.n];
  const double w = s.plan_w[t.n];
  const T* Xk = (const T*)s.x[a.xin] + (long long)k * s.image_stride;
  const T* Xk1 = Xk - s.image_stride;
  T* O = (T*)s.x[a.xout] + (long long)t.n * s.image_stride;
  // A fixed and shared by every image stays as it is, in both buffers.
  const int p0 = s.a_shared ? 2 : 0;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
    for (int p = p0; p < s.nplanes; ++p) {
      const long long o = p * s.plane_stride + off;
      O[o] = (T)fma(w, (double)Xk[o], (1 - w) * (double)Xk1[o]);
    }
  }
}

// Copies the fixed A planes of a NO_SELF_FIELD string from buffer xout to
// buffer xin after a redistribution, so both field buffers hold the same A.
template <class T>
__global__ __launch_bounds__(kBlock) void k_copy_a(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (!s.nsf || s.a_shared) return;
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const T* S = (const T*)s.x[a.xout] + (long long)t.n * s.image_stride;
  T* O = (T*)s.x[a.xin] + (long long)t.n * s.image_stride;
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * s.pitch + t.i;
    O[off] = S[off];
    O[s.plane_stride + off] = S[s.plane_stride + off];
  }
}

// redistributeCubic (interpolation.cu:292-481) in double, one spline system
// per (site, plane): divided differences and end third differences
// (computeDeltaX/Y), the Thomas solve of the system along the image axis (in
// place of cusparseSgtsv2), and the spline evaluation
// (generateNewFramesCubic).  textbook = 0 keeps the reference's
// (w-1)^3-(w-1) weight on sigma_{k-1}; 1 uses (1-w)^3-(1-w).
//
// The systems run along the image axis, so the F work items of a tile share
// its (row, plane) pairs: item n takes pairs n, n + F, ..., and a thread
// solves about rows * planes / F systems.  Every system's arithmetic and its
// order are those of a sequential solve.
template <class T>
__global__ __launch_bounds__(kBlock) void k_redistribute_cubic(AuxArgs a, int textbook) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  // Everything read from the descriptor is read before the first store.
  // A fixed and shared by every image is not interpolated (planes from p0).
  const int F = s.F, p0 = s.a_shared ? 2 : 0, NP = s.nplanes - p0;
  const long long IS = s.image_stride, PS = s.plane_stride, PT = s.pitch;
  const double* cb = s.cubic;
  const int* ck = s.cplan_k;
  const double* cw = s.cplan_w;
  const T* X = (const T*)s.x[a.xin];
  T* Y = (T*)s.x[a.xout];
  double* Dl = (double*)s.cubic_rhs;  // in the fields' layout, in double
  const int pairs = (t.j1 - t.j0) * NP;
  for (int q = t.n; q < pairs; q += F) {
    const int j = t.j0 + q / NP, p = p0 + q % NP;
    const long long off = p * PS + (long long)j * PT + t.i;
    const T* Xp = X + off;
    auto f = [&](int n) { return (double)Xp[(long long)n * IS]; };
    auto D = [&](int n) -> double& { return Dl[(long long)n * IS + off]; };
    auto h = [&](int n) { return cb[4 * n + 3]; };
    // First differences dl(n) = (f(n) - f(n-1)) / h(n); the end conditions
    // need only the first and last three.
    auto dl = [&](int n) { return (f(n) - f(n - 1)) / h(n); };
    const double d1 = dl(1), d2 = dl(2), d3 = dl(3);
    const double e1 = dl(F - 1), e2 = dl(F - 2), e3 = dl(F - 3);
    const double d3s = ((d3 - d2) / (h(3) + h(2)) - (d2 - d1) / (h(2) + h(1))) / (h(3) + h(2) + h(1));
    const double d3e =
        ((e1 - e2) / (h(F - 1) + h(F - 2)) - (e2 - e3) / (h(F - 2) + h(F - 3))) /
        (h(F - 1) + h(F - 2) + h(F - 3));
    // One pass: the right-hand side dl(n+1) - dl(n) formed on the fly, and
    // the Thomas forward sweep; then the back substitution.
    double prev = h(1) * h(1) * d3s;
    prev = prev / cb[1];
    D(0) = prev;
    double dprev = d1;
    for (int n = 1; n < F - 1; ++n) {
      const double dnext = dl(n + 1);
      const double rhs = dnext - dprev;
      prev = (rhs - cb[4 * n + 2] * prev) / cb[4 * n + 1];
      D(n) = prev;
      dprev = dnext;
    }
    {
      const double rhs = -(h(F - 1) * h(F - 1)) * d3e;
      prev = (rhs - cb[4 * (F - 1) + 2] * prev) / cb[4 * (F - 1) + 1];
      D(F - 1) = prev;
    }
    double next = prev;
    for (int n = F - 2; n >= 0; --n) {
      next = D(n) - cb[4 * n + 0] * next;
      D(n) = next;
    }
    T* O = Y + off;
    for (int n = 0; n < F; ++n) {
      const int k = ck[n];
      const double w = cw[n];
      const double wm = textbook ? ((1 - w) * (1 - w) * (1 - w) - (1 - w))
                                 : ((w - 1) * (w - 1) * (w - 1) - (w - 1));
      const double v = w * f(k) + (1 - w) * f(k - 1) +
                       h(k) * h(k) * ((w * w * w - w) * D(k) + wm * D(k - 1));
      O[(long long)n * IS] = (T)v;
    }
  }
}

// ---------------------------------------------------------------------------
// Constant-phase alignment of neighbouring images (no reference function; the
// reference relies on the gauge-invariant s_jb metric).  Partials of
// sum conj(psi_n) psi_{n-1} over sites and components, image n >= 1.
template <class T, int NC>
__device__ void align_body(const AuxArgs& a, const DevString& s, const WorkItem& it) {
  const Strip t = strip_of(s, it);
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  if (t.n > 0 && t.i < s.N) {
    const T* X1 = (const T*)s.x[a.xin] + (long long)t.n * s.image_stride;
    const T* X0 = X1 - s.image_stride;
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      double re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const long long pu = (PU1 + 2 * c) * s.plane_stride + off;
        const long long pv = (PV1 + 2 * c) * s.plane_stride + off;
        const double u1 = X1[pu], v1 = X1[pv], u0 = X0[pu], v0 = X0[pv];
        re += u1 * u0 + v1 * v0;
        im += u1 * v0 - v1 * u0;
      }
      two_sum_acc(acc_s[0], acc_c[0], re);
      two_sum_acc(acc_s[1], acc_c[1], im);
    }
  }
  publish_partial<2, true>(acc_s, acc_c, s.part_a, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_align(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  if (s.nc == 2)
    align_body<T, 2>(a, s, it);
  else
    align_body<T, 1>(a, s, it);
}

// Theta_n = sum_{m <= n} arg(z_m), one workgroup per string: the angles in
// parallel, then the running sum in image order by one thread, from LDS.
__global__ __launch_bounds__(kArcBlock) void k_align_finalize(AuxArgs a, int nstrings) {
  __shared__ double th[kMaxImages];
  const int si = blockIdx.x;
  if (si >= nstrings) return;
  const DevString& s = a.strings[si];
  ImageScalars* sc = a.sc + s.img0;
  const int F = s.F;
  if (F > kMaxImages) return;
  for (int n = threadIdx.x; n < F; n += kArcBlock)
    th[n] = n == 0 ? 0.0 : atan2(sc[n].align_im, sc[n].align_re);
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int n = 1; n < F; ++n) {
      run = run + th[n];
      th[n] = run;
    }
  }
  __syncthreads();
  for (int n = threadIdx.x; n < F; n += kArcBlock) sc[n].theta = th[n];
}

// Rotates psi of image n by theta_n in the fields, the direction and both
// gradient buffers (the energy and the NLCG inner products are invariant).
template <class T>
__global__ __launch_bounds__(kBlock) void k_align_apply(AuxArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.n == 0 || t.i >= s.N) return;
  const double th = a.sc[s.img0 + t.n].theta;
  const T cs = (T)cos(th), sn = (T)sin(th);
  T* bufs[4] = {(T*)s.x[a.xin], (T*)s.d[a.dcur], (T*)s.g[0], (T*)s.g[1]};
  for (int b = 0; b < 4; ++b) {
    T* B = bufs[b] + (long long)t.n * s.image_stride;
    for (int j = t.j0; j < t.j1; ++j) {
      const long long off = (long long)j * s.pitch + t.i;
      for (int c = 0; c < s.nc; ++c) {
        const long long pu = (PU1 + 2 * c) * s.plane_stride + off;
        const long long pv = (PV1 + 2 * c) * s.plane_stride + off;
        const T u = B[pu], v = B[pv];
        B[pu] = u * cs - v * sn;
        B[pv] = u * sn + v * cs;
      }
    }
  }
}

// Phase winding around the grid boundary of one image (gl_utils.py:10-41,
// post.py:140-144): the boundary is walked along row 0 with x increasing,
// column N-1 with y increasing, row N-1 with x decreasing and column 0 with
// y decreasing; each edge adds (u dv - v du) / |psi|^2 at its start point.
// One wavefront per image; lanes take boundary points in turn and the lane
// sums are combined in lane order.
template <class T>
__global__ __launch_bounds__(kBlock) void k_winding(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  const int N = s.N, L = N - 1;
  const T* X = (const T*)s.x[f.xin] + (long long)n * s.image_stride;
  double acc[2] = {0, 0}, comp[2] = {0, 0};
  for (int k = threadIdx.x; k < 4 * L; k += kWave) {
    const int seg = k / L, m = k % L;
    int i0, j0, i1, j1;
    if (seg == 0) { i0 = m; j0 = 0; i1 = m + 1; j1 = 0; }
    else if (seg == 1) { i0 = L; j0 = m; i1 = L; j1 = m + 1; }
    else if (seg == 2) { i0 = L - m; j0 = L; i1 = L - m - 1; j1 = L; }
    else { i0 = 0; j0 = L - m; i1 = 0; j1 = L - m - 1; }
    const long long o0 = (long long)j0 * s.pitch + i0, o1 = (long long)j1 * s.pitch + i1;
    for (int c = 0; c < s.nc; ++c) {
      const long long pu = (PU1 + 2 * c) * s.plane_stride, pv = (PV1 + 2 * c) * s.plane_stride;
      const double u = X[pu + o0], v = X[pv + o0], un = X[pu + o1], vn = X[pv + o1];
      acc[c] += (u * (vn - v) - v * (un - u)) / (u * u + v * v);
    }
  }
  wave_reduce<2, true>(acc, comp);
  if (threadIdx.x == 0) {
    f.sc[blockIdx.x].winding[0] = acc[0] + comp[0];
    f.sc[blockIdx.x].winding[1] = s.nc == 2 ? acc[1] + comp[1] : 0.0;
  }
}

// After the alignment partials.
__global__ __launch_bounds__(kBlock) void k_finish_align(FinishArgs f) {
  int n;
  const DevString& s = string_of_image(f, n);
  double fin[2];
  reduce_partials<2, true>(s.part_a, n, s.tiles_x * s.tiles_y, fin);
  if (threadIdx.x == 0) {
    f.sc[blockIdx.x].align_re = fin[0];
    f.sc[blockIdx.x].align_im = fin[1];
  }
}

// ---------------------------------------------------------------------------
// buildNearestNeighbourMap (nlcg.cu:48-103) for the superconducting (low
// byte) and computational (high byte) domains of one string.
__global__ void k_nn_map(const int* sc_dom, const int* comp_dom, int N, int pitch,
                         uint16_t* nn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= N) return;
  auto bits = [&](const int* dom) {
    const int idx = j * N + i;
    int m = 0;
    if (dom[idx] == 0) return 0;
    const bool xr = i != N - 1, xl = i != 0, yu = j != N - 1, yd = j != 0;
    if (xr && dom[idx + 1]) m |= 1 << 0;
    if (xr && yu && dom[idx + 1 + N]) m |= 1 << 1;
    if (yu && dom[idx + N]) m |= 1 << 2;
    if (yu && xl && dom[idx - 1 + N]) m |= 1 << 3;
    if (xl && dom[idx - 1]) m |= 1 << 4;
    if (xl && yd && dom[idx - 1 - N]) m |= 1 << 5;
    if (yd && dom[idx - N]) m |= 1 << 6;
    if (yd && xr && dom[idx - N + 1]) m |= 1 << 7;
    return m;
  };
  nn[(long long)j * pitch + i] = (uint16_t)(bits(sc_dom) | (bits(comp_dom) << 8));
}

// Per-image scalar maintenance (graph-safe, no memset nodes).
__global__ void k_reset_norms(AuxArgs a, int nstrings, int interior_only) {
  const int si = blockIdx.x;
  if (si >= nstrings || (a.only_string >= 0 && si != a.only_string)) return;
  const DevString& s = a.strings[si];
  for (int n = threadIdx.x; n < s.F; n += blockDim.x) {
    if (interior_only && (n == 0 || n == s.F - 1)) continue;
    ImageScalars& x = a.sc[s.img0 + n];
    x.g_norm2[0] = INFINITY;
    x.g_norm2[1] = INFINITY;
    if (!interior_only) {
      // A full reset (loop()'s initial state) also clears the last step's
      // totals and line-search coefficients.
      x.g_delta_g = 0;
      x.beta = 0;
      x.fenergy = 0;
      for (int k = 0; k < 4; ++k) x.poly[k] = 0;
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T* p, long long n, T v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += stride) p[e] = v;
}

// ---------------------------------------------------------------------------
// Host launchers.
hipError_t launch_step(const StepArgs& a, int which, bool fp32, hipStream_t st) {
  const bool two = which != kKernelUpdate && a.nc == 2;
  if (fp32) return two ? launch_step_nc<float, 2>(a, which, st) : launch_step_nc<float, 1>(a, which, st);
  return two ? launch_step_nc<double, 2>(a, which, st) : launch_step_nc<double, 1>(a, which, st);
}

template <class T>
static hipError_t launch_aux_impl(const AuxArgs& a, int which, int arg, hipStream_t st) {
  const dim3 g((unsigned)a.nitems), b(kBlock);
  switch (which) {
    case kKernelEnergy:
      if (a.nitems > 0) hipLaunchKernelGGL(k_energy<T>, g, b, 0, st, a);
      break;
    case kKernelJB:
      if (a.nitems > 0) hipLaunchKernelGGL(k_jb<T>, g, b, 0, st, a);
      break;
    case kKernelArclen:
      if (a.nitems > 0) launch_arclen<T>(a, st);
      break;
    case kKernelArcFinalize:
      hipLaunchKernelGGL(k_arc_finalize, dim3(arg), dim3(kArcBlock), 0, st, a, arg);
      break;
    case kKernelCubicPlan:
      hipLaunchKernelGGL(k_cubic_plan, dim3(arg), dim3(kArcBlock), 0, st, a, arg);
      break;
    case kKernelLinear:
      if (a.nitems > 0) hipLaunchKernelGGL(k_redistribute_linear<T>, g, b, 0, st, a);
      break;
    case kKernelCubic:
      if (a.nitems > 0) hipLaunchKernelGGL(k_redistribute_cubic<T>, g, b, 0, st, a, arg);
      break;
    case kKernelCopyA:
      if (a.nitems > 0) hipLaunchKernelGGL(k_copy_a<T>, g, b, 0, st, a);
      break;
    case kKernelResetNorms:
      hipLaunchKernelGGL(k_reset_norms, dim3(a.nstrings), dim3(64), 0, st, a, a.nstrings, arg);
      break;
    case kKernelAlign:
      if (a.nitems > 0) hipLaunchKernelGGL(k_align<T>, g, b, 0, st, a);
      break;
    case kKernelAlignFinalize:
      hipLaunchKernelGGL(k_align_finalize, dim3(arg), dim3(kArcBlock), 0, st, a, arg);
      break;
    case kKernelAlignApply:
      if (a.nitems > 0) hipLaunchKernelGGL(k_align_apply<T>, g, b, 0, st, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_aux(const AuxArgs& a, int which, int arg, bool fp32, hipStream_t st) {
  // The arclength pass applies a pending update only for mixed-precision fields.
  if (a.upd && (which != kKernelArclen || !fp32)) return hipErrorInvalidValue;
  return fp32 ? launch_aux_impl<float>(a, which, arg, st)
              : launch_aux_impl<double>(a, which, arg, st);
}

hipError_t launch_finish(const FinishArgs& f, int which, bool fp32, hipStream_t st) {
  if (f.nimages <= 0) return hipSuccess;
  const dim3 g((unsigned)f.nimages), b(kBlock);
  switch (which) {
    case kKernelPoly:
      if (fp32)
        hipLaunchKernelGGL(k_finish_poly<true>, g, dim3(kFinishPolyBlock), 0, st, f);
      else
        hipLaunchKernelGGL(k_finish_poly<false>, g, dim3(kFinishPolyBlock), 0, st, f);
      break;
    case kKernelEnergy:
      hipLaunchKernelGGL(k_finish_energy, g, b, 0, st, f);
      break;
    case kKernelArclen:
      hipLaunchKernelGGL(k_finish_arclen, g, b, 0, st, f);
      break;
    case kKernelAlign:
      hipLaunchKernelGGL(k_finish_align, g, b, 0, st, f);
      break;
    case kKernelWinding:
      if (fp32)
        hipLaunchKernelGGL(k_winding<float>, g, b, 0, st, f);
      else
        hipLaunchKernelGGL(k_winding<double>, g, b, 0, st, f);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_nn_map(const int* sc_dom, const int* comp_dom, int N, int pitch, uint16_t* nn,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_nn_map, dim3((N + 63) / 64, N), dim3(64), 0, st, sc_dom, comp_dom, N,
                     pitch, nn);
  return hipGetLastError();
}

hipError_t launch_fill(void* p, long long n, bool fp32, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
  if (fp32)
    hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, st, (float*)p, n, (float)v);
  else
    hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, st, (double*)p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill<unsigned>, dim3(blocks), dim3(256), 0, st, p, n, 0u);
  return hipGetLastError();
}

}  // namespace mepgl
