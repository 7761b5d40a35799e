// The G and P kernels of one NLCG step of the MI355X gauged-string-method
// relaxation loop (gfx950) and the in-place update, with their launcher;
// instantiated once per storage type and component count (step_f32_1.hip,
// step_f32_2.hip, step_f64_1.hip, step_f64_2.hip), so that the four parts of
// the kernel code build concurrently.
//
// Work decomposition: one wavefront per work item = one 64-column strip of
// one row segment of one image of one string.  Lane l owns column
// 64*strip + l; the wavefront marches along the rows of its segment (down in
// even segments, up in odd ones), keeping the rows j-1, j, j+1 it needs in a
// register window and prefetching the next-but-one row while it computes
// row j.  Neighbours along x come from the adjacent lanes through DPP wave
// shifts; the two columns just outside the strip are loaded by lanes 0 and
// 63 only.  Every field value is therefore read from HBM once per launch,
// plus the halo rows and columns the strip shares with its neighbours,
// which L2 serves: adjacent work items run at the same time on the same
// XCD, and neighbouring segments reach their shared rows together.
//
// One NLCG step (nlcgSteps, nlcg.cu:1478-1583) is two launches:
//   G  fused pending update x += alpha d (updateFields), gradient
//      (gradientPsi + gradientA), |g|^2 and g.(g - g_old) partials
//      (gradientDensity).
//   P  beta from G's partials (every wavefront totals its image's partials
//      itself), direction d = -g + beta d (computeStepDirection), evaluated
//      on the stencil, line-search quartic partials (polynomialExpansion).
// P is followed by a finishing launch with one workgroup per image that
// totals the quartic partials and runs the line search (lineSearch, a
// parallel scan over a host-built table).  Reductions are deterministic
// (fixed partial order, fixed trees), with two-sum compensation in the
// mixed-precision path.
#pragma once

#include "row_window.hpp"

namespace mepgl {

// ---------------------------------------------------------------------------
// G kernel.
// Own-site values of a row that only the row's own computation needs: the
// old gradient (Polak-Ribiere), the frozen psi gradient (Maxwell mode) and,
// with parameter planes, the row's parameters.
template <class T, int NC, bool PL>
struct GSide {
  static constexpr int P = 2 + 2 * NC;
  T go[P];
  T gst[NC][2];
  RowParams<T, NC, PL> rp;
};

// The old gradient is read unconditionally (the diagnostic entry point that
// does not need it still reads a valid buffer), so the batch has no
// data-dependent branches.
template <class T, int NC, bool NSF, bool PL, bool MW>
__device__ __forceinline__ void g_side_load(GSide<T, NC, PL>& sd, const GBufs& b, int row, int P_,
                                            const Strip& t, const ParamView<NC>& pv,
                                            gptr<const double> hp, int N) {
  const unsigned o = (unsigned)((row * P_ + t.ic) * (int)sizeof(T));
#pragma unroll
  for (int p = 0; p < GSide<T, NC, PL>::P; ++p)
    sd.go[p] = !(NSF && p < 2) ? bload<T>(b.go, o, p * b.ps) : (T)0;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    sd.gst[q][0] = MW ? bload<T>(b.g, o, (PU1 + 2 * q) * b.ps) : (T)0;
    sd.gst[q][1] = MW ? bload<T>(b.g, o, (PV1 + 2 * q) * b.ps) : (T)0;
  }
  sd.rp.load(pv, hp, (long long)row * P_ + t.ic, row, N, P_, t);
}

template <class T, int NC, bool PL>
__device__ __forceinline__ void pin_side(GSide<T, NC, PL>& sd) {
#pragma unroll
  for (int p = 0; p < GSide<T, NC, PL>::P; ++p) pin(sd.go[p]);
#pragma unroll
  for (int q = 0; q < NC; ++q) pin(sd.gst[q][0]), pin(sd.gst[q][1]);
  if constexpr (PL) {
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) pin(sd.rp.prm[q][k]);
#pragma unroll
    for (int k = 0; k < 5; ++k) pin(sd.rp.h[k]);
  }
}

// In mixed precision G and P run as workgroups of kStepWaves wavefronts on
// consecutive work items, which are horizontally adjacent strips of one
// segment (the work list runs strip by strip, then segment by segment).  A
// workgroup starts on one CU as a unit, so neighbouring strips march
// together in every dispatch generation, and each strip's halo columns are
// read while its neighbour's wavefront still has those cache lines in L2.
// With one wavefront per workgroup, neighbours dispatched in later
// generations start at unrelated times and the halo lines are fetched
// twice: two_band_512 mixed G read 1.35x its algorithmic bytes, 1.06x
// grouped.  FP64 keeps one wavefront per workgroup: at two wavefronts per
// SIMD, four-wavefront groups leave slots idle while a group drains
// (two_band_512 FP64 2.3 % and thin_film_disk_1024 FP64 5.5 % slower).
constexpr int kStepWaves = 4;

template <class T>
constexpr int step_waves() {
  return sizeof(T) == 4 ? kStepWaves : 1;
}

// Work item of this wavefront (possibly past the end of the list); groups of
// W items are spread over the XCDs as single items are elsewhere.
template <int W>
__device__ __forceinline__ int step_item(int nitems) {
  const int groups = (nitems + W - 1) / W;
  return xcd_item(blockIdx.x, groups) * W + (W > 1 ? (int)(threadIdx.x / kWave) : 0);
}

// Edge columns traded inside a workgroup.  The W items of a workgroup are
// consecutive in the work list: horizontally adjacent strips of one segment
// (or of two, when a row has fewer than W strips).  A strip's halo columns
// are then its neighbours' edge columns, which the neighbours have loaded
// and processed already: they pass through LDS, with one barrier per row.
// That needs every wavefront of the workgroup to have an item and to march
// (the same rows, so the same barriers).  A halo column that is the strip's
// own column (lane 0 of the first strip, lane 63 of the last, where the
// halo index is clamped into the grid) takes the lane's own values.  Other
// halo columns are loaded as before.
struct EdgePlan {
  bool lds = false;                            // the workgroup trades edges
  bool nb_left = false, nb_right = false;      // neighbour strip in the group
  bool own_left = false, own_right = false;    // halo column is the own one
  int wave = 0;                                // wavefront in the group
};

template <int W>
__device__ __forceinline__ EdgePlan edge_plan(const StepArgs& a, const DevString& s,
                                              const WorkItem& it) {
  EdgePlan e;
  if constexpr (W > 1) {
    if (!a.edges) return e;
    const int groups = (a.nitems + W - 1) / W;
    const int kb = xcd_item(blockIdx.x, groups) * W;
    if (kb + W > a.nitems) return e;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    // Decided from the group's items alone, so every wavefront of the group
    // takes the same decision: all items march (they have bonds) over the
    // same rows in the same direction (one segment of one image).
    const WorkItem o0 = load_item(a.items, kb);
    const int seg0 = o0.tile / s.tiles_x;
    bool all = !(o0.flags & kItemNoBonds), left = false, right = false;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const WorkItem o = q == 0 ? o0 : load_item(a.items, kb + q);
      all = all && !(o.flags & kItemNoBonds) && o.string == o0.string && o.image == o0.image &&
            o.tile / s.tiles_x == seg0;
      left = left || (q == wave - 1 && o.tile == it.tile - 1);
      right = right || (q == wave + 1 && o.tile == it.tile + 1);
    }
    if (!all) return e;
    const int sx = it.tile % s.tiles_x;
    e.lds = true;
    e.wave = wave;
    e.own_left = sx == 0;
    e.own_right = (sx + 1) * kTileX >= s.N;
    e.nb_left = !e.own_left && left;
    e.nb_right = !e.own_right && right;
  }
  return e;
}

// Whether this lane loads its halo column from memory.
__device__ __forceinline__ bool halo_from_memory(const Strip& t, const EdgePlan& e) {
  if (!t.edge) return false;
  if (!e.lds) return true;
  return t.lane == 0 ? !(e.nb_left || e.own_left) : !(e.nb_right || e.own_right);
}

// The barrier of the edge exchange: LDS writes before it are visible to the
// workgroup after it.  Only LDS operations are waited for, so loads issued
// ahead stay in flight.
__device__ __forceinline__ void edge_barrier() {
  memory_fence_compiler();
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left as they are
  __builtin_amdgcn_s_barrier();
  memory_fence_compiler();
}

template <class T, int NC, bool UPD, bool NSF, bool PL, bool MW>
__device__ __forceinline__ void grad_body(const StepArgs& a, const DevString& s, const WorkItem& it) {
  constexpr bool COMP = sizeof(T) == 4;
  const Strip t = strip_of(s, it);
  const int N = s.N, P_ = s.pitch;
  const long long ibytes = s.image_stride * (long long)sizeof(T);
  const long long ib = (long long)t.n * s.image_stride;
  GBufs b;
  b.x = rsrc_of((const T*)s.x[a.xin] + ib, ibytes);
  b.xa = NSF && s.a_shared ? rsrc_of(s.x[a.xin], ibytes) : b.x;
  b.xo = rsrc_of((const T*)s.x[a.xout] + ib, ibytes);
  b.d = rsrc_of((const T*)s.d[a.din] + ib, ibytes);
  b.g = rsrc_of((const T*)(a.scratch ? s.gs : s.g[a.gcur]) + ib, ibytes);
  b.go = rsrc_of((const T*)s.g[a.gold] + ib, ibytes);
  b.nn = rsrc_of(s.nn, (long long)s.N * s.pitch * (long long)sizeof(uint16_t));
  b.ps = (unsigned)(s.plane_stride * (long long)sizeof(T));
  ImageScalars& sc = a.sc[s.img0 + t.n];
  T al = (T)0;
  if (UPD && !a.ls_fold) al = (T)sc.alpha;
  constexpr bool do_a = !NSF;
  const bool want_go = !a.scratch;
  const bool store = t.i < N;
  const T idx = (T)s.inv_dx, dx2 = (T)s.dx2;
  const T eta = (T)s.eta, gam = (T)s.gamma, del = (T)s.delta;
  CompConst<T> kconst[NC];
  if (!PL) load_comp_const<T, NC>(s, 0, kconst);
  const ParamView<NC> pv = param_view<NC>(s);
  gptr<const double> __restrict__ hp = as_global(s.h.p);
  const T hv = (T)s.h.v;
  double acc_s[2] = {0, 0}, acc_c[2] = {0, 0};
  auto nnoff = [&](int row) { return (unsigned)((row * P_ + t.ic) * (int)sizeof(uint16_t)); };
  // Edge exchange (EdgePlan): in mixed precision a row's A values outside
  // the strip are final only once the neighbours' edges are in, so A's
  // x-neighbours are formed one row later (g_finish), when the row becomes
  // the next row of the march.
  constexpr int W = step_waves<T>();
  const EdgePlan ep = edge_plan<W>(a, s, it);
  const bool hmem = halo_from_memory(t, ep);
  constexpr int EV = 2 + 2 * NC;  // A and psi
  __shared__ T edges[W > 1 ? 3 : 1][W][2][EV];
  auto put_edges = [&](const GRow<T, NC>& w, auto slot_tag) {
    constexpr int SL = decltype(slot_tag)::value;
    if constexpr (W > 1) {
      if (ep.lds && t.edge) {
        T* dst = edges[SL][ep.wave][t.lane == 0 ? 0 : 1];
        dst[0] = w.ax;
        dst[1] = w.ay;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
          dst[2 + 2 * q] = w.u[q];
          dst[3 + 2 * q] = w.v[q];
        }
      }
    }
  };
  auto finish = [&](GRow<T, NC>& w, auto slot_tag) {
    constexpr int SL = decltype(slot_tag)::value;
    if constexpr (W > 1) {
      if (ep.lds && t.edge) {
        const bool left = t.lane == 0;
        if (left ? ep.nb_left : ep.nb_right) {
          const T* src = edges[SL][left ? ep.wave - 1 : ep.wave + 1][left ? 1 : 0];
          w.eax = src[0];
          w.eay = src[1];
#pragma unroll
          for (int q = 0; q < NC; ++q) {
            w.eu[q] = src[2 + 2 * q];
            w.ev[q] = src[3 + 2 * q];
          }
        } else if (left ? ep.own_left : ep.own_right) {
          w.eax = w.ax;
          w.eay = w.ay;
#pragma unroll
          for (int q = 0; q < NC; ++q) {
            w.eu[q] = w.u[q];
            w.ev[q] = w.v[q];
          }
        }
      }
      g_finish<T, NC>(w);
    }
  };
  auto process = [&](GRow<T, NC>& w, const GRaw<T, NC>& r) {
    if constexpr (W > 1)
      g_process_own<T, NC>(w, r, al, UPD);
    else
      g_process<T, NC>(w, r, al, UPD);
  };

  // The march is the only part that depends on the direction
  // (marches_up); everything it needs from the string's descriptor is
  // read above, before the branch and before any store, so it stays in
  // scalar registers.
  const bool up = marches_up(s, it);
  if (UPD && a.ls_fold) {
    // The previous P's line search, done here instead of in a finishing
    // launch: every wavefront of the image scans the table itself, and the
    // first tile's stores the results that launch would have stored (after
    // every read of the descriptor, which thus stays in scalar registers).
    double q[5];
    const int first = ls_scan_wave<COMP>(a.ls, s.part_p, t.n, s.tiles_x * s.tiles_y, q);
    const double alpha = a.ls->alpha[first];
    al = (T)alpha;
    if (it.tile == 0 && lane_id() == 0) {
      sc.fenergy = q[0];
      sc.poly[0] = q[1];
      sc.poly[1] = q[2];
      sc.poly[2] = q[3];
      sc.poly[3] = q[4];
      sc.alpha = alpha;
      sc.ls_index = first;
      if (!__builtin_isfinite(q[0]) || !__builtin_isfinite(q[1]) || !__builtin_isfinite(q[2]) ||
          !__builtin_isfinite(q[3]) || !__builtin_isfinite(q[4]))
        *a.health = 1;
    }
  }
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, 2>;
  // Sites without bonds (an empty neighbour map: outside both domains)
  // have exact zeros for g and d, so an update leaves their fields as they
  // are.  Both field buffers hold the same values there: only uploads (which
  // write both buffers), the reparametrisation and the gauge alignments
  // change such a site's fields, the last two leave no update pending, and
  // the next G, one without the update, copies those sites' fields to the
  // other buffer (sync_sites).  G with the update then skips tiles without
  // bonds (kItemNoBonds) altogether, and writes x' = x at the other such
  // sites with the rest of their row.
  const bool sync_sites = !UPD && !a.scratch;
  auto copy_tile = [&]() {
    if (!sync_sites || !store) return;
    constexpr int P0 = NSF ? 2 : 0;  // fixed A is not stored
    for (int j = t.j0; j < t.j1; ++j) {
      const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
#pragma unroll
      for (int p = P0; p < GRaw<T, NC>::P; ++p)
        bstore(bload<T>(b.x, o, p * b.ps), b.xo, o, p * b.ps);
    }
  };
  auto march = [&](auto up_tag) {
    constexpr bool UP = decltype(up_tag)::value;
    constexpr int D = UP ? -1 : 1;  // march direction in rows
    // The window is three row slots used in rotation: (previous, current,
    // next) = (w[IP], w[IC], w[IN]) in the order of the march, with the side
    // values of each row in the slot of the same index.  Each row writes the
    // newly loaded row into the previous row's slot once the current row is
    // computed, and the next row uses the slots shifted by one, so no register
    // copies advance the window.  A segment marches down (rows j0 .. j1-1) or,
    // when UP, up (rows j1-1 .. j0); the previous row of the march is then
    // row j + 1.
    const int jb = UP ? t.j1 - 1 : t.j0;  // first row of the march
    GRaw<T, NC> raw;
    GRow<T, NC> w[3];
    GSide<T, NC, PL> sd[3];
    g_load<T, NC, UPD, NSF>(raw, b, clampr(jb - D, N), P_, t, nnoff(clampr(jb - D, N)), hmem);
    process(w[0], raw);
    g_load<T, NC, UPD, NSF>(raw, b, clampr(jb, N), P_, t, nnoff(clampr(jb, N)), hmem);
    process(w[1], raw);
    g_load<T, NC, UPD, NSF>(raw, b, clampr(jb + D, N), P_, t, nnoff(clampr(jb + D, N)), hmem);
    process(w[2], raw);
    g_side_load<T, NC, NSF, PL, MW>(sd[1], b, jb, P_, t, pv, hp, N);
    // The first rows' edges; row jb + D is finished at the first row's start.
    put_edges(w[0], S0());
    put_edges(w[1], S1());
    put_edges(w[2], S2());
    if (ep.lds) edge_barrier();
    finish(w[0], S0());
    finish(w[1], S1());

    // Row j: issue one batch of loads (the fields of row j + 2D, the own-site
    // values of row j + D), compute row j from registers, turn the batch into
    // window form, then store row j.  Every load is consumed before this
    // row's stores are issued, and the last row (no batch) is a separate copy
    // of the body, so the loads in flight at each point are known exactly.
    // Rows past the end of the segment (padding of the last rotation) are
    // computed but neither stored nor summed.
    auto row = [&](int j, auto ip_tag, auto ic_tag, auto in_tag, auto batch_tag) {
      constexpr int IP = decltype(ip_tag)::value, IC = decltype(ic_tag)::value;
      constexpr int IN = decltype(in_tag)::value;
      constexpr bool BATCH = decltype(batch_tag)::value;
      GRow<T, NC>& wm = w[UP ? IN : IP];  // row j - 1
      GRow<T, NC>& wc = w[IC];
      GRow<T, NC>& wpl = w[UP ? IP : IN];  // row j + 1
      const bool live = store && (UP ? j >= t.j0 : j < t.j1);
      if constexpr (BATCH) {
        g_load<T, NC, UPD, NSF>(raw, b, clampr(j + 2 * D, N), P_, t,
                                nnoff(clampr(j + 2 * D, N)), hmem);
        g_side_load<T, NC, NSF, PL, MW>(sd[IN], b, clampr(j + D, N), P_, t, pv, hp, N);
      }
      // Row j + D, processed one row ago, takes its neighbours' edges.
      if (ep.lds) edge_barrier();
      finish(w[IN], in_tag);
      __builtin_amdgcn_sched_barrier(0);

      Nbhd<T, NC> nb;
      nb.ax[0][0] = wm.axl; nb.ax[0][1] = wm.ax; nb.ax[0][2] = wm.axr;
      nb.ax[1][0] = wc.axl; nb.ax[1][1] = wc.ax; nb.ax[1][2] = wc.axr;
      nb.ax[2][0] = wpl.axl; nb.ax[2][1] = wpl.ax; nb.ax[2][2] = wpl.axr;
      nb.ay[0][0] = wm.ayl; nb.ay[0][1] = wm.ay; nb.ay[0][2] = wm.ayr;
      nb.ay[1][0] = wc.ayl; nb.ay[1][1] = wc.ay; nb.ay[1][2] = wc.ayr;
      nb.ay[2][0] = wpl.ayl; nb.ay[2][1] = wpl.ay; nb.ay[2][2] = wpl.ayr;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        nb.u[q][0] = wc.u[q];
        nb.v[q][0] = wc.v[q];
        nb.u[q][1] = from_right(wc.u[q], wc.eu[q]);
        nb.v[q][1] = from_right(wc.v[q], wc.ev[q]);
        nb.u[q][2] = from_left(wc.u[q], wc.eu[q]);
        nb.v[q][2] = from_left(wc.v[q], wc.ev[q]);
        nb.u[q][3] = wpl.u[q];
        nb.v[q][3] = wpl.v[q];
        nb.u[q][4] = wm.u[q];
        nb.v[q][4] = wm.v[q];
      }
      const GSide<T, NC, PL>& cur = sd[IC];
      CompConst<T> k[NC];
      if constexpr (PL) {
        comp_from_params<T, NC>(cur.rp.prm, pv.q, k);
#pragma unroll
        for (int q = 0; q < 5; ++q) nb.h[q] = hp ? cur.rp.h[q] : hv;
      } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) k[q] = kconst[q];
#pragma unroll
        for (int q = 0; q < 5; ++q) nb.h[q] = hv;
      }

      T gp[NC][2], ga[2] = {0, 0};
      site_gradient<T, NC>(nb, k, eta, gam, del, idx, dx2, wc.nn & 0xff, wc.nn >> 8, do_a,
                           gp, ga);
      if constexpr (MW) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
          gp[q][0] = cur.gst[q][0];
          gp[q][1] = cur.gst[q][1];
        }
      }
      if (live && want_go) {
        // gradientDensity, nlcg.cu:919-949
        double n2 = 0, dg = 0;
        if (do_a) {
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const double gv = (double)ga[q], go_ = (double)cur.go[q];
            n2 += gv * gv;
            dg += gv * (gv - go_);
          }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            const double gv = (double)gp[q][r], go_ = (double)cur.go[PU1 + 2 * q + r];
            n2 += gv * gv;
            dg += gv * (gv - go_);
          }
        }
        // Per-lane running sums over the segment's rows in double; the
        // compensation (mixed) starts at the cross-lane and cross-tile sums.
        acc_s[0] += n2;
        acc_s[1] += dg;
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (BATCH) {
        process(w[IP], raw);
        put_edges(w[IP], ip_tag);
        if constexpr (W > 1)
          pin_row_own<T, NC>(w[IP]);
        else
          pin_row<T, NC>(w[IP]);
        pin_side<T, NC, PL>(sd[IN]);
        memory_fence_compiler();
      }
      __builtin_amdgcn_sched_barrier(0);
      if (live) {
        const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
        if constexpr (!MW) {
#pragma unroll
          for (int q = 0; q < NC; ++q) {
            bstore(gp[q][0], b.g, o, (PU1 + 2 * q) * b.ps);
            bstore(gp[q][1], b.g, o, (PV1 + 2 * q) * b.ps);
          }
        }
        if (do_a) {
          bstore(ga[0], b.g, o, PAX * b.ps);
          bstore(ga[1], b.g, o, PAY * b.ps);
        }
        if (UPD || (sync_sites && wc.nn == 0)) {
          if (do_a) {
            bstore(wc.ax, b.xo, o, PAX * b.ps);
            bstore(wc.ay, b.xo, o, PAY * b.ps);
          }
#pragma unroll
          for (int q = 0; q < NC; ++q) {
            bstore(wc.u[q], b.xo, o, (PU1 + 2 * q) * b.ps);
            bstore(wc.v[q], b.xo, o, (PV1 + 2 * q) * b.ps);
          }
        }
      }
    };
    // All rows of the segment but the last carry a batch; they run in
    // rotations of three, the last rotation padded past the segment's end when
    // their count is not a multiple of three.
    const int groups = (t.j1 - 1 - t.j0 + 2) / 3;
    int j = jb;
#pragma clang loop unroll(disable)
    for (int g = 0; g < groups; ++g, j += 3 * D) {
      row(j, S0(), S1(), S2(), std::true_type());
      row(j + D, S1(), S2(), S0(), std::true_type());
      row(j + 2 * D, S2(), S0(), S1(), std::true_type());
    }
    row(j, S0(), S1(), S2(), std::false_type());
  };
  if (it.flags & kItemNoBonds)
    copy_tile();
  else if (up)
    march(std::true_type());
  else
    march(std::false_type());
  if (a.scratch) return;

  publish_partial<2, COMP>(acc_s, acc_c, s.part_g, t.n, t.tile, s.tiles_x * s.tiles_y);
}

template <class T, int NC, bool UPD, bool NSF, bool PL, bool MW>
__global__ __launch_bounds__(kWave * step_waves<T>()) void k_grad(StepArgs a) {
  const int k = step_item<step_waves<T>()>(a.nitems);
  if (k >= a.nitems) return;
  const WorkItem it = load_item(a.items, k);
  grad_body<T, NC, UPD, NSF, PL, MW>(a, a.strings[it.string], it);
}

// ---------------------------------------------------------------------------
// P kernel.
template <class T, int NC>
struct PRaw {
  static constexpr int P = 2 + 2 * NC;
  T x[P], g[P], d[P], hx[P], hg[P], hd[P];
  uint16_t nn;
};

template <class T, int NC>
struct PRow {
  T ax, ay, dax, day;              // A and its new direction
  T eax, eay, edax, eday;          // the same outside the strip (lanes 0 / 63)
  T u[NC], v[NC], du[NC], dv[NC];  // psi and its new direction
  T eu[NC], ev[NC], edu[NC], edv[NC];
  uint16_t nn;
};

struct PBufs {
  Rsrc x, xa, g, d, dout, nn;  // xa: where the A planes are read (GBufs)
  unsigned ps;
};

template <class T, int NC, bool NSF>
__device__ __forceinline__ void p_load(PRaw<T, NC>& r, const PBufs& b, int row, int P_,
                                       const Strip& t, unsigned nnoff, bool halo) {
  const RowOff o = row_off<T>(row, P_, t);
#pragma unroll
  for (int p = 0; p < PRaw<T, NC>::P; ++p) {
    const bool dir = !(NSF && p < 2);
    r.x[p] = bload<T>(p < 2 ? b.xa : b.x, o.own, p * b.ps);
    r.g[p] = dir ? bload<T>(b.g, o.own, p * b.ps) : (T)0;
    r.d[p] = dir ? bload<T>(b.d, o.own, p * b.ps) : (T)0;
  }
#pragma unroll
  for (int p = 0; p < PRaw<T, NC>::P; ++p) r.hx[p] = r.hg[p] = r.hd[p] = (T)0;
  if (halo) {
#pragma unroll
    for (int p = 0; p < PRaw<T, NC>::P; ++p) {
      const bool dir = !(NSF && p < 2);
      r.hx[p] = bload<T>(p < 2 ? b.xa : b.x, o.halo, p * b.ps);
      r.hg[p] = dir ? bload<T>(b.g, o.halo, p * b.ps) : (T)0;
      r.hd[p] = dir ? bload<T>(b.d, o.halo, p * b.ps) : (T)0;
    }
  }
  r.nn = bload_u16(b.nn, nnoff);
}

// d = -g + beta d  (computeStepDirection, nlcg.cu:985-993)
template <class T, int NC>
__device__ __forceinline__ void pin_row(PRow<T, NC>& w) {
  pin(w.ax), pin(w.ay), pin(w.dax), pin(w.day), pin(w.eax), pin(w.eay), pin(w.edax), pin(w.eday);
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    pin(w.u[q]), pin(w.v[q]), pin(w.du[q]), pin(w.dv[q]);
    pin(w.eu[q]), pin(w.ev[q]), pin(w.edu[q]), pin(w.edv[q]);
  }
  pin(w.nn);
}

template <class T, int NC>
__device__ __forceinline__ void p_process(PRow<T, NC>& w, const PRaw<T, NC>& r, T beta) {
  T dn[PRaw<T, NC>::P], en[PRaw<T, NC>::P];
#pragma unroll
  for (int p = 0; p < PRaw<T, NC>::P; ++p) {
    dn[p] = fma(beta, r.d[p], -r.g[p]);
    en[p] = fma(beta, r.hd[p], -r.hg[p]);
  }
  w.ax = r.x[PAX];
  w.ay = r.x[PAY];
  w.eax = r.hx[PAX];
  w.eay = r.hx[PAY];
  w.dax = dn[PAX];
  w.day = dn[PAY];
  w.edax = en[PAX];
  w.eday = en[PAY];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.u[q] = r.x[PU1 + 2 * q];
    w.v[q] = r.x[PV1 + 2 * q];
    w.eu[q] = r.hx[PU1 + 2 * q];
    w.ev[q] = r.hx[PV1 + 2 * q];
    w.du[q] = dn[PU1 + 2 * q];
    w.dv[q] = dn[PV1 + 2 * q];
    w.edu[q] = en[PU1 + 2 * q];
    w.edv[q] = en[PV1 + 2 * q];
  }
  w.nn = r.nn;
}

// A P row's values traded at the strip edges: A, its new direction, psi and
// its new direction.
template <int NC>
constexpr int p_edge_values() {
  return 4 + 4 * NC;
}

template <class T, int NC>
__device__ __forceinline__ void p_put_edges(const PRow<T, NC>& w, T* dst) {
  dst[0] = w.ax;
  dst[1] = w.ay;
  dst[2] = w.dax;
  dst[3] = w.day;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    dst[4 + 4 * q] = w.u[q];
    dst[5 + 4 * q] = w.v[q];
    dst[6 + 4 * q] = w.du[q];
    dst[7 + 4 * q] = w.dv[q];
  }
}

template <class T, int NC>
__device__ __forceinline__ void p_get_edges(PRow<T, NC>& w, const T* src) {
  w.eax = src[0];
  w.eay = src[1];
  w.edax = src[2];
  w.eday = src[3];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.eu[q] = src[4 + 4 * q];
    w.ev[q] = src[5 + 4 * q];
    w.edu[q] = src[6 + 4 * q];
    w.edv[q] = src[7 + 4 * q];
  }
}

template <class T, int NC>
__device__ __forceinline__ void p_own_edges(PRow<T, NC>& w) {
  w.eax = w.ax;
  w.eay = w.ay;
  w.edax = w.dax;
  w.eday = w.day;
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.eu[q] = w.u[q];
    w.ev[q] = w.v[q];
    w.edu[q] = w.du[q];
    w.edv[q] = w.dv[q];
  }
}

template <class T, int NC, bool NSF, bool PL>
__device__ __forceinline__ void poly_body(const StepArgs& a, const DevString& s, const WorkItem& it) {
  constexpr bool COMP = sizeof(T) == 4;
  const Strip t = strip_of(s, it);
  const int N = s.N, P_ = s.pitch;
  const long long ibytes = s.image_stride * (long long)sizeof(T);
  const long long ib = (long long)t.n * s.image_stride;
  PBufs b;
  b.x = rsrc_of((const T*)s.x[a.xin] + ib, ibytes);
  b.xa = NSF && s.a_shared ? rsrc_of(s.x[a.xin], ibytes) : b.x;
  b.g = rsrc_of((const T*)s.g[a.gcur] + ib, ibytes);
  b.d = rsrc_of((const T*)s.d[a.din] + ib, ibytes);
  b.dout = rsrc_of((const T*)s.d[a.dout] + ib, ibytes);
  b.nn = rsrc_of(s.nn, (long long)s.N * s.pitch * (long long)sizeof(uint16_t));
  b.ps = (unsigned)(s.plane_stride * (long long)sizeof(T));
  // The Polak-Ribiere beta (computeStepDirection, nlcg.cu:972-982): every
  // wavefront totals the gradient kernel's partials of its image in the same
  // fixed order, so all of them hold the same beta.  The old norm read here
  // sits in the other slot than the one the image's first tile stores at the
  // end of the body.
  ImageScalars* const scp = a.sc + s.img0 + t.n;
  auto totals = [&](double& gnorm2, double& gdelta, double& beta_d) {
    double fin[2];
    reduce_partials<2, COMP>(s.part_g, t.n, s.tiles_x * s.tiles_y, fin);
    gnorm2 = fin[0];
    gdelta = fin[1];
    beta_d = a.nlcg ? fmax(gdelta / scp->g_norm2[a.sold], 0.0) : 0.0;
  };
  auto nnoff = [&](int row) { return (unsigned)((row * P_ + t.ic) * (int)sizeof(uint16_t)); };
  constexpr int W = step_waves<T>();
  const EdgePlan ep = edge_plan<W>(a, s, it);
  const bool hmem = halo_from_memory(t, ep);
  // Edge values per row slot of the window, wavefront and side (lane 0,
  // lane 63).
  constexpr int EV = p_edge_values<NC>();
  __shared__ T edges[W > 1 ? 3 : 1][W][2][EV];
  auto put_edges = [&](const PRow<T, NC>& w, auto slot_tag) {
    constexpr int SL = decltype(slot_tag)::value;
    if constexpr (W > 1) {
      if (ep.lds && t.edge) p_put_edges<T, NC>(w, edges[SL][ep.wave][t.lane == 0 ? 0 : 1]);
    }
  };
  auto get_edges = [&](PRow<T, NC>& w, auto slot_tag) {
    constexpr int SL = decltype(slot_tag)::value;
    if constexpr (W > 1) {
      if (ep.lds && t.edge) {
        const bool left = t.lane == 0;
        if (left ? ep.nb_left : ep.nb_right)
          p_get_edges<T, NC>(w, edges[SL][left ? ep.wave - 1 : ep.wave + 1][left ? 1 : 0]);
        else if (left ? ep.own_left : ep.own_right)
          p_own_edges<T, NC>(w);
      }
    }
  };
  // The march is the only part that depends on the direction
  // (marches_up).  Its rows are loaded after beta is totalled: holding a row
  // across the partials' reduction costs the mixed kernels a wave per SIMD
  // (129 VGPRs), more than the overlap gains.
  const bool up = marches_up(s, it);
  const bool bonds = !(it.flags & kItemNoBonds);
  double gnorm2, gdelta, beta_d;
  totals(gnorm2, gdelta, beta_d);
  const T beta = uniform_value((T)beta_d);
  constexpr bool nsf = NSF;
  const bool exact = s.poly_exact;
  const bool store = t.i < N;
  const T idx = (T)s.inv_dx, dx2 = (T)s.dx2;
  const T eta = (T)s.eta, gam = (T)s.gamma, del = (T)s.delta;
  CompConst<T> kconst[NC];
  if (!PL) load_comp_const<T, NC>(s, 0, kconst);
  const ParamView<NC> pv = param_view<NC>(s);
  gptr<const double> __restrict__ hp = as_global(s.h.p);
  const T hv = (T)s.h.v;
  double acc_s[5] = {0, 0, 0, 0, 0}, acc_c[5] = {0, 0, 0, 0, 0};
  // Everything the march needs from the string's descriptor is read above,
  // before the branch and before any store, so it stays in scalar registers.
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, 2>;
  auto march = [&](auto up_tag) {
    constexpr bool UP = decltype(up_tag)::value;
    constexpr int D = UP ? -1 : 1;  // march direction in rows
    // Three row slots in rotation, as in the G kernel: w[IP] is the previous
    // row of the march, w[IC] the current, w[IN] the next; the parameters of a
    // row (parameter planes) sit in the slot of the same index.  Of row j - 1
    // only A_x and its direction are read.
    const int jb = UP ? t.j1 - 1 : t.j0;  // first row of the march
    PRaw<T, NC> raw;
    PRow<T, NC> w[3];
    RowParams<T, NC, PL> rp[3];
    p_load<T, NC, NSF>(raw, b, clampr(jb - D, N), P_, t, nnoff(clampr(jb - D, N)), hmem);
    p_process<T, NC>(w[0], raw, beta);
    p_load<T, NC, NSF>(raw, b, clampr(jb, N), P_, t, nnoff(clampr(jb, N)), hmem);
    p_process<T, NC>(w[1], raw, beta);
    p_load<T, NC, NSF>(raw, b, clampr(jb + D, N), P_, t, nnoff(clampr(jb + D, N)), hmem);
    p_process<T, NC>(w[2], raw, beta);
    rp[1].load(pv, hp, (long long)jb * P_ + t.ic, jb, N, P_, t);
    // The first rows' edges; row jb + D's are read at the first row's start.
    put_edges(w[0], S0());
    put_edges(w[1], S1());
    put_edges(w[2], S2());
    if (ep.lds) edge_barrier();
    get_edges(w[0], S0());
    get_edges(w[1], S1());

    // Row j: one batch of loads (the fields of row j + 2D, the parameters of
    // row j + D), the quartic of row j from registers, the batch into the
    // previous row's slot, then the stores of row j (see the G loop).
    auto row = [&](int j, auto ip_tag, auto ic_tag, auto in_tag, auto batch_tag) {
      constexpr int IP = decltype(ip_tag)::value, IC = decltype(ic_tag)::value;
      constexpr int IN = decltype(in_tag)::value;
      constexpr bool BATCH = decltype(batch_tag)::value;
      PRow<T, NC>& wm = w[UP ? IN : IP];  // row j - 1
      PRow<T, NC>& wc = w[IC];
      PRow<T, NC>& wpl = w[UP ? IP : IN];  // row j + 1
      const bool live = store && (UP ? j >= t.j0 : j < t.j1);
      if constexpr (BATCH) {
        p_load<T, NC, NSF>(raw, b, clampr(j + 2 * D, N), P_, t, nnoff(clampr(j + 2 * D, N)),
                           hmem);
        const int r1 = clampr(j + D, N);
        rp[IN].load(pv, hp, (long long)r1 * P_ + t.ic, r1, N, P_, t);
      }
      // Row j + D's edges, stored by the neighbours when they processed it
      // one row ago: row j + D is the current row at the next row.
      if (ep.lds) edge_barrier();
      get_edges(w[IN], in_tag);
      __builtin_amdgcn_sched_barrier(0);

      PolyNbhd<T, NC> nb;
      nb.ax[0] = wc.ax;
      nb.ax[1] = from_right(wc.ax, wc.eax);
      nb.ax[2] = 0;
      nb.ax[3] = wpl.ax;
      nb.ax[4] = wm.ax;
      nb.ay[0] = wc.ay;
      nb.ay[1] = from_right(wc.ay, wc.eay);
      nb.ay[2] = from_left(wc.ay, wc.eay);
      nb.ay[3] = wpl.ay;
      nb.ay[4] = 0;
      nb.dax[0] = wc.dax;
      nb.dax[1] = from_right(wc.dax, wc.edax);
      nb.dax[2] = 0;
      nb.dax[3] = wpl.dax;
      nb.dax[4] = wm.dax;
      nb.day[0] = wc.day;
      nb.day[1] = from_right(wc.day, wc.eday);
      nb.day[2] = from_left(wc.day, wc.eday);
      nb.day[3] = wpl.day;
      nb.day[4] = 0;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        nb.u[q][0] = wc.u[q];
        nb.v[q][0] = wc.v[q];
        nb.du[q][0] = wc.du[q];
        nb.dv[q][0] = wc.dv[q];
        nb.u[q][1] = from_right(wc.u[q], wc.eu[q]);
        nb.v[q][1] = from_right(wc.v[q], wc.ev[q]);
        nb.du[q][1] = from_right(wc.du[q], wc.edu[q]);
        nb.dv[q][1] = from_right(wc.dv[q], wc.edv[q]);
        nb.u[q][2] = wpl.u[q];
        nb.v[q][2] = wpl.v[q];
        nb.du[q][2] = wpl.du[q];
        nb.dv[q][2] = wpl.dv[q];
      }
      CompConst<T> k[NC];
      if constexpr (PL) {
        comp_from_params<T, NC>(rp[IC].prm, pv.q, k);
        nb.h0 = hp ? rp[IC].h[0] : hv;
      } else {
#pragma unroll
        for (int q = 0; q < NC; ++q) k[q] = kconst[q];
        nb.h0 = hv;
      }

      if (live) {
        const Quartic<T> qv =
            site_poly<T, NC>(nb, k, eta, gam, del, idx, dx2, wc.nn & 0xff, wc.nn >> 8, nsf, exact);
#pragma unroll
        for (int r = 0; r < 5; ++r) acc_s[r] += (double)qv.c[r];
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (BATCH) {
        p_process<T, NC>(w[IP], raw, beta);
        put_edges(w[IP], ip_tag);
        pin_row<T, NC>(w[IP]);
        if constexpr (PL) {
#pragma unroll
          for (int q = 0; q < NC; ++q)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) pin(rp[IN].prm[q][k2]);
#pragma unroll
          for (int k2 = 0; k2 < 5; ++k2) pin(rp[IN].h[k2]);
        }
        memory_fence_compiler();
      }
      __builtin_amdgcn_sched_barrier(0);
      if (live) {
        const unsigned o = (unsigned)((j * P_ + t.ic) * (int)sizeof(T));
        if (!nsf) {
          bstore(wc.dax, b.dout, o, PAX * b.ps);
          bstore(wc.day, b.dout, o, PAY * b.ps);
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
          bstore(wc.du[q], b.dout, o, (PU1 + 2 * q) * b.ps);
          bstore(wc.dv[q], b.dout, o, (PV1 + 2 * q) * b.ps);
        }
      }
    };
    const int groups = (t.j1 - 1 - t.j0 + 2) / 3;
    int j = jb;
#pragma clang loop unroll(disable)
    for (int g = 0; g < groups; ++g, j += 3 * D) {
      row(j, S0(), S1(), S2(), std::true_type());
      row(j + D, S1(), S2(), S0(), std::true_type());
      row(j + 2 * D, S2(), S0(), S1(), std::true_type());
    }
    row(j, S0(), S1(), S2(), std::false_type());
  };
  // Tiles without bonds: the new direction is -0 + beta 0 = 0, which the
  // other direction buffer already holds, and every quartic term is zero.
  if (bonds) {
    if (up)
      march(std::true_type());
    else
      march(std::false_type());
  }

  publish_partial<5, COMP>(acc_s, acc_c, s.part_p, t.n, t.tile, s.tiles_x * s.tiles_y);
  // The first tile's wavefront recomputes the totals here (same order, same
  // values) rather than keeping them in registers across the march, and
  // stores them only now: a store before the march would make the compiler
  // read the descriptor with vector loads.
  if (it.tile == 0) {
    totals(gnorm2, gdelta, beta_d);
    if (lane_id() == 0) {
      scp->g_norm2[a.scur] = gnorm2;
      scp->g_delta_g = gdelta;
      scp->beta = beta_d;
      // A gradient total that is not finite (NaN or +-Inf) marks the step;
      // the host reads the flag once per outer iteration.
      if (!__builtin_isfinite(gnorm2) || !__builtin_isfinite(gdelta)) *a.health = 1;
    }
  }
}

template <class T, int NC, bool NSF, bool PL>
__global__ __launch_bounds__(kWave * step_waves<T>()) void k_poly(StepArgs a) {
  const int k = step_item<step_waves<T>()>(a.nitems);
  if (k >= a.nitems) return;
  const WorkItem it = load_item(a.items, k);
  poly_body<T, NC, NSF, PL>(a, a.strings[it.string], it);
}

// ---------------------------------------------------------------------------
// Pending update applied in place (updateFields, nlcg.cu:1372-1396), used
// before anything other than the next gradient reads the fields.  The planes
// of a row are loaded together (NP known at compile time), and the strides
// are read before the first store, so they stay in scalar registers.
template <class T, int NP>
__device__ __forceinline__ void update_rows(T* X, const T* D, T al, bool nsf, long long pitch,
                                            long long ps, const Strip& t) {
  for (int j = t.j0; j < t.j1; ++j) {
    const long long off = (long long)j * pitch + t.i;
    T xv[NP], dv[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const bool own = !(nsf && p < 2);  // the fixed A takes no update
      xv[p] = own ? X[p * ps + off] : (T)0;
      dv[p] = own ? D[p * ps + off] : (T)0;
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (!(nsf && p < 2)) X[p * ps + off] = fma(al, dv[p], xv[p]);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_update(StepArgs a) {
  const WorkItem it = load_item(a.items, xcd_item(blockIdx.x, a.nitems));
  const DevString& s = a.strings[it.string];
  const Strip t = strip_of(s, it);
  if (t.i >= s.N) return;
  const long long ib = (long long)t.n * s.image_stride;
  T* X = (T*)s.x[a.xin] + ib;
  const T* D = (const T*)s.d[a.din] + ib;
  const T al = (T)a.sc[s.img0 + t.n].alpha;
  const bool nsf = s.nsf;
  const long long pitch = s.pitch, ps = s.plane_stride;
  if (s.nplanes == 6)
    update_rows<T, 6>(X, D, al, nsf, pitch, ps, t);
  else
    update_rows<T, 4>(X, D, al, nsf, pitch, ps, t);
}

// ---------------------------------------------------------------------------
// Host launchers of the step kernels.

// One wavefront per work item, step_waves<T>() per workgroup.
template <class T, class K>
static void launch_items(K kernel, const StepArgs& a, hipStream_t st) {
  constexpr int W = step_waves<T>();
  const unsigned groups = (unsigned)((a.nitems + W - 1) / W);
  hipLaunchKernelGGL(kernel, dim3(groups), dim3(kWave * W), 0, st, a);
}

template <class T, int NC, bool NSF, bool PL>
static void launch_gp_pl(const StepArgs& a, int which, hipStream_t st) {
  if (which == kKernelPoly)
    launch_items<T>(k_poly<T, NC, NSF, PL>, a, st);
  else if (a.maxwell && a.upd)
    launch_items<T>(k_grad<T, NC, true, NSF, PL, true>, a, st);
  else if (a.maxwell)
    launch_items<T>(k_grad<T, NC, false, NSF, PL, true>, a, st);
  else if (a.upd)
    launch_items<T>(k_grad<T, NC, true, NSF, PL, false>, a, st);
  else
    launch_items<T>(k_grad<T, NC, false, NSF, PL, false>, a, st);
}

template <class T, int NC, bool NSF>
static void launch_gp(const StepArgs& a, int which, hipStream_t st) {
  if (a.pl)
    launch_gp_pl<T, NC, NSF, true>(a, which, st);
  else
    launch_gp_pl<T, NC, NSF, false>(a, which, st);
}

// Launches the step kernel `which` for items of NC components (the in-place
// update, which does not depend on NC, is launched by the NC = 1 unit).
template <class T, int NC>
hipError_t launch_step_nc(const StepArgs& a, int which, hipStream_t st) {
  if (a.nitems <= 0) return hipSuccess;
  const dim3 g((unsigned)a.nitems), b(kBlock);
  switch (which) {
    case kKernelGrad:
    case kKernelPoly:
      a.nsf ? launch_gp<T, NC, true>(a, which, st) : launch_gp<T, NC, false>(a, which, st);
      break;
    case kKernelUpdate:
      if constexpr (NC != 1) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_update<T>, g, b, 0, st, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mepgl
