// The register row window of a strip: the raw loads of one strip row of
// the fields (own column, and the halo columns loaded by lanes 0 and 63),
// optionally with the pending update x + alpha d applied, turned into the
// values a 5-point or 3x3 stencil needs at row j (A with its x-neighbours,
// psi with its values outside the strip).  Used by the G kernel and the
// arclength pass.
#pragma once

#include "kernels_common.hpp"

namespace mepgl {

template <class T, int NC>
struct GRaw {
  static constexpr int P = 2 + 2 * NC;
  T x[P], d[P], hx[P], hd[P];
  uint16_t nn;
};

template <class T, int NC>
struct GRow {
  T ax, ay, axl, axr, ayl, ayr;  // A and its x-neighbours
  T u[NC], v[NC];                // psi
  T eu[NC], ev[NC];              // psi outside the strip (lanes 0 / 63)
  T eax, eay;                    // A outside the strip, until g_finish
  uint16_t nn;
};

// Buffers of one work item of the G kernel or the arclength pass.  xa is
// where the A planes are read: the image's own field set, or image 0's for a
// string whose fixed A is shared by every image.
struct GBufs {
  Rsrc x, xa, xo, d, g, go, nn;
  unsigned ps;  // plane stride in bytes
};

template <class T, int NC, bool UPD, bool NSF>
__device__ __forceinline__ void g_load(GRaw<T, NC>& r, const GBufs& b, int row, int P_,
                                       const Strip& t, unsigned nnoff, bool halo) {
  const RowOff o = row_off<T>(row, P_, t);
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) {
    r.x[p] = bload<T>(p < 2 ? b.xa : b.x, o.own, p * b.ps);
    r.d[p] = (UPD && !(NSF && p < 2)) ? bload<T>(b.d, o.own, p * b.ps) : (T)0;
  }
  // Halo columns: only lanes 0 and 63 read the columns outside the strip
  // (exec-masked loads); the other lanes' halo values are never used.
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) r.hx[p] = r.hd[p] = (T)0;
  if (halo) {
#pragma unroll
    for (int p = 0; p < GRaw<T, NC>::P; ++p) {
      r.hx[p] = bload<T>(p < 2 ? b.xa : b.x, o.halo, p * b.ps);
      r.hd[p] = (UPD && !(NSF && p < 2)) ? bload<T>(b.d, o.halo, p * b.ps) : (T)0;
    }
  }
  r.nn = bload_u16(b.nn, nnoff);
}

template <class T, int NC, bool UPD, bool NSF>
__device__ __forceinline__ void g_load(GRaw<T, NC>& r, const GBufs& b, int row, int P_,
                                       const Strip& t, unsigned nnoff) {
  g_load<T, NC, UPD, NSF>(r, b, row, P_, t, nnoff, t.edge);
}

template <class T, int NC>
__device__ __forceinline__ void pin_row(GRow<T, NC>& w) {
  pin(w.ax), pin(w.ay), pin(w.axl), pin(w.axr), pin(w.ayl), pin(w.ayr);
#pragma unroll
  for (int q = 0; q < NC; ++q) pin(w.u[q]), pin(w.v[q]), pin(w.eu[q]), pin(w.ev[q]);
  pin(w.nn);
}

template <class T, int NC>
__device__ __forceinline__ void g_process(GRow<T, NC>& w, const GRaw<T, NC>& r, T al, bool upd) {
  T c[GRaw<T, NC>::P], e[GRaw<T, NC>::P];
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) {
    c[p] = upd ? fma(al, r.d[p], r.x[p]) : r.x[p];
    e[p] = upd ? fma(al, r.hd[p], r.hx[p]) : r.hx[p];
  }
  w.ax = c[PAX];
  w.ay = c[PAY];
  w.axl = from_left(c[PAX], e[PAX]);
  w.axr = from_right(c[PAX], e[PAX]);
  w.ayl = from_left(c[PAY], e[PAY]);
  w.ayr = from_right(c[PAY], e[PAY]);
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.u[q] = c[PU1 + 2 * q];
    w.v[q] = c[PV1 + 2 * q];
    w.eu[q] = e[PU1 + 2 * q];
    w.ev[q] = e[PV1 + 2 * q];
  }
  w.nn = r.nn;
}

// The same in two parts, for kernels whose halo columns may arrive later
// (from the neighbouring strip's wavefront): g_process_own keeps A's values
// outside the strip, and g_finish forms A's x-neighbours once they are final.
// Together they give g_process's values bit for bit.
template <class T, int NC>
__device__ __forceinline__ void g_process_own(GRow<T, NC>& w, const GRaw<T, NC>& r, T al,
                                              bool upd) {
  T c[GRaw<T, NC>::P], e[GRaw<T, NC>::P];
#pragma unroll
  for (int p = 0; p < GRaw<T, NC>::P; ++p) {
    c[p] = upd ? fma(al, r.d[p], r.x[p]) : r.x[p];
    e[p] = upd ? fma(al, r.hd[p], r.hx[p]) : r.hx[p];
  }
  w.ax = c[PAX];
  w.ay = c[PAY];
  w.eax = e[PAX];
  w.eay = e[PAY];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    w.u[q] = c[PU1 + 2 * q];
    w.v[q] = c[PV1 + 2 * q];
    w.eu[q] = e[PU1 + 2 * q];
    w.ev[q] = e[PV1 + 2 * q];
  }
  w.nn = r.nn;
}

template <class T, int NC>
__device__ __forceinline__ void pin_row_own(GRow<T, NC>& w) {
  pin(w.ax), pin(w.ay), pin(w.eax), pin(w.eay);
#pragma unroll
  for (int q = 0; q < NC; ++q) pin(w.u[q]), pin(w.v[q]), pin(w.eu[q]), pin(w.ev[q]);
  pin(w.nn);
}

template <class T, int NC>
__device__ __forceinline__ void g_finish(GRow<T, NC>& w) {
  w.axl = from_left(w.ax, w.eax);
  w.axr = from_right(w.ax, w.eax);
  w.ayl = from_left(w.ay, w.eay);
  w.ayr = from_right(w.ay, w.eay);
}

}  // namespace mepgl
