// Per-site Ginzburg-Landau arithmetic of the relaxation loop, shared by the
// HIP kernels.  Each function restates one reference kernel's per-site body
// (mepgl/src/nlcg.cu) on values already gathered into registers, in the
// compute type T (double for the FP64 path, float for the mixed path).
//
// Neighbour slots: 0 = own site, 1 = +x, 2 = -x, 3 = +y, 4 = -y.
// A is also needed on the diagonals for the curl gradient: ax[dy+1][dx+1].
#pragma once

#include "common.hpp"

namespace mepgl {

__device__ __forceinline__ int bitof(int m, int k) { return (m >> k) & 1; }

template <class T>
__device__ __forceinline__ T sq(T v) { return v * v; }

// Neighbour-pair and quadrant presence of an 8-bit neighbour mask, as bit
// masks: bit k of pair_bits is set when neighbours k and k + 1 (mod 8) are
// both present; bit 2q of quad_bits when quadrant q (neighbours 2q, 2q + 1,
// 2q + 2 mod 8) is complete.  Weights and flags below are popcounts and bits
// of these, so they are the same small integers as the products of single
// bits they replace.
__device__ __forceinline__ unsigned pair_bits(int m) {
  const unsigned w = (unsigned)m | ((unsigned)m << 8);
  return w & (w >> 1);
}
__device__ __forceinline__ unsigned quad_bits(int m) {
  const unsigned p = pair_bits(m);
  return p & (p >> 1) & 0x55u;
}
__device__ __forceinline__ int popc(unsigned v) { return __builtin_popcount(v); }

// Quadrant area weight (S0S1S2 + S2S3S4 + S4S5S6 + S6S7S0)/4, nlcg.cu:140-143.
__device__ __forceinline__ int quad_count(int m) { return popc(quad_bits(m)); }

// Bond weights of nlcg.cu:188-189 and 552-571 as pair counts:
// (S1S2 + S6S7)/2 (forward x), (S2S3 + S5S6)/2 (backward x),
// (S0S1 + S3S4)/2 (forward y), (S4S5 + S7S0)/2 (backward y).
__device__ __forceinline__ int wx_fwd(unsigned p) { return popc(p & 0x42u); }
__device__ __forceinline__ int wx_bwd(unsigned p) { return popc(p & 0x24u); }
__device__ __forceinline__ int wy_fwd(unsigned p) { return popc(p & 0x09u); }
__device__ __forceinline__ int wy_bwd(unsigned p) { return popc(p & 0x90u); }

// Per-site constants of one component.
template <class T>
struct CompConst {
  T q;       // charge
  T two_a;   // 2 a
  T two_b;   // 2 b
  T b;       // b
  T ab;      // a / b
  T hx, hy;  // 0.5 / m_xx, 0.5 / m_yy
};

template <class T>
__device__ __forceinline__ double param_at(const DevParam& p, long long off) {
  // Parameter planes live in global memory: a global (not flat) load.
  using gd = __attribute__((address_space(1))) const double*;
  return p.p ? ((gd)p.p)[off] : p.v;
}

template <class T, int NC>
__device__ __forceinline__ void load_comp_const(const DevString& s, long long off,
                                                CompConst<T> (&k)[NC]) {
  if (!s.planes) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      k[c].q = (T)s.kc[c][0];
      k[c].two_a = (T)s.kc[c][1];
      k[c].two_b = (T)s.kc[c][2];
      k[c].b = (T)s.kc[c][3];
      k[c].ab = (T)s.kc[c][4];
      k[c].hx = (T)s.kc[c][5];
      k[c].hy = (T)s.kc[c][6];
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double a = param_at<T>(s.a[c], off), b = param_at<T>(s.b[c], off);
    const double mx = param_at<T>(s.mxx[c], off), my = param_at<T>(s.myy[c], off);
    k[c].q = (T)s.q[c];
    k[c].two_a = (T)(2.0 * a);
    k[c].two_b = (T)(2.0 * b);
    k[c].b = (T)b;
    k[c].ab = (T)(a / b);
    k[c].hx = (T)(0.5 / mx);
    k[c].hy = (T)(0.5 / my);
  }
}

// The parameter descriptors of a string, read once into registers before a
// row loop (reading them from the descriptor inside the loop would be a
// memory load per row).
template <int NC>
struct ParamView {
  int planes;
  DevParam a[NC], b[NC], mxx[NC], myy[NC];
  double q[NC];
};

template <int NC>
__device__ __forceinline__ ParamView<NC> param_view(const DevString& s) {
  ParamView<NC> v;
  v.planes = s.planes;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    v.a[c] = s.a[c];
    v.b[c] = s.b[c];
    v.mxx[c] = s.mxx[c];
    v.myy[c] = s.myy[c];
    v.q[c] = s.q[c];
  }
  return v;
}

// Per-site constants of plane parameters (ParamView::planes != 0).
template <class T, int NC>
__device__ __forceinline__ void load_comp_planes(const ParamView<NC>& v, long long off,
                                                 CompConst<T> (&k)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double a = param_at<T>(v.a[c], off), b = param_at<T>(v.b[c], off);
    const double mx = param_at<T>(v.mxx[c], off), my = param_at<T>(v.myy[c], off);
    k[c].q = (T)v.q[c];
    k[c].two_a = (T)(2.0 * a);
    k[c].two_b = (T)(2.0 * b);
    k[c].b = (T)b;
    k[c].ab = (T)(a / b);
    k[c].hx = (T)(0.5 / mx);
    k[c].hy = (T)(0.5 / my);
  }
}

// Gathered neighbourhood of one site.
template <class T, int NC>
struct Nbhd {
  T ax[3][3], ay[3][3];  // A on the 3x3 stencil, [dy+1][dx+1]
  T u[NC][5], v[NC][5];  // order parameters on the 5-point stencil
  T h[5];                // applied field on the 5-point stencil
};

// ---------------------------------------------------------------------------
// Kinetic bond of one component, gradient with respect to psi and A.
// Forward bond (+x or +y): nlcg.cu:529-549 / 573-593 and 769-776 / 787-794.
// Backward bond (-x or -y): nlcg.cu:551-571 / 595-615 and 778-785 / 796-803.
// u0,v0,A0: own site; un,vn,An: neighbour; hm = 0.5/m; w = bond weight.
template <class T>
__device__ __forceinline__ void bond_fwd(T u0, T v0, T A0, T un, T vn, T An, T q, T hm, T w,
                                         T idx, T& gu, T& gv, T& gA) {
  const T du = (un - u0) * idx, dv = (vn - v0) * idx;
  const T t1 = -du - q * A0 * v0;
  const T t2 = -du - q * An * vn;
  const T t3 = dv - q * A0 * u0;
  const T t4 = dv - q * An * un;
  const T hw = hm * w;
  gu += hw * (idx * t1 + idx * t2 - q * A0 * t3);
  gv += hw * (-idx * t3 - idx * t4 - q * A0 * t1);
  gA += -q * hw * (v0 * t1 + u0 * t3);
}

template <class T>
__device__ __forceinline__ void bond_bwd(T u0, T v0, T A0, T um, T vm, T Am, T q, T hm, T w,
                                         T idx, T& gu, T& gv, T& gA) {
  const T du = (u0 - um) * idx, dv = (v0 - vm) * idx;
  const T s1 = -du - q * A0 * v0;
  const T s2 = -du - q * Am * vm;
  const T s3 = dv - q * A0 * u0;
  const T s4 = dv - q * Am * um;
  const T hw = hm * w;
  gu += hw * (-idx * s1 - idx * s2 - q * A0 * s3);
  gv += hw * (idx * s3 + idx * s4 - q * A0 * s1);
  gA += -q * hw * (v0 * s1 + u0 * s3);
}

// Full gradient of one site (gradientPsi + gradientA, nlcg.cu:504-888),
// already multiplied by dx^2.  do_a = false skips the A part
// (NO_SELF_FIELD).  The psi part is always evaluated: its kinetic bonds also
// give the bond terms of the A gradient, and in Maxwell mode (nlcg.cu:1521)
// the caller keeps the frozen psi gradient instead of gpsi.
template <class T, int NC>
__device__ __forceinline__ void site_gradient(const Nbhd<T, NC>& s, const CompConst<T> (&k)[NC],
                                              T eta, T gamma, T delta, T idx, T dx2, int sc,
                                              int cm, bool do_a, T (&gpsi)[NC][2], T (&ga)[2]) {
  const unsigned sp = pair_bits(sc);
  const T af = (T)quad_count(sc) * (T)0.25;
  const T wxf = (T)wx_fwd(sp) * (T)0.5;
  const T wxb = (T)wx_bwd(sp) * (T)0.5;
  const T wyf = (T)wy_fwd(sp) * (T)0.5;
  const T wyb = (T)wy_bwd(sp) * (T)0.5;
  const T AX0 = s.ax[1][1], AY0 = s.ay[1][1];
  T gax = 0, gay = 0;

  // Each bond contributes only when its neighbour exists (bits 0, 4, 2, 6).
  // Its presence is folded into its weight once per site: an absent bond
  // gets weight 0, and every term it adds is then an exact zero (the fields
  // are finite), so no lane diverges and no term needs a select.
  const T wxf0 = bitof(sc, 0) ? wxf : (T)0, wxb0 = bitof(sc, 4) ? wxb : (T)0;
  const T wyf0 = bitof(sc, 2) ? wyf : (T)0, wyb0 = bitof(sc, 6) ? wyb : (T)0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const T u0 = s.u[c][0], v0 = s.v[c][0];
    const T n2 = u0 * u0 + v0 * v0;
    T gu = (k[c].two_a * u0 + k[c].two_b * u0 * n2) * af;
    T gv = (k[c].two_a * v0 + k[c].two_b * v0 * n2) * af;
    T bu, bv, bA;
    bu = bv = bA = 0;
    bond_fwd(u0, v0, AX0, s.u[c][1], s.v[c][1], s.ax[1][2], k[c].q, k[c].hx, wxf0, idx, bu, bv, bA);
    gu += bu;
    gv += bv;
    gax += bA;
    bu = bv = bA = 0;
    bond_bwd(u0, v0, AX0, s.u[c][2], s.v[c][2], s.ax[1][0], k[c].q, k[c].hx, wxb0, idx, bu, bv, bA);
    gu += bu;
    gv += bv;
    gax += bA;
    bu = bv = bA = 0;
    bond_fwd(u0, v0, AY0, s.u[c][3], s.v[c][3], s.ay[2][1], k[c].q, k[c].hy, wyf0, idx, bu, bv, bA);
    gu += bu;
    gv += bv;
    gay += bA;
    bu = bv = bA = 0;
    bond_bwd(u0, v0, AY0, s.u[c][4], s.v[c][4], s.ay[0][1], k[c].q, k[c].hy, wyb0, idx, bu, bv, bA);
    gu += bu;
    gv += bv;
    gay += bA;
    gpsi[c][0] = gu;
    gpsi[c][1] = gv;
  }
  if (NC == 2) {
    // Couplings, nlcg.cu:623-650.
    const T u1 = s.u[0][0], v1 = s.v[0][0], u2 = s.u[NC - 1][0], v2 = s.v[NC - 1][0];
    const T n1 = u1 * u1 + v1 * v1, n2 = u2 * u2 + v2 * v2;
    const T P = u1 * u2 + v1 * v2, Q = u1 * v2 - v1 * u2;
    const T ea = eta * af, ga_ = gamma * af, da = delta * af;
    gpsi[0][0] += ea * u2 + ga_ * u1 * n2 + da * (u2 * P - v2 * Q);
    gpsi[0][1] += ea * v2 + ga_ * v1 * n2 + da * (v2 * P + u2 * Q);
    gpsi[NC - 1][0] += ea * u1 + ga_ * u2 * n1 + da * (u1 * P + v1 * Q);
    gpsi[NC - 1][1] += ea * v1 + ga_ * v2 * n1 + da * (v1 * P - u1 * Q);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    gpsi[c][0] *= dx2;
    gpsi[c][1] *= dx2;
  }

  if (do_a) {
    // Curl terms, nlcg.cu:852-883: 8 own-quadrant and 8 neighbour-quadrant
    // contributions, gated on the own site's quadrant bits.
    const T c4 = (T)0.25 * idx;
    const T(&X)[3][3] = s.ax;
    const T(&Y)[3][3] = s.ay;
    const T dyf = (Y[1][2] - Y[1][1]) * idx;  // forward x-difference of A_y
    const T dyb = (Y[1][1] - Y[1][0]) * idx;
    const T dxf = (X[2][1] - X[1][1]) * idx;  // forward y-difference of A_x
    const T dxb = (X[1][1] - X[0][1]) * idx;
    const T h0 = s.h[0];
    // Quadrant presence folded into the quadrant's weight, as for the bonds.
    const unsigned cq = quad_bits(cm);
    const T c0 = (cq & 0x01u) ? c4 : (T)0;
    const T c1 = (cq & 0x04u) ? c4 : (T)0;
    const T c2 = (cq & 0x10u) ? c4 : (T)0;
    const T c3 = (cq & 0x40u) ? c4 : (T)0;
    {
      const T e = dyf - dxf - h0;
      const T n1 = dyf - (X[2][2] - X[1][2]) * idx - s.h[1];
      const T n2 = (Y[2][2] - Y[2][1]) * idx - dxf - s.h[3];
      gax += c0 * e + c0 * n2;
      gay += -c0 * e - c0 * n1;
    }
    {
      const T e = dyb - dxf - h0;
      const T n1 = (Y[2][1] - Y[2][0]) * idx - dxf - s.h[3];
      const T n2 = dyb - (X[2][0] - X[1][0]) * idx - s.h[2];
      gax += c1 * e + c1 * n1;
      gay += c1 * e + c1 * n2;
    }
    {
      const T e = dyb - dxb - h0;
      const T n1 = dyb - (X[1][0] - X[0][0]) * idx - s.h[2];
      const T n2 = (Y[0][1] - Y[0][0]) * idx - dxb - s.h[4];
      gax += -c2 * e - c2 * n2;
      gay += c2 * e + c2 * n1;
    }
    {
      const T e = dyf - dxb - h0;
      const T n1 = (Y[0][2] - Y[0][1]) * idx - dxb - s.h[4];
      const T n2 = dyf - (X[1][2] - X[0][2]) * idx - s.h[1];
      gax += -c3 * e - c3 * n1;
      gay += -c3 * e - c3 * n2;
    }
    ga[0] = gax * dx2;
    ga[1] = gay * dx2;
  }
}

// ---------------------------------------------------------------------------
// Line-search quartic of one site, polynomialExpansion nlcg.cu:997-1370.
// x and direction d on the 5-point stencil (A and psi); returns A..E.
template <class T, int NC>
struct PolyNbhd {
  T ax[5], ay[5];            // slots 0 own, 1 +x, 2 -x, 3 +y, 4 -y
  T dax[5], day[5];
  T u[NC][3], v[NC][3];      // slots 0 own, 1 +x, 2 +y
  T du[NC][3], dv[NC][3];
  T h0;
};

template <class T>
struct Quartic {
  T c[5];
};

// One kinetic bond (+x: A = A_x, +y: A = A_y) of one component.
// c1 multiplies d_A psi in the first-order term, c2 the cubic coefficient.
template <class T>
__device__ __forceinline__ void poly_bond(T u0, T v0, T un, T vn, T du0, T dv0, T dun, T dvn,
                                          T A0, T An, T dA0, T dAn, T q, T c1, T c2, T f,
                                          T idx, Quartic<T>& r) {
  const T Du = (un - u0) * idx, Dv = (vn - v0) * idx;
  const T Ddu = (dun - du0) * idx, Ddv = (dvn - dv0) * idx;
  T k0[4], k1[4], k2[4];
  k0[0] = -Du - q * A0 * v0;
  k0[1] = -Du - q * An * vn;
  k0[2] = Dv - q * A0 * u0;
  k0[3] = Dv - q * An * un;
  k1[0] = -Ddu - q * A0 * dv0 + c1 * dA0 * v0;
  k1[1] = -Ddu - q * An * dvn + c1 * dAn * vn;
  k1[2] = Ddv - q * A0 * du0 + c1 * dA0 * u0;
  k1[3] = Ddv - q * An * dun + c1 * dAn * un;
  k2[0] = -q * dA0 * dv0;
  k2[1] = -q * dAn * dvn;
  k2[2] = -q * dA0 * du0;
  k2[3] = -q * dAn * dun;
  T sA = 0, sB = 0, sC = 0, sD = 0, sE = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    sA += k0[t] * k0[t];
    sB += k0[t] * k1[t];
    sC += (T)0.5 * k1[t] * k1[t] + k0[t] * k2[t];
    sD += k1[t] * k2[t];
    sE += k2[t] * k2[t];
  }
  r.c[0] += (T)0.5 * sA * f;
  r.c[1] += sB * f;
  r.c[2] += sC * f;
  r.c[3] += sD * f * c2;
  r.c[4] += (T)0.5 * sE * f;
}

template <class T, int NC>
__device__ __forceinline__ Quartic<T> site_poly(const PolyNbhd<T, NC>& s,
                                                const CompConst<T> (&k)[NC], T eta, T gamma,
                                                T delta, T idx, T dx2, int sc, int cm, bool nsf,
                                                bool exact) {
  Quartic<T> r;
#pragma unroll
  for (int t = 0; t < 5; ++t) r.c[t] = 0;
  const T af = (T)quad_count(sc) * (T)0.25;
  const T afx = af * dx2;
  // Nonlinear terms, nlcg.cu:1024-1112.
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const T u = s.u[c][0], v = s.v[c][0], ud = s.du[c][0], vd = s.dv[c][0];
    const T n0 = k[c].ab + u * u + v * v;
    const T n1 = (T)2 * (u * ud + v * vd);
    const T n2 = ud * ud + vd * vd;
    const T kb = k[c].b * afx;
    r.c[0] += kb * (T)0.5 * n0 * n0;
    r.c[1] += kb * n0 * n1;
    r.c[2] += kb * ((T)0.5 * n1 * n1 + n0 * n2);
    r.c[3] += kb * n1 * n2;
    r.c[4] += kb * (T)0.5 * n2 * n2;
  }
  if (NC == 2) {
    const T u1 = s.u[0][0], v1 = s.v[0][0], u2 = s.u[NC - 1][0], v2 = s.v[NC - 1][0];
    const T ud1 = s.du[0][0], vd1 = s.dv[0][0], ud2 = s.du[NC - 1][0], vd2 = s.dv[NC - 1][0];
    const T e = eta * afx;
    r.c[0] += e * (u1 * u2 + v1 * v2);
    r.c[1] += e * (ud1 * u2 + vd1 * v2 + u1 * ud2 + v1 * vd2);
    r.c[2] += e * (ud1 * ud2 + vd1 * vd2);
    const T n1 = u1 * u1 + v1 * v1, n2 = u2 * u2 + v2 * v2;
    const T pd1 = u1 * ud1 + v1 * vd1, pd2 = u2 * ud2 + v2 * vd2;
    const T dd1 = ud1 * ud1 + vd1 * vd1, dd2 = ud2 * ud2 + vd2 * vd2;
    const T g = gamma * (T)0.5 * afx;
    r.c[0] += g * n1 * n2;
    r.c[1] += g * ((T)2 * pd1 * n2 + (T)2 * pd2 * n1);
    r.c[2] += g * ((T)4 * pd1 * pd2 + n1 * dd2 + n2 * dd1);
    r.c[3] += g * ((T)2 * pd1 * dd2 + (T)2 * pd2 * dd1);
    r.c[4] += g * (dd1 * dd2);
    const T P = u1 * u2 + v1 * v2, Q = u1 * v2 - v1 * u2;
    const T P1 = u1 * ud2 + ud1 * u2 + v1 * vd2 + vd1 * v2;
    const T Q1 = u1 * vd2 + ud1 * v2 - v1 * ud2 - vd1 * u2;
    const T P2 = ud1 * ud2 + vd1 * vd2, Q2 = ud1 * vd2 - vd1 * ud2;
    const T dl = delta * (T)0.5 * afx;
    r.c[0] += dl * (P * P - Q * Q);
    r.c[1] += dl * ((T)2 * P1 * P - (T)2 * Q1 * Q);
    r.c[2] += dl * (P1 * P1 - Q1 * Q1 + (T)2 * P * P2 - (T)2 * Q * Q2);
    r.c[3] += dl * ((T)2 * P1 * P2 - (T)2 * Q1 * Q2);
    r.c[4] += dl * (P2 * P2 - Q2 * Q2);
  }
  // Magnetic term, nlcg.cu:1116-1157.
  if (!nsf) {
    const T f = (T)0.25 * dx2;
    const T yf = (s.ay[1] - s.ay[0]) * idx, yb = (s.ay[0] - s.ay[2]) * idx;
    const T xf = (s.ax[3] - s.ax[0]) * idx, xb = (s.ax[0] - s.ax[4]) * idx;
    const T dyf = (s.day[1] - s.day[0]) * idx, dyb = (s.day[0] - s.day[2]) * idx;
    const T dxf = (s.dax[3] - s.dax[0]) * idx, dxb = (s.dax[0] - s.dax[4]) * idx;
    // Quadrant presence folded into the quadrant's weight (absent: 0).
    T fq[4];
    const unsigned cq = quad_bits(cm);
    fq[0] = (cq & 0x01u) ? f : (T)0;
    fq[1] = (cq & 0x04u) ? f : (T)0;
    fq[2] = (cq & 0x10u) ? f : (T)0;
    fq[3] = (cq & 0x40u) ? f : (T)0;
    T e0[4], e1[4];
    e0[0] = yf - xf - s.h0;  e1[0] = dyf - dxf;
    e0[1] = yb - xf - s.h0;  e1[1] = dyb - dxf;
    e0[2] = yb - xb - s.h0;  e1[2] = dyb - dxb;
    e0[3] = yf - xb - s.h0;  e1[3] = dyf - dxb;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      r.c[0] += (T)0.5 * e0[t] * e0[t] * fq[t];
      r.c[1] += e0[t] * e1[t] * fq[t];
      r.c[2] += (T)0.5 * e1[t] * e1[t] * fq[t];
    }
  }
  // Kinetic bonds, nlcg.cu:1160-1366.
  const unsigned sp = pair_bits(sc);
  const T cxf = (T)wx_fwd(sp) * (T)0.5;
  const T cyf = (T)wy_fwd(sp) * (T)0.5;
  const T c2y = exact ? (T)1 : (T)0.5;
  // Bond presence (bits 0 and 2) folded into the bond weights (absent: 0).
  const T bx = bitof(sc, 0) ? (T)1 : (T)0, by = bitof(sc, 2) ? (T)1 : (T)0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const T q = k[c].q;
    const T c1x = exact ? -q : (T)1;
    Quartic<T> rx, ry;
#pragma unroll
    for (int t = 0; t < 5; ++t) rx.c[t] = ry.c[t] = 0;
    poly_bond(s.u[c][0], s.v[c][0], s.u[c][1], s.v[c][1], s.du[c][0], s.dv[c][0], s.du[c][1],
              s.dv[c][1], s.ax[0], s.ax[1], s.dax[0], s.dax[1], q, c1x, (T)1,
              k[c].hx * cxf * dx2 * bx, idx, rx);
    poly_bond(s.u[c][0], s.v[c][0], s.u[c][2], s.v[c][2], s.du[c][0], s.dv[c][0], s.du[c][2],
              s.dv[c][2], s.ay[0], s.ay[3], s.day[0], s.day[3], q, -q, c2y,
              k[c].hy * cyf * dx2 * by, idx, ry);
#pragma unroll
    for (int t = 0; t < 5; ++t) r.c[t] += rx.c[t] + ry.c[t];
  }
  return r;
}

// ---------------------------------------------------------------------------
// Energy density of one site in double, freeEnergyDensity nlcg.cu:105-214.
template <int NC>
__device__ __forceinline__ double site_energy(const PolyNbhd<double, NC>& s,
                                              const CompConst<double> (&k)[NC], double eta,
                                              double gamma, double delta, double idx,
                                              double dx2, int sc, int cm, bool nsf) {
  const double af = quad_count(sc) * 0.25;
  double nl = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double u = s.u[c][0], v = s.v[c][0];
    nl += 0.5 * k[c].b * sq(k[c].ab + u * u + v * v);
  }
  if (NC == 2) {
    const double u1 = s.u[0][0], v1 = s.v[0][0], u2 = s.u[NC - 1][0], v2 = s.v[NC - 1][0];
    nl += eta * (u1 * u2 + v1 * v2);
    nl += gamma * 0.5 * (u1 * u1 + v1 * v1) * (u2 * u2 + v2 * v2);
    nl += delta * 0.5 * (sq(u1 * u2 + v1 * v2) - sq(u1 * v2 - v1 * u2));
  }
  nl *= af;
  double kx = 0, ky = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double q = k[c].q;
    if (bitof(sc, 0)) {
      const double Du = (s.u[c][1] - s.u[c][0]) * idx, Dv = (s.v[c][1] - s.v[c][0]) * idx;
      kx += 0.5 * k[c].hx *
            (sq(-Du - q * s.ax[0] * s.v[c][0]) + sq(-Du - q * s.ax[1] * s.v[c][1]) +
             sq(Dv - q * s.ax[0] * s.u[c][0]) + sq(Dv - q * s.ax[1] * s.u[c][1]));
    }
    if (bitof(sc, 2)) {
      const double Du = (s.u[c][2] - s.u[c][0]) * idx, Dv = (s.v[c][2] - s.v[c][0]) * idx;
      ky += 0.5 * k[c].hy *
            (sq(-Du - q * s.ay[0] * s.v[c][0]) + sq(-Du - q * s.ay[3] * s.v[c][2]) +
             sq(Dv - q * s.ay[0] * s.u[c][0]) + sq(Dv - q * s.ay[3] * s.u[c][2]));
    }
  }
  kx *= (bitof(sc, 1) * bitof(sc, 2) + bitof(sc, 6) * bitof(sc, 7)) * 0.5;
  ky *= (bitof(sc, 0) * bitof(sc, 1) + bitof(sc, 3) * bitof(sc, 4)) * 0.5;
  double em = 0;
  if (!nsf) {
    const double yf = (s.ay[1] - s.ay[0]) * idx, yb = (s.ay[0] - s.ay[2]) * idx;
    const double xf = (s.ax[3] - s.ax[0]) * idx, xb = (s.ax[0] - s.ax[4]) * idx;
    if (bitof(cm, 0) && bitof(cm, 1) && bitof(cm, 2)) em += 0.125 * sq(yf - xf - s.h0);
    if (bitof(cm, 2) && bitof(cm, 3) && bitof(cm, 4)) em += 0.125 * sq(yb - xf - s.h0);
    if (bitof(cm, 4) && bitof(cm, 5) && bitof(cm, 6)) em += 0.125 * sq(yb - xb - s.h0);
    if (bitof(cm, 6) && bitof(cm, 7) && bitof(cm, 0)) em += 0.125 * sq(yf - xb - s.h0);
  }
  return ((kx + ky) + nl + em) * dx2;
}

// Magnetic field and supercurrents of one site, computeJBKernel
// nlcg.cu:243-364, in R (double for the JB output and the FP64 arclength;
// float in the mixed-precision arclength, as the mixed step kernels' own
// stencils).  Uses own-site A and no bond weights.
template <class R, int NC>
__device__ __forceinline__ void site_jb(const PolyNbhd<R, NC>& s, const R (&um)[NC],
                                        const R (&vm)[NC], const R (&uym)[NC],
                                        const R (&vym)[NC], const CompConst<R> (&k)[NC], R idx,
                                        int sc, int cm, R& b, R (&j)[NC][2]) {
  const R yf = (s.ay[1] - s.ay[0]) * idx, yb = (s.ay[0] - s.ay[2]) * idx;
  const R xf = (s.ax[3] - s.ax[0]) * idx, xb = (s.ax[0] - s.ax[4]) * idx;
  const R qr = (R)0.25;
  R bb = 0;
  if (bitof(cm, 0) && bitof(cm, 1) && bitof(cm, 2)) bb += qr * (yf - xf);
  if (bitof(cm, 2) && bitof(cm, 3) && bitof(cm, 4)) bb += qr * (yb - xf);
  if (bitof(cm, 4) && bitof(cm, 5) && bitof(cm, 6)) bb += qr * (yb - xb);
  if (bitof(cm, 6) && bitof(cm, 7) && bitof(cm, 0)) bb += qr * (yf - xb);
  if (bb != 0) bb /= (R)quad_count(cm) * qr;
  b = bb;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const R q = k[c].q, u0 = s.u[c][0], v0 = s.v[c][0];
    const R ax0 = s.ax[0], ay0 = s.ay[0];
    R jx = 0, jy = 0;
    const R fx = q * k[c].hx, fy = q * k[c].hy;  // 0.5 q / m
    if (bitof(sc, 0)) {
      const R Du = (s.u[c][1] - u0) * idx, Dv = (s.v[c][1] - v0) * idx;
      jx += fx * (v0 * (-Du - q * ax0 * v0) + u0 * (Dv - q * ax0 * u0));
    }
    if (bitof(sc, 4)) {
      const R Du = (u0 - um[c]) * idx, Dv = (v0 - vm[c]) * idx;
      jx += fx * (v0 * (-Du - q * ax0 * v0) + u0 * (Dv - q * ax0 * u0));
    }
    if (bitof(sc, 2)) {
      const R Du = (s.u[c][2] - u0) * idx, Dv = (s.v[c][2] - v0) * idx;
      jy += fy * (v0 * (-Du - q * ay0 * v0) + u0 * (Dv - q * ay0 * u0));
    }
    if (bitof(sc, 6)) {
      const R Du = (u0 - uym[c]) * idx, Dv = (v0 - vym[c]) * idx;
      jy += fy * (v0 * (-Du - q * ay0 * v0) + u0 * (Dv - q * ay0 * u0));
    }
    j[c][0] = jx;
    j[c][1] = jy;
  }
}

}  // namespace mepgl
