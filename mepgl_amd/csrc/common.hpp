// Shared host/device definitions of the MI355X gauged-string-method library.
//
// Data layout in HBM (one allocation set per string):
//   field sets x[2], d[2], g[2]: [image n][row j][plane p][col i].  A plane
//   row is N rounded up to a multiple of 64 elements (plane_stride); a grid
//   row holds the rows of every plane one after another (pitch = nplanes *
//   plane_stride); an image is N + kPlanePadRows grid rows (the padding row
//   breaks power-of-two image strides).  Element (n, p, j, i) sits at
//   n * image_stride + p * plane_stride + j * pitch + i.  Planes p = 0 A_x,
//   1 A_y, 2 u_1, 3 v_1, 4 u_2, 5 v_2 (2 + 2*ncomp planes).  Column i runs
//   along x, so a 64-lane wavefront reads one contiguous row segment, and the
//   planes it reads at one row lie in one contiguous span.
//   JB output (b, j_1, j_2 in double) and the cubic right-hand sides use the
//   same layout.
//   nn: uint16 per site, low byte = superconducting-domain neighbour bits,
//   high byte = computational-domain neighbour bits (bit k: neighbour k,
//   3 2 1 / 4 X 0 / 5 6 7).
//   parameter planes (double) only for non-uniform parameters.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mepgl {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kTileX = 64;         // strip width: one wavefront row segment
constexpr int kBlock = kWave;      // one wavefront per workgroup / work item
constexpr int kLsSteps = 2000;     // line-search scan length (nlcg.cu:1421)
constexpr int kNumXcd = 8;         // MI355X accelerator complex dies
constexpr int kMaxImages = 2048;   // images per string (reparametrisation in LDS)
constexpr int kArcBlock = 256;     // threads of the per-string arclength pass
constexpr int kPlanePadRows = 1;   // unused grid rows after each image

enum Plane { PAX = 0, PAY = 1, PU1 = 2, PV1 = 3, PU2 = 4, PV2 = 5 };

struct DevParam {
  double v;          // uniform value (used when p == nullptr)
  const double* p;   // per-site plane [j][i] with the string's pitch, or null
};

// Per-image scalar block (double), indexed by the global image id.
struct ImageScalars {
  double g_norm2[2];   // [0]: caller's g_norm2, [1]: caller's g_norm2_old
  double g_delta_g;
  double beta;
  double alpha;
  double fenergy;      // A coefficient of the last line search
  double poly[4];      // B, C, D, E
  double s, s_jb;      // arclength (plain metric, gauge-invariant metric)
  double ds2, ds2_jb;  // squared distance to the previous image
  double align_re, align_im;  // sum of conj(psi_n) psi_{n-1} (alignment)
  double theta;        // constant phase applied by the last alignment
  double winding[2];   // boundary phase winding of each component
  int ls_index;        // scan index at which the line search stopped
};

struct DevString {
  int N, pitch, F, nc;      // grid side, row pitch, images, components
  int nplanes;              // 2 + 2*nc
  int nsf;                  // NO_SELF_FIELD: A is fixed, no A gradient
  int a_shared;             // NO_SELF_FIELD with the same A in every image:
                            // the step kernels read image 0's A planes and
                            // the reparametrisation leaves A untouched
  int img0;                 // global image id of image 0
  int tiles_x, tiles_y;     // 64-column strips x row segments per image
  int seg_rows;             // rows per segment (one wavefront marches them)
  int poly_exact;           // 0: reference line-search quartic (Q1, Q2)
  int planes;               // some of a, b, m_xx, m_yy is a per-site plane
  long long plane_stride;   // elements between planes (one plane row)
  long long image_stride;   // elements per image ((N + kPlanePadRows) * pitch)
  double dx, inv_dx, dx2;
  double q[2], eta, gamma, delta;
  // Per-component constants of uniform parameters: q, 2a, 2b, b, a/b,
  // 0.5/m_xx, 0.5/m_yy (used when planes == 0).
  double kc[2][7];
  DevParam a[2], b[2], mxx[2], myy[2], h;
  const uint16_t* nn;
  void* x[2];
  void* d[2];
  void* g[2];
  void* gs;                 // scratch gradient (diagnostic entry point)
  double* part_g;           // [F][ntiles][kPartG] workgroup partials
  double* part_p;           // [F][ntiles][kPartP]
  double* part_a;           // [F][ntiles][kPartA] arclength partials
  double* cubic;            // [F][4] Thomas factors c', m, sub-diagonal, h
  void* cubic_rhs;          // cubic right-hand sides, the fields' layout in double
  int* plan_k;              // [F] linear redistribution: segment end node
  double* plan_w;           // [F] linear redistribution: weight of node k
  int* cplan_k;             // [F] cubic redistribution: segment end node
  double* cplan_w;          // [F] cubic redistribution: weight of node k
};

// One work item = one wavefront's strip segment of one image of one string.
struct WorkItem {
  int string;
  int image;
  int tile;   // strip + tiles_x * segment
  int flags;  // kItemNoBonds
};

// Every site of the item's tile has an empty neighbour map: it lies outside
// both domains and has no bond, quadrant or condensation term, so its
// gradient, direction and line-search terms are exact zeros.
constexpr int kItemNoBonds = 1;

constexpr int kPartG = 4;   // g_norm2, g_delta_g (+ compensation terms)
constexpr int kPartP = 10;  // A..E (+ compensation terms)
constexpr int kPartA = 4;   // ds2, ds2_jb (+ compensation terms)

struct LineSearchTable {
  double x[kLsSteps + 1];      // x_k = 1e-10 * m^k, sequential product
  double x4[kLsSteps + 1];     // pow(x_k, 4)
  double alpha[kLsSteps];      // x_{k+1} / m
};

}  // namespace mepgl
