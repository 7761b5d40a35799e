// Launch interface between the host context and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace mepgl {

enum KernelId {
  kKernelGrad = 0,
  kKernelPoly,
  kKernelUpdate,
  kKernelEnergy,
  kKernelJB,
  kKernelArclen,
  kKernelArcFinalize,
  kKernelLinear,
  kKernelCubic,
  kKernelCopyA,
  kKernelResetNorms,
  kKernelAlign,        // alignment partials
  kKernelAlignFinalize,
  kKernelAlignApply,
  kKernelWinding,      // boundary phase winding per image (finishing-style launch)
  kKernelCubicPlan,    // spline factors and plan of the cubic redistribution
  kKernelCount
};

// Arguments of one NLCG-step launch (G, P or the in-place update U).
struct StepArgs {
  const DevString* strings;
  const WorkItem* items;
  int nitems;
  ImageScalars* sc;
  const LineSearchTable* ls;
  int xin, xout;    // field buffers read / written
  int din, dout;    // direction buffers read / written
  int gcur, gold;   // gradient buffers (current, previous)
  int scur, sold;   // g_norm2 slots (current, previous)
  int upd;          // G applies the pending update x += alpha d
  int ls_fold;      // ... with alpha from the previous P's partials (its line
                    // search done by G; no finishing launch after that P)
  int maxwell;      // Maxwell mode: psi gradient not recomputed
  int nlcg;         // Polak-Ribiere beta (1) or steepest descent (0)
  int scratch;      // G writes the diagnostic gradient buffer only
  int nc, nsf;      // class of every item in `items` (components, fixed A)
  int pl;           // ... and whether its strings have parameter planes
  int edges;        // mixed precision: workgroups trade edge columns through LDS
  int* health;      // set to 1 when a step's gradient totals are not finite
};

// Arguments of the auxiliary launches (energy, JB, arclength,
// redistribution, scalar maintenance).
struct AuxArgs {
  const DevString* strings;
  const WorkItem* items;
  int nitems;
  int nstrings;
  ImageScalars* sc;
  int xin, xout;
  int dcur;             // current direction buffer
  int upd;              // arclength applies the pending update x += alpha d
                        // and writes the new fields to buffer xout
  int reset_interior;   // arclength finalize also resets the interior images'
                        // gradient norms (beta = 0 on the next step)
  double* out;          // per global image (energy)
  double* const* jb;    // per string JB planes (device array of pointers)
  int only_string;      // scalar maintenance: string index, or -1 for all
  int nc, pl;           // arclength: class of every item in `items`
                        // (components, parameter planes)
};

// Arguments of the per-image finishing launches (one wavefront per image).
struct FinishArgs {
  const DevString* strings;
  const int2* img;   // global image id -> (string, image)
  const int* gids;   // the images of a launch (global ids); null: every image, in order
  int nimages;
  ImageScalars* sc;
  const LineSearchTable* ls;
  double* out;
  int xin;           // current field buffer (winding)
  int* health;       // set to 1 when a line-search quartic is not finite
};

hipError_t launch_step(const StepArgs& a, int which, bool fp32, hipStream_t st);
// The step kernels of one storage type and component count, instantiated in
// step_{f32,f64}_{1,2}.hip.
template <class T, int NC>
hipError_t launch_step_nc(const StepArgs& a, int which, hipStream_t st);
hipError_t launch_finish(const FinishArgs& f, int which, bool fp32, hipStream_t st);
hipError_t launch_aux(const AuxArgs& a, int which, int arg, bool fp32, hipStream_t st);
hipError_t launch_nn_map(const int* sc_dom, const int* comp_dom, int N, int pitch, uint16_t* nn,
                         hipStream_t st);
hipError_t launch_fill(void* p, long long n, bool fp32, double v, hipStream_t st);
hipError_t launch_fill_u32(unsigned* p, long long n, hipStream_t st);

}  // namespace mepgl
