// fx_march_launch.hpp -- host launchers of the column-march kernels.
//
// Each (solver family RS, real type R) pair is compiled in a translation
// unit of its own (fx_march_inst.hip built once per pair), so that the
// march's many template instances build in parallel; the main unit calls
// the launchers declared here.
#pragma once
#include "fx_march.hpp"

namespace fx {

// column-tile width code of a multi-class march whose width is chosen per
// class (16 where the class's x extent allows it, else 8)
constexpr int kPerClassTX = 1;

// host state a launch needs
struct MarchHost {
  hipStream_t stream;
  int tuneOcc;               // autotuned waves-per-SIMD hint of the FP32 3-D march (0 or 4)
  int numCU;                 // compute units of the device
  int persist;               // multi-class march: multiple of the resident workgroups (0 = 1)
  const DevParams* dcls;     // multi-class: parameters per class (device)
  const MarchItem* ditems;   // multi-class: column-tile work items (device)
  int nItems;
  const long* dfaceStart;    // multi-class: faces before each tile (device)
};

// One workgroup per column tile of a batch of equal-shape tiles (TX = 8 or 16).
template <int RS, typename R>
void launch_march_r(const DevParams& P, const MarchLaunch& M, int TX, dim3 grid,
                    const MarchHost& H);
// Tiles of several classes in one persistent launch (TX = 8, 16 or kPerClassTX).
template <int RS, typename R>
void launch_march_multi_r(int ndim, int TX, const MarchLaunch& M, const MarchHost& H);

#define FX_MARCH_DECLARE(RS_, R_)                                                        \
  extern template void launch_march_r<RS_, R_>(const DevParams&, const MarchLaunch&, int, \
                                               dim3, const MarchHost&);                  \
  extern template void launch_march_multi_r<RS_, R_>(int, int, const MarchLaunch&,        \
                                                     const MarchHost&);
FX_MARCH_DECLARE(FX_HLLC, double)
FX_MARCH_DECLARE(FX_HLLC, float)
FX_MARCH_DECLARE(FX_HYBR, double)
FX_MARCH_DECLARE(FX_HYBR, float)
FX_MARCH_DECLARE(RS_STATES, double)
FX_MARCH_DECLARE(RS_STATES, float)
#undef FX_MARCH_DECLARE

}  // namespace fx
