// fx_march_inst.hip -- the column-march kernels of one solver family and
// real type, and their launchers.  Built once per pair with
// -DFX_MARCH_RS=<family> -DFX_MARCH_R=<double|float> (flash5_amd/build.py).
#include <hip/hip_runtime.h>

#include "fx_march_launch.hpp"

#if !defined(FX_MARCH_RS) || !defined(FX_MARCH_R)
#error "build with -DFX_MARCH_RS=... -DFX_MARCH_R=..."
#endif

namespace fx {

template <int RS, typename R>
void launch_march_r(const DevParams& P, const MarchLaunch& M, int TX, dim3 grid,
                    const MarchHost& H) {
  if constexpr (sizeof(R) == 4) {  // FP32 3-D: the autotuner may ask for 4 waves per SIMD
    if (P.ndim == 3 && H.tuneOcc == 4) {
      if (TX == 16)
        hipLaunchKernelGGL((k_march<3, 16, 8, 512, RS, R, 4>), grid, dim3(512), 0, H.stream, P, M);
      else
        hipLaunchKernelGGL((k_march<3, 8, 8, 256, RS, R, 4>), grid, dim3(256), 0, H.stream, P, M);
      return;
    }
  }
  if constexpr (sizeof(R) == 8) {  // FP64 3-D 16x8: 768 threads, 3 waves per SIMD
    if (P.ndim == 3 && TX == 16) {
      hipLaunchKernelGGL((k_march<3, 16, 8, 768, RS, R, 3>), grid, dim3(768), 0, H.stream, P, M);
      return;
    }
  }
  if (P.ndim == 3 && TX == 16) hipLaunchKernelGGL((k_march<3, 16, 8, 512, RS, R>), grid, dim3(512), 0, H.stream, P, M);
  else if (P.ndim == 3) hipLaunchKernelGGL((k_march<3, 8, 8, 256, RS, R>), grid, dim3(256), 0, H.stream, P, M);
  else if (TX == 16) hipLaunchKernelGGL((k_march<2, 16, 8, 256, RS, R>), grid, dim3(256), 0, H.stream, P, M);
  else hipLaunchKernelGGL((k_march<2, 8, 8, 256, RS, R>), grid, dim3(256), 0, H.stream, P, M);
}

// Persistent multi-class march: as many workgroups as stay resident (or
// H.persist per CU), each taking every G-th column tile.
template <int NDIM, int TX, int TY, int NT, int RS, typename R, int OCC = 0>
void launch_march_multi_k(const MarchLaunch& M, const MarchHost& H) {
  static int resident = -1;
  if (resident < 0) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, k_march_multi<NDIM, TX, TY, NT, RS, R, OCC>, NT, 0) != hipSuccess ||
        per < 1)
      per = 1;
    resident = per;
  }
  long G = (long)H.numCU * resident * (H.persist > 0 ? H.persist : 1);
  if (G > H.nItems) G = H.nItems;
  if (G < 1) G = 1;
  hipLaunchKernelGGL((k_march_multi<NDIM, TX, TY, NT, RS, R, OCC>), dim3((unsigned)G), dim3(NT), 0,
                     H.stream, H.dcls, M, H.ditems, H.nItems, H.dfaceStart);
}

template <int NDIM, int NT, int RS, typename R, int OCC = 0>
void launch_march_mixw_k(const MarchLaunch& M, const MarchHost& H) {
  static int resident = -1;
  if (resident < 0) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_march_mixw<NDIM, NT, RS, R, OCC>, NT,
                                                     0) != hipSuccess ||
        per < 1)
      per = 1;
    resident = per;
  }
  long G = (long)H.numCU * resident * (H.persist > 0 ? H.persist : 1);
  if (G > H.nItems) G = H.nItems;
  if (G < 1) G = 1;
  hipLaunchKernelGGL((k_march_mixw<NDIM, NT, RS, R, OCC>), dim3((unsigned)G), dim3(NT), 0,
                     H.stream, H.dcls, M, H.ditems, H.nItems, H.dfaceStart);
}

template <int RS, typename R>
void launch_march_multi_r(int ndim, int TX, const MarchLaunch& M, const MarchHost& H) {
  if (TX == kPerClassTX) {  // 16- and 8-wide column tiles in one launch
    if constexpr (sizeof(R) == 4) {
      if (ndim == 3 && H.tuneOcc == 4) {
        launch_march_mixw_k<3, 512, RS, R, 4>(M, H);
        return;
      }
    }
    if constexpr (sizeof(R) == 8) {  // FP64: 768 threads, as k_march (quadrant grid
      if (ndim == 3) {               // march 15.4 -> 14.8 ms, though 26 registers spill)
        launch_march_mixw_k<3, 768, RS, R, 3>(M, H);
        return;
      }
    }
    if (ndim == 3) launch_march_mixw_k<3, 512, RS, R>(M, H);
    else launch_march_mixw_k<2, 256, RS, R>(M, H);
    return;
  }
  if constexpr (sizeof(R) == 4) {  // FP32 3-D: the autotuner may ask for 4 waves per SIMD
    if (ndim == 3 && H.tuneOcc == 4) {
      if (TX == 16) launch_march_multi_k<3, 16, 8, 512, RS, R, 4>(M, H);
      else launch_march_multi_k<3, 8, 8, 256, RS, R, 4>(M, H);
      return;
    }
  }
  if constexpr (sizeof(R) == 8) {  // FP64 3-D 16x8: 768 threads, as k_march
    if (ndim == 3 && TX == 16) {
      launch_march_multi_k<3, 16, 8, 768, RS, R, 3>(M, H);
      return;
    }
  }
  if (ndim == 3 && TX == 16) launch_march_multi_k<3, 16, 8, 512, RS, R>(M, H);
  else if (ndim == 3) launch_march_multi_k<3, 8, 8, 256, RS, R>(M, H);
  else if (TX == 16) launch_march_multi_k<2, 16, 8, 256, RS, R>(M, H);
  else launch_march_multi_k<2, 8, 8, 256, RS, R>(M, H);
}

template void launch_march_r<FX_MARCH_RS, FX_MARCH_R>(const DevParams&, const MarchLaunch&, int,
                                                      dim3, const MarchHost&);
template void launch_march_multi_r<FX_MARCH_RS, FX_MARCH_R>(int, int, const MarchLaunch&,
                                                            const MarchHost&);

}  // namespace fx
