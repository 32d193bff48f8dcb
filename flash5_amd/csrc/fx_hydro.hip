// fx_hydro.hip -- MI355X (gfx950) kernels and C ABI of the unsplit hydro
// block update on the ideal gamma-law EOS.
//
// Pipeline per call (one launch each, all tiles of the call batched):
//   k_states   one thread per (tile, direction, state cell of that
//              direction's band): flattening coefficient, PPM with
//              characteristic tracing, 2-D CTU transverse corrections,
//              3-D full-CTU corner terms, positivity fallback
//              (hy_getRiemannState.F90:356-1327, hy_dataReconstOneStep.F90)
//   k_fluxes   one thread per (tile, direction, face): Riemann solver and
//              aux internal-energy fluxes (hy_getFaceFlux.F90:289-500)
//   k_update   one thread per interior cell: conservative update, aux
//              internal-energy update, energy fix, EOS
//              (hy_unsplitUpdate.F90:440-925, hy_energyFix.F90:188-251,
//              Eos_wrapped MODE_DENS_EI); 3-D tiles of one class take
//              k_update_cols, one thread per column, instead
// The column march (fx_march.hpp) replaces k_states + k_fluxes where the
// block shape allows it (kernel path 0).
// State cells of direction d are the cells whose d-coordinate lies in
// [lo-1, hi+1] and whose other coordinates are interior: exactly the cells
// the reference stores face states for (hy_getRiemannState.F90:553-591).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fx_device.hpp"
#include "fx_mlp.hpp"
#include "fx_hydro.h"
#include "fx_common.hpp"
#include "fx_march_launch.hpp"

using namespace fx;

// Scratch of fx_hydro_advance_batch_noalloc: the status record, the tile
// table, and the face fluxes and face-averaged pressure of every block of a
// call, carved from one fixed device array.  Each GPU the library runs on
// holds its own copy until the process ends.  Host code never uses the
// array's name as an address: it takes the current device's address from
// hipGetSymbolAddress (noalloc_scratch below).
constexpr size_t kNoAllocBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {


// Face states of one state cell of direction D (hy_getRiemannState.F90:470-901
// for that direction, then :952-1327 in 3-D).  Returns false on an
// imaginary sound speed.
template <int D, int NDIM>
__device__ __forceinline__ bool state_cell(const DevParams& P, const gdouble* __restrict__ U,
                                           const double del[3], double dt, long c, double Wp[8],
                                           double Wn[8], DevStatus* status) {
  constexpr int O1 = (D + 1) % 3, O2 = (D + 2) % 3;
  constexpr bool A1 = O1 < NDIM, A2 = O2 < NDIM;
  const long sd = axis_stride(P, D);
  double Vc[8];
  {
    double Vm[8], Vp[8], Vmm[8], Vpp[8];
    load_prim(P, U, c, Vc);
    load_prim(P, U, c - sd, Vm);
    load_prim(P, U, c + sd, Vp);
    load_prim(P, U, c - 2 * sd, Vmm);
    load_prim(P, U, c + 2 * sd, Vpp);
    Eig e0;
    const double flat = P.flat ? flat_coeff<D>(P, U, c, sd) : 0.;
    if (!eig_make<D>(Vc, e0)) return false;
    if (!ppm_full<D>(P, dt / del[D], del[D], Vmm, Vm, Vc, Vp, Vpp, flat, e0, Wp, Wn)) return false;
  }
  // transverse sigma at this cell and 2-D CTU corrections, in ascending
  // order of the transverse direction (hy_dataReconstOneStep.F90:906-951)
  double s1[8], s2[8];
  Eig e1, e2;
  if constexpr (A1) {
    if (!sigma_at<O1>(P, U, c, s1, &e1)) return false;
  }
  if constexpr (A2) {
    if (!sigma_at<O2>(P, U, c, s2, &e2)) return false;
  }
  auto correct = [&](const double* sg, int o) {
    const double hdt = 0.5 * dt / del[o];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      Wn[v] = Wn[v] - hdt * sg[v];
      Wp[v] = Wp[v] - hdt * sg[v];
    }
  };
  if constexpr (A1 && A2) {
    if constexpr (O1 < O2) {
      correct(s1, O1);
      correct(s2, O2);
    } else {
      correct(s2, O2);
      correct(s1, O1);
    }
  } else if constexpr (A1) {
    correct(s1, O1);
  } else if constexpr (A2) {
    correct(s2, O2);
  }
  if (Wn[P_DENS] < 0. || Wp[P_DENS] < 0. || Wn[P_PRES] < 0. || Wp[P_PRES] < 0.) {
    atomicAdd(&status->fallbacks, 1ULL);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      Wn[v] = Vc[v];
      Wp[v] = Vc[v];
    }
  }
  // 3-D full CTU: term A = sigma of O2 at the +-1 neighbours along O1,
  // upwinded with O1's eigen system here; term B = sigma of O1 along O2
  // with O2's eigen system (X: YZ then ZY, Y: ZX then XZ, Z: XY then YX).
  if constexpr (NDIM == 3) {
    if (P.ctu3d) {
      double tA[8], tB[8];
      {
        const long sn = axis_stride(P, O1);
        double sm[8], sp[8];
        if (!sigma_at<O2>(P, U, c - sn, sm, nullptr)) return false;
        if (!sigma_at<O2>(P, U, c + sn, sp, nullptr)) return false;
        trans_sigma<O1>(P, sm, s2, sp, e1, tA);
      }
      {
        const long sn = axis_stride(P, O2);
        double sm[8], sp[8];
        if (!sigma_at<O1>(P, U, c - sn, sm, nullptr)) return false;
        if (!sigma_at<O1>(P, U, c + sn, sp, nullptr)) return false;
        trans_sigma<O2>(P, sm, s1, sp, e2, tB);
      }
      const double fac = dt * dt / (6. * del[O1] * del[O2]);
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        Wp[v] = Wp[v] + (tA[v] + tB[v]) * fac;
        Wn[v] = Wn[v] + (tA[v] + tB[v]) * fac;
      }
      if (Wp[P_DENS] <= 0. || Wn[P_DENS] <= 0. || Wp[P_PRES] <= 0. || Wn[P_PRES] <= 0.) {
        atomicAdd(&status->fallbacks, 1ULL);
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          Wn[v] = Vc[v];
          Wp[v] = Vc[v];
        }
      }
    }
  }
  return true;
}

template <int NDIM>
__global__ void __launch_bounds__(kThreads) k_states(DevParams P, Launch L) {
  const long perTile = P.scellsPerTile;
  const long total = perTile * L.nTiles;
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int t = (int)(gid / perTile);
    long r = gid - (long)t * perTile;
    int d = 0;
    while (d < NDIM - 1 && r >= P.bcells[d]) {
      r -= P.bcells[d];
      ++d;
    }
    // band-local coordinates (x fastest)
    const int x0 = P.bext[d][0], x1 = P.bext[d][1];
    const int t0 = (int)(r % x0);
    const int t1 = (int)((r / x0) % x1);
    const int t2 = (int)(r / ((long)x0 * x1));
    int ci[3] = {t0 + P.lo[0], t1 + P.lo[1], t2 + P.lo[2]};
    if (d == 0) ci[0] -= 1;
    else if (d == 1) ci[1] -= 1;
    else ci[2] -= 1;
    const fx_tile& T = L.tiles[t];
    const long c = cidx(P, ci[0], ci[1], ci[2]);
    double Wp[8], Wn[8];
    bool ok;
    const double dt = step_dt(L.dt, L.dtp);
    const gdouble* Ut = as_global(T.U);
    if (d == 0) ok = state_cell<0, NDIM>(P, Ut, T.del, dt, c, Wp, Wn, L.status);
    else if (d == 1) ok = state_cell<1, NDIM>(P, Ut, T.del, dt, c, Wp, Wn, L.status);
    else ok = state_cell<2, NDIM>(P, Ut, T.del, dt, c, Wp, Wn, L.status);
    if (!ok) {
      report_error(L.status, 1, t, ci[0], ci[1], ci[2], d);
      continue;
    }
    // store Wp (comps 0..5) and Wn (6..11): dens, vx, vy, vz, pres, eint
    double* S = L.states + (long)t * perTile * kStateComps + P.boff[d] * kStateComps;
    const long bc = P.bcells[d];
    S[0 * bc + r] = Wp[P_DENS];
    S[1 * bc + r] = Wp[P_VELX];
    S[2 * bc + r] = Wp[P_VELY];
    S[3 * bc + r] = Wp[P_VELZ];
    S[4 * bc + r] = Wp[P_PRES];
    S[5 * bc + r] = Wp[P_EINT];
    S[6 * bc + r] = Wn[P_DENS];
    S[7 * bc + r] = Wn[P_VELX];
    S[8 * bc + r] = Wn[P_VELY];
    S[9 * bc + r] = Wn[P_VELZ];
    S[10 * bc + r] = Wn[P_PRES];
    S[11 * bc + r] = Wn[P_EINT];
  }
}

// band-local linear index of block cell (i,j,k) in direction d's band
__device__ __forceinline__ long band_index(const DevParams& P, int d, int i, int j, int k) {
  int a = i - P.lo[0], b = j - P.lo[1], cc = k - P.lo[2];
  if (d == 0) a += 1;
  else if (d == 1) b += 1;
  else cc += 1;
  return ((long)cc * P.bext[d][1] + b) * P.bext[d][0] + a;
}


template <int RS>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_fluxes(DevParams P, Launch L) {
  const long perTile = P.nface[0] + P.nface[1] + P.nface[2];
  const long total = perTile * L.nTiles;
  ExactStats es{0ULL, 0ULL};
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int t = (int)(gid / perTile);
    long r = gid - (long)t * perTile;
    int d = 0;
    while (d < P.ndim - 1 && r >= P.nface[d]) {
      r -= P.nface[d];
      ++d;
    }
    const int f0 = P.fext[d][0], f1 = P.fext[d][1];
    const int a0 = (int)(r % f0);
    const int a1 = (int)((r / f0) % f1);
    const int a2 = (int)(r / ((long)f0 * f1));
    const int i = a0 + P.lo[0], j = a1 + P.lo[1], k = a2 + P.lo[2];
    int im = i, jm = j, km = k;
    if (d == 0) im -= 1;
    else if (d == 1) jm -= 1;
    else km -= 1;
    const fx_tile& T = L.tiles[t];
    const gdouble* __restrict__ U = as_global(T.U);
    const double* S = L.states + (long)t * P.scellsPerTile * kStateComps + P.boff[d] * kStateComps;
    const long bc = P.bcells[d];
    const long bl = band_index(P, d, im, jm, km);
    const long br = band_index(P, d, i, j, k);
    const long cl = cidx(P, im, jm, km), cr = cidx(P, i, j, k);
    double F[8];
    int ierr;
    if constexpr (RS == FX_EXACT) {
      ierr = exact_face_flux<false, double>(P, d, S + bl, S + 6 * bc + br, bc,
                                            U[P.voff[FX_GAMC] + cl], U[P.voff[FX_GAME] + cl],
                                            U[P.voff[FX_GAME] + cr], -1.0, F, es);
    } else {
      double VL[8], VR[8];
      VL[P_DENS] = S[0 * bc + bl];
      VL[P_VELX] = S[1 * bc + bl];
      VL[P_VELY] = S[2 * bc + bl];
      VL[P_VELZ] = S[3 * bc + bl];
      VL[P_PRES] = S[4 * bc + bl];
      VL[P_EINT] = S[5 * bc + bl];
      VR[P_DENS] = S[6 * bc + br];
      VR[P_VELX] = S[7 * bc + br];
      VR[P_VELY] = S[8 * bc + br];
      VR[P_VELZ] = S[9 * bc + br];
      VR[P_PRES] = S[10 * bc + br];
      VR[P_EINT] = S[11 * bc + br];
      // face-state gammas equal the cell's (PPM.F90:733-736 under
      // FLASH_EOS_GAMMA, and the first-order fallback)
      VL[P_GAMC] = U[P.voff[FX_GAMC] + cl];
      VL[P_GAME] = U[P.voff[FX_GAME] + cl];
      VR[P_GAMC] = U[P.voff[FX_GAMC] + cr];
      VR[P_GAME] = U[P.voff[FX_GAME] + cr];
      const bool useHll = use_hll<RS>(P, U, cl, cr);
      if (d == 0) ierr = face_flux<0, RS>(P, VL, VR, useHll, F, es);
      else if (d == 1) ierr = face_flux<1, RS>(P, VL, VR, useHll, F, es);
      else ierr = face_flux<2, RS>(P, VL, VR, useHll, F, es);
    }
    if (ierr) {
      report_error(L.status, 2, t, i, j, k, d);
      continue;
    }
    gdouble* Fo = as_global(T.flux[d]);
    const long nf = P.nface[d];
#pragma unroll
    for (int q = 0; q < FX_NFLUX; ++q) Fo[q * nf + r] = F[q];
  }
  flush_exact(L.status, es);
}

// hy_shockDetectBlk.F90:93-229, one workgroup per tile.  SHOK_VAR := 0 on
// guard cells; interior cells take the Balsara-Spicer flag (undivided
// pressure jump >= 0.5 of the 3^d minimum pressure and undivided velocity
// divergence <= -0.1 of the 3^d minimum sound speed).  The workgroup marches
// along z: each plane's pressure and sound speed over the interior plus one
// guard layer go to LDS once (one division and root per cell); each thread
// takes the 3x3 minima of its interior columns and keeps those of the two
// planes below in LDS slots only it touches, so a cell's 3^d minimum is the
// minimum of three.  Minima are exact, so the result does not depend on the
// order.  Reads the interior and the first guard layer only; the SHOK_VAR
// slot is distinct from the slots read.  cls == nullptr: every tile has the
// parameters P0.  Dynamic LDS: shock_lds_doubles() of the largest class.
__host__ __device__ inline long shock_lds_doubles(const DevParams& P) {
  const long nx2 = P.nxb + 2, ny2 = P.ndim > 1 ? P.nyb + 2 : 1;
  const long ncol = (long)P.nxb * (P.ndim > 1 ? P.nyb : 1);
  return 2 * nx2 * ny2 + (P.ndim > 2 ? 4 * ncol : 0);
}
constexpr long kShockLdsMax = 65536 / sizeof(double);   // dynamic LDS of one workgroup

__global__ void __launch_bounds__(kThreads) k_shock_detect(Launch L, DevParams P0,
                                                           const DevParams* __restrict__ cls,
                                                           const unsigned char* __restrict__ tcls) {
  extern __shared__ double sm[];
  const int t = blockIdx.x;
  const DevParams& P = cls ? cls[tcls[t]] : P0;
  gdouble* __restrict__ U = as_global(L.tiles[t].U);
  const int k2 = P.ndim > 1, k3 = P.ndim > 2;
  const long sy = P.nx, sz = (long)P.nx * P.ny;
  const long pr = P.voff[FX_PRES], de = P.voff[FX_DENS], gc = P.voff[FX_GAMC];
  const long vx = P.voff[FX_VELX], vy = P.voff[FX_VELY], vz = P.voff[FX_VELZ];
  const int nxb = P.nxb, nyb = k2 ? P.nyb : 1;
  const int nx2 = nxb + 2, ny2 = k2 ? nyb + 2 : 1;
  const int ncol = nxb * nyb, nplane = nx2 * ny2;
  double* __restrict__ plane = sm;                 // [2][nplane]: p, then c
  double* __restrict__ hist = sm + 2 * nplane;     // 3-D: [2 planes][2][ncol] 3x3 minima
  // guard cells of SHOK_VAR (hy_shockDetectBlk.F90:117)
  for (long c = threadIdx.x; c < P.ncell; c += blockDim.x) {
    const int i = (int)(c % P.nx), j = (int)((c / P.nx) % P.ny), k = (int)(c / sz);
    if (i < P.lo[0] || i > P.hi[0] || j < P.lo[1] || j > P.hi[1] || k < P.lo[2] || k > P.hi[2])
      U[P.shokOff + c] = 0.;
  }
  int any = 0;
  const int klo = P.lo[2] - k3, khi = P.hi[2] + k3;
  for (int k = klo; k <= khi; ++k) {
    for (int q = threadIdx.x; q < nplane; q += blockDim.x) {
      const int i = P.lo[0] - 1 + q % nx2, j = P.lo[1] - k2 + q / nx2;
      const long c = ((long)k * P.ny + j) * P.nx + i;
      const double p = U[pr + c];
      plane[q] = p;
      plane[nplane + q] = sqrt(U[gc + c] * p / U[de + c]);   // :135
    }
    __syncthreads();
    for (int q = threadIdx.x; q < ncol; q += blockDim.x) {
      const int a = q % nxb, b = q / nxb;
      double minP = plane[(b + k2) * nx2 + a + 1], minC = plane[nplane + (b + k2) * nx2 + a + 1];
      for (int y = 0; y <= 2 * k2; ++y)
        for (int x = 0; x <= 2; ++x) {
          minP = fmin(minP, plane[(b + y) * nx2 + a + x]);
          minC = fmin(minC, plane[nplane + (b + y) * nx2 + a + x]);
        }
      int kc = k;
      if (k3) {   // the cell below: planes k-2, k-1 from the history, k from above
        const double p1 = hist[q], c1 = hist[ncol + q];
        const double p2 = hist[2 * ncol + q], c2 = hist[3 * ncol + q];
        hist[2 * ncol + q] = p1;
        hist[3 * ncol + q] = c1;
        hist[q] = minP;
        hist[ncol + q] = minC;
        if (k < klo + 2) continue;
        minP = fmin(fmin(minP, p1), p2);
        minC = fmin(fmin(minC, c1), c2);
        kc = k - 1;
      }
      const long c = ((long)kc * P.ny + P.lo[1] + b) * P.nx + P.lo[0] + a;
      double divv = U[vx + c + 1] - U[vx + c - 1];               // :157-164
      if (k2) divv = divv + U[vy + c + sy] - U[vy + c - sy];
      if (k3) divv = divv + U[vz + c + sz] - U[vz + c - sz];
      divv = 0.5 * divv;
      const double gradPx = 0.5 * (U[pr + c + 1] - U[pr + c - 1]);  // :167-174
      double gradPy = 0., gradPz = 0.;
      if (k2) gradPy = 0.5 * (U[pr + c + sy] - U[pr + c - sy]);
      if (k3) gradPz = 0.5 * (U[pr + c + sz] - U[pr + c - sz]);
      const bool sw1 = fabs(gradPx) + fabs(gradPy) + fabs(gradPz) >= 0.5 * minP;  // :175-178
      const bool sw2 = -(0.1 * minC) >= divv;                                     // :180-182
      const bool f = sw1 && sw2;
      any |= f;
      U[P.shokOff + c] = f ? 1. : 0.;
    }
    __syncthreads();   // the next plane reuses `plane`
  }
  if (__syncthreads_or(any) && threadIdx.x == 0) atomicAdd(&L.status->shockTiles, 1ULL);
}


#include "fx_learned.hpp"
#include "fx_amr.hpp"

// FROM_STATES: face-averaged pressure from this call's states; otherwise
// read from the tile's pavg array (hy_updateSolution path).
// Interior cell r of tile t: conservative update, energy fix and EOS
// (hy_unsplitUpdate.F90:440-915, hy_energyFix.F90:188-251, Eos_wrapped).
template <bool FROM_STATES, bool WRITE_PAVG_ONLY>
__device__ __forceinline__ void update_cell(const DevParams& P, const Launch& L, int t, long r,
                                            const unsigned char* __restrict__ inPlace) {
  const int nxi = P.nxb, nyi = P.ndim > 1 ? P.nyb : 1;
  {
    const fx_tile& T = L.tiles[t];
    const bool doUpdate = WRITE_PAVG_ONLY ? false : (inPlace ? inPlace[t] != 0 : true);
    const int ri = (int)r;  // a tile holds < 2^31 cells
    const int i = ri % nxi + P.lo[0];
    const int j = (ri / nxi) % nyi + P.lo[1];
    const int k = ri / (nxi * nyi) + P.lo[2];
    double pres = 0.;
    if (P.aux) {
      if (FROM_STATES) {
        const double* Sb = L.states + (long)t * P.scellsPerTile * kStateComps;
        for (int d = 0; d < P.ndim; ++d) {
          const double* S = Sb + P.boff[d] * kStateComps;
          const long bc = P.bcells[d];
          const long bi = band_index(P, d, i, j, k);
          const double pN = S[10 * bc + bi], pP = S[4 * bc + bi];
          if (d == 0) pres = 0.5 * (pN + pP);
          else if (d == 1) pres = pres * 0.5 + (pN + pP) * 0.25;
          else pres = div_const<6>(pres * 4.0 + (pN + pP));
        }
        if (!doUpdate && T.pavg) as_global(T.pavg)[r] = pres;
      } else {
        pres = as_global(T.pavg)[r];
      }
    }
    if (!doUpdate) return;
    gdouble* U = as_global(T.U);
    const long c = cidx(P, i, j, k);
    const double dx = T.del[0], dy = P.ndim > 1 ? T.del[1] : 1., dz = P.ndim > 2 ? T.del[2] : 1.;
    const double dt = step_dt(L.dt, L.dtp);
    // One direction at a time: its face pair, the conserved-variable update,
    // the aux e_int term and the pressure source.  Each quantity still takes
    // its x, y, z terms in that order, with only one face pair live.
    const double dens0 = U[P.voff[FX_DENS] + c];
    double U0[5];
    U0[0] = dens0;
    U0[1] = U[P.voff[FX_VELX] + c] * dens0;
    U0[2] = U[P.voff[FX_VELY] + c] * dens0;
    U0[3] = U[P.voff[FX_VELZ] + c] * dens0;
    U0[4] = dens0 * U[P.voff[FX_ENER] + c];
    double IntEner = 0.;
    if (P.aux) IntEner = dens0 * U[P.voff[FX_EINT] + c];
    double Sphys[5] = {0., 0., 0., 0., 0.};
    auto apply = [&](int d, double del, long fL, long fR) {
      const gdouble* Fd = as_global(T.flux[d]);
      const long nf = P.nface[d];
      double FL[8], FR[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        FL[q] = Fd[q * nf + fL];
        FR[q] = Fd[q * nf + fR];
      }
      Sphys[1 + d] = (FL[5] - FR[5]) * (1.0 - P.fP) / del;
      if (P.aux) IntEner = IntEner + dt / del * (FL[6] - FR[6] + pres * (FL[7] - FR[7]));
#pragma unroll
      for (int q = 0; q < 5; ++q) U0[q] = U0[q] - dt / del * (FR[q] - FL[q]);
    };
    {
      const long f0 = face_index(P, 0, i, j, k);
      apply(0, dx, f0, f0 + 1);
    }
    if (P.ndim >= 2) {
      const long g0 = face_index(P, 1, i, j, k);
      apply(1, dy, g0, g0 + P.fext[1][0]);
    }
    if (P.ndim == 3) {
      const long h0 = face_index(P, 2, i, j, k);
      apply(2, dz, h0, h0 + (long)P.fext[2][0] * P.fext[2][1]);
    }
    if (P.aux) IntEner = IntEner + dt * 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q) U0[q] = U0[q] + dt * (0. + Sphys[q]);
    const double densNew = fmax(U0[0], P.smlrho);
    const double vx = U0[1] / densNew, vy = U0[2] / densNew, vz = U0[3] / densNew;
    const double ener = U0[4] / densNew;
    double eintAux = P.aux ? IntEner / densNew : U[P.voff[FX_EINT] + c];
    // energy fix (hy_energyFix.F90:188-251)
    const double ekin = 0.5 * (vx * vx + vy * vy + vz * vz);
    const double eint = ener - ekin - 0.0;
    double newEint;
    if (!P.aux || eint > P.eintSwitch * (ekin + 0.0)) newEint = fmax(P.smallE, eint);
    else newEint = fmax(P.smallE, eintAux);
    const double enerFix = newEint + ekin;
    const EosOut eo = eos_dens_ei(P, densNew, vx, vy, vz, newEint, enerFix);
    U[P.voff[FX_DENS] + c] = densNew;
    U[P.voff[FX_VELX] + c] = vx;
    U[P.voff[FX_VELY] + c] = vy;
    U[P.voff[FX_VELZ] + c] = vz;
    U[P.voff[FX_ENER] + c] = eo.ener;
    U[P.voff[FX_EINT] + c] = eo.eint;
    U[P.voff[FX_PRES] + c] = eo.pres;
    U[P.voff[FX_TEMP] + c] = eo.temp;
    U[P.voff[FX_GAMC] + c] = P.gamma;
    U[P.voff[FX_GAME] + c] = eo.game;
  }
}

template <bool FROM_STATES, bool WRITE_PAVG_ONLY>
__global__ void __launch_bounds__(kThreads) k_update(DevParams P, Launch L,
                                                     const unsigned char* __restrict__ inPlace) {
  const long perTile = P.icells;
  const long total = perTile * L.nTiles;
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int t = (int)(gid / perTile);
    update_cell<FROM_STATES, WRITE_PAVG_ONLY>(P, L, t, gid - (long)t * perTile, inPlace);
  }
}

// The update of 3-D tiles by columns: one thread per (tile, i, j) column
// marches k upward and keeps the z face between two cells in registers, so
// each z-face flux is read once (per cell, k_update reads both of its z
// faces): 1.79 against 1.84 ms on 4096 x 16^3.  upd_start /
// upd_finish take each quantity's x, y and z terms in update_cell's order,
// so the bits are update_cell's.
__global__ void __launch_bounds__(kThreads)
k_update_cols(DevParams P, Launch L, const unsigned char* __restrict__ inPlace) {
  const int nx = P.nxb, ny = P.nyb;
  const long ncol = (long)nx * ny;
  const long total = ncol * L.nTiles;
  const double dt = step_dt(L.dt, L.dtp);
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int t = (int)(gid / ncol);
    if (inPlace && !inPlace[t]) continue;
    const int col = (int)(gid - (long)t * ncol);
    const int i = col % nx + P.lo[0], j = col / nx + P.lo[1];
    const fx_tile& T = L.tiles[t];
    gdouble* const U = as_global(T.U);
    const gdouble* const Fx = as_global(T.flux[0]);
    const gdouble* const Fy = as_global(T.flux[1]);
    const gdouble* const Fz = as_global(T.flux[2]);
    const long nfx = P.nface[0], nfy = P.nface[1], nfz = P.nface[2];
    const long zStep = (long)P.fext[2][0] * P.fext[2][1];
    const double dx = T.del[0], dy = T.del[1], dz = T.del[2];
    long hz = face_index(P, 2, i, j, P.lo[2]);
    double FLz[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) FLz[q] = Fz[q * nfz + hz];
    for (int k = P.lo[2]; k <= P.hi[2]; ++k) {
      const long c = cidx(P, i, j, k);
      const long f0 = face_index(P, 0, i, j, k), g0 = face_index(P, 1, i, j, k);
      const long r = ((long)(k - P.lo[2]) * ny + (j - P.lo[1])) * nx + (i - P.lo[0]);
      double FLx[8], FRx[8], FLy[8], FRy[8], FRz[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        FLx[q] = Fx[q * nfx + f0];
        FRx[q] = Fx[q * nfx + f0 + 1];
        FLy[q] = Fy[q * nfy + g0];
        FRy[q] = Fy[q * nfy + g0 + P.fext[1][0]];
        FRz[q] = Fz[q * nfz + hz + zStep];
      }
      const UpdPart part = upd_start(P, dt, dx, dy, U[P.voff[FX_DENS] + c], U[P.voff[FX_VELX] + c],
                                     U[P.voff[FX_VELY] + c], U[P.voff[FX_VELZ] + c],
                                     U[P.voff[FX_ENER] + c], U[P.voff[FX_EINT] + c],
                                     P.aux ? as_global(T.pavg)[r] : 0., FLx, FRx, FLy, FRy);
      double out[10];
      upd_finish(P, dt, dz, part, FLz, FRz, P.aux ? 0. : U[P.voff[FX_EINT] + c], out);
#pragma unroll
      for (int v = 0; v < 10; ++v) U[P.voff[upd_var(v)] + c] = out[v];
#pragma unroll
      for (int q = 0; q < 8; ++q) FLz[q] = FRz[q];
      hz += zStep;
    }
  }
}

// the update over tiles of several block classes in one launch
__global__ void __launch_bounds__(kThreads) k_update_multi(Launch L, ClassMap M,
                                                           const unsigned char* __restrict__ inPlace) {
  const long total = M.start[L.nTiles];
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int t = find_tile_w(M.start, M.wtab, L.nTiles, gid);
    update_cell<false, false>(M.cls[M.tcls[t]], L, t, gid - M.start[t], inPlace);
  }
}

// the CFL dt of each block (one workgroup per block), as dt_pick_block
// describes; the CFL is cflDev[0] when cflDev is set (a lowered CFL kept on
// the device), else cfl
__global__ void __launch_bounds__(kThreads) k_dt(DevParams P, const double* __restrict__ Ubase,
                                                 long stride, double idx, double idy, double idz,
                                                 double cfl, const double* cflDev,
                                                 double* __restrict__ dtBlock) {
  const double c = cflDev ? cflDev[0] : cfl;
  const double* U = Ubase + (long)blockIdx.x * stride;
  const int nxi = P.nxb, nyi = P.ndim > 1 ? P.nyb : 1, n = (int)P.icells;
  DtPick best = dt_pick_none();
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const int i = r % nxi + P.lo[0];
    const int j = (r / nxi) % nyi + P.lo[1];
    const int k = r / (nxi * nyi) + P.lo[2];
    const long cc = cidx(P, i, j, k);
    const double snd2 = U[P.voff[FX_GAMC] + cc] * U[P.voff[FX_PRES] + cc] / U[P.voff[FX_DENS] + cc];
    const double cs = sqrt(snd2);
    double dtl = (fabs(U[P.voff[FX_VELX] + cc] - 0.) + cs) * idx;
    if (P.ndim > 1) dtl = fmax(dtl, (fabs(U[P.voff[FX_VELY] + cc] - 0.) + cs) * idy);
    if (P.ndim > 2) dtl = fmax(dtl, (fabs(U[P.voff[FX_VELZ] + cc] - 0.) + cs) * idz);
    dt_pick_cell(best, dtl, c, r);
  }
  const double dt = dt_pick_block(best, c);
  if (threadIdx.x == 0) dtBlock[blockIdx.x] = dt;
}

// dt = the minimum of the blocks' dts (Hydro_computeDt.F90:284-288), kept on
// the device; t += dt
__global__ void __launch_bounds__(kThreads) k_dt_final(const double* __restrict__ dtBlock, int n,
                                                       double* dtDev, double* tDev) {
  __shared__ double red[kThreads];
  double m = 1.e300;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, dtBlock[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    dtDev[0] = red[0];
    tDev[0] = tDev[0] + red[0];
  }
}

// hy_cfl from one graph-captured step to the next, launched at the end of
// each step.  cfl[0] holds the CFL of the next dt, cfl[1] the run's fallback
// count at the end of the previous step.  A step resets the CFL to orig
// before the detector when the detector runs (Hydro.F90:104-108), a flagged
// cell lowers it to shockCap under shockLowerCFL (hy_shockDetectBlk.F90:
// 228-240), the recovery before the fluxes raises it by 1.2 up to orig
// (Hydro.F90:128-141), and under fallbackLowerCFL a state that fell back to
// first order lowers it to fallbackCap (hy_getRiemannState.F90:934-941,
// 1303-1319).  A cap of 0 is off.
__global__ void k_step_cfl(const DevStatus* s, double orig, int detect, double shockCap,
                           double fallbackCap, double* cfl) {
  double c = detect ? orig : cfl[0];
  if (shockCap > 0. && s->shockTiles) c = fmin(c, shockCap);
  c = 1.2 * c < orig ? 1.2 * c : orig;
  const double fb = (double)s->fallbacks;
  if (fallbackCap > 0. && fb != cfl[1]) c = fmin(c, fallbackCap);
  cfl[1] = fb;
  cfl[0] = c;
}

__global__ void k_zero_shock_tiles(DevStatus* s) { s->shockTiles = 0; }

__global__ void k_reset_status(DevStatus* s) {
  s->code = 0;
  s->tile = s->i = s->j = s->k = s->dir = -1;
  s->fallbacks = 0;
  s->newtonIters = 0;
  s->newtonFallbacks = 0;
  s->shockTiles = 0;
}

// Tile table of a strided batch built on the device (the no-allocation
// batch step): block b's U at U + b*stride, its flux arrays (directions
// below ndim) and pavg carved from flux / pavg at fixed per-block strides.
__global__ void __launch_bounds__(kThreads) k_batch_tiles(fx_tile* __restrict__ tiles, double* U,
                                                          long n, long stride, double dx, double dy,
                                                          double dz, double* flux, long nf0,
                                                          long nf1, long nf2, int ndim,
                                                          double* pavg, long icells) {
  const long perBlock = (nf0 + (ndim > 1 ? nf1 : 0) + (ndim > 2 ? nf2 : 0)) * FX_NFLUX;
  for (long b = (long)blockIdx.x * blockDim.x + threadIdx.x; b < n;
       b += (long)gridDim.x * blockDim.x) {
    fx_tile t;
    t.U = U + b * stride;
    double* f = flux + b * perBlock;
    t.flux[0] = f;
    t.flux[1] = ndim > 1 ? f + nf0 * FX_NFLUX : nullptr;
    t.flux[2] = ndim > 2 ? f + (nf0 + nf1) * FX_NFLUX : nullptr;
    t.pavg = pavg + b * icells;
    t.del[0] = dx;
    t.del[1] = dy;
    t.del[2] = dz;
    t.level = 1;
    t.block_class = 0;
    tiles[b] = t;
  }
}

// ---------------------------------------------------------------------
// host state
// ---------------------------------------------------------------------
// ---- flux register (fx_fluxreg_put / fx_fluxreg_conserve) ----
struct RegLaunch {
  const fx_tile* tiles;
  const fx_fluxreg_face* faces;
  int nFaces;
  double* reg;
  const DevParams* cls;        // block classes, or null: every tile has P's shape
  const unsigned char* tcls;   // class of each tile of the list
};

__device__ __forceinline__ long reg_face(const DevParams& P, int d, int a, int b, int c) {
  return ((long)c * P.fext[d][1] + b) * P.fext[d][0] + a;
}

// reg = sum over the covering fine faces of coef * F, transverse offsets
// ascending with the higher axis outermost (AMReX fluxreg_fineadd order)
__global__ void __launch_bounds__(kThreads) k_fluxreg_put(DevParams P, RegLaunch R, double coef) {
  const long total = (long)R.nFaces * FX_NFLUX;
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int q = (int)(gid / R.nFaces);
    const int e = (int)(gid - (long)q * R.nFaces);
    const fx_fluxreg_face f = R.faces[e];
    const int d = f.dir;
    const DevParams& Pt = R.cls ? R.cls[R.tcls[f.fine_tile]] : P;
    const gdouble* F = as_global(R.tiles[f.fine_tile].flux[d]) + (long)q * Pt.nface[d];
    const int inner = d == 0 ? 1 : 0, outer = d == 2 ? 1 : 2;
    const int nIn = P.ndim > 1 ? 2 : 1, nOut = P.ndim > 2 ? 2 : 1;
    double acc = 0.0;
    for (int o = 0; o < nOut; ++o)
      for (int in = 0; in < nIn; ++in) {
        int idx[3] = {f.fi, f.fj, f.fk};
        idx[inner] += in;
        idx[outer] += o;
        acc = acc + coef * F[reg_face(Pt, d, idx[0], idx[1], idx[2])];
      }
    R.reg[gid] = acc;
  }
}

__global__ void __launch_bounds__(kThreads) k_fluxreg_conserve(DevParams P, RegLaunch R) {
  const long total = (long)R.nFaces * FX_NFLUX;
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int q = (int)(gid / R.nFaces);
    const int e = (int)(gid - (long)q * R.nFaces);
    const fx_fluxreg_face f = R.faces[e];
    const int d = f.dir;
    const DevParams& Pt = R.cls ? R.cls[R.tcls[f.coarse_tile]] : P;
    gdouble* F = as_global(R.tiles[f.coarse_tile].flux[d]) + (long)q * Pt.nface[d];
    F[reg_face(Pt, d, f.ci, f.cj, f.ck)] = R.reg[gid] * 1.0;
  }
}

// ---- guard-cell fill on one level (Grid_fillGuardCells for a uniform block
// grid, Grid/GridMain/.../Grid_fillGuardCells.F90) ----
// Pass `axis` fills the guard slabs along that axis: rows interior in the
// later axes and guard-inclusive in the earlier ones, so three passes also
// fill edges and corners from the neighbours' already filled guards.
// nbr[b*6 + 2*axis + side]: neighbour block, or FX_BC_OUTFLOW (zero
// gradient) or FX_BC_REFLECT (mirror, normal velocity negated).  onlyVar >= 0
// fills that variable slot alone.
__global__ void __launch_bounds__(kThreads) k_fill_guards(DevParams P, double* __restrict__ U,
                                                          long nBlocks, long stride,
                                                          const int* __restrict__ nbr, int axis,
                                                          int onlyVar) {
  const int ng = P.ng;
  const int ext[3] = {P.nx, P.ny, P.nz};
  // extent of the slab along the two other axes
  int r0 = 0, n0 = 1, r1 = 0, n1 = 1;   // ranges of the other axes (start, count)
  const int oa = axis == 0 ? 1 : 0, ob = axis == 2 ? 1 : 2;
  if (oa < P.ndim) {
    r0 = oa < axis ? 0 : P.lo[oa];
    n0 = oa < axis ? ext[oa] : (P.hi[oa] - P.lo[oa] + 1);
  }
  if (ob < P.ndim) {
    r1 = ob < axis ? 0 : P.lo[ob];
    n1 = ob < axis ? ext[ob] : (P.hi[ob] - P.lo[ob] + 1);
  }
  const long perBlock = 2L * ng * n0 * n1;
  const long total = perBlock * nBlocks;
  const long sAx = axis == 0 ? 1L : (axis == 1 ? (long)P.nx : (long)P.nx * P.ny);
  const long sA = oa == 0 ? 1L : (oa == 1 ? (long)P.nx : (long)P.nx * P.ny);
  const long sB = ob == 0 ? 1L : (ob == 1 ? (long)P.nx : (long)P.nx * P.ny);
  const int lo = P.lo[axis], hi = P.hi[axis];
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const long b = gid / perBlock;
    long r = gid - b * perBlock;
    const int ia = (int)(r % n0);
    r /= n0;
    const int ib = (int)(r % n1);
    r /= n1;
    const int gg = (int)(r % ng) + 1;          // 1..ng away from the interior
    const int side = (int)(r / ng);
    const long base = (long)(r0 + ia) * sA + (long)(r1 + ib) * sB;
    const long dst = base + (long)(side ? hi + gg : lo - gg) * sAx;
    const int nb = nbr[b * 6 + 2 * axis + side];
    const double* src;
    long sidx;
    double vsign = 1.0;
    if (nb >= 0) {
      src = U + (long)nb * stride;
      sidx = base + (long)(side ? lo + gg - 1 : hi - gg + 1) * sAx;
    } else {
      src = U + b * stride;
      if (nb == FX_BC_REFLECT) {
        sidx = base + (long)(side ? hi - gg + 1 : lo + gg - 1) * sAx;
        vsign = -1.0;
      } else {
        sidx = base + (long)(side ? hi : lo) * sAx;
      }
    }
    double* dstU = U + b * stride;
    const int v0 = onlyVar >= 0 ? onlyVar : 0, v1 = onlyVar >= 0 ? onlyVar + 1 : P.nunk;
    for (int v = v0; v < v1; ++v) {
      const long off = (long)v * P.ncell;
      double val = src[off + sidx];
      if (v == P.velSlot[axis]) val *= vsign;
      dstU[off + dst] = val;
    }
  }
}

struct Lib {
  bool init = false;
  fx_params prm{};
  DevParams dp{};
  hipStream_t stream = nullptr;
  int async = 0;
  double* states = nullptr;
  size_t statesCap = 0;   // doubles
  double* flux = nullptr; // library-owned flux arrays for tiles without them
  size_t fluxCap = 0;
  double* pavg = nullptr; // library-owned face-averaged pressure arrays
  size_t pavgCap = 0;
  fx_tile* dtiles = nullptr;
  size_t dtilesCap = 0;
  fx_tile* dregTiles = nullptr;        // tile list of a register call
  size_t dregTilesCap = 0;
  fx_fluxreg_face* dfaces = nullptr;   // face list of a register call
  size_t dfacesCap = 0;
  unsigned char* dregCls = nullptr;    // tile classes of a register call
  size_t dregClsCap = 0;
  std::vector<unsigned char> hregCls;
  MlpPacked* mlpPacked = nullptr;      // BF16 fragments for fx_mlp_* / fx_riemann_batch
  MlpPacked* seedPacked = nullptr;     // network seeding the exact solver in hydro steps
  bool seedReady = false;
  double seedGate = 0.2;               // fx_mlp_set_gate
  float* seedY = nullptr;              // network outputs per interface / face
  size_t seedYCap = 0;
  unsigned* strongList = nullptr;      // faces at or above the gate (k_face_gate)
  size_t strongListCap = 0;
  unsigned* strongCount = nullptr;
  double* fstates = nullptr;           // face states for the separate Riemann stage
  size_t fstatesCap = 0;
  int solveUpdOff = 0;                 // 1: exact batch steps keep the separate update kernel
  double* seedFlux = nullptr;          // fluxes of the seeded strong faces (batch steps), in R
  size_t seedFluxCap = 0;
  unsigned* seedBits = nullptr;        // one bit per face: flux in seedFlux (kept all zero between steps)
  size_t seedBitsCap = 0;
  const double* dtDev = nullptr;       // set while capturing steps whose dt lives on the device
  double* dtBuf = nullptr;             // [dt, t, next dt's CFL, fallbacks so far] of graph runs
  double runCfl = 0.;                  // the CFL argument of the graph-captured run
  hipStream_t capStream = nullptr;     // library stream for graph capture (default priority)
  std::vector<int> hnbr;               // neighbour table last uploaded (host copy)
  int* dnbr = nullptr;
  size_t dnbrCap = 0;
  long lastFaces = 0;                  // faces of the last exact-solver step
  int lastTiles = 0;
  unsigned char* dinplace = nullptr;
  size_t dinplaceCap = 0;
  DevStatus* status = nullptr;
  double* partial = nullptr;
  int partialCap = 0;
  std::vector<fx_tile> htiles;
  int numCU = 256;
  // block classes (fx_hydro_set_block_classes); empty: the fx_hydro_init shape
  std::vector<DevParams> hcls;
  DevParams* dcls = nullptr;
  size_t dclsCap = 0;
  // the uploaded tile list: one shape (cur) or several classes (multi)
  DevParams cur{};
  bool multi = false;
  bool lastMulti = false;              // multi state of the last exact-solver step
  int lastPrec = FX_PREC_FP64;         // its precision (the face states' type)
  std::vector<unsigned char> htcls;
  std::vector<long> hcellStart, hfaceStart;
  unsigned char* dtcls = nullptr;
  size_t dtclsCap = 0;
  long* dcellStart = nullptr;
  size_t dcellStartCap = 0;
  long* dfaceStart = nullptr;
  size_t dfaceStartCap = 0;
  std::vector<int> hwCell, hwFace;     // find_tile_w tables of the cell and face prefixes
  int* dwCell = nullptr;
  size_t dwCellCap = 0;
  int* dwFace = nullptr;
  size_t dwFaceCap = 0;
  std::vector<int2> hgateChunks;       // multi-class: (tile, first face) per kThreads faces
  int2* dgateChunks = nullptr;
  size_t gateChunksCap = 0;
  int nGateChunks = 0;
  MarchItem* ditems = nullptr;         // column tiles of a multi-class march
  size_t ditemsCap = 0;
  int nItems = 0;
  int itemsTX = 0;                     // column-tile width the items were built for
  int itemsOrder = 0;                  // item order they were built in
  int persist = 0;                     // multi-class march: multiple of the resident workgroups (0 = 1)
  int tuneTX = 0;                      // autotuned column-tile width (0: built-in choice)
  int tuneOcc = 0;                     // autotuned waves-per-SIMD hint of the FP32 3-D march (0 or 4)
  int itemOrder = 0;                   // multi-class items: 0 longest first, 1 tile by tile
  unsigned long uploadGen = 0;         // counts tile-list uploads
  // multi-level grid (fx_amr_*)
  struct Amr {
    bool ready = false;
    int nBlocks = 0, nLevels = 0;
    int periodic[3] = {0, 0, 0};
    unsigned long long perMass = 0;
    std::vector<fx_tile> tiles;
    std::vector<AmrBlock> hblocks;
    AmrBlock* dblocks = nullptr;
    AmrLevelDev* dlevels = nullptr;
    std::vector<int*> dindex;                 // per level
    long* dboxStart = nullptr;                // guard-inclusive cells before each block
    int* dwBox = nullptr;                     // find_tile_w table of the prefix
    int2* dchunks = nullptr;                  // guard-fill work: (block, first guard cell)
    int2* dplan = nullptr;                    // guard-fill source of every work slot (k_amr_plan)
    const double** dsrc = nullptr;            // the copy slots' source addresses
    int* drest = nullptr;                     // plan slots that are not same-level copies
    unsigned* dnRest = nullptr;               // their count (on the device)
    unsigned nRest = 0;
    AmrCoarseRec* drecs = nullptr;            // distinct coarse cells under prolonged guard cells
    int nRec = 0;
    int* drecOf = nullptr;                    // per rest slot: its record, or -1
    double* dcv = nullptr;                    // per record and variable: q0, s_a * alpha
    int* drecOk = nullptr;
    int* dguard = nullptr;                    // guard-cell box indices, class by class
    int* dguardOff = nullptr;                 // per class: first entry in dguard (+ total)
    long nChunks = 0;
    std::vector<int*> dlist;                  // per fine level: its blocks
    std::vector<long*> drstart;               // per fine level: covered parent cells before each
    std::vector<int> nlist;
    std::vector<fx_fluxreg_face*> dfaces;     // per level pair (coarse level L, L+1)
    std::vector<int> nfaces;
    std::vector<double*> dreg;
    long boxCells = 0;
    double* flux = nullptr;                   // the grid's own flux and pavg arrays
    double* pavg = nullptr;
    unsigned long gen = 0;                    // upload generation of the grid's tile list
    int TX = 8, TY = 8;
  } amr;
  // batch tile-list cache (fx_hydro_advance_batch)
  double* batchU = nullptr;
  long batchN = -1, batchStride = -1;
  double batchDel[3] = {0, 0, 0};
  bool batchValid = false;
  // per-stage timing (fx_hydro_set_timing)
  int timing = 0;
  // true while a graph run captures its steps: the captured steps keep the
  // status of the whole run (first error, summed counters) instead of
  // resetting it per step; only the per-step shock-tile count is cleared
  bool capturing = false;
  int legacy = 0;  // 1: force the three-kernel path (fx_hydro_set_kernel_path)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float stageMs[3] = {0.f, 0.f, 0.f};
};
Lib g;

int hip_fail(hipError_t e, const char* what, fx_status* st) {
  if (e == hipSuccess) return 0;
  fprintf(stderr, "fx_hydro: %s failed: %s\n", what, hipGetErrorString(e));
  if (st) st->code = 4;
  return 4;
}
#define HIPCALL(x, st)                                    \
  do {                                                    \
    int rc__ = hip_fail((x), #x, (st));                   \
    if (rc__) return rc__;                                \
  } while (0)

template <typename T>
int grow(T** ptr, size_t* cap, size_t need, fx_status* st) {
  if (*cap >= need) return 0;
  if (*ptr) HIPCALL(hipFree(*ptr), st);
  *ptr = nullptr;
  *cap = 0;
  HIPCALL(hipMalloc((void**)ptr, need * sizeof(T)), st);
  *cap = need;
  return 0;
}

// Hydro_init.F90:117-124, 387-395: HYBR and shockLowerCFL turn detection on
bool shock_on(const fx_params& p) {
  return p.shockDetect || p.shockLowerCFL || p.riemann == FX_HYBR;
}

DevParams make_dev_params(const fx_params& p) {
  DevParams d{};
  d.ndim = p.ndim;
  d.nxb = p.nxb;
  d.nyb = p.ndim > 1 ? p.nyb : 1;
  d.nzb = p.ndim > 2 ? p.nzb : 1;
  d.ng = p.ng;
  d.nunk = p.nunk;
  d.nx = p.nxb + 2 * p.ng;
  d.ny = p.ndim > 1 ? p.nyb + 2 * p.ng : 1;
  d.nz = p.ndim > 2 ? p.nzb + 2 * p.ng : 1;
  d.ncell = (long)d.nx * d.ny * d.nz;
  for (int v = 0; v < FX_NVAR_USED; ++v) d.voff[v] = (long)p.var[v] * d.ncell;
  d.lo[0] = p.ng;
  d.hi[0] = p.ng + p.nxb - 1;
  d.lo[1] = p.ndim > 1 ? p.ng : 0;
  d.hi[1] = p.ndim > 1 ? p.ng + p.nyb - 1 : 0;
  d.lo[2] = p.ndim > 2 ? p.ng : 0;
  d.hi[2] = p.ndim > 2 ? p.ng + p.nzb - 1 : 0;
  d.gamma = p.gamma;
  d.limiter = p.limiter;
  d.beta = p.limitedSlopeBeta;
  d.charLim = p.charLimiting;
  d.flat = p.flattening;
  d.steep = p.steepening;
  d.ctu3d = p.use3dFullCTU;
  d.aux = p.useAuxEint;
  d.riemann = p.riemann;
  d.precision = p.precision;
  d.learned = p.learnedSeed;
  for (int a = 0; a < 3; ++a) d.velSlot[a] = p.var[FX_VELX + a];
  d.fP = p.fP;
  d.smlrho = p.smlrho;
  d.smallE = p.smallE;
  d.eintSwitch = p.eintSwitch;
  d.tiny = p.tiny;
  d.gammam1 = 1.0 / (p.gamma - 1.0);
  const double ggprod = d.gammam1 * p.gasConstant;
  d.ggprodinv = 1. / ggprod;
  d.gam1inv = 1. / d.gammam1;
  d.abar = p.singleSpeciesA;
  d.exactMaxIter = p.exactMaxIter;
  d.exactTol = p.exactTol;
  d.shokOff = shock_on(p) ? (long)p.shokVar * d.ncell : -1;
  const int n[3] = {d.nxb, d.nyb, d.nzb};
  long off = 0;
  for (int dir = 0; dir < 3; ++dir) {
    for (int a = 0; a < 3; ++a) {
      d.bext[dir][a] = n[a] + ((a == dir && dir < p.ndim) ? 2 : 0);
      d.fext[dir][a] = n[a] + ((a == dir && dir < p.ndim) ? 1 : 0);
    }
    if (dir < p.ndim) {
      d.bcells[dir] = (long)d.bext[dir][0] * d.bext[dir][1] * d.bext[dir][2];
      d.nface[dir] = (long)d.fext[dir][0] * d.fext[dir][1] * d.fext[dir][2];
    } else {
      d.bcells[dir] = 0;
      d.nface[dir] = 0;
    }
    d.boff[dir] = off;
    off += d.bcells[dir];
  }
  d.scellsPerTile = off;
  d.icells = (long)d.nxb * d.nyb * d.nzb;
  return d;
}

// Column-tile shape of the fused kernel for this block shape, or false
// when only the three-kernel path applies (1-D, or sizes not tileable).
bool march_shape(const DevParams& P, int* TX, int* TY, int* NT) {
  if (P.ndim < 2) return false;
  const int nx = P.nxb, ny = P.nyb;
  if (nx % 16 == 0 && ny % 8 == 0 && g.tuneTX != 8) {
    *TX = 16;
    *TY = 8;
  } else if (nx % 8 == 0 && ny % 8 == 0) {
    *TX = 8;
    *TY = 8;
  } else {
    return false;
  }
  *NT = (P.ndim == 3 && *TX == 16) ? 512 : 256;
  return true;
}

int grid_for(long items) {
  long blocks = (items + kThreads - 1) / kThreads;
  const long cap = (long)g.numCU * 64;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

int validate(const fx_params* p) {
  if (!p) return 3;
  if (p->ndim < 1 || p->ndim > 3 || p->ng != 4 || p->nxb < 2) return 3;
  if (p->ndim > 1 && p->nyb < 2) return 3;
  if (p->ndim > 2 && p->nzb < 2) return 3;
  for (int v = 0; v < FX_NVAR_USED; ++v)
    if (p->var[v] < 0 || p->var[v] >= p->nunk) return 3;
  if (!(p->gamma > 1.0)) return 3;
  if (p->riemann != FX_HLL && p->riemann != FX_HLLC && p->riemann != FX_HYBR &&
      p->riemann != FX_EXACT)
    return 3;
  if (p->precision != FX_PREC_FP64 && p->precision != FX_PREC_MIXED) return 3;
  if (p->learnedSeed && p->riemann != FX_EXACT) return 3;
  if (p->fallbackLowerCFL && !(p->cflFallbackFactor > 0.0)) return 3;
  if (shock_on(*p)) {  // SHOK_VAR: its own slot of U (hy_shockDetectBlk.F90:119-122)
    if (p->shokVar < 0 || p->shokVar >= p->nunk) return 3;
    for (int v = 0; v < FX_NVAR_USED; ++v)
      if (p->var[v] == p->shokVar) return 3;
    // the detector's planes fit one workgroup's LDS (x-y extents up to 36 x 36)
    if (shock_lds_doubles(make_dev_params(*p)) > kShockLdsMax) return 3;
  }
  return 0;
}

// hy_shockDetectBlk.F90:228-235: the CFL cap of a step with a flagged cell
double shock_cfl_max(int ndim) { return ndim == 1 ? 0.60 : (ndim == 2 ? 0.45 : 0.25); }

// hy_fallbackLowerCFL: the CFL cap of a step in which a state fell back to
// first order, hy_cflFallbackFactor / NDIM (hy_dataReconstOneStep.F90:
// 1125-1129, hy_getRiemannState.F90:1473); 0 when off.  In 1-D the
// reference never copies the lowered cell CFL into hy_cfl (the assignment
// at hy_getRiemannState.F90:934-941 sits under NDIM > 1).
double fallback_cfl_cap(const fx_params& p) {
  return p.fallbackLowerCFL && p.ndim > 1 ? p.cflFallbackFactor / double(p.ndim) : 0.0;
}

// a CFL lowered by the step's flags or fallbacks
bool cfl_varies(const fx_params& p) { return p.shockLowerCFL || fallback_cfl_cap(p) > 0.0; }

// Synchronise and copy the device status record ds into st (nothing in
// async mode, where the call returns before its kernels end).
int finish_from(const DevStatus* ds, fx_status* st) {
  if (g.async) return 0;
  HIPCALL(hipStreamSynchronize(g.stream), st);
  DevStatus h{};
  HIPCALL(hipMemcpy(&h, ds, sizeof h, hipMemcpyDeviceToHost), st);
  if (st) {
    st->code = h.code;
    st->first_tile = h.tile;
    st->first_i = h.i;
    st->first_j = h.j;
    st->first_k = h.k;
    st->first_dir = h.dir;
    st->fallbacks = (long long)h.fallbacks;
    st->newton_iters = (long long)h.newtonIters;
    st->newton_fallbacks = (long long)h.newtonFallbacks;
    // hy_shockDetectBlk.F90:211-227: the CFL cap of a step with a flagged cell
    if (g.prm.shockLowerCFL && h.shockTiles) st->min_cell_cfl = shock_cfl_max(g.prm.ndim);
    if (h.fallbacks) st->fallback_cfl = fallback_cfl_cap(g.prm);
  }
  return h.code;
}

int finish(fx_status* st) { return finish_from(g.status, st); }

// ---- the batch step without allocation ----
// Device address of the scratch array on each device, taken once per device
// (with that device current) by hipGetSymbolAddress.
constexpr int kMaxDevices = 64;
void* g_scratchAddr[kMaxDevices] = {};

int noalloc_scratch(void** out, fx_status* st) {
  int dev = 0;
  HIPCALL(hipGetDevice(&dev), st);
  if (dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "fx_hydro: device %d beyond the %d devices the no-allocation scratch serves\n",
            dev, kMaxDevices);
    if (st) st->code = 3;
    return 3;
  }
  if (!g_scratchAddr[dev]) {
    void* p = nullptr;
    HIPCALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), st);
    g_scratchAddr[dev] = p;
  }
  *out = g_scratchAddr[dev];
  return 0;
}

// Byte offsets of the no-allocation step's parts inside the scratch array.
struct NoAllocLayout {
  size_t tiles, flux, pavg, total;  // the status record sits at offset 0
};
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
NoAllocLayout noalloc_layout(const DevParams& P, long nBlocks) {
  NoAllocLayout L;
  long facesPerBlock = 0;
  for (int d = 0; d < P.ndim; ++d) facesPerBlock += P.nface[d];
  L.tiles = align256(sizeof(DevStatus));
  L.flux = align256(L.tiles + sizeof(fx_tile) * (size_t)nBlocks);
  L.pavg = align256(L.flux + sizeof(double) * FX_NFLUX * (size_t)facesPerBlock * (size_t)nBlocks);
  L.total = L.pavg + sizeof(double) * (size_t)P.icells * (size_t)nBlocks;
  return L;
}

// Inside a graph-captured run under shockLowerCFL or fallbackLowerCFL: the
// device slot holding the next dt's CFL, or null (dt takes the CFL argument).
const double* step_cfl_dev() { return cfl_varies(g.prm) && g.dtDev ? g.dtDev + 2 : nullptr; }

// Inside a graph-captured run under shockLowerCFL or fallbackLowerCFL, at
// the end of a step: the next dt's CFL from this step's flags and fallbacks
// (k_step_cfl).  Eager callers get the caps in fx_status.min_cell_cfl and
// fallback_cfl instead.
void launch_step_cfl() {
  if (!cfl_varies(g.prm) || !g.dtDev) return;   // g.dtDev == g.dtBuf while capturing
  hipLaunchKernelGGL(k_step_cfl, dim3(1), dim3(1), 0, g.stream, g.status, g.runCfl,
                     shock_on(g.prm) ? 1 : 0,
                     g.prm.shockLowerCFL ? shock_cfl_max(g.prm.ndim) : 0.0,
                     fallback_cfl_cap(g.prm), g.dtBuf + 2);
}

// k_shock_detect over the uploaded tiles, one workgroup each, with the LDS
// of the largest class present
void launch_shock_detect(int nTiles) {
  Launch L{g.dtiles, nTiles, nullptr, g.status, 0.0, nullptr};
  if (g.multi) {
    long lds = 0;
    for (const DevParams& P : g.hcls) lds = std::max(lds, shock_lds_doubles(P));
    hipLaunchKernelGGL(k_shock_detect, dim3(nTiles), dim3(kThreads), sizeof(double) * lds,
                       g.stream, L, g.cur, g.dcls, g.dtcls);
  } else {
    hipLaunchKernelGGL(k_shock_detect, dim3(nTiles), dim3(kThreads),
                       sizeof(double) * shock_lds_doubles(g.cur), g.stream, L, g.cur,
                       (const DevParams*)nullptr, (const unsigned char*)nullptr);
  }
}

// parameters of tile t of the uploaded list
const DevParams& tile_params(int t) { return g.multi ? g.hcls[g.htcls[t]] : g.cur; }

// Classifies the tiles (useClasses: by fx_tile.block_class once classes are
// set; otherwise every tile has the fx_hydro_init shape) and builds the
// per-tile cell and face prefix offsets.  Returns 3 for a class out of range.
int classify_tiles(const fx_tile* tiles, int nTiles, bool useClasses) {
  g.multi = false;
  g.cur = g.dp;
  g.htcls.assign((size_t)nTiles, 0);
  if (useClasses && !g.hcls.empty()) {
    for (int t = 0; t < nTiles; ++t) {
      const int c = tiles[t].block_class;
      if (c < 0 || c >= (int)g.hcls.size()) return 3;
      g.htcls[t] = (unsigned char)c;
      if (c != g.htcls[0]) g.multi = true;
    }
    if (nTiles > 0) g.cur = g.hcls[g.htcls[0]];
  }
  g.hcellStart.assign((size_t)nTiles + 1, 0);
  g.hfaceStart.assign((size_t)nTiles + 1, 0);
  for (int t = 0; t < nTiles; ++t) {
    const DevParams& P = tile_params(t);
    g.hcellStart[t + 1] = g.hcellStart[t] + P.icells;
    g.hfaceStart[t + 1] = g.hfaceStart[t] + P.nface[0] + P.nface[1] + P.nface[2];
  }
  return 0;
}

// tile of every 64th item of a prefix (find_tile_w); every tile must hold
// at least 64 items
std::vector<int> wave_table(const std::vector<long>& start) {
  const int n = (int)start.size() - 1;
  const long total = start[n];
  std::vector<int> w((size_t)((total + 63) / 64));
  int t = 0;
  for (size_t c = 0; c < w.size(); ++c) {
    const long item = (long)c * 64;
    while (t + 1 < n && start[t + 1] <= item) ++t;
    w[c] = t;
  }
  return w;
}

// Uploads the tile list, substituting library-owned flux arrays where a
// tile has none.
int upload_tiles(const fx_tile* tiles, int nTiles, bool needFlux, fx_status* st,
                 bool useClasses = true) {
  g.batchValid = false;
  ++g.uploadGen;
  if (classify_tiles(tiles, nTiles, useClasses)) {
    if (st) st->code = 3;
    return 3;
  }
  g.htiles.assign(tiles, tiles + nTiles);
  const int ndim = g.dp.ndim;
  size_t missing = 0;
  for (int t = 0; t < nTiles; ++t)
    for (int d = 0; d < ndim; ++d)
      if (!g.htiles[t].flux[d]) ++missing;
  if (needFlux && missing) {
    if (grow(&g.flux, &g.fluxCap, (size_t)g.hfaceStart[nTiles] * FX_NFLUX, st)) return 4;
    for (int t = 0; t < nTiles; ++t) {
      const DevParams& P = tile_params(t);
      double* base = g.flux + (size_t)g.hfaceStart[t] * FX_NFLUX;
      for (int d = 0; d < ndim; ++d) {
        if (!g.htiles[t].flux[d]) g.htiles[t].flux[d] = base;
        base += (size_t)P.nface[d] * FX_NFLUX;
      }
    }
  }
  if (needFlux) {
    size_t missingP = 0;
    for (int t = 0; t < nTiles; ++t)
      if (!g.htiles[t].pavg) ++missingP;
    if (missingP) {
      if (grow(&g.pavg, &g.pavgCap, (size_t)g.hcellStart[nTiles], st)) return 4;
      for (int t = 0; t < nTiles; ++t)
        if (!g.htiles[t].pavg) g.htiles[t].pavg = g.pavg + (size_t)g.hcellStart[t];
    }
  }
  if (grow(&g.dtiles, &g.dtilesCap, (size_t)nTiles, st)) return 4;
  HIPCALL(hipMemcpyAsync(g.dtiles, g.htiles.data(), sizeof(fx_tile) * nTiles, hipMemcpyHostToDevice,
                         g.stream),
          st);
  if (g.multi) {
    if (grow(&g.dtcls, &g.dtclsCap, (size_t)nTiles, st)) return 4;
    if (grow(&g.dcellStart, &g.dcellStartCap, (size_t)nTiles + 1, st)) return 4;
    if (grow(&g.dfaceStart, &g.dfaceStartCap, (size_t)nTiles + 1, st)) return 4;
    HIPCALL(hipMemcpyAsync(g.dtcls, g.htcls.data(), (size_t)nTiles, hipMemcpyHostToDevice, g.stream),
            st);
    HIPCALL(hipMemcpyAsync(g.dcellStart, g.hcellStart.data(), sizeof(long) * (nTiles + 1),
                           hipMemcpyHostToDevice, g.stream),
            st);
    HIPCALL(hipMemcpyAsync(g.dfaceStart, g.hfaceStart.data(), sizeof(long) * (nTiles + 1),
                           hipMemcpyHostToDevice, g.stream),
            st);
    g.hwCell = wave_table(g.hcellStart);
    g.hwFace = wave_table(g.hfaceStart);
    if (grow(&g.dwCell, &g.dwCellCap, g.hwCell.size(), st)) return 4;
    if (grow(&g.dwFace, &g.dwFaceCap, g.hwFace.size(), st)) return 4;
    HIPCALL(hipMemcpyAsync(g.dwCell, g.hwCell.data(), sizeof(int) * g.hwCell.size(),
                           hipMemcpyHostToDevice, g.stream),
            st);
    HIPCALL(hipMemcpyAsync(g.dwFace, g.hwFace.data(), sizeof(int) * g.hwFace.size(),
                           hipMemcpyHostToDevice, g.stream),
            st);
    g.nItems = 0;  // march items are built for the tile shape of the call
    // the learned seed's gate pass: each tile's faces in chunks of kThreads
    std::vector<int2>& ch = g.hgateChunks;
    ch.clear();
    for (int t = 0; t < nTiles; ++t)
      for (long f = 0; f < g.hfaceStart[t + 1] - g.hfaceStart[t]; f += kThreads)
        ch.push_back(make_int2(t, (int)f));
    g.nGateChunks = (int)ch.size();
    if (grow(&g.dgateChunks, &g.gateChunksCap, ch.size(), st)) return 4;
    HIPCALL(hipMemcpyAsync(g.dgateChunks, ch.data(), sizeof(int2) * ch.size(),
                           hipMemcpyHostToDevice, g.stream),
            st);
  }
  return 0;
}

// Column-tile width of a multi-class march: 16 across x where every class
// allows it, else 16 for the classes that allow it and 8 for the others
// (returned as kPerClassTX), unless the autotuner chose one width for all.
bool multi_march_shape(int* TX) {
  bool all16 = true, any16 = false;
  for (const DevParams& P : g.hcls) {
    if (P.ndim < 2 || P.nxb % 8 || P.nyb % 8) return false;
    if (P.nxb % 16) all16 = false;
    else any16 = true;
  }
  if (g.tuneTX == 8 || !any16) *TX = 8;
  else if (all16) *TX = 16;
  else *TX = kPerClassTX;
  return true;
}

// Column-tile work items of the uploaded multi-class list, longest columns
// first (stable in tile order within a class).
int build_march_items(int nTiles, int TX, fx_status* st) {
  if (g.nItems > 0 && g.itemsTX == TX && g.itemsOrder == g.itemOrder) return 0;
  const int TY = 8;
  std::vector<MarchItem> items;
  for (int t = 0; t < nTiles; ++t) {
    const DevParams& P = tile_params(t);
    const int W = TX == kPerClassTX ? (P.nxb % 16 ? 8 : 16) : TX;
    for (int ty = 0; ty < P.nyb / TY; ++ty)
      for (int tx = 0; tx < P.nxb / W; ++tx)
        items.push_back(MarchItem{t, (int)g.htcls[t], P.lo[0] + tx * W, P.lo[1] + ty * TY, W});
  }
  if (g.itemOrder == 0)
    std::stable_sort(items.begin(), items.end(), [](const MarchItem& a, const MarchItem& b) {
      return g.hcls[a.cls].nzb > g.hcls[b.cls].nzb;
    });
  if (grow(&g.ditems, &g.ditemsCap, items.size(), st)) return 4;
  HIPCALL(hipMemcpyAsync(g.ditems, items.data(), sizeof(MarchItem) * items.size(),
                         hipMemcpyHostToDevice, g.stream),
          st);
  g.nItems = (int)items.size();
  g.itemsTX = TX;
  g.itemsOrder = g.itemOrder;
  return 0;
}

MarchHost march_host() {
  return MarchHost{g.stream, g.tuneOcc, g.numCU, g.persist, g.dcls, g.ditems, g.nItems,
                   g.dfaceStart};
}

// FX_PREC_MIXED reconstructs and solves HLLC in FP32; the update kernel and
// the exact solver's Newton iteration stay FP64
template <int RS>
void launch_march(const DevParams& P, const MarchLaunch& M, int TX, dim3 grid) {
  if (P.precision == FX_PREC_MIXED) launch_march_r<RS, float>(P, M, TX, grid, march_host());
  else launch_march_r<RS, double>(P, M, TX, grid, march_host());
}

template <int RS>
void launch_march_multi(int ndim, int precision, int TX, const MarchLaunch& M) {
  if (precision == FX_PREC_MIXED) launch_march_multi_r<RS, float>(ndim, TX, M, march_host());
  else launch_march_multi_r<RS, double>(ndim, TX, M, march_host());
}
dim3 wave_grid(long n);

int record(int n) {
  if (!g.timing) return 0;
  if (!g.ev[n]) {
    if (hipEventCreate(&g.ev[n]) != hipSuccess) return 4;
  }
  return hipEventRecord(g.ev[n], g.stream) == hipSuccess ? 0 : 4;
}

// The exact solver's face stage over `faces` faces: with the learned seed,
// the network over the listed strong faces first.  R is the type of the
// stored face states and of the solve: float in the mixed-precision mode.
template <typename R>
void launch_face_solve(const DevParams& P, const FaceLaunch& F, long faces, bool multi,
                       bool withUpdate, double dt) {
  const dim3 gf(grid_for(faces));
  if (P.learned) {  // the gate lists the strong faces for the seed kernel
    float* y = withUpdate ? nullptr : g.seedY;
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, g.stream, g.strongCount);
    if (multi)
      hipLaunchKernelGGL((k_face_gate<true, R>), dim3((unsigned)g.nGateChunks), dim3(kThreads), 0,
                         g.stream, P, F, make_seed_gate(g.seedGate), y, g.strongList,
                         g.strongCount, (unsigned)faces, g.dgateChunks);
    else
      hipLaunchKernelGGL((k_face_gate<false, R>), gf, dim3(kThreads), 0, g.stream, P, F,
                         make_seed_gate(g.seedGate), y, g.strongList, g.strongCount,
                         (unsigned)faces, nullptr);
  }
  if (multi && P.learned)
    hipLaunchKernelGGL((k_face_seed<true, R>), wave_grid(faces), dim3(kLearnThreads), 0, g.stream,
                       P, F, g.seedPacked, g.seedY, g.strongList, g.strongCount, faces);
  else if (P.learned)
    hipLaunchKernelGGL((k_face_seed<false, R>), wave_grid(faces), dim3(kLearnThreads), 0,
                       g.stream, P, F, g.seedPacked, g.seedY, g.strongList, g.strongCount, faces);
  if (withUpdate) {  // the face stage and the update in one kernel (batch steps)
    if (P.learned) {   // the listed strong faces first, seeded; the rest unseeded
      R* over = reinterpret_cast<R*>(g.seedFlux);
      hipLaunchKernelGGL((k_seeded_solve<R>), dim3((unsigned)std::min<long>(grid_for(faces), 4L * g.numCU)),
                         dim3(kThreads), 0, g.stream, P, F, g.strongList, g.strongCount, faces,
                         faces, over, g.seedBits);
      hipLaunchKernelGGL((k_solve_update<true, R>), dim3((unsigned)F.nTiles), dim3(kSolveThreads),
                         0, g.stream, P, F, dt, g.dtDev, over, faces, g.seedBits);
    } else {
      hipLaunchKernelGGL((k_solve_update<false, R>), dim3((unsigned)F.nTiles), dim3(kSolveThreads),
                         0, g.stream, P, F, dt, g.dtDev, nullptr, 0L, nullptr);
    }
    return;
  }
  if (multi && P.learned)
    hipLaunchKernelGGL((k_face_solve<true, true, R>), gf, dim3(kThreads), 0, g.stream, P, F);
  else if (multi)
    hipLaunchKernelGGL((k_face_solve<false, true, R>), gf, dim3(kThreads), 0, g.stream, P, F);
  else if (P.learned)
    hipLaunchKernelGGL((k_face_solve<true, false, R>), gf, dim3(kThreads), 0, g.stream, P, F);
  else
    hipLaunchKernelGGL((k_face_solve<false, false, R>), gf, dim3(kThreads), 0, g.stream, P, F);
}

// Flux stages of a step on the uploaded tiles (march path): the column march,
// then the exact solver's face stage (records events 1 and 2).  Launches
// only: scratch and work items are in place.
int launch_flux_stages(int nTiles, double dt, int TX, int TY, bool solveUpdate = false) {
  const DevParams& P = g.cur;
  MarchLaunch M{g.dtiles, nTiles, g.multi ? 0 : P.nxb / TX,
                g.multi ? 0 : (P.ndim > 1 ? P.nyb : 1) / TY, g.status, dt, nullptr, g.dtDev};
  const dim3 grid((unsigned)(g.multi ? 1 : nTiles * M.tilesX * M.tilesY));
  if (P.riemann == FX_EXACT) {
    // exact solver: the march stores face states, a separate stage solves
    // (with the learned seed, after the gate and the seed kernel)
    M.fstates = g.fstates;
    if (g.multi) launch_march_multi<RS_STATES>(P.ndim, P.precision, TX, M);
    else launch_march<RS_STATES>(P, M, TX, grid);
    if (record(1)) return 4;
    const long faces = g.hfaceStart[nTiles];
    FaceLaunch F{g.dtiles, nTiles, g.fstates, P.learned ? g.seedY : nullptr, g.status,
                 g.multi ? g.dcls : nullptr, g.multi ? g.dtcls : nullptr,
                 g.multi ? g.dfaceStart : nullptr, g.multi ? g.dwFace : nullptr};
    g.lastFaces = faces;
    g.lastTiles = nTiles;
    g.lastMulti = g.multi;
    g.lastPrec = P.precision;
    if (P.precision == FX_PREC_MIXED) launch_face_solve<float>(P, F, faces, g.multi, solveUpdate, dt);
    else launch_face_solve<double>(P, F, faces, g.multi, solveUpdate, dt);
    if (record(2)) return 4;
  } else if (P.riemann == FX_HLLC) {
    if (g.multi) launch_march_multi<FX_HLLC>(P.ndim, P.precision, TX, M);
    else launch_march<FX_HLLC>(P, M, TX, grid);
    if (record(1)) return 4;
    if (record(2)) return 4;
  } else {  // HLL, or the hybrid switch
    if (g.multi) launch_march_multi<FX_HYBR>(P.ndim, P.precision, TX, M);
    else launch_march<FX_HYBR>(P, M, TX, grid);
    if (record(1)) return 4;
    if (record(2)) return 4;
  }
  return 0;
}

// update stage of a step on the uploaded tiles (dIn: per-tile in-place flags
// on the device, or null for all)
void launch_update(int nTiles, double dt, const unsigned char* dIn) {
  Launch L{g.dtiles, nTiles, nullptr, g.status, dt, g.dtDev};
  if (g.multi) {
    const ClassMap C{g.dcls, g.dtcls, g.dcellStart, g.dwCell};
    hipLaunchKernelGGL(k_update_multi, dim3(grid_for(g.hcellStart[nTiles])), dim3(kThreads), 0,
                       g.stream, L, C, dIn);
  } else if (g.cur.ndim == 3) {
    hipLaunchKernelGGL(k_update_cols,
                       dim3(grid_for((long)g.cur.nxb * g.cur.nyb * nTiles)), dim3(kThreads), 0,
                       g.stream, g.cur, L, dIn);
  } else {
    hipLaunchKernelGGL((k_update<false, false>), dim3(grid_for(g.cur.icells * nTiles)),
                       dim3(kThreads), 0, g.stream, g.cur, L, dIn);
  }
}

// scratch of the separate Riemann stage, grown before (never during) a step
int ensure_step_scratch(int nTiles, fx_status* st) {
  const DevParams& P = g.dp;
  if (P.riemann != FX_EXACT) return 0;
  const long faces = (long)nTiles <= (long)g.hfaceStart.size() - 1 ? g.hfaceStart[nTiles]
                                                                     : 0;
  if (grow(&g.fstates, &g.fstatesCap, (size_t)kFaceComps * faces, st)) return 4;
  if (P.learned && faces > 0xffffffffL) {  // the strong-face list holds 32-bit face numbers
    if (st) st->code = 3;
    return 3;
  }
  if (P.learned && grow(&g.seedY, &g.seedYCap, (size_t)faces, st)) return 4;
  if (P.learned && grow(&g.strongList, &g.strongListCap, (size_t)faces, st)) return 4;
  if (P.learned && !g.strongCount) HIPCALL(hipMalloc((void**)&g.strongCount, sizeof(unsigned)), st);
  return 0;
}

// the seeded strong faces' fluxes and bits of a learned batch step; the
// bits start at zero and every step clears those it sets
__global__ void k_zero_words(unsigned* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int ensure_seed_patch(long faces, int nTiles, fx_status* st) {
  if (grow(&g.seedFlux, &g.seedFluxCap, (size_t)8 * faces, st)) return 4;
  const size_t words = (size_t)seed_bit_words(g.cur) * (size_t)nTiles;
  if (g.seedBitsCap < words) {
    if (grow(&g.seedBits, &g.seedBitsCap, words, st)) return 4;
    hipLaunchKernelGGL(k_zero_words, dim3(grid_for((long)words)), dim3(kThreads), 0, g.stream,
                       g.seedBits, (long)words);
  }
  return 0;
}

int run_compute(const fx_tile* tiles, int nTiles, double dt, const unsigned char* inPlace,
                fx_status* st, bool tilesUploaded = false) {
  if (!g.init) {
    if (st) st->code = 3;
    return 3;
  }
  if (st) memset(st, 0, sizeof *st);
  if (nTiles <= 0) return 0;
  if (!tilesUploaded) {
    const int rc = upload_tiles(tiles, nTiles, true, st);
    if (rc) return rc;
  }
  const DevParams& P = g.cur;
  const unsigned char* dIn = nullptr;
  bool anyFluxOnly = false;
  if (inPlace) {
    for (int t = 0; t < nTiles; ++t)
      if (!inPlace[t]) {
        anyFluxOnly = true;
        if (P.aux && !tiles[t].pavg) {
          if (st) st->code = 3;
          return 3;
        }
      }
    if (grow(&g.dinplace, &g.dinplaceCap, (size_t)nTiles, st)) return 4;
    HIPCALL(hipMemcpyAsync(g.dinplace, inPlace, nTiles, hipMemcpyHostToDevice, g.stream), st);
    dIn = g.dinplace;
  }
  int TX = 0, TY = 8, NT = 0;
  const bool march = g.multi ? multi_march_shape(&TX) : (!g.legacy && march_shape(P, &TX, &TY, &NT));
  if (P.learned && (!march || !g.seedReady)) {  // the seed needs the march path and a network
    if (st) st->code = 3;
    return 3;
  }
  if (g.multi && !march) {  // classes admit only the march path
    if (st) st->code = 3;
    return 3;
  }
  if (march) {
    // batch steps with the exact solver: its face stage and the update run
    // as one kernel (k_solve_update) for 3-D blocks up to 16 x 16 across
    const bool solveUpdate = tilesUploaded && !tiles && !inPlace && !g.multi && !g.solveUpdOff &&
                             P.riemann == FX_EXACT && P.ndim == 3 &&
                             P.nxb * P.nyb <= kSolveMaxCols &&
                             2 * (P.nxb + 1) * P.nyb + P.nxb * P.nyb <= kSolveMaxFaces &&
                             P.nxb + 1 <= 17 && P.nyb + 1 <= 17 &&
                             (!P.learned || seed_bit_words(P) <= kSolveMaxBitWords);
    if (g.multi && build_march_items(nTiles, TX, st)) return 4;
    if (ensure_step_scratch(nTiles, st)) return 4;
    if (solveUpdate && P.learned && ensure_seed_patch(g.hfaceStart[nTiles], nTiles, st)) return 4;
    if (!g.capturing) hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
    if (record(0)) return 4;
    if (launch_flux_stages(nTiles, dt, TX, TY, solveUpdate)) return 4;
    if (!solveUpdate) launch_update(nTiles, dt, dIn);
    if (record(3)) return 4;
  } else {
  if (grow(&g.states, &g.statesCap, (size_t)P.scellsPerTile * kStateComps * nTiles, st)) return 4;
  Launch L{g.dtiles, nTiles, g.states, g.status, dt, g.dtDev};
  if (!g.capturing) hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  if (record(0)) return 4;
  {
    const dim3 gs(grid_for(P.scellsPerTile * nTiles));
    if (P.ndim == 3) hipLaunchKernelGGL(k_states<3>, gs, dim3(kThreads), 0, g.stream, P, L);
    else if (P.ndim == 2) hipLaunchKernelGGL(k_states<2>, gs, dim3(kThreads), 0, g.stream, P, L);
    else hipLaunchKernelGGL(k_states<1>, gs, dim3(kThreads), 0, g.stream, P, L);
  }
  if (record(1)) return 4;
  {
    const dim3 gf(grid_for((P.nface[0] + P.nface[1] + P.nface[2]) * nTiles));
    if (P.riemann == FX_EXACT) hipLaunchKernelGGL(k_fluxes<FX_EXACT>, gf, dim3(kThreads), 0, g.stream, P, L);
    else if (P.riemann == FX_HLLC) hipLaunchKernelGGL(k_fluxes<FX_HLLC>, gf, dim3(kThreads), 0, g.stream, P, L);
    else hipLaunchKernelGGL(k_fluxes<FX_HYBR>, gf, dim3(kThreads), 0, g.stream, P, L);
  }
  if (record(2)) return 4;
  (void)anyFluxOnly;
  hipLaunchKernelGGL((k_update<true, false>), dim3(grid_for(P.icells * nTiles)), dim3(kThreads), 0,
                     g.stream, P, L, dIn);
  if (record(3)) return 4;
  }
  HIPCALL(hipGetLastError(), st);
  int rc = finish(st);
  if (g.timing && !g.async) {
    for (int k = 0; k < 3; ++k)
      if (hipEventElapsedTime(&g.stageMs[k], g.ev[k], g.ev[k + 1]) != hipSuccess) g.stageMs[k] = -1.f;
  }
  return rc;
}

// Check every register face against the tile list and flux extents before
// anything is launched; the kernels index without further checks.
int check_reg_faces(const fx_tile* tiles, int nTiles, const fx_fluxreg_face* faces, int nFaces,
                    bool fine) {
  for (int t = 0; t < nTiles; ++t)
    if (!g.hcls.empty() && (tiles[t].block_class < 0 || tiles[t].block_class >= (int)g.hcls.size()))
      return 3;
  for (int e = 0; e < nFaces; ++e) {
    const fx_fluxreg_face& f = faces[e];
    const int t = fine ? f.fine_tile : f.coarse_tile;
    if (t < 0 || t >= nTiles) return 3;
    const DevParams& P = g.hcls.empty() ? g.dp : g.hcls[tiles[t].block_class];
    if (f.dir < 0 || f.dir >= P.ndim) return 3;
    if (!tiles[t].flux[f.dir]) return 3;
    const int* ext = P.fext[f.dir];
    int idx[3] = {fine ? f.fi : f.ci, fine ? f.fj : f.cj, fine ? f.fk : f.ck};
    int span[3] = {1, 1, 1};
    if (fine)
      for (int a = 0; a < P.ndim; ++a)
        if (a != f.dir) span[a] = 2;
    for (int a = 0; a < 3; ++a)
      if (idx[a] < 0 || idx[a] + span[a] > ext[a]) return 3;
  }
  return 0;
}

int run_fluxreg(const fx_tile* tiles, int nTiles, const fx_fluxreg_face* faces, int nFaces,
                double* reg, bool put, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.init) {
    if (st) st->code = 3;
    return 3;
  }
  if (nFaces <= 0) return 0;
  if (!tiles || !faces || !reg || nTiles <= 0 || check_reg_faces(tiles, nTiles, faces, nFaces, put)) {
    if (st) st->code = 3;
    return 3;
  }
  if (grow(&g.dregTiles, &g.dregTilesCap, (size_t)nTiles, st)) return 4;
  if (grow(&g.dfaces, &g.dfacesCap, (size_t)nFaces, st)) return 4;
  HIPCALL(hipMemcpyAsync(g.dregTiles, tiles, sizeof(fx_tile) * nTiles, hipMemcpyHostToDevice, g.stream),
          st);
  HIPCALL(hipMemcpyAsync(g.dfaces, faces, sizeof(fx_fluxreg_face) * nFaces, hipMemcpyHostToDevice,
                         g.stream),
          st);
  const DevParams& P = g.dp;
  RegLaunch R{g.dregTiles, g.dfaces, nFaces, reg, nullptr, nullptr};
  if (!g.hcls.empty()) {
    g.hregCls.resize((size_t)nTiles);
    for (int t = 0; t < nTiles; ++t) g.hregCls[t] = (unsigned char)tiles[t].block_class;
    if (grow(&g.dregCls, &g.dregClsCap, (size_t)nTiles, st)) return 4;
    HIPCALL(hipMemcpyAsync(g.dregCls, g.hregCls.data(), (size_t)nTiles, hipMemcpyHostToDevice,
                           g.stream),
            st);
    R.cls = g.dcls;
    R.tcls = g.dregCls;
  }
  const dim3 grid(grid_for((long)nFaces * FX_NFLUX));
  if (put) {
    const double coef = P.ndim == 3 ? 0.25 : (P.ndim == 2 ? 0.5 : 1.0);
    hipLaunchKernelGGL(k_fluxreg_put, grid, dim3(kThreads), 0, g.stream, P, R, coef);
  } else {
    hipLaunchKernelGGL(k_fluxreg_conserve, grid, dim3(kThreads), 0, g.stream, P, R);
  }
  HIPCALL(hipGetLastError(), st);
  if (g.async) return 0;
  HIPCALL(hipStreamSynchronize(g.stream), st);
  return 0;
}

// ---- learned seed: host helpers ----
int ensure_status(fx_status* st) {
  if (g.status) return 0;
  HIPCALL(hipMalloc((void**)&g.status, sizeof(DevStatus)), st);
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    g.numCU = prop.multiProcessorCount;
  return 0;
}

int pack_mlp(const float* params, MlpPacked** dst, fx_status* st) {
  if (!*dst) HIPCALL(hipMalloc((void**)dst, sizeof(MlpPacked)), st);
  hipLaunchKernelGGL(k_mlp_pack, dim3(2), dim3(256), 0, g.stream, params, *dst);
  HIPCALL(hipGetLastError(), st);
  return 0;
}

dim3 wave_grid(long n) {
  const long chunks = (n + 63) / 64;
  long blocks = (chunks + kLearnWaves - 1) / kLearnWaves;
  const long cap = (long)g.numCU * 8;
  if (blocks > cap) blocks = cap;
  return dim3((unsigned)(blocks < 1 ? 1 : blocks));
}

int sync_or_async(fx_status* st) {
  HIPCALL(hipGetLastError(), st);
  if (g.async) return 0;
  HIPCALL(hipStreamSynchronize(g.stream), st);
  return 0;
}

// tile list of a strided batch on the device (kept while the batch is unchanged)
int prepare_batch(double* U, long nBlocks, long blockStride, const double del[3], fx_status* st) {
  if (!g.init || !U || nBlocks < 0 || blockStride < g.dp.ncell * g.dp.nunk) return 3;
  const bool same = g.batchValid && g.batchU == U && g.batchN == nBlocks &&
                    g.batchStride == blockStride && g.batchDel[0] == del[0] &&
                    g.batchDel[1] == del[1] && g.batchDel[2] == del[2];
  if (!same) {
    std::vector<fx_tile> tl((size_t)nBlocks);
    for (long b = 0; b < nBlocks; ++b) {
      fx_tile& t = tl[(size_t)b];
      memset(&t, 0, sizeof t);
      t.U = U + b * blockStride;
      for (int d = 0; d < 3; ++d) t.del[d] = del[d];
      t.level = 1;
    }
    if (upload_tiles(tl.data(), (int)nBlocks, true, st, false)) return 4;
    g.batchU = U;
    g.batchN = nBlocks;
    g.batchStride = blockStride;
    for (int d = 0; d < 3; ++d) g.batchDel[d] = del[d];
    g.batchValid = true;
  }
  return 0;
}

// checked neighbour table, kept on the device until it changes
int upload_neighbors(double* U, long nBlocks, long blockStride, const int* nbr, fx_status* st) {
  const DevParams& P = g.dp;
  bool ok = g.init && U && nbr && nBlocks > 0 && blockStride >= P.ncell * P.nunk;
  for (int a = 0; ok && a < P.ndim; ++a)
    if (P.hi[a] - P.lo[a] + 1 < P.ng) ok = false;   // neighbours supply ng interior layers
  for (long e = 0; ok && e < nBlocks * 6; ++e) {
    const int axis = (int)((e % 6) / 2);
    if (axis >= P.ndim) continue;
    if (!(nbr[e] >= 0 ? nbr[e] < nBlocks : (nbr[e] == FX_BC_OUTFLOW || nbr[e] == FX_BC_REFLECT)))
      ok = false;
  }
  if (!ok) {
    if (st) st->code = 3;
    return 3;
  }
  const size_t n6 = (size_t)nBlocks * 6;
  if (g.hnbr.size() != n6 || memcmp(g.hnbr.data(), nbr, n6 * sizeof(int)) != 0) {
    if (grow(&g.dnbr, &g.dnbrCap, n6, st)) return 4;
    g.hnbr.assign(nbr, nbr + n6);
    HIPCALL(hipMemcpy(g.dnbr, g.hnbr.data(), n6 * sizeof(int), hipMemcpyHostToDevice), st);
  }
  return 0;
}

// onlyVar >= 0: that variable slot alone (the shock flags after detection)
void launch_fill(double* U, long nBlocks, long blockStride, int onlyVar = -1) {
  const DevParams& P = g.dp;
  for (int axis = 0; axis < P.ndim; ++axis) {
    const long slab = 2L * P.ng * (P.nx + P.ny + P.nz) * (P.nx + P.ny + P.nz);
    hipLaunchKernelGGL(k_fill_guards, dim3(grid_for(slab * nBlocks)), dim3(kThreads), 0, g.stream,
                       P, U, nBlocks, blockStride, g.dnbr, axis, onlyVar);
  }
}

}  // namespace

// ---------------------------------------------------------------------
// autotuner: launch shapes of the flux stages timed on the caller's batch
// ---------------------------------------------------------------------
int tune_flux_stages(int nTiles, double dt, int reps, fx_tune_result* out, int cap, int* n,
                     int* chosenTX, fx_status* st) {
  struct Cand {
    int tx, persist, order, occ;
  };
  const DevParams& P = g.cur;
  bool can16 = true, any16 = false;
  if (g.multi) {
    for (const DevParams& C : g.hcls) {
      if (C.nxb % 16 || C.nyb % 8) can16 = false;
      else any16 = true;
    }
  } else {
    can16 = P.nxb % 16 == 0 && P.nyb % 8 == 0;
  }
  std::vector<Cand> cands;
  // waves-per-SIMD hint 4 for the FP32 3-D march: at most 128 VGPRs (a few
  // spills), so that two 512-thread workgroups share a CU
  const bool occ4 = P.precision == FX_PREC_MIXED && P.ndim == 3;
  for (int occ : {0, 4}) {
    if (occ == 4 && !occ4) continue;
    for (int tx : {16, kPerClassTX, 8}) {
      if (tx == 16 && !can16) continue;
      if (tx == kPerClassTX && (!g.multi || can16 || !any16)) continue;
      if (g.multi) {
        for (int persist : {1, 2, 4})
          for (int order : {0, 1}) cands.push_back(Cand{tx, persist, order, occ});
      } else {
        cands.push_back(Cand{tx, 0, 0, occ});
      }
    }
  }
  if (ensure_step_scratch(nTiles, st)) return 4;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCALL(hipEventCreate(&e0), st);
  HIPCALL(hipEventCreate(&e1), st);
  const int savedTiming = g.timing;
  g.timing = 0;
  int best = -1, rc = 0;
  float bestMs = 0.f;
  std::vector<float> ms(cands.size(), -1.f);
  for (size_t c = 0; c < cands.size() && !rc; ++c) {
    g.tuneTX = cands[c].tx;
    g.persist = cands[c].persist;
    g.itemOrder = cands[c].order;
    g.tuneOcc = cands[c].occ;
    int TX = 8, TY = 8, NT = 0;
    const bool ok = g.multi ? multi_march_shape(&TX) : march_shape(P, &TX, &TY, &NT);
    if (!ok) continue;
    if (g.multi && build_march_items(nTiles, TX, st)) {
      rc = 4;
      break;
    }
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
    if (launch_flux_stages(nTiles, dt, TX, TY)) rc = 4;  // warm-up
    for (int r = 0; r < reps && !rc; ++r) {
      float t = 0.f;
      if (hipEventRecord(e0, g.stream) != hipSuccess || launch_flux_stages(nTiles, dt, TX, TY) ||
          hipEventRecord(e1, g.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
        rc = 4;
        break;
      }
      if (ms[c] < 0.f || t < ms[c]) ms[c] = t;
    }
    if (!rc && ms[c] >= 0.f && (best < 0 || ms[c] < bestMs)) {
      best = (int)c;
      bestMs = ms[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g.timing = savedTiming;
  if (rc || best < 0) {
    g.tuneTX = g.persist = g.itemOrder = g.tuneOcc = 0;
    if (st) st->code = rc ? rc : 3;
    return rc ? rc : 3;
  }
  g.tuneTX = cands[best].tx;
  g.persist = cands[best].persist;
  g.itemOrder = cands[best].order;
  g.tuneOcc = cands[best].occ;
  int TX = 8;
  if (g.multi) {
    multi_march_shape(&TX);
    if (build_march_items(nTiles, TX, st)) return 4;
  } else {
    TX = cands[best].tx;
  }
  if (chosenTX) *chosenTX = TX;
  int k = 0;
  for (size_t c = 0; c < cands.size(); ++c) {
    if (ms[c] < 0.f) continue;
    if (out && k < cap)
      out[k] = fx_tune_result{cands[c].tx, cands[c].persist, cands[c].order, cands[c].occ,
                              (int)c == best, ms[c]};
    ++k;
  }
  if (n) *n = k;
  return finish(st);
}

// ---------------------------------------------------------------------
// multi-level grids (fx_amr_*)
// ---------------------------------------------------------------------
void amr_free() {
  Lib::Amr& A = g.amr;
  void* bufs[] = {A.dblocks, A.dlevels, A.dboxStart, A.dwBox,
                  A.flux,    A.pavg,    A.dchunks,   A.dplan,      (void*)A.dsrc, A.drest,
                  A.dnRest,  A.drecs,   A.drecOf,    A.dcv,        A.drecOk, A.dguard,
                  A.dguardOff};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int* q : A.dindex) if (q) (void)hipFree(q);
  for (int* q : A.dlist) if (q) (void)hipFree(q);
  for (long* q : A.drstart) if (q) (void)hipFree(q);
  for (fx_fluxreg_face* q : A.dfaces) if (q) (void)hipFree(q);
  for (double* q : A.dreg) if (q) (void)hipFree(q);
  A = Lib::Amr{};
}

template <typename T>
int upload_vec(T** dst, const std::vector<T>& v, fx_status* st) {
  *dst = nullptr;
  if (v.empty()) return 0;
  HIPCALL(hipMalloc((void**)dst, sizeof(T) * v.size()), st);
  HIPCALL(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), st);
  return 0;
}

AmrLaunch amr_launch() {
  const Lib::Amr& A = g.amr;
  AmrLaunch L{};
  L.blocks = A.dblocks;
  L.nBlocks = A.nBlocks;
  L.levels = A.dlevels;
  L.nLevels = A.nLevels;
  L.cls = g.dcls;
  L.list = nullptr;
  L.nList = A.nBlocks;
  L.start = A.dboxStart;
  L.wtab = A.dwBox;
  for (int a = 0; a < 3; ++a) L.periodic[a] = A.periodic[a];
  L.perMass = A.perMass;
  L.status = g.status;
  L.onlyVar = -1;
  L.guard = A.dguard;
  L.guardOff = A.dguardOff;
  return L;
}

// the grid's tile list on the device (another call may have replaced it)
int amr_ensure_uploaded(fx_status* st) {
  Lib::Amr& A = g.amr;
  if (A.gen == g.uploadGen) return 0;
  const int rc = upload_tiles(A.tiles.data(), A.nBlocks, true, st);
  if (rc) return rc;
  A.gen = g.uploadGen;
  if (g.multi && build_march_items(A.nBlocks, A.TX, st)) return 4;
  return ensure_step_scratch(A.nBlocks, st);
}

// onlyVar >= 0: that variable slot alone, no EOS (the shock flags after
// detection; the other variables are unchanged since the full fill)
void amr_launch_fill(int onlyVar = -1) {
  AmrLaunch L = amr_launch();
  L.onlyVar = onlyVar;
  hipLaunchKernelGGL(k_amr_fill, dim3((unsigned)g.amr.nChunks), dim3(kThreads), 0, g.stream, L,
                     g.amr.dchunks, g.amr.dsrc);
  const Lib::Amr& A = g.amr;
  if (A.nRec)
    hipLaunchKernelGGL(k_amr_coarse, dim3(grid_for(A.nRec)), dim3(kThreads), 0, g.stream, L,
                       A.drecs, A.nRec, A.dcv, A.drecOk);
  if (A.nRest)
    hipLaunchKernelGGL(k_amr_prolong, dim3(grid_for(A.nRest)), dim3(kThreads), 0, g.stream, L,
                       A.dchunks, A.drest, A.drecOf, (int)A.nRest, A.dcv, A.drecOk, A.nRec);
}

// fine-to-coarse flux correction, finest level pair first
void amr_launch_register() {
  const Lib::Amr& A = g.amr;
  const DevParams& P = g.dp;
  const double coef = P.ndim == 3 ? 0.25 : (P.ndim == 2 ? 0.5 : 1.0);
  for (int L = A.nLevels - 2; L >= 0; --L) {
    const int n = A.nfaces[L];
    if (n <= 0) continue;
    RegLaunch R{g.dtiles, A.dfaces[L], n, A.dreg[L], g.dcls, g.multi ? g.dtcls : nullptr};
    if (!g.multi) R.cls = nullptr;
    const dim3 grid(grid_for((long)n * FX_NFLUX));
    hipLaunchKernelGGL(k_fluxreg_put, grid, dim3(kThreads), 0, g.stream, g.cur, R, coef);
    hipLaunchKernelGGL(k_fluxreg_conserve, grid, dim3(kThreads), 0, g.stream, g.cur, R);
  }
}

// restriction, finest level first
void amr_launch_restrict() {
  const Lib::Amr& A = g.amr;
  for (int L = A.nLevels - 1; L >= 1; --L) {
    if (A.nlist[L] <= 0) continue;
    AmrLaunch R = amr_launch();
    R.list = A.dlist[L];
    R.nList = A.nlist[L];
    R.start = A.drstart[L];
    R.wtab = nullptr;
    long cells = 0;
    for (int b = 0; b < A.nBlocks; ++b)
      if (A.hblocks[b].level == L) {
        const DevParams& P = g.hcls[A.hblocks[b].cls];
        cells += (long)(P.nxb / 2) * (P.ndim > 1 ? P.nyb / 2 : 1) * (P.ndim > 2 ? P.nzb / 2 : 1);
      }
    hipLaunchKernelGGL(k_amr_restrict, dim3(grid_for(cells)), dim3(kThreads), 0, g.stream, R);
  }
}

// dt over the leaf cells of every level (the minimum of the leaf blocks'
// dts, as k_amr_dt picks them), into g.dtBuf[0]; t += dt
int amr_launch_dt(double cfl, fx_status* st) {
  const int grid = g.amr.nBlocks;
  if (g.partialCap < grid) {
    if (g.partial) HIPCALL(hipFree(g.partial), st);
    g.partial = nullptr;
    HIPCALL(hipMalloc((void**)&g.partial, sizeof(double) * grid), st);
    g.partialCap = grid;
  }
  AmrLaunch L = amr_launch();
  hipLaunchKernelGGL(k_amr_dt, dim3(grid), dim3(kThreads), 0, g.stream, L, cfl, step_cfl_dev(),
                     g.partial);
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kThreads), 0, g.stream, g.partial, grid, g.dtBuf,
                     g.dtBuf + 1);
  return 0;
}

// one step of every level: guard fill, fluxes, flux correction, update,
// restriction (launches only; dt from the argument or, when dtDev is set,
// from the device)
int amr_launch_step(double dt) {
  if (!g.capturing) hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  amr_launch_fill();
  if (g.dp.shokOff >= 0) {  // Hydro.F90:94-123: flag, then fill SHOK_VAR's guard cells
    if (g.capturing)
      hipLaunchKernelGGL(k_zero_shock_tiles, dim3(1), dim3(1), 0, g.stream, g.status);
    launch_shock_detect(g.amr.nBlocks);
    amr_launch_fill(g.prm.shokVar);
  }
  if (launch_flux_stages(g.amr.nBlocks, dt, g.amr.TX, g.amr.TY)) return 4;
  amr_launch_register();
  launch_update(g.amr.nBlocks, dt, nullptr);
  amr_launch_restrict();
  launch_step_cfl();
  return 0;
}

extern "C" {

int fx_hydro_init(const fx_params* p) {
  int rc = validate(p);
  if (rc) return rc;
  if (!g.init || memcmp(&g.prm, p, sizeof *p) != 0) {
    g.batchValid = false;
    g.hcls.clear();  // classes carry the knobs of the init they were set under
    g.amr.ready = false;
  }
  g.prm = *p;
  g.dp = make_dev_params(*p);
  if (!g.status) {
    if (hipMalloc((void**)&g.status, sizeof(DevStatus)) != hipSuccess) return 4;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      g.numCU = prop.multiProcessorCount;
  }
  void* scratchAddr = nullptr;  // the no-allocation step's scratch on this device
  rc = noalloc_scratch(&scratchAddr, nullptr);
  if (rc) return rc;
  g.init = true;
  return 0;
}

void fx_hydro_finalize(void) {
  void* bufs[] = {g.strongList, g.strongCount, g.dwCell, g.dwFace, g.dgateChunks, g.states, g.flux, g.pavg, g.dtiles, g.dinplace, g.status, g.partial,
                  g.dregTiles, g.dfaces, g.mlpPacked, g.seedPacked, g.seedY, g.fstates,
                  g.dnbr, g.dtBuf, g.dcls, g.dtcls, g.dcellStart, g.dfaceStart,
                  g.ditems, g.dregCls, g.seedFlux, g.seedBits};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : g.ev)
    if (e) (void)hipEventDestroy(e);
  if (g.capStream) (void)hipStreamDestroy(g.capStream);
  amr_free();
  g = Lib{};
}

int fx_hydro_compute_fluxes(const fx_tile* tiles, int nTiles, double dt, double dtOld,
                            const unsigned char* updateInPlace, fx_status* st) {
  (void)dtOld;
  return run_compute(tiles, nTiles, dt, updateInPlace, st);
}

int fx_hydro_update_solution(const fx_tile* tiles, int nTiles, double dt, double dtOld,
                             fx_status* st) {
  (void)dtOld;
  if (!g.init) return 3;
  if (st) memset(st, 0, sizeof *st);
  if (nTiles <= 0) return 0;
  for (int t = 0; t < nTiles; ++t) {
    for (int d = 0; d < g.dp.ndim; ++d)
      if (!tiles[t].flux[d]) return 3;
    if (g.dp.aux && !tiles[t].pavg) return 3;
  }
  {
    const int rc = upload_tiles(tiles, nTiles, false, st);
    if (rc) return rc;
  }
  const DevParams& P = g.cur;
  Launch L{g.dtiles, nTiles, nullptr, g.status, dt, g.dtDev};
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  if (g.multi) {
    const ClassMap C{g.dcls, g.dtcls, g.dcellStart, g.dwCell};
    hipLaunchKernelGGL(k_update_multi, dim3(grid_for(g.hcellStart[nTiles])), dim3(kThreads), 0,
                       g.stream, L, C, (const unsigned char*)nullptr);
  } else {
    hipLaunchKernelGGL((k_update<false, false>), dim3(grid_for(P.icells * nTiles)), dim3(kThreads),
                       0, g.stream, P, L, (const unsigned char*)nullptr);
  }
  HIPCALL(hipGetLastError(), st);
  return finish(st);
}

int fx_hydro_advance_batch(double* U, long nBlocks, long blockStride, const double del[3], double dt,
                           fx_status* st) {
  const int rc = prepare_batch(U, nBlocks, blockStride, del, st);
  if (rc) return rc;
  return run_compute(nullptr, (int)nBlocks, dt, nullptr, st, true);
}

int fx_hydro_advance_batch_noalloc(double* U, long nBlocks, long blockStride,
                                   const double del[3], double dt, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.init || !U || !del || nBlocks < 0 || nBlocks > 0x7fffffffL ||
      blockStride < g.dp.ncell * g.dp.nunk) {
    if (st) st->code = 3;
    return 3;
  }
  if (nBlocks == 0) return 0;
  const DevParams& P = g.dp;
  int TX = 0, TY = 0, NT = 0;
  if (P.riemann == FX_EXACT || !march_shape(P, &TX, &TY, &NT)) {
    fprintf(stderr, "fx_hydro_advance_batch_noalloc: takes the HLLC, HLL and hybrid solvers on "
                    "2-D/3-D blocks with x and y extents multiples of 8\n");
    if (st) st->code = 3;
    return 3;
  }
  const NoAllocLayout Lay = noalloc_layout(P, nBlocks);
  if (Lay.total > kNoAllocBytes) {
    fprintf(stderr, "fx_hydro_advance_batch_noalloc: %ld blocks need %zu bytes of scratch, "
                    "more than its fixed %zu\n", nBlocks, Lay.total, kNoAllocBytes);
    if (st) st->code = 3;
    return 3;
  }
  void* base = nullptr;
  const int rc = noalloc_scratch(&base, st);
  if (rc) return rc;
  char* s = static_cast<char*>(base);
  DevStatus* status = reinterpret_cast<DevStatus*>(s);
  fx_tile* tiles = reinterpret_cast<fx_tile*>(s + Lay.tiles);
  hipLaunchKernelGGL(k_batch_tiles, dim3(grid_for(nBlocks)), dim3(kThreads), 0, g.stream, tiles,
                     U, nBlocks, blockStride, del[0], del[1], del[2],
                     reinterpret_cast<double*>(s + Lay.flux), P.nface[0], P.nface[1], P.nface[2],
                     P.ndim, reinterpret_cast<double*>(s + Lay.pavg), P.icells);
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, status);
  MarchLaunch M{tiles, (int)nBlocks, P.nxb / TX, (P.ndim > 1 ? P.nyb : 1) / TY, status, dt,
                nullptr, nullptr};
  const dim3 grid((unsigned)(nBlocks * M.tilesX * M.tilesY));
  if (P.riemann == FX_HLLC) launch_march<FX_HLLC>(P, M, TX, grid);
  else launch_march<FX_HYBR>(P, M, TX, grid);
  Launch L{tiles, (int)nBlocks, nullptr, status, dt, nullptr};
  if (P.ndim == 3)
    hipLaunchKernelGGL(k_update_cols, dim3(grid_for((long)P.nxb * P.nyb * nBlocks)),
                       dim3(kThreads), 0, g.stream, P, L, nullptr);
  else
    hipLaunchKernelGGL((k_update<false, false>), dim3(grid_for(P.icells * nBlocks)),
                       dim3(kThreads), 0, g.stream, P, L, nullptr);
  HIPCALL(hipGetLastError(), st);
  return finish_from(status, st);
}

int fx_hydro_shock_detect_batch(double* U, long nBlocks, long blockStride, const double del[3],
                                fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.init || g.dp.shokOff < 0 || !del) {
    if (st) st->code = 3;
    return 3;
  }
  const int rc = prepare_batch(U, nBlocks, blockStride, del, st);
  if (rc) return rc;
  if (nBlocks == 0) return 0;
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  launch_shock_detect((int)nBlocks);
  HIPCALL(hipGetLastError(), st);
  return finish(st);
}

int fx_hydro_shock_detect(const fx_tile* tiles, int nTiles, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.init || !tiles || nTiles < 0 || g.dp.shokOff < 0) {
    if (st) st->code = 3;
    return 3;
  }
  if (nTiles == 0) return 0;
  const int rc = upload_tiles(tiles, nTiles, true, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  launch_shock_detect(nTiles);
  HIPCALL(hipGetLastError(), st);
  return finish(st);
}


int fx_hydro_compute_dt(const double* U, long nBlocks, long blockStride, const double del[3],
                        double cfl, double* dtOut) {
  if (!g.init || !U || nBlocks <= 0 || !dtOut) return 3;
  if (nBlocks > 0x7fffffffL) return 3;
  const DevParams& P = g.dp;
  const int grid = (int)nBlocks;  // one workgroup per block
  if (g.partialCap < grid) {
    if (g.partial) (void)hipFree(g.partial);
    g.partial = nullptr;
    if (hipMalloc((void**)&g.partial, sizeof(double) * grid) != hipSuccess) return 4;
    g.partialCap = grid;
  }
  hipLaunchKernelGGL(k_dt, dim3(grid), dim3(kThreads), 0, g.stream, P, U, blockStride,
                     1. / del[0], 1. / del[1], 1. / del[2], cfl, nullptr, g.partial);
  std::vector<double> h((size_t)grid);
  if (hipMemcpyAsync(h.data(), g.partial, sizeof(double) * grid, hipMemcpyDeviceToHost, g.stream) !=
      hipSuccess)
    return 4;
  if (hipStreamSynchronize(g.stream) != hipSuccess) return 4;
  double m = 1.e300;  // Hydro_computeDt.F90:284-288: the smallest block dt
  for (double v : h) m = v < m ? v : m;
  *dtOut = m;
  return 0;
}

int fx_fluxreg_put(const fx_tile* fine, int nFine, const fx_fluxreg_face* faces, int nFaces,
                   double* reg, fx_status* st) {
  return run_fluxreg(fine, nFine, faces, nFaces, reg, true, st);
}

int fx_fluxreg_conserve(const fx_tile* coarse, int nCoarse, const fx_fluxreg_face* faces,
                        int nFaces, const double* reg, fx_status* st) {
  return run_fluxreg(coarse, nCoarse, faces, nFaces, const_cast<double*>(reg), false, st);
}

int fx_mlp_features(const double* states, long n, double gamma, float* x) {
  if (n <= 0) return 0;
  if (!states || !x || !(gamma > 1.0)) return 3;
  if (ensure_status(nullptr)) return 4;
  hipLaunchKernelGGL(k_mlp_features, dim3(grid_for(n)), dim3(kThreads), 0, g.stream, states, n,
                     gamma, x);
  return sync_or_async(nullptr);
}

int fx_mlp_forward(const float* params, const float* x, long n, float* y) {
  if (n <= 0) return 0;
  if (!params || !x || !y) return 3;
  if (ensure_status(nullptr) || pack_mlp(params, &g.mlpPacked, nullptr)) return 4;
  hipLaunchKernelGGL(k_mlp_forward, wave_grid(n), dim3(kLearnThreads), 0, g.stream, g.mlpPacked,
                     x, n, y);
  return sync_or_async(nullptr);
}

int fx_mlp_grad(const float* params, const float* x, const float* t, long n, float scale,
                float* grad, float* loss) {
  if (n <= 0) return 0;
  if (!params || !x || !t || !grad || !loss) return 3;
  if (ensure_status(nullptr) || pack_mlp(params, &g.mlpPacked, nullptr)) return 4;
  hipLaunchKernelGGL(k_mlp_grad, wave_grid(n), dim3(kLearnThreads), 0, g.stream, g.mlpPacked, x,
                     t, n, scale, grad, loss);
  return sync_or_async(nullptr);
}

int fx_mlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr) {
  if (!params || !grad || !m || !v || step < 1) return 3;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     g.stream, params, grad, m, v, lr, b1, b2, c1, c2, 1e-8f);
  return sync_or_async(nullptr);
}

int fx_riemann_batch(const double* states, long n, double gamma, const float* params, int maxit,
                     double tol, double* pstar, int* iters, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (n <= 0) return 0;
  if (!states || !pstar || !iters || !(gamma > 1.0) || maxit < 1 || !(tol > 0.0)) {
    if (st) st->code = 3;
    return 3;
  }
  if (ensure_status(st)) return 4;
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  if (params) {
    if (pack_mlp(params, &g.mlpPacked, st)) return 4;
    if (grow(&g.seedY, &g.seedYCap, (size_t)n, st)) return 4;
    if (record(0)) return 4;
    hipLaunchKernelGGL(k_mlp_seed_states, wave_grid(n), dim3(kLearnThreads), 0, g.stream, states,
                       n, gamma, g.mlpPacked, g.seedY, make_seed_gate(g.seedGate));
    if (record(1)) return 4;
    hipLaunchKernelGGL(k_riemann_batch<true>, dim3(grid_for(n)), dim3(kThreads), 0, g.stream,
                       states, n, gamma, g.seedY, tol, maxit, pstar, iters, g.status, make_seed_gate(g.seedGate));
    if (record(2)) return 4;
  } else {
    if (record(0) || record(1)) return 4;
    hipLaunchKernelGGL(k_riemann_batch<false>, dim3(grid_for(n)), dim3(kThreads), 0, g.stream,
                       states, n, gamma, (const float*)nullptr, tol, maxit, pstar, iters, g.status,
                       make_seed_gate(g.seedGate));
    if (record(2)) return 4;
  }
  if (record(3)) return 4;
  HIPCALL(hipGetLastError(), st);
  int rc = finish(st);
  if (g.timing && !g.async)
    for (int k = 0; k < 3; ++k)
      if (hipEventElapsedTime(&g.stageMs[k], g.ev[k], g.ev[k + 1]) != hipSuccess) g.stageMs[k] = -1.f;
  return rc;
}

int fx_grid_fill_guards(double* U, long nBlocks, long blockStride, const int* nbr,
                        fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  int rc = upload_neighbors(U, nBlocks, blockStride, nbr, st);
  if (rc) return rc;
  launch_fill(U, nBlocks, blockStride);
  HIPCALL(hipGetLastError(), st);
  if (g.async) return 0;
  HIPCALL(hipStreamSynchronize(g.stream), st);
  return 0;
}

int fx_grid_run_graph(double* U, long nBlocks, long blockStride, const int* nbr,
                      const double del[3], double cfl, int stepsPerGraph, int nReplays,
                      double* tOut, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (stepsPerGraph < 1 || nReplays < 1 || !(cfl > 0.) || !del) {
    if (st) st->code = 3;
    return 3;
  }
  if (g.dp.learned) {            // the seeded step is not offered under capture
    if (st) st->code = 3;
    return 3;
  }
  int rc = upload_neighbors(U, nBlocks, blockStride, nbr, st);
  if (!rc) rc = prepare_batch(U, nBlocks, blockStride, del, st);
  if (!rc) rc = ensure_step_scratch((int)nBlocks, st);
  if (rc) return rc;
  const DevParams& P = g.dp;
  const int dtGrid = (int)nBlocks;  // one workgroup per block
  if (g.partialCap < dtGrid) {
    if (g.partial) HIPCALL(hipFree(g.partial), st);
    g.partial = nullptr;
    HIPCALL(hipMalloc((void**)&g.partial, sizeof(double) * dtGrid), st);
    g.partialCap = dtGrid;
  }
  if (!g.dtBuf) HIPCALL(hipMalloc((void**)&g.dtBuf, 4 * sizeof(double)), st);
  if (!g.capStream) HIPCALL(hipStreamCreateWithFlags(&g.capStream, hipStreamNonBlocking), st);
  const hipStream_t saved = g.stream;
  const int savedAsync = g.async, savedTiming = g.timing;
  // the clock starts at zero; dt and t stay on the device for the whole run
  if (saved) HIPCALL(hipStreamSynchronize(saved), st);
  g.stream = g.capStream;
  // t = 0, and a graph run starts at the original CFL with no fallbacks
  const double init4[4] = {0.0, 0.0, cfl, 0.0};
  HIPCALL(hipMemcpy(g.dtBuf, init4, sizeof init4, hipMemcpyHostToDevice), st);
  g.runCfl = cfl;
  g.async = 1;
  g.timing = 0;
  g.dtDev = g.dtBuf;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int err = 0;
  g.capturing = true;
  if (hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) err = 4;
  for (int s = 0; !err && s < stepsPerGraph; ++s) {
    // dt from the interior cells (the guard fill does not change them), with
    // the CFL the previous step's flags and fallbacks left
    hipLaunchKernelGGL(k_dt, dim3(dtGrid), dim3(kThreads), 0, g.stream, P, U, blockStride,
                       1. / del[0], 1. / del[1], 1. / del[2], cfl, step_cfl_dev(), g.partial);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kThreads), 0, g.stream, g.partial, dtGrid,
                       g.dtBuf, g.dtBuf + 1);
    launch_fill(U, nBlocks, blockStride);
    if (P.shokOff >= 0) {  // Hydro.F90:94-123: flag, then fill SHOK_VAR's guard cells
      hipLaunchKernelGGL(k_zero_shock_tiles, dim3(1), dim3(1), 0, g.stream, g.status);
      launch_shock_detect((int)nBlocks);
      launch_fill(U, nBlocks, blockStride, g.prm.shokVar);
    }
    if (run_compute(nullptr, (int)nBlocks, 0.0, nullptr, nullptr, true)) err = 4;
    launch_step_cfl();
  }
  if (hipStreamEndCapture(g.stream, &graph) != hipSuccess) err = 4;
  g.capturing = false;
  if (!err && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) err = 4;
  // one status for the whole run: the replays keep the first error and sum
  // the counters of every step
  if (!err) hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  for (int r = 0; !err && r < nReplays; ++r)
    if (hipGraphLaunch(exec, g.stream) != hipSuccess) err = 4;
  if (!err && hipStreamSynchronize(g.stream) != hipSuccess) err = 4;
  double td[2] = {0.0, 0.0};
  if (!err && hipMemcpy(td, g.dtBuf, sizeof td, hipMemcpyDeviceToHost) != hipSuccess) err = 4;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  g.dtDev = nullptr;
  g.async = savedAsync;
  g.timing = savedTiming;
  g.stream = saved;
  if (err) {
    if (st) st->code = 4;
    return 4;
  }
  if (tOut) *tOut = td[1];
  return finish(st);
}

int fx_hydro_face_states(double* S, long cap, long* n) {
  if (!n) return 3;
  *n = g.lastFaces;
  if (!g.init || g.lastFaces <= 0 || !S || cap < g.lastFaces) return 3;
  FaceLaunch F{g.dtiles, g.lastTiles, g.fstates, nullptr, g.status,
               g.dcls, g.dtcls, g.dfaceStart, g.dwFace};
  const dim3 gc(grid_for(g.lastFaces));
  if (g.lastPrec == FX_PREC_MIXED) {
    if (g.lastMulti)
      hipLaunchKernelGGL((k_collect_faces<true, float>), gc, dim3(kThreads), 0, g.stream, g.cur, F,
                         g.lastFaces, S);
    else
      hipLaunchKernelGGL((k_collect_faces<false, float>), gc, dim3(kThreads), 0, g.stream, g.cur,
                         F, g.lastFaces, S);
  } else if (g.lastMulti) {
    hipLaunchKernelGGL((k_collect_faces<true, double>), gc, dim3(kThreads), 0, g.stream, g.cur, F,
                       g.lastFaces, S);
  } else {
    hipLaunchKernelGGL((k_collect_faces<false, double>), gc, dim3(kThreads), 0, g.stream, g.cur, F,
                       g.lastFaces, S);
  }
  return sync_or_async(nullptr);
}

int fx_mlp_set_gate(double gate) {
  if (!(gate >= 0.0)) return 3;
  g.seedGate = gate;
  return 0;
}

int fx_hydro_set_seed_params(const float* params) {
  if (ensure_status(nullptr)) return 4;
  if (!params) {
    g.seedReady = false;
    return 0;
  }
  if (pack_mlp(params, &g.seedPacked, nullptr)) return 4;
  if (hipStreamSynchronize(g.stream) != hipSuccess) return 4;
  g.seedReady = true;
  return 0;
}

void fx_hydro_set_stream(void* stream) { g.stream = (hipStream_t)stream; }
void fx_hydro_set_timing(int on) { g.timing = on; }
void fx_hydro_set_kernel_path(int path) {
  g.legacy = (path == 1);
  g.solveUpdOff = (path == 2);
  g.batchValid = false;
}
int fx_hydro_stage_times(float* ms, int n) {
  for (int k = 0; k < n && k < 3; ++k) ms[k] = g.stageMs[k];
  return 3;
}
void fx_hydro_set_async(int a) { g.async = a; }

int fx_hydro_set_block_classes(const int* nb, int nClasses) {
  if (!g.init || nClasses < 0 || nClasses > FX_MAX_CLASSES || (nClasses > 0 && !nb)) return 3;
  std::vector<DevParams> cls;
  for (int c = 0; c < nClasses; ++c) {
    fx_params p = g.prm;
    p.nxb = nb[3 * c];
    p.nyb = p.ndim > 1 ? nb[3 * c + 1] : 1;
    p.nzb = p.ndim > 2 ? nb[3 * c + 2] : 1;
    if (validate(&p)) return 3;
    if (p.ndim < 2 || p.nxb % 8 || p.nyb % 8) return 3;  // the column march's tiles
    cls.push_back(make_dev_params(p));
  }
  if (nClasses > 0) {
    if (grow(&g.dcls, &g.dclsCap, (size_t)nClasses, nullptr)) return 4;
    if (hipMemcpy(g.dcls, cls.data(), sizeof(DevParams) * nClasses, hipMemcpyHostToDevice) !=
        hipSuccess)
      return 4;
  }
  bool same = cls.size() == g.hcls.size();
  for (size_t c = 0; same && c < cls.size(); ++c)
    same = cls[c].nxb == g.hcls[c].nxb && cls[c].nyb == g.hcls[c].nyb && cls[c].nzb == g.hcls[c].nzb;
  g.hcls = cls;
  if (!same) {  // work items and a grid name classes by index
    g.nItems = 0;
    g.batchValid = false;
    g.amr.ready = false;
  }
  return 0;
}

int fx_amr_setup(const fx_tile* tiles, const fx_amr_block* blocks, int nBlocks, int nLevels,
                 const int* levelExtent, const int* periodic, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  amr_free();
  auto bad = [&]() {
    if (st) st->code = 3;
    return 3;
  };
  if (!g.init || g.hcls.empty() || !tiles || !blocks || nBlocks <= 0 || nLevels < 1 ||
      nLevels > 16 || !levelExtent || g.prm.nunk > 64)
    return bad();
  const int ndim = g.dp.ndim;
  // every level has one block shape, its cell width, and twice the
  // resolution of the one below
  std::vector<int> lcls(nLevels, -1);
  std::vector<std::array<double, 3>> ldel(nLevels);
  for (int b = 0; b < nBlocks; ++b) {
    const int L = blocks[b].level - 1, c = tiles[b].block_class;
    if (L < 0 || L >= nLevels || c < 0 || c >= (int)g.hcls.size()) return bad();
    if (lcls[L] < 0) {
      lcls[L] = c;
      for (int a = 0; a < 3; ++a) ldel[L][a] = tiles[b].del[a];
    } else if (g.hcls[lcls[L]].nxb != g.hcls[c].nxb || g.hcls[lcls[L]].nyb != g.hcls[c].nyb ||
               g.hcls[lcls[L]].nzb != g.hcls[c].nzb) {
      return bad();
    }
  }
  std::vector<AmrLevelDev> lev(nLevels);
  std::vector<std::vector<int>> index(nLevels);
  for (int L = 0; L < nLevels; ++L) {
    if (lcls[L] < 0) return bad();
    const DevParams& P = g.hcls[lcls[L]];
    const int nb[3] = {P.nxb, P.nyb, P.nzb};
    for (int a = 0; a < 3; ++a) {
      lev[L].nb[a] = nb[a];
      lev[L].ext[a] = a < ndim ? levelExtent[3 * L + a] : 1;
      lev[L].del[a] = ldel[L][a];
      if (lev[L].ext[a] < 1) return bad();
      if (L > 0 && a < ndim &&
          (long)lev[L].ext[a] * nb[a] != 2L * lev[L - 1].ext[a] * lev[L - 1].nb[a])
        return bad();
    }
    index[L].assign((size_t)lev[L].ext[0] * lev[L].ext[1] * lev[L].ext[2], -1);
  }
  Lib::Amr& A = g.amr;
  A.hblocks.resize(nBlocks);
  for (int b = 0; b < nBlocks; ++b) {
    const int L = blocks[b].level - 1;
    AmrBlock& B = A.hblocks[b];
    B.U = tiles[b].U;
    B.level = L;
    B.cls = tiles[b].block_class;
    for (int a = 0; a < 3; ++a) {
      B.pos[a] = a < ndim ? blocks[b].pos[a] : 0;
      if (B.pos[a] < 0 || B.pos[a] >= lev[L].ext[a]) return bad();
    }
    int& slot = index[L][((size_t)B.pos[2] * lev[L].ext[1] + B.pos[1]) * lev[L].ext[0] + B.pos[0]];
    if (slot >= 0) return bad();  // two blocks at one position
    slot = b;
  }
  // Each block of a level is either covered by finer blocks over its whole
  // area or not at all, as FLASH's AMReX grids are (blocking factor 2 NXB,
  // gr_amrexInit.F90:139-145).  The leaf-cell dt (a block with any finer
  // block over it leaves the LEAF iteration, block_1lev_iterator.F90:293-330)
  // then equals the device's covered-cell rule.
  for (int L = 0; L + 1 < nLevels; ++L) {
    const AmrLevelDev& C = lev[L];
    const AmrLevelDev& F = lev[L + 1];
    for (int pz = 0; pz < C.ext[2]; ++pz)
      for (int py = 0; py < C.ext[1]; ++py)
        for (int px = 0; px < C.ext[0]; ++px) {
          if (index[L][((size_t)pz * C.ext[1] + py) * C.ext[0] + px] < 0) continue;
          const int p[3] = {px, py, pz};
          int lo[3], hi[3];
          for (int a = 0; a < 3; ++a) {
            if (a >= ndim) {
              lo[a] = hi[a] = 0;
              continue;
            }
            lo[a] = 2 * p[a] * C.nb[a] / F.nb[a];
            hi[a] = (2 * (p[a] + 1) * C.nb[a] - 1) / F.nb[a];
          }
          long have = 0, all = 0;
          for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
              for (int x = lo[0]; x <= hi[0]; ++x) {
                ++all;
                if (index[L + 1][((size_t)z * F.ext[1] + y) * F.ext[0] + x] >= 0) ++have;
              }
          if (have > 0 && have < all) return bad();
        }
  }
  A.nBlocks = nBlocks;
  A.nLevels = nLevels;
  for (int a = 0; a < 3; ++a) A.periodic[a] = periodic ? periodic[a] != 0 : 0;
  for (int v : {FX_VELX, FX_VELY, FX_VELZ, FX_ENER, FX_EINT}) A.perMass |= 1ULL << g.prm.var[v];
  // the grid's tiles, with flux and pavg arrays of the grid's own where absent
  // (the library's shared ones may be regrown by other calls)
  {
    std::vector<fx_tile> own(tiles, tiles + nBlocks);
    size_t fluxNeed = 0, pavgNeed = 0;
    for (int b = 0; b < nBlocks; ++b) {
      const DevParams& P = g.hcls[own[b].block_class];
      for (int d = 0; d < ndim; ++d)
        if (!own[b].flux[d]) fluxNeed += (size_t)P.nface[d] * FX_NFLUX;
      if (!own[b].pavg) pavgNeed += (size_t)P.icells;
    }
    if (fluxNeed && hipMalloc((void**)&A.flux, sizeof(double) * fluxNeed) != hipSuccess) {
      amr_free();
      return 4;
    }
    if (pavgNeed && hipMalloc((void**)&A.pavg, sizeof(double) * pavgNeed) != hipSuccess) {
      amr_free();
      return 4;
    }
    double *fp = A.flux, *pp = A.pavg;
    for (int b = 0; b < nBlocks; ++b) {
      const DevParams& P = g.hcls[own[b].block_class];
      for (int d = 0; d < ndim; ++d)
        if (!own[b].flux[d]) {
          own[b].flux[d] = fp;
          fp += (size_t)P.nface[d] * FX_NFLUX;
        }
      if (!own[b].pavg) {
        own[b].pavg = pp;
        pp += (size_t)P.icells;
      }
    }
    const int rc = upload_tiles(own.data(), nBlocks, true, st);
    if (rc) {
      amr_free();
      return rc;
    }
  }
  A.tiles = g.htiles;
  A.gen = g.uploadGen;
  if (g.multi) {
    if (!multi_march_shape(&A.TX)) return bad();
    A.TY = 8;
  } else {
    int NT = 0;
    if (!march_shape(g.cur, &A.TX, &A.TY, &NT)) return bad();
  }
  if (g.multi && build_march_items(nBlocks, A.TX, st)) return 4;
  if (ensure_step_scratch(nBlocks, st)) return 4;
  // prefix of guard-inclusive box cells per block
  std::vector<long> box(nBlocks + 1, 0);
  for (int b = 0; b < nBlocks; ++b) box[b + 1] = box[b] + g.hcls[A.hblocks[b].cls].ncell;
  A.boxCells = box[nBlocks];
  int rc = upload_vec(&A.dboxStart, box, st);
  if (!rc) rc = upload_vec(&A.dwBox, wave_table(box), st);
  {  // guard-fill work: each class's guard cells in box order, and chunks of
     // kThreads guard cells of one block
    std::vector<int> glist, goff(1, 0);
    for (const DevParams& P : g.hcls) {
      for (long r = 0; r < P.ncell; ++r) {
        const int loc[3] = {(int)(r % P.nx), (int)((r / P.nx) % P.ny),
                            (int)(r / ((long)P.nx * P.ny))};
        bool inside = true;
        for (int a = 0; a < P.ndim; ++a)
          if (loc[a] < P.lo[a] || loc[a] > P.hi[a]) inside = false;
        if (!inside) glist.push_back((int)r);
      }
      goff.push_back((int)glist.size());
    }
    std::vector<int2> ch;
    for (int b = 0; b < nBlocks; ++b) {
      const int c = A.hblocks[b].cls;
      for (int o = 0; o < goff[c + 1] - goff[c]; o += kThreads) ch.push_back(make_int2(b, o));
    }
    A.nChunks = (long)ch.size();
    if (!rc) rc = upload_vec(&A.dchunks, ch, st);
    if (!rc) rc = upload_vec(&A.dguard, glist, st);
    if (!rc) rc = upload_vec(&A.dguardOff, goff, st);
  }
  for (int L = 0; !rc && L < nLevels; ++L) {
    int* d = nullptr;
    rc = upload_vec(&d, index[L], st);
    A.dindex.push_back(d);
    lev[L].index = d;
  }
  if (!rc) rc = upload_vec(&A.dlevels, lev, st);
  if (!rc) rc = upload_vec(&A.dblocks, A.hblocks, st);
  // restriction lists: the blocks of each level above the first, with their
  // covered parent cells
  A.dlist.assign(nLevels, nullptr);
  A.drstart.assign(nLevels, nullptr);
  A.nlist.assign(nLevels, 0);
  for (int L = 1; !rc && L < nLevels; ++L) {
    std::vector<int> list;
    std::vector<long> start(1, 0);
    for (int b = 0; b < nBlocks; ++b)
      if (A.hblocks[b].level == L) {
        const DevParams& P = g.hcls[A.hblocks[b].cls];
        list.push_back(b);
        start.push_back(start.back() + (long)(P.nxb / 2) * (ndim > 1 ? P.nyb / 2 : 1) *
                                           (ndim > 2 ? P.nzb / 2 : 1));
      }
    A.nlist[L] = (int)list.size();
    rc = upload_vec(&A.dlist[L], list, st);
    if (!rc) rc = upload_vec(&A.drstart[L], start, st);
  }
  // flux-register faces of each level pair: every coarse copy of each face
  // on the boundary of the finer level (periodic wrap included, none on a
  // non-periodic domain boundary)
  A.dfaces.assign(nLevels, nullptr);
  A.dreg.assign(nLevels, nullptr);
  A.nfaces.assign(nLevels, 0);
  for (int L = 0; !rc && L + 1 < nLevels; ++L) {
    const AmrLevelDev &C = lev[L], &F = lev[L + 1];
    std::vector<fx_fluxreg_face> faces;
    auto at = [&](int Lv, const int p[3]) {
      const AmrLevelDev& V = lev[Lv];
      return index[Lv][((size_t)p[2] * V.ext[1] + p[1]) * V.ext[0] + p[0]];
    };
    for (int fb = 0; fb < nBlocks; ++fb) {
      const AmrBlock& B = A.hblocks[fb];
      if (B.level != L + 1) continue;
      for (int d = 0; d < ndim; ++d)
        for (int side = 0; side < 2; ++side) {
          int np[3] = {B.pos[0], B.pos[1], B.pos[2]};
          np[d] += side ? 1 : -1;
          if (np[d] < 0 || np[d] >= F.ext[d]) {
            if (!A.periodic[d]) continue;
            np[d] = (np[d] + F.ext[d]) % F.ext[d];
          }
          if (at(L + 1, np) >= 0) continue;  // a fine neighbour: no register face
          const int ncell = C.ext[d] * C.nb[d];
          const int gcf = (B.pos[d] * F.nb[d] + (side ? F.nb[d] : 0)) / 2;  // coarse face
          const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
          const int n1 = t1 < ndim ? F.nb[t1] / 2 : 1, n2 = t2 < ndim ? F.nb[t2] / 2 : 1;
          for (int c2 = 0; c2 < n2; ++c2)
            for (int c1 = 0; c1 < n1; ++c1) {
              int gct[3] = {0, 0, 0};  // coarse transverse cells
              if (t1 < ndim) gct[t1] = B.pos[t1] * (F.nb[t1] / 2) + c1;
              if (t2 < ndim) gct[t2] = B.pos[t2] * (F.nb[t2] / 2) + c2;
              int fi[3] = {0, 0, 0};
              fi[d] = side ? F.nb[d] : 0;
              if (t1 < ndim) fi[t1] = 2 * c1;
              if (t2 < ndim) fi[t2] = 2 * c2;
              // the coarse cells on either side of the face and their blocks
              int last = -1;
              for (int k = 0; k < 2; ++k) {
                int gc[3] = {gct[0], gct[1], gct[2]};
                gc[d] = k ? gcf : gcf - 1;
                if (gc[d] < 0 || gc[d] >= ncell) {
                  if (!A.periodic[d]) continue;
                  gc[d] = (gc[d] + ncell) % ncell;
                }
                int bp[3];
                for (int a = 0; a < 3; ++a) bp[a] = gc[a] / C.nb[a];
                const int cb = at(L, bp);
                if (cb < 0) { rc = 3; break; }  // the coarse level does not cover the face
                if (cb == last) continue;       // both cells in one block: one copy
                last = cb;
                fx_fluxreg_face f{};
                f.coarse_tile = cb;
                f.fine_tile = fb;
                f.dir = d;
                int ci[3];
                for (int a = 0; a < 3; ++a) ci[a] = gc[a] - bp[a] * C.nb[a];
                if (!k) ci[d] += 1;  // the face is the high face of the low cell
                f.ci = ci[0];
                f.cj = ci[1];
                f.ck = ci[2];
                f.fi = fi[0];
                f.fj = fi[1];
                f.fk = fi[2];
                faces.push_back(f);
              }
            }
        }
    }
    if (rc) break;
    if (check_reg_faces(A.tiles.data(), nBlocks, faces.data(), (int)faces.size(), true) ||
        check_reg_faces(A.tiles.data(), nBlocks, faces.data(), (int)faces.size(), false)) {
      rc = 3;
      break;
    }
    A.nfaces[L] = (int)faces.size();
    rc = upload_vec(&A.dfaces[L], faces, st);
    if (!rc && !faces.empty())
      if (hipMalloc((void**)&A.dreg[L], sizeof(double) * FX_NFLUX * faces.size()) != hipSuccess)
        rc = 4;
  }
  // the guard fill's sources, fixed by the layout: looked up once here
  if (!rc && A.nChunks > 0) {
    const size_t slots = (size_t)A.nChunks * kThreads;
    if (hipMalloc((void**)&A.dplan, sizeof(int2) * slots) != hipSuccess ||
        hipMalloc((void**)&A.dsrc, sizeof(double*) * slots) != hipSuccess ||
        hipMalloc((void**)&A.drest, sizeof(int) * slots) != hipSuccess ||
        hipMalloc((void**)&A.dnRest, sizeof(unsigned)) != hipSuccess) {
      rc = 4;
    } else {
      hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, g.stream, A.dnRest);
      hipLaunchKernelGGL(k_amr_plan, dim3((unsigned)A.nChunks), dim3(kThreads), 0, g.stream,
                         amr_launch(), A.dchunks, A.dplan, A.dsrc, A.drest, A.dnRest);
      if (hipMemcpyAsync(&A.nRest, A.dnRest, sizeof(unsigned), hipMemcpyDeviceToHost,
                         g.stream) != hipSuccess ||
          hipStreamSynchronize(g.stream) != hipSuccess)
        rc = 4;
    }
    // the distinct coarse cells under the prolonged guard cells, and each
    // prolonged slot's record
    if (!rc && A.nRest) {
      std::vector<int> rest(A.nRest);
      std::vector<int2> plan(slots);
      if (hipMemcpy(rest.data(), A.drest, sizeof(int) * A.nRest, hipMemcpyDeviceToHost) !=
              hipSuccess ||
          hipMemcpy(plan.data(), A.dplan, sizeof(int2) * slots, hipMemcpyDeviceToHost) !=
              hipSuccess) {
        rc = 4;
      } else {
        std::vector<std::pair<long long, int>> keys;
        keys.reserve(A.nRest);
        for (unsigned q = 0; q < A.nRest; ++q) {
          const int2 pe = plan[rest[q]];
          if (pe.x <= kPlanCoarse)
            keys.emplace_back(((long long)(kPlanCoarse - pe.x) << 32) | (unsigned)pe.y, (int)q);
        }
        std::sort(keys.begin(), keys.end());
        std::vector<int> recOf(A.nRest, -1);
        std::vector<AmrCoarseRec> recs;
        for (size_t e = 0; e < keys.size(); ++e) {
          if (e == 0 || keys[e].first != keys[e - 1].first)
            recs.push_back(AmrCoarseRec{(int)(keys[e].first >> 32),
                                        (int)(keys[e].first & 0xffffffffLL)});
          recOf[keys[e].second] = (int)recs.size() - 1;
        }
        A.nRec = (int)recs.size();
        rc = upload_vec(&A.drecs, recs, st);
        if (!rc) rc = upload_vec(&A.drecOf, recOf, st);
        if (!rc && A.nRec &&
            (hipMalloc((void**)&A.dcv, sizeof(double) * 4 * (size_t)g.prm.nunk * A.nRec) !=
                 hipSuccess ||
             hipMalloc((void**)&A.drecOk, sizeof(int) * (size_t)A.nRec) != hipSuccess))
          rc = 4;
      }
    }
  }
  if (rc) {
    amr_free();
    if (st) st->code = rc;
    return rc;
  }
  if (!g.dtBuf) HIPCALL(hipMalloc((void**)&g.dtBuf, 4 * sizeof(double)), st);
  A.ready = true;
  return 0;
}

int fx_hydro_autotune(const fx_tile* tiles, int nTiles, double dt, int reps, fx_tune_result* out,
                      int cap, int* n) {
  fx_status st{};
  if (n) *n = 0;
  if (!g.init || !tiles || nTiles <= 0 || !(dt > 0.) || reps < 1) return 3;
  if (g.cur.learned && !g.seedReady) return 3;
  const int rc = upload_tiles(tiles, nTiles, true, &st);
  if (rc) return rc;
  int TX = 8, TY = 8, NT = 0;
  if (g.multi ? !multi_march_shape(&TX) : (g.legacy || !march_shape(g.cur, &TX, &TY, &NT)))
    return 3;  // the tuned stages are the column march's
  if (g.cur.learned && !g.seedReady) return 3;
  return tune_flux_stages(nTiles, dt, reps, out, cap, n, nullptr, &st);
}

int fx_amr_autotune(double dt, int reps, fx_tune_result* out, int cap, int* n) {
  fx_status st{};
  if (n) *n = 0;
  if (!g.amr.ready || !(dt > 0.) || reps < 1) return 3;
  if (g.cur.learned && !g.seedReady) return 3;
  if (amr_ensure_uploaded(&st)) return 4;
  int TX = g.amr.TX;
  const int rc = tune_flux_stages(g.amr.nBlocks, dt, reps, out, cap, n, &TX, &st);
  if (rc) return rc;
  g.amr.TX = TX;
  return 0;
}

void fx_hydro_autotune_reset(void) {
  g.tuneTX = g.persist = g.itemOrder = g.tuneOcc = 0;
  g.nItems = 0;
  if (g.amr.ready) {  // back to the built-in tile of the grid
    int TX = 8, TY = 8, NT = 0;
    if (g.multi ? multi_march_shape(&TX) : march_shape(g.cur, &TX, &TY, &NT)) g.amr.TX = TX;
  }
}

int fx_amr_fill_guards(fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.amr.ready) return 3;
  if (amr_ensure_uploaded(st)) return 4;
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  amr_launch_fill();
  HIPCALL(hipGetLastError(), st);
  return finish(st);
}

int fx_amr_restrict(fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.amr.ready) return 3;
  if (amr_ensure_uploaded(st)) return 4;
  hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  amr_launch_restrict();
  HIPCALL(hipGetLastError(), st);
  return finish(st);
}

int fx_amr_compute_dt(double cfl, double* dtOut) {
  if (!g.amr.ready || !(cfl > 0.) || !dtOut) return 3;
  if (amr_ensure_uploaded(nullptr)) return 4;
  if (amr_launch_dt(cfl, nullptr)) return 4;
  double td[2];
  if (hipMemcpyAsync(td, g.dtBuf, sizeof td, hipMemcpyDeviceToHost, g.stream) != hipSuccess ||
      hipStreamSynchronize(g.stream) != hipSuccess)
    return 4;
  *dtOut = td[0];
  return 0;
}

int fx_amr_step(double dt, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.amr.ready || !(dt > 0.)) {
    if (st) st->code = 3;
    return 3;
  }
  if (g.cur.learned && !g.seedReady) {
    if (st) st->code = 3;
    return 3;
  }
  if (amr_ensure_uploaded(st)) return 4;
  if (amr_launch_step(dt)) return 4;
  HIPCALL(hipGetLastError(), st);
  return finish(st);
}

int fx_amr_run_graph(double cfl, int stepsPerGraph, int nReplays, double* tOut, fx_status* st) {
  if (st) memset(st, 0, sizeof *st);
  if (!g.amr.ready || stepsPerGraph < 1 || nReplays < 1 || !(cfl > 0.) ||
      (g.dp.learned && !g.seedReady)) {
    if (st) st->code = 3;
    return 3;
  }
  if (amr_ensure_uploaded(st)) return 4;
  {  // the blocks' dts of the dt reduction, sized before capture (amr_launch_dt)
    const int grid = g.amr.nBlocks;
    if (g.partialCap < grid) {
      if (g.partial) HIPCALL(hipFree(g.partial), st);
      g.partial = nullptr;
      HIPCALL(hipMalloc((void**)&g.partial, sizeof(double) * grid), st);
      g.partialCap = grid;
    }
  }
  if (!g.capStream) HIPCALL(hipStreamCreateWithFlags(&g.capStream, hipStreamNonBlocking), st);
  const hipStream_t saved = g.stream;
  const int savedAsync = g.async, savedTiming = g.timing;
  if (saved) HIPCALL(hipStreamSynchronize(saved), st);
  HIPCALL(hipDeviceSynchronize(), st);
  g.stream = g.capStream;
  // t = 0, and a graph run starts at the original CFL with no fallbacks
  const double init4[4] = {0.0, 0.0, cfl, 0.0};
  HIPCALL(hipMemcpy(g.dtBuf, init4, sizeof init4, hipMemcpyHostToDevice), st);
  g.runCfl = cfl;
  g.async = 1;
  g.timing = 0;
  g.dtDev = g.dtBuf;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int err = 0;
  g.capturing = true;
  if (hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) err = 4;
  for (int s = 0; !err && s < stepsPerGraph; ++s) {
    if (amr_launch_dt(cfl, nullptr)) err = 4;
    if (!err && amr_launch_step(0.0)) err = 4;
  }
  if (hipStreamEndCapture(g.stream, &graph) != hipSuccess) err = 4;
  g.capturing = false;
  if (!err && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) err = 4;
  // one status for the whole run: the replays keep the first error and sum
  // the counters of every step
  if (!err) hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(1), 0, g.stream, g.status);
  for (int r = 0; !err && r < nReplays; ++r)
    if (hipGraphLaunch(exec, g.stream) != hipSuccess) err = 4;
  if (!err && hipStreamSynchronize(g.stream) != hipSuccess) err = 4;
  double td[2] = {0.0, 0.0};
  if (!err && hipMemcpy(td, g.dtBuf, sizeof td, hipMemcpyDeviceToHost) != hipSuccess) err = 4;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  g.dtDev = nullptr;
  g.async = savedAsync;
  g.timing = savedTiming;
  g.stream = saved;
  if (err) {
    if (st) st->code = 4;
    return 4;
  }
  if (tOut) *tOut = td[1];
  return finish(st);
}
const char* fx_hydro_version(void) { return "flash5_amd fx_hydro 0.1 (gfx950)"; }

}  // extern "C"
