// fx_common.hpp -- device structures and helpers shared by the kernels of
// every translation unit of the library (the main unit fx_hydro.hip and the
// column-march units built from fx_march_inst.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fx_device.hpp"
#include "fx_hydro.h"

namespace fx {
constexpr int kStateComps = 12;  // Wp and Wn of dens, vx, vy, vz, pres, eint
constexpr int kThreads = 256;

struct DevStatus {
  int code;
  int tile, i, j, k, dir;
  unsigned long long fallbacks;
  unsigned long long newtonIters;
  unsigned long long newtonFallbacks;
  unsigned long long shockTiles;  // tiles with a flagged cell (k_shock_detect)
};

struct Launch {
  const fx_tile* tiles;  // device array
  int nTiles;
  double* states;        // scratch: [tile][band][comp][cell]
  DevStatus* status;
  double dt;
  const double* dtp;     // device-resident dt (graph-captured steps), or null
};

// Tiles of several block classes in one launch: each tile's class selects its
// parameters, and per-tile prefix offsets map a flat cell (or face) index to
// its tile.
struct ClassMap {
  const DevParams* cls;        // parameters per class
  const unsigned char* tcls;   // class of each tile
  const long* start;           // items before tile t; start[nTiles] = all items
  const int* wtab;             // tile of every 64th item (find_tile_w)
};

// t with start[t] <= gid < start[t + 1] (start non-decreasing, start[0] = 0)
__device__ __forceinline__ int find_tile(const long* __restrict__ start, int n, long gid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= gid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The same through a table of 64-item chunks: wtab[c] is the tile holding
// item 64 c.  Every tile holds at least 64 items, so an item lies in that
// tile or the next; one load and one compare instead of a search.  wtab
// null: the search.
__device__ __forceinline__ int find_tile_w(const long* __restrict__ start,
                                          const int* __restrict__ wtab, int n, long gid) {
  if (!wtab) return find_tile(start, n, gid);
  int t = wtab[gid >> 6];
  if (t + 1 < n && start[t + 1] <= gid) ++t;
  return t;
}

// the step's dt: from device memory when the host does not know it
__device__ __forceinline__ double step_dt(double dt, const double* dtp) { return dtp ? *dtp : dt; }

__device__ __forceinline__ void report_error(DevStatus* s, int code, int tile, int i, int j, int k, int dir) {
  if (atomicCAS(&s->code, 0, code) == 0) {
    s->tile = tile;
    s->i = i;
    s->j = j;
    s->k = k;
    s->dir = dir;
  }
}

// Adds a thread's Newton counters to the status.  A whole wavefront sums
// its lanes first and adds once per counter: one atomic per lane on one
// address serialises (a face stage with a thread per face spent a second
// on them).
__device__ __forceinline__ void flush_exact(DevStatus* s, const ExactStats& es) {
  unsigned long long it = es.iters, fb = es.fallbacks;
  if (__activemask() == ~0ULL) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      it += __shfl_xor(it, o);
      fb += __shfl_xor(fb, o);
    }
    if ((threadIdx.x & 63) != 0) return;
  }
  if (it) atomicAdd(&s->newtonIters, it);
  if (fb) atomicAdd(&s->newtonFallbacks, fb);
}

__device__ __forceinline__ long cidx(const DevParams& P, int i, int j, int k) {
  return ((long)k * P.ny + j) * P.nx + i;
}

// primitive vector of cell c in the reconstruction type R (a conversion
// from the FP64 cell data when R is float)
template <typename R>
__device__ __forceinline__ void load_prim(const DevParams& P, const gdouble* __restrict__ U, long c,
                                          R V[8]) {
  V[P_DENS] = R(U[P.voff[FX_DENS] + c]);
  V[P_VELX] = R(U[P.voff[FX_VELX] + c]);
  V[P_VELY] = R(U[P.voff[FX_VELY] + c]);
  V[P_VELZ] = R(U[P.voff[FX_VELZ] + c]);
  V[P_PRES] = R(U[P.voff[FX_PRES] + c]);
  V[P_GAMC] = R(U[P.voff[FX_GAMC] + c]);
  V[P_GAME] = R(U[P.voff[FX_GAME] + c]);
  V[P_EINT] = R(U[P.voff[FX_EINT] + c]);
}

// FlatTilde of one cell along DIR (hy_getRiemannState.F90:365-404)
template <int DIR, typename R = double>
__device__ __forceinline__ R flat_tilde(const DevParams& P, const gdouble* __restrict__ U, long c,
                                        long s) {
  const gdouble* pr = U + P.voff[FX_PRES];
  const gdouble* vn = U + P.voff[FX_VELX + DIR];
  const R dp1 = (R(pr[c + s]) - R(pr[c - s]));
  const R dp2 = (R(pr[c + 2 * s]) - R(pr[c - 2 * s]));
  const R dv1 = R(vn[c + s]) - R(vn[c - s]);
  const R presL = R(pr[c - s]), presR = R(pr[c + s]);
  const R Sp = (fabs(dp2) > R(1.e-15)) ? dp1 / dp2 - R(0.75) : R(0.);
  R ft = fmax(R(0.0), fmin(R(1.0), R(10.0) * Sp));
  if ((fabs(dp1) / fmin(presL, presR) < R(1.) / R(3.)) || dv1 > R(0.)) ft = R(0.);
  return ft;
}

// FlatCoeff (hy_getRiemannState.F90:408-453)
template <int DIR, typename R = double>
__device__ __forceinline__ R flat_coeff(const DevParams& P, const gdouble* __restrict__ U, long c,
                                        long s) {
  const gdouble* pr = U + P.voff[FX_PRES];
  const R dp1 = (R(pr[c + s]) - R(pr[c - s]));
  const R f0 = flat_tilde<DIR, R>(P, U, c, s);
  if (dp1 < R(0.0)) return fmax(f0, flat_tilde<DIR, R>(P, U, c + s, s));
  if (dp1 == R(0.)) return f0;
  return fmax(f0, flat_tilde<DIR, R>(P, U, c - s, s));
}

__device__ __forceinline__ long axis_stride(const DevParams& P, int a) {
  return a == 0 ? 1L : (a == 1 ? (long)P.nx : (long)P.nx * P.ny);
}

// sigma of cell c along direction SD (the transverse-only PPM call)
template <class T> struct NoDeduce { using type = T; };

template <int SD, typename R>
__device__ __forceinline__ bool sigma_at(const DevParams& P, const gdouble* __restrict__ U, long c,
                                         R sig[8], typename NoDeduce<EigT<R>>::type* eout) {
  const long s = axis_stride(P, SD);
  R Vm[8], Vc[8], Vp[8];
  load_prim(P, U, c - s, Vm);
  load_prim(P, U, c, Vc);
  load_prim(P, U, c + s, Vp);
  EigT<R> e;
  if (!eig_make<SD>(Vc, e)) return false;
  trans_sigma<SD>(P, Vm, Vc, Vp, e, sig);
  if (eout) *eout = e;
  return true;
}

// aux internal-energy fluxes F[6], F[7] from the mass flux f0 (upwind face
// state)
template <typename R>
__device__ __forceinline__ void aux_fluxes(const DevParams& P, R f0, const R* VL, const R* VR,
                                           double F[8]) {
  F[6] = 0.;
  F[7] = 0.;
  if (P.aux) {
    if (f0 > R(0.)) {
      F[7] = double(f0 / VL[P_DENS]);
      F[6] = double(f0 * VL[P_EINT]);
    } else {
      F[7] = double(f0 / VR[P_DENS]);
      F[6] = double(f0 * VR[P_EINT]);
    }
  }
}

// Whether the face between cells cl and cr takes HLL: RiemannSolver = HLL,
// or HYBR with a flagged cell on either side (hy_getFaceFlux.F90:353-367).
__device__ __forceinline__ bool hll_face(const DevParams& P, const gdouble* __restrict__ U, long cl,
                                         long cr) {
  if (P.riemann == FX_HLL) return true;
  if (P.riemann != FX_HYBR) return false;
  return U[P.shokOff + cl] + U[P.shokOff + cr] > 0.;
}

// The kernels' solver family: RS = FX_HLLC compiles HLLC alone; RS = FX_HYBR
// serves RiemannSolver = HLL and HYBR, choosing per face at run time.
template <int RS>
__device__ __forceinline__ bool use_hll(const DevParams& P, const gdouble* __restrict__ U, long cl,
                                        long cr) {
  if constexpr (RS == FX_HYBR) return hll_face(P, U, cl, cr);
  else return false;
}

// Face flux densities F[0..7] (hy_getFaceFlux.F90:324-442): HLL (useHll) or
// HLLC on the face states VL, VR in the reconstruction type R, then the aux
// internal-energy fluxes.  The exact solver takes exact_face_flux.
template <int D, int RS, typename R>
__device__ __forceinline__ int face_flux(const DevParams& P, const R* VL, const R* VR, bool useHll,
                                         double F[8], ExactStats& es) {
  static_assert(RS == FX_HLLC || RS == FX_HYBR, "face_flux serves HLL, HLLC and the hybrid switch");
  R f0;
  if constexpr (sizeof(R) == sizeof(double)) {
    const int ierr = (RS == FX_HYBR && useHll) ? hll<D>(P, VL, VR, F) : hllc<D>(P, VL, VR, F);
    if (ierr) return ierr;
    f0 = F[0];
  } else {
    R Fr[6];
    const int ierr = (RS == FX_HYBR && useHll) ? hll<D>(P, VL, VR, Fr) : hllc<D>(P, VL, VR, Fr);
    if (ierr) return ierr;
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] = double(Fr[q]);
    f0 = Fr[0];
  }
  aux_fluxes(P, f0, VL, VR, F);
  return 0;
}

__device__ __forceinline__ long face_index(const DevParams& P, int d, int i, int j, int k) {
  return ((long)(k - P.lo[2]) * P.fext[d][1] + (j - P.lo[1])) * P.fext[d][0] + (i - P.lo[0]);
}

// ---- the conservative update of one 3-D cell in two parts ----
// hy_unsplitUpdate.F90:440-925, hy_energyFix.F90:188-251 and the EOS, as
// update_cell does them, split where a plane-by-plane sweep knows a cell's
// faces: the x and y terms once the cell's plane has its fluxes
// (upd_start), the z term (which needs the face above), the energy fix and
// the EOS one plane later (upd_finish).  Each quantity still takes its x, y
// and z terms in that order, so the bits are update_cell's.
struct UpdPart {
  double U0[5];     // dens, momenta, total energy after the x and y terms
  double IntEner;   // aux internal energy after the x and y terms
  double S1, S2;    // pressure sources of x and y
  double pres;      // the cell's face-averaged pressure
};

// one direction's terms (update_cell's apply)
__device__ __forceinline__ void upd_dir(const DevParams& P, double dt, double del, double pres,
                                        const double FL[8], const double FR[8], double U0[5],
                                        double& IntEner, double& S) {
  S = (FL[5] - FR[5]) * (1.0 - P.fP) / del;
  if (P.aux) IntEner = IntEner + dt / del * (FL[6] - FR[6] + pres * (FL[7] - FR[7]));
#pragma unroll
  for (int q = 0; q < 5; ++q) U0[q] = U0[q] - dt / del * (FR[q] - FL[q]);
}

// the conserved state before any flux term, from the cell's primitives
__device__ __forceinline__ UpdPart upd_init(const DevParams& P, double dens0, double vx, double vy,
                                            double vz, double ener, double eint, double pres) {
  UpdPart u;
  u.U0[0] = dens0;
  u.U0[1] = vx * dens0;
  u.U0[2] = vy * dens0;
  u.U0[3] = vz * dens0;
  u.U0[4] = dens0 * ener;
  u.IntEner = 0.;
  if (P.aux) u.IntEner = dens0 * eint;
  u.pres = pres;
  u.S1 = u.S2 = 0.;
  return u;
}

// x and y terms from the cell's primitive state and its four side faces
__device__ __forceinline__ UpdPart upd_start(const DevParams& P, double dt, double dx, double dy,
                                             double dens0, double vx, double vy, double vz,
                                             double ener, double eint, double pres,
                                             const double FLx[8], const double FRx[8],
                                             const double FLy[8], const double FRy[8]) {
  UpdPart u = upd_init(P, dens0, vx, vy, vz, ener, eint, pres);
  upd_dir(P, dt, dx, pres, FLx, FRx, u.U0, u.IntEner, u.S1);
  upd_dir(P, dt, dy, pres, FLy, FRy, u.U0, u.IntEner, u.S2);
  return u;
}

// z term, sources, energy fix and EOS; out = dens, velx, vely, velz, ener,
// eint, pres, temp, gamc, game (eintOld: the cell's e_int before the step,
// used when the aux equation is off)
__device__ __forceinline__ void upd_finish(const DevParams& P, double dt, double dz, UpdPart u,
                                           const double FLz[8], const double FRz[8],
                                           double eintOld, double out[10]) {
  double Sz;
  upd_dir(P, dt, dz, u.pres, FLz, FRz, u.U0, u.IntEner, Sz);
  if (P.aux) u.IntEner = u.IntEner + dt * 0.0;
  const double Sphys[5] = {0., u.S1, u.S2, Sz, 0.};
#pragma unroll
  for (int q = 0; q < 5; ++q) u.U0[q] = u.U0[q] + dt * (0. + Sphys[q]);
  const double densNew = fmax(u.U0[0], P.smlrho);
  const double vx = u.U0[1] / densNew, vy = u.U0[2] / densNew, vz = u.U0[3] / densNew;
  const double ener = u.U0[4] / densNew;
  const double eintAux = P.aux ? u.IntEner / densNew : eintOld;
  const double ekin = 0.5 * (vx * vx + vy * vy + vz * vz);
  const double eint = ener - ekin - 0.0;
  double newEint;
  if (!P.aux || eint > P.eintSwitch * (ekin + 0.0)) newEint = fmax(P.smallE, eint);
  else newEint = fmax(P.smallE, eintAux);
  const double enerFix = newEint + ekin;
  const EosOut eo = eos_dens_ei(P, densNew, vx, vy, vz, newEint, enerFix);
  out[0] = densNew;
  out[1] = vx;
  out[2] = vy;
  out[3] = vz;
  out[4] = eo.ener;
  out[5] = eo.eint;
  out[6] = eo.pres;
  out[7] = eo.temp;
  out[8] = P.gamma;
  out[9] = eo.game;
}

// U slots of upd_finish's outputs
__device__ __forceinline__ int upd_var(int v) {
  constexpr int slots[10] = {FX_DENS, FX_VELX, FX_VELY, FX_VELZ, FX_ENER, FX_EINT, FX_PRES,
                             FX_TEMP, FX_GAMC, FX_GAME};
  return slots[v];
}

// ---- the CFL dt of one block (Hydro_computeDt.F90:186-283) ----
// The reference's cell loop keeps a cell's (|v_d| + c) / del_d when its
// product with the CFL exceeds the kept one's product, so a block's value
// is that of the first cell in loop order (i fastest) whose product rounds
// to the largest product, and the block's dt is the CFL over that value.
// Of two cells whose products round alike the later one is never taken,
// even when its own value is larger; an exact maximum would differ from
// the reference in the last place of dt there.
struct DtPick {
  double prod, dtl;
  int idx;  // interior cell index in loop order; INT_MAX: none taken
};

__device__ __forceinline__ DtPick dt_pick_none() { return DtPick{0., 0., 0x7fffffff}; }

// one lane's cells come in increasing order: the strict test keeps the first
__device__ __forceinline__ void dt_pick_cell(DtPick& a, double dtl, double cfl, int idx) {
  const double prod = dtl * cfl;
  if (prod > a.prod) a = DtPick{prod, dtl, idx};
}

// the workgroup's pick (kThreads lanes): the largest product, then the
// lowest cell index; lane 0 returns the block's dt (1e300 when no cell had
// a positive value, as the reference's huge(1.0))
__device__ __forceinline__ double dt_pick_block(DtPick a, double cfl) {
  __shared__ double sp[kThreads], sd[kThreads];
  __shared__ int si[kThreads];
  const int t = threadIdx.x;
  sp[t] = a.prod;
  sd[t] = a.dtl;
  si[t] = a.idx;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (t < s && (sp[t + s] > sp[t] || (sp[t + s] == sp[t] && si[t + s] < si[t]))) {
      sp[t] = sp[t + s];
      sd[t] = sd[t + s];
      si[t] = si[t + s];
    }
    __syncthreads();
  }
  return sp[0] > 0. ? cfl / sd[0] : 1.e300;
}

}  // namespace fx
