// fx_device.hpp -- device-side math of the unsplit hydro block update.
//
// Reference semantics (paths relative to
// source/physics/Hydro/HydroMain/unsplit):
//   primitive eigen system      hy_eigenParameters.F90:42-89,
//                               hy_eigenValue.F90:33-50,
//                               OnlyHydro/hy_eigenVector.F90:46-151
//   slope limiters              hy_slopeLimiters.F90:38-60
//   TVD slopes                  hy_TVDslope.F90:54-265
//   transverse flux sigma       hy_upwindTransverseFlux.F90:43-211
//   PPM + characteristic trace  hy_DataReconstructNormalDir_PPM.F90:73-958
//   HLLC                        hy_HLLC.F90:53-324, hy_prim2con.F90,
//                               hy_prim2flx.F90
//
// The eigenvectors are sparse (each left/right vector has at most three
// non-zero entries); the products below keep only those entries, in the
// reference's summation order, so that every non-zero result rounds the
// same way as the dense 5x5 form.
//
// Primitive vectors are double[8]: dens, velx, vely, velz, pres, gamc, game,
// eint.  Waves: 0 fast-left, 1 slow-left, 2 entropy, 3 slow-right,
// 4 fast-right.
#pragma once
#include <hip/hip_runtime.h>

#include "fx_hydro.h"

namespace fx {

enum { P_DENS = 0, P_VELX = 1, P_VELY = 2, P_VELZ = 3, P_PRES = 4, P_GAMC = 5,
       P_GAME = 6, P_EINT = 7 };

// Pointers read from device memory (a tile's U, flux and pavg arrays) are
// generic, and accesses through generic pointers compile to flat
// instructions, which also count against the LDS wait counter.  The arrays
// are device memory, so the kernels hold them as global-address-space
// pointers (gdouble*) from the point they are read; loads and stores through
// them are global instructions, which also take a scalar base address.
typedef __attribute__((address_space(1))) double gdouble;
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

struct DevParams {
  int ndim, nxb, nyb, nzb, ng, nunk;
  int nx, ny, nz;            // guard-inclusive extents
  long ncell;                // cells per tile
  long voff[FX_NVAR_USED];   // var slot * ncell
  int lo[3], hi[3];          // interior limits, 0-based inclusive
  double gamma;
  int limiter;
  double beta;
  int charLim, flat, steep, ctu3d, aux, riemann, precision, learned;
  int velSlot[3];            // var slots of velx, vely, velz
  double fP, smlrho, smallE, eintSwitch, tiny;
  double gammam1, gam1inv, ggprodinv, abar;
  int exactMaxIter;
  double exactTol;
  long shokOff;              // SHOK_VAR slot * ncell, or -1 without shock detection
  int bext[3][3];            // [dir][axis] extent of the dir-band of state cells
  long bcells[3];            // cells per dir-band
  long boff[3];              // offset of each band in a tile's state scratch
  long scellsPerTile;        // sum of band cells
  int fext[3][3];            // [dir][axis] flux-array extents
  long nface[3];
  long icells;               // interior cells per tile
};

// Every routine below is templated on the real type R of the reconstruction:
// double in the FP64 modes, float in the mixed mode.  Literal constants are
// written R(c) so that the double instantiation performs exactly the FP64
// operations of the reference order.
template <typename R> struct EigT {
  R un, cf, ahinv, lfp, a2inv, rp, rho;
};
using Eig = EigT<double>;

template <int DIR, typename R>
__device__ __forceinline__ bool eig_make(const R* V, EigT<R>& e) {
  e.rho = V[P_DENS];
  e.un = V[P_VELX + DIR];
  const R a2 = V[P_GAMC] * V[P_PRES] / V[P_DENS];
  if (!(a2 > R(0.0))) return false;
  e.cf = sqrt(a2);
  const R dinv = R(1.0) / V[P_DENS];
  const R a2e = e.cf * e.cf;
  e.a2inv = R(1.0) / a2e;
  e.ahinv = R(0.5) / e.cf;
  const R Na = R(0.5) * e.a2inv;
  e.lfp = dinv * Na;
  e.rp = V[P_DENS] * a2e;
  return true;
}

template <typename R>
__device__ __forceinline__ void eig_lambda(const EigT<R>& e, R lam[5]) {
  lam[0] = e.un - e.cf;
  lam[1] = e.un;
  lam[2] = e.un;
  lam[3] = e.un;
  lam[4] = e.un + e.cf;
}

// velocity slots normal / first / second transverse to DIR
template <int DIR> struct Vel {
  static constexpr int v1 = 1 + DIR, v2 = 1 + (DIR + 1) % 3, v3 = 1 + (DIR + 2) % 3;
};

// L_n . d over the five primitive slots
template <int DIR, int N, typename R>
__device__ __forceinline__ R eig_left1(const EigT<R>& e, const R* d) {
  constexpr int v1 = Vel<DIR>::v1, v2 = Vel<DIR>::v2, v3 = Vel<DIR>::v3;
  if constexpr (N == 0) return (-e.ahinv) * d[v1] + e.lfp * d[P_PRES];
  else if constexpr (N == 1) return d[v2];
  else if constexpr (N == 2) return d[P_DENS] + (-e.a2inv) * d[P_PRES];
  else if constexpr (N == 3) return d[v3];
  else return e.ahinv * d[v1] + e.lfp * d[P_PRES];
}

template <int DIR, typename R>
__device__ __forceinline__ void eig_left(const EigT<R>& e, const R* d, R w[5]) {
  w[0] = eig_left1<DIR, 0>(e, d);
  w[1] = eig_left1<DIR, 1>(e, d);
  w[2] = eig_left1<DIR, 2>(e, d);
  w[3] = eig_left1<DIR, 3>(e, d);
  w[4] = eig_left1<DIR, 4>(e, d);
}

// out[0..4] = sum_n R_n * w[n]
template <int DIR, typename R>
__device__ __forceinline__ void eig_right(const EigT<R>& e, const R w[5], R* out) {
  constexpr int v1 = Vel<DIR>::v1, v2 = Vel<DIR>::v2, v3 = Vel<DIR>::v3;
  out[P_DENS] = e.rho * w[0] + w[2] + e.rho * w[4];
  out[v1] = (-e.cf) * w[0] + e.cf * w[4];
  out[v2] = w[1];
  out[v3] = w[3];
  out[P_PRES] = e.rp * w[0] + e.rp * w[4];
}

template <typename R> __device__ __forceinline__ R sgn1(R a) { return copysign(R(1.0), a); }

template <typename R>
__device__ __forceinline__ R limit(int kind, R a, R b, R beta) {
  switch (kind) {
    case FX_MINMOD: return R(.5) * (sgn1(a) + sgn1(b)) * fmin(fabs(a), fabs(b));
    case FX_VANLEER: return (a * b <= R(0.)) ? R(0.) : R(2.) * a * b / (a + b);
    case FX_LIMITED:
      if (b > R(0.)) return fmax(fmax(R(0.), fmin(beta * a, b)), fmin(a, beta * b));
      return fmin(fmin(R(0.), fmax(beta * a, b)), fmax(a, beta * b));
    default:  // FX_MC
      return (sgn1(a) + sgn1(b)) * fmin(fmin(fabs(a), R(.25) * fabs(a + b)), fabs(b));
  }
}

// Limited slope of the cell whose eigen system is e (hy_TVDslope), projected
// back to primitive slopes when characteristic limiting is on
// (PPM.F90:523-570).  out[0..7].
template <int DIR, typename R>
__device__ __forceinline__ void tvd_slope(const DevParams& P, const R* VL, const R* V0,
                                          const R* VR, const EigT<R>& e, R out[8]) {
  const R beta = R(P.beta);
  R dL[8], dR[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    dL[k] = V0[k] - VL[k];
    dR[k] = VR[k] - V0[k];
  }
  if (P.limiter == FX_HYBRID) {
    if (P.charLim) {
      R a[5], b[5], w[5];
      eig_left<DIR>(e, dL, a);
      eig_left<DIR>(e, dR, b);
#pragma unroll
      for (int n = 0; n < 5; ++n)
        w[n] = (n == 2) ? limit(FX_MC, a[n], b[n], beta) : limit(FX_MINMOD, a[n], b[n], beta);
      eig_right<DIR>(e, w, out);
      for (int k = P_GAMC; k <= P_EINT; ++k) out[k] = limit(FX_MC, dL[k], dR[k], beta);
    } else {
      out[P_DENS] = limit(FX_MC, dL[P_DENS], dR[P_DENS], beta);
      for (int k = P_VELX; k <= P_PRES; ++k) out[k] = limit(FX_MINMOD, dL[k], dR[k], beta);
      out[P_GAMC] = limit(FX_MC, dL[P_GAMC], dR[P_GAMC], beta);
      out[P_GAME] = limit(FX_MC, dL[P_GAME], dR[P_GAME], beta);
      out[P_EINT] = R(0.);
    }
    return;
  }
  if (P.charLim) {
    R a[5], b[5], w[5];
    eig_left<DIR>(e, dL, a);
    eig_left<DIR>(e, dR, b);
#pragma unroll
    for (int n = 0; n < 5; ++n) w[n] = limit(P.limiter, a[n], b[n], beta);
    eig_right<DIR>(e, w, out);
    for (int k = P_GAMC; k <= P_EINT; ++k) out[k] = limit(P.limiter, dL[k], dR[k], beta);
  } else {
    for (int k = 0; k < 8; ++k) out[k] = limit(P.limiter, dL[k], dR[k], beta);
  }
}

// Transverse flux sigma of one cell (hy_upwindTransverseFlux, order 1).
template <int DIR, typename R>
__device__ __forceinline__ void trans_sigma(const DevParams& P, const R* vm1, const R* vc0,
                                            const R* vp1, const EigT<R>& e, R sig[8]) {
  R lam[5];
  eig_lambda(e, lam);
  constexpr int v1 = Vel<DIR>::v1, v2 = Vel<DIR>::v2, v3 = Vel<DIR>::v3;
#pragma unroll
  for (int v = 0; v < 8; ++v) sig[v] = R(0.);
  R dm[5], dp[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    dm[v] = vc0[v] - vm1[v];
    dp[v] = vp1[v] - vc0[v];
  }
  // upwinded characteristic amplitude of wave N
#define FX_TS_DB(N) (lam[N] > R(0.) ? eig_left1<DIR, N>(e, dm) : eig_left1<DIR, N>(e, dp))
  {
    const R db = FX_TS_DB(0);
    sig[P_DENS] = sig[P_DENS] + lam[0] * e.rho * db;
    sig[v1] = sig[v1] + lam[0] * (-e.cf) * db;
    sig[P_PRES] = sig[P_PRES] + lam[0] * e.rp * db;
  }
  { const R db = FX_TS_DB(1); sig[v2] = sig[v2] + lam[1] * db; }
  { const R db = FX_TS_DB(2); sig[P_DENS] = sig[P_DENS] + lam[2] * db; }
  { const R db = FX_TS_DB(3); sig[v3] = sig[v3] + lam[3] * db; }
  {
    const R db = FX_TS_DB(4);
    sig[P_DENS] = sig[P_DENS] + lam[4] * e.rho * db;
    sig[v1] = sig[v1] + lam[4] * e.cf * db;
    sig[P_PRES] = sig[P_PRES] + lam[4] * e.rp * db;
  }
#undef FX_TS_DB
  if (P.aux) {
    const R d = (lam[2] > R(0.)) ? vc0[P_EINT] - vm1[P_EINT] : vp1[P_EINT] - vc0[P_EINT];
    sig[P_EINT] = lam[2] * d;
  }
}

// x / C for C = 3, 6, 12 with the IEEE quotient's bits, in four FP64
// instructions instead of the division sequence (a quarter-rate reciprocal
// and about ten instructions).  q = x RN(1/C) is faithful, so rho = x - q C
// is exact in one FMA; q + rho RN(1/C), rounded once by a second FMA, lies
// within 2^-52 ulp of x/C.  As C = 3 * 2^k, a normal quotient x/C sits 0,
// 1/3 or 2/3 of an ulp past a double, never within 1/6 ulp of a rounding
// midpoint, so both round alike; copysign keeps the sign of a zero.  This
// holds for every finite x whose quotient is normal or zero (checked against
// IEEE division by tests/cdiv/const_div_check.c); for C = 6 and 12 a
// quotient below 2^-1022 may differ in its last place, and an infinite x
// gives NaN (fma(-inf, C, inf)) where IEEE division gives inf.  Those are the
// limits of the bitwise-parity claim: states whose PPM differences are
// subnormal or infinite.  FP32 divides.
template <int C>
__device__ __forceinline__ double div_const(double x) {
  static_assert(C == 3 || C == 6 || C == 12, "constants of the form 3 * 2^k");
  constexpr double r = 1.0 / C;
  const double q = x * r;
  const double rho = fma(-q, double(C), x);
  return copysign(fma(rho, r, q), x);
}
template <int C>
__device__ __forceinline__ float div_const(float x) {
  return x / float(C);
}

// PPM interface values, limiting and characteristic tracing from the three
// projected slopes of cells i-1, i, i+1 (PPM.F90:578-842).  Vm/Vc/Vp are the
// primitive vectors of i-1, i, i+1; densMM/densPP the densities of i-2, i+2
// (contact steepening only).
template <int DIR, typename R>
__device__ __forceinline__ void ppm_trace(const DevParams& P, R dtn, R delta, const R* Vm,
                                          const R* Vc, const R* Vp, R densMM, R densPP,
                                          const R* dbN, const R* db0, const R* dbP, R flatC,
                                          const EigT<R>& e0, R Wp[8], R Wm[8]) {
  R Vmm[1] = {densMM}, Vpp[1] = {densPP};
  const int endVar = P.aux ? P_EINT : P_PRES;
  R vL[8], vR[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    vL[v] = R(0.);
    vR[v] = R(0.);
    if (v <= endVar) {
      vL[v] = R(0.5) * (Vc[v] + Vm[v]) - div_const<6>(db0[v] - dbN[v]);
      vR[v] = R(0.5) * (Vc[v] + Vp[v]) - div_const<6>(dbP[v] - db0[v]);
    }
  }
  if (P.steep) {
    const R temp1 = Vp[P_DENS] - Vm[P_DENS];
    if (fabs(temp1) > R(P.tiny)) {
      const R d2R = (Vpp[0] - R(2.) * Vp[P_DENS] + Vc[P_DENS]) / (R(6.) * delta * delta);
      const R d2L = (Vc[P_DENS] - R(2.) * Vm[P_DENS] + Vmm[0]) / (R(6.) * delta * delta);
      R eta = (d2L - d2R) * (delta * delta) / temp1;
      if (d2R * d2L >= R(0.)) eta = R(0.);
      if (R(0.01) * fmin(Vp[P_DENS], Vm[P_DENS]) - fabs(Vp[P_DENS] - Vm[P_DENS]) >= R(0.)) eta = R(0.);
      eta = fmax(R(0.), fmin(R(1.), R(20.0) * (eta - R(0.05))));
      const R t2 = fabs(Vp[P_PRES] - Vm[P_PRES]) / fmin(Vp[P_PRES], Vm[P_PRES]);
      const R t3 = fabs(Vp[P_DENS] - Vm[P_DENS]) / fmin(Vp[P_DENS], Vm[P_DENS]);
      if (Vc[P_GAME] * R(0.1) * t3 - t2 < R(0.0)) eta = R(0.);
      vL[P_DENS] = vL[P_DENS] * (R(1.) - eta) + (Vm[P_DENS] + R(0.5) * dbN[P_DENS]) * eta;
      vR[P_DENS] = vR[P_DENS] * (R(1.) - eta) + (Vp[P_DENS] - R(0.5) * dbP[P_DENS]) * eta;
    }
  }
  R dW[8], W6[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    dW[n] = R(0.);
    W6[n] = R(0.);
    if (n > endVar) continue;
    if (P.flat) {
      vL[n] = flatC * Vc[n] + (R(1.0) - flatC) * vL[n];
      vR[n] = flatC * Vc[n] + (R(1.0) - flatC) * vR[n];
    }
    if ((vR[n] - Vc[n]) * (Vc[n] - vL[n]) <= R(0.)) {
      vL[n] = Vc[n];
      vR[n] = Vc[n];
    }
    R dd = vR[n] - vL[n];
    if (R(6.) * (vR[n] - vL[n]) * (Vc[n] - R(0.5) * (vL[n] + vR[n])) > dd * dd)
      vL[n] = R(3.) * Vc[n] - R(2.) * vR[n];
    dd = vR[n] - vL[n];
    if (R(6.) * (vR[n] - vL[n]) * (Vc[n] - R(0.5) * (vL[n] + vR[n])) < -(dd * dd))
      vR[n] = R(3.) * Vc[n] - R(2.) * vL[n];
    dW[n] = vR[n] - vL[n];
    W6[n] = R(6.) * (Vc[n] - R(0.5) * (vR[n] + vL[n]));
  }
  R lam[5];
  eig_lambda(e0, lam);
  const R hdtn = R(0.5) * dtn;
  const R four3 = R(4.) / R(3.);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    Wp[v] = R(0.);
    Wm[v] = R(0.);
  }
  {
    const R lmax = fmax(lam[4], R(0.)), lmin = -fmin(lam[0], R(0.));
    const R emax = fmax(lam[2], R(0.)), emin = -fmin(lam[2], R(0.));
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      if (v > endVar) continue;
      const R a = v <= P_PRES ? lmax : emax;
      const R b = v <= P_PRES ? lmin : emin;
      Wp[v] = vR[v] - a * hdtn * (dW[v] - (R(1.0) - div_const<3>(a * hdtn * R(4.))) * W6[v]);
      Wm[v] = vL[v] + b * hdtn * (dW[v] + (R(1.0) - div_const<3>(b * hdtn * R(4.))) * W6[v]);
    }
  }
  Wp[P_GAMC] = Vc[P_GAMC];
  Wp[P_GAME] = Vc[P_GAME];
  Wm[P_GAMC] = Vc[P_GAMC];
  Wm[P_GAME] = Vc[P_GAME];
  // characteristic tracing (PPM.F90:751-800), HLL-type branch for all
  // solvers but Roe
  R sL[5] = {R(0.), R(0.), R(0.), R(0.), R(0.)}, sR[5] = {R(0.), R(0.), R(0.), R(0.), R(0.)};
  R mW6[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) mW6[v] = -W6[v];
  constexpr int v1 = Vel<DIR>::v1, v2 = Vel<DIR>::v2, v3 = Vel<DIR>::v3;
  const bool roe = (P.riemann == FX_ROE);
  R cA[5], cB[5];
  eig_left<DIR>(e0, dW, cA);
  eig_left<DIR>(e0, mW6, cB);
  // contribution of wave N through its right-eigenvector entry RV at slot V
#define FX_TRACE(N, V, RV)                                                              \
  {                                                                                     \
    const R x = dtn * lam[N];                                                           \
    if (!roe || lam[N] < R(0.))                                                         \
      sL[V] = sL[V] + (R(.5) * (R(-1.) - x) * (RV) * cA[N] +                            \
                       R(.25) * (R(1.) + R(2.) * dtn * lam[N] + four3 * (x * x)) * (RV) * cB[N]); \
    if (!roe || lam[N] > R(0.))                                                         \
      sR[V] = sR[V] + (R(.5) * (R(1.) - x) * (RV) * cA[N] +                             \
                       R(.25) * (R(1.) - R(2.) * dtn * lam[N] + four3 * (x * x)) * (RV) * cB[N]); \
  }
  FX_TRACE(0, P_DENS, e0.rho)
  FX_TRACE(0, v1, -e0.cf)
  FX_TRACE(0, P_PRES, e0.rp)
  FX_TRACE(1, v2, R(1.))
  FX_TRACE(2, P_DENS, R(1.))
  FX_TRACE(3, v3, R(1.))
  FX_TRACE(4, P_DENS, e0.rho)
  FX_TRACE(4, v1, e0.cf)
  FX_TRACE(4, P_PRES, e0.rp)
#undef FX_TRACE
#pragma unroll
  for (int v = 0; v <= P_PRES; ++v) {
    Wm[v] = Vc[v] + div_const<12>(W6[v]) + sL[v];
    Wp[v] = Vc[v] + div_const<12>(W6[v]) + sR[v];
  }
}

template <int DIR, typename R>
__device__ __forceinline__ bool ppm_full(const DevParams& P, R dtn, R delta, const R* Vmm,
                                         const R* Vm, const R* Vc, const R* Vp, const R* Vpp,
                                         R flatC, const EigT<R>& e0, R Wp[8], R Wm[8]) {
  EigT<R> eP, eN;
  if (!eig_make<DIR>(Vp, eP)) return false;
  if (!eig_make<DIR>(Vm, eN)) return false;
  R dbN[8], db0[8], dbP[8];
  tvd_slope<DIR>(P, Vmm, Vm, Vc, eN, dbN);
  tvd_slope<DIR>(P, Vm, Vc, Vp, e0, db0);
  tvd_slope<DIR>(P, Vc, Vp, Vpp, eP, dbP);
  ppm_trace<DIR>(P, dtn, delta, Vm, Vc, Vp, Vmm[P_DENS], Vpp[P_DENS], dbN, db0, dbP, flatC, e0, Wp,
                 Wm);
  return true;
}

// ---- HLLC (hy_HLLC.F90:53-324) ----
template <int DIR, typename R>
__device__ __forceinline__ void prim2flx(const R* V, R fP, R F[6]) {
  const R u2 = V[P_VELX] * V[P_VELX] + V[P_VELY] * V[P_VELY] + V[P_VELZ] * V[P_VELZ];
  const R E = R(0.5) * V[P_DENS] * u2 + V[P_PRES] / (V[P_GAME] - R(1.));
  const R un = V[P_VELX + DIR];
  F[5] = V[P_PRES];
  F[0] = V[P_DENS] * un;
  F[1] = F[0] * V[P_VELX];
  F[2] = F[0] * V[P_VELY];
  F[3] = F[0] * V[P_VELZ];
  F[1 + DIR] = F[1 + DIR] + fP * V[P_PRES];
  F[4] = (E + V[P_PRES]) * un;
}

// returns 0 ok, 1 imaginary sound speed
template <int DIR, typename R>
__device__ __forceinline__ int hllc(const DevParams& P, const R* Vm, const R* Vp, R F[6]) {
  const R tiny = R(P.tiny), fP = R(P.fP);
  const R velNL = Vm[P_VELX + DIR], velNR = Vp[P_VELX + DIR];
  const R aL2 = Vm[P_GAMC] * Vm[P_PRES] / Vm[P_DENS];
  const R aR2 = Vp[P_GAMC] * Vp[P_PRES] / Vp[P_DENS];
  if ((Vm[P_DENS] < tiny && Vm[P_DENS] > R(0.)) || (Vp[P_DENS] < tiny && Vp[P_DENS] > R(0.)) ||
      (Vm[P_PRES] < tiny && Vm[P_PRES] > R(0.)) || (Vp[P_PRES] < tiny && Vp[P_PRES] > R(0.))) {
#pragma unroll
    for (int f = 0; f < 6; ++f) F[f] = R(0.);
    return 0;
  }
  if (aL2 < R(0.) || aR2 < R(0.)) return 1;
  const R cfL = sqrt(aL2), cfR = sqrt(aR2);
  const R SL = fmin(velNL - cfL, velNR - cfR);
  const R SR = fmax(velNL + cfL, velNR + cfR);
  const R tpL = Vm[P_PRES], tpR = Vp[P_PRES];
  R qStar = (Vp[P_DENS] * velNR * (SR - velNR) - Vm[P_DENS] * velNL * (SL - velNL) + tpL - tpR -
             R(0.) * R(0.) + R(0.) * R(0.));
  qStar = qStar / (Vp[P_DENS] * (SR - velNR) - Vm[P_DENS] * (SL - velNL));
  if (SL >= R(0.)) {
    prim2flx<DIR>(Vm, fP, F);
    return 0;
  }
  if (!(qStar < R(0.) && SR >= R(0.)) && !(SL < R(0.) && qStar >= R(0.))) {
    prim2flx<DIR>(Vp, fP, F);
    return 0;
  }
  const bool left = (SL < R(0.) && qStar >= R(0.));
  // element-wise select keeps both inputs in registers (a pointer select
  // between two private arrays would put them in scratch memory)
  R V[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) V[v] = left ? Vm[v] : Vp[v];
  const R S = left ? SL : SR;
  const R velN = left ? velNL : velNR;
  const R tp = left ? tpL : tpR;
  const R numer = S - velN, denom = S - qStar;
  const R pStar = Vm[P_DENS] * (SL - velNL) * (qStar - velNL) + tpL;
  // the chosen side's flux, built in F itself (a second private array here
  // ends up in scratch memory after inlining)
  prim2flx<DIR>(V, fP, F);
  // conserved state of the chosen side (hy_prim2con.F90)
  const R u2 = V[P_VELX] * V[P_VELX] + V[P_VELY] * V[P_VELY] + V[P_VELZ] * V[P_VELZ];
  R Uc[5];
  Uc[0] = V[P_DENS];
  Uc[1] = V[P_DENS] * V[P_VELX];
  Uc[2] = V[P_DENS] * V[P_VELY];
  Uc[3] = V[P_DENS] * V[P_VELZ];
  Uc[4] = R(0.5) * V[P_DENS] * u2 + V[P_PRES] / (V[P_GAME] - R(1.)) + R(0.5) * R(0.);
  R Us[5];
  const R dStar = Uc[0] * numer / denom;
  Us[0] = dStar;
  Us[4] = Uc[4] * numer / denom + ((pStar * qStar - tp * velN)) / denom;
#pragma unroll
  for (int m = 0; m < 3; ++m) Us[1 + m] = (m == DIR) ? dStar * qStar : Uc[1 + m] * numer / denom;
#pragma unroll
  for (int f = 0; f < 5; ++f) F[f] = F[f] + S * (Us[f] - Uc[f]);
  return 0;
}

// ---- HLL (hy_HLL.F90:82-165) ----
// conserved state (hy_prim2con.F90): dens, three momenta, total energy
template <typename R>
__device__ __forceinline__ void prim2con(const R* V, R Uc[5]) {
  const R u2 = V[P_VELX] * V[P_VELX] + V[P_VELY] * V[P_VELY] + V[P_VELZ] * V[P_VELZ];
  Uc[0] = V[P_DENS];
  Uc[1] = V[P_DENS] * V[P_VELX];
  Uc[2] = V[P_DENS] * V[P_VELY];
  Uc[3] = V[P_DENS] * V[P_VELZ];
  Uc[4] = R(0.5) * V[P_DENS] * u2 + V[P_PRES] / (V[P_GAME] - R(1.)) + R(0.5) * R(0.);
}

// returns 0 ok, 1 imaginary sound speed.  The fan average runs over all six
// slots; the pressure slot's state jump is zero (hy_HLL.F90:146).
template <int DIR, typename R>
__device__ __forceinline__ int hll(const DevParams& P, const R* Vm, const R* Vp, R F[6]) {
  const R tiny = R(P.tiny), fP = R(P.fP);
  const R velNL = Vm[P_VELX + DIR], velNR = Vp[P_VELX + DIR];
  const R aL2 = Vm[P_GAMC] * Vm[P_PRES] / Vm[P_DENS];
  const R aR2 = Vp[P_GAMC] * Vp[P_PRES] / Vp[P_DENS];
  if ((Vm[P_DENS] < tiny && Vm[P_DENS] > R(0.)) || (Vp[P_DENS] < tiny && Vp[P_DENS] > R(0.)) ||
      (Vm[P_PRES] < tiny && Vm[P_PRES] > R(0.)) || (Vp[P_PRES] < tiny && Vp[P_PRES] > R(0.))) {
#pragma unroll
    for (int f = 0; f < 6; ++f) F[f] = R(0.);
    return 0;
  }
  if (aL2 < R(0.) || aR2 < R(0.)) return 1;
  const R cfL = sqrt(aL2), cfR = sqrt(aR2);
  const R SL = fmin(velNL - cfL, velNR - cfR);
  const R SR = fmax(velNL + cfL, velNR + cfR);
  if (SL > R(0.)) {
    prim2flx<DIR>(Vm, fP, F);
    return 0;
  }
  if (SR < R(0.)) {
    prim2flx<DIR>(Vp, fP, F);
    return 0;
  }
  R FL[6], FR[6], UL[5], UR[5];
  prim2flx<DIR>(Vm, fP, FL);
  prim2flx<DIR>(Vp, fP, FR);
  prim2con(Vm, UL);
  prim2con(Vp, UR);
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const R dU = f < 5 ? UR[f] - UL[f] : R(0.0) - R(0.0);
    F[f] = (SR * FL[f] - SL * FR[f] + SR * SL * dU) / (SR - SL);
  }
  return 0;
}

// ---- exact Riemann solver for the gamma-law gas (no counterpart in the
// reference; it takes hy_HLLC.F90's contract).  Newton iteration on the
// exact pressure function (Toro, Riemann Solvers and Numerical Methods, 3rd
// ed., eqs. 4.6-4.9, derivative 4.37), seeded by a caller guess or by the
// two-shock guess (Toro eq. 9.42), sampled at x/t = 0 (Toro sec. 4.5).  The
// operation order follows the CPU oracle's restatement, except where a power
// follows from one already taken: the rarefaction derivative reuses the
// value's power, the density behind a rarefaction is gamma p / c^2, and the
// fan pressure's power is the density's times (c/ck)^2; and powers are taken
// as exp2(e log2 x) (pow_pos) where the oracle calls pow.

// R = double is the FP64 solver.  R = float serves the mixed-precision
// mode: the pressure function, sampling and flux run in FP32, while the
// Newton iterate and its convergence test stay in FP64 (the iterate is
// rounded to FP32 only to evaluate the function).  The R = double
// instantiation performs the FP64 operations in the order described above.

// x^e for x > 0 as exp2(e log2 x).  In FP64 this is about 130 instructions
// against pow's 230 (pow also handles signs, integer exponents and extra
// precision that positive pressure and sound-speed ratios do not need); its
// error, a few ulp, is far below the solver's 1e-12 convergence tolerance.
// In FP32 it is the hardware log2 and exp2.  x = 0 gives 0, as pow does for
// e > 0.
__device__ __forceinline__ double pow_pos(double x, double e) { return exp2(e * log2(x)); }
__device__ __forceinline__ float pow_pos(float x, float e) { return exp2f(e * log2f(x)); }

// Tolerance floor of the FP32 solve.  FP32 evaluations of the pressure
// function leave rounding noise of up to a few 1e-7 in p*, where Newton's
// relative change can stall; a change of 1e-5 already puts the next iterate
// (quadratic convergence) below FP32 resolution.
constexpr double kExactTolF32 = 1.e-5;

template <typename R>
__device__ __forceinline__ void exact_fk(R p, R rk, R pk, R ck, R g, R& f, R& fd) {
  if (p <= pk) {
    // rarefaction: the derivative's power is the value's divided by pratio,
    // since (g - 1)/(2g) + (g + 1)/(2g) = 1 (one pow instead of two)
    const R pratio = p / pk;
    const R q = pow_pos(pratio, (g - R(1.)) / (R(2.) * g));
    f = R(2.) / (g - R(1.)) * ck * (q - R(1.));
    fd = (R(1.) / (rk * ck)) * (q / pratio);
  } else {
    const R ak = R(2.) / ((g + R(1.)) * rk);
    const R bk = (g - R(1.)) / (g + R(1.)) * pk;
    const R qrt = sqrt(ak / (bk + p));
    f = (p - pk) * qrt;
    fd = (R(1.) - R(0.5) * (p - pk) / (bk + p)) * qrt;
  }
}

template <typename R>
__device__ __forceinline__ R two_shock_guess(R g, R rl, R ul, R pl, R cl, R rr, R ur, R pr, R cr) {
  const R ppv = R(0.5) * (pl + pr) - R(0.125) * (ur - ul) * (rl + rr) * (cl + cr);
  const R p0 = fmax(R(1.e-30), ppv);
  const R gl = sqrt((R(2.) / ((g + R(1.)) * rl)) / ((g - R(1.)) / (g + R(1.)) * pl + p0));
  const R gr = sqrt((R(2.) / ((g + R(1.)) * rr)) / ((g - R(1.)) / (g + R(1.)) * pr + p0));
  const R pts = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
  return fmax(R(1.e-30), pts);
}

// Newton from the better of two starts: the pressure function is evaluated
// at both, and the iteration begins at the one with the smaller residual
// |f(p)| (its f and f' serve the first step).  pb <= 0 means no second
// start: plain Newton from pa, in the CPU oracle's operation order.
// Returns iterations, negative when maxit was reached; pick = 0 when pa was
// taken.  The iterate p is FP64 whatever R is.
template <typename R>
__device__ __forceinline__ int exact_newton_pick(R g, R rl, R pl, R cl, R rr, R pr, R cr, R du,
                                                 double pa, double pb, double tol, int maxit,
                                                 double& pstar, int& pick) {
  R fl, fld, fr, frd;
  exact_fk(R(pa), rl, pl, cl, g, fl, fld);
  exact_fk(R(pa), rr, pr, cr, g, fr, frd);
  double p = pa;
  R f = fl + fr + du, fd = fld + frd;
  pick = 0;
  if (pb > 0.) {
    R gl, gld, gr, grd;
    exact_fk(R(pb), rl, pl, cl, g, gl, gld);
    exact_fk(R(pb), rr, pr, cr, g, gr, grd);
    const R fb = gl + gr + du;
    if (!(fabs(f) <= fabs(fb))) {
      pick = 1;
      p = pb;
      f = fb;
      fd = gld + grd;
    }
  }
  for (int it = 1; it <= maxit; ++it) {
    if (it > 1) {
      exact_fk(R(p), rl, pl, cl, g, fl, fld);
      exact_fk(R(p), rr, pr, cr, g, fr, frd);
      f = fl + fr + du;
      fd = fld + frd;
    }
    double pn = p - double(f / fd);
    if (!(pn > 0.)) pn = 1.e-30;
    const double change = 2. * fabs((pn - p) / (pn + p));
    p = pn;
    if (change <= tol) {
      pstar = p;
      return it;
    }
  }
  pstar = p;
  return -maxit;
}

struct ExactStats {
  unsigned long long iters, fallbacks;
};

// state at x/t = 0 (Toro sec. 4.5); leftside tells which input supplies the
// tangential velocities and gammas
template <typename R>
__device__ __forceinline__ void exact_sample(R g, R rl, R ul, R pl, R cl, R rr, R ur, R pr, R cr,
                                             R pm, R um, R& r, R& u, R& pp, bool& leftside) {
  if (R(0.) <= um) {
    leftside = true;
    if (pm <= pl) {
      const R shl = ul - cl;
      if (R(0.) <= shl) { r = rl; u = ul; pp = pl; return; }
      const R cml = cl * pow_pos(pm / pl, (g - R(1.)) / (R(2.) * g));
      const R stl = um - cml;
      // isentropic: rho = gamma p / c^2 behind the rarefaction
      if (R(0.) > stl) { r = g * pm / (cml * cml); u = um; pp = pm; return; }
      const R c = R(2.) / (g + R(1.)) * (cl + (g - R(1.)) / R(2.) * (ul - R(0.)));
      u = R(2.) / (g + R(1.)) * (cl + (g - R(1.)) / R(2.) * ul + R(0.));
      // in the fan, (c/cl)^(2g/(g-1)) = (c/cl)^(2/(g-1)) (c/cl)^2
      const R q = pow_pos(c / cl, R(2.) / (g - R(1.))), s2 = (c / cl) * (c / cl);
      r = rl * q;
      pp = pl * (q * s2);
      return;
    }
    const R pml = pm / pl;
    const R sl = ul - cl * sqrt((g + R(1.)) / (R(2.) * g) * pml + (g - R(1.)) / (R(2.) * g));
    if (R(0.) <= sl) { r = rl; u = ul; pp = pl; return; }
    r = rl * (pml + (g - R(1.)) / (g + R(1.))) / (pml * (g - R(1.)) / (g + R(1.)) + R(1.));
    u = um; pp = pm;
    return;
  }
  leftside = false;
  if (pm > pr) {
    const R pmr = pm / pr;
    const R sr = ur + cr * sqrt((g + R(1.)) / (R(2.) * g) * pmr + (g - R(1.)) / (R(2.) * g));
    if (R(0.) >= sr) { r = rr; u = ur; pp = pr; return; }
    r = rr * (pmr + (g - R(1.)) / (g + R(1.))) / (pmr * (g - R(1.)) / (g + R(1.)) + R(1.));
    u = um; pp = pm;
    return;
  }
  const R shr = ur + cr;
  if (R(0.) >= shr) { r = rr; u = ur; pp = pr; return; }
  const R cmr = cr * pow_pos(pm / pr, (g - R(1.)) / (R(2.) * g));
  const R str = um + cmr;
  if (R(0.) <= str) { r = g * pm / (cmr * cmr); u = um; pp = pm; return; }
  const R c = R(2.) / (g + R(1.)) * (cr - (g - R(1.)) / R(2.) * (ur - R(0.)));
  u = R(2.) / (g + R(1.)) * (-cr + (g - R(1.)) / R(2.) * ur + R(0.));
  const R q = pow_pos(c / cr, R(2.) / (g - R(1.))), s2 = (c / cr) * (c / cr);
  r = rr * q;
  pp = pr * (q * s2);
}

// Exact solution at x/t = 0 of the face problem (rl, ul, pl | rr, ur, pr)
// with the face-normal velocities ul, ur; returns 0 ok, 1 non-positive state,
// 2 vacuum generated.  guess > 0 seeds Newton: it starts from the better of
// the seed and the two-shock guess, and a seeded start that does not
// converge restarts from the two-shock guess (counted as fallback).  One
// Newton loop serves both passes, which keeps a seeded kernel's register
// budget that of the unseeded one.  SEEDED = false ignores guess.  The FP32
// solve (R = float) converges to max(exactTol, kExactTolF32), and a seed is
// held at 1e-30 or above, inside FP32's normal range.
template <bool SEEDED, typename R>
__device__ __forceinline__ int exact_riemann(const DevParams& P, R g, R rl, R ul, R pl, R rr, R ur,
                                             R pr, double guess, ExactStats& es, R& r, R& u,
                                             R& pp, bool& left) {
  constexpr bool kF32 = sizeof(R) == sizeof(float);
  if (!(rl > R(0.) && rr > R(0.) && pl > R(0.) && pr > R(0.))) return 1;
  const R cl = sqrt(g * pl / rl), cr = sqrt(g * pr / rr);
  if (R(2.) / (g - R(1.)) * (cl + cr) <= ur - ul) return 2;
  const R du = ur - ul;
  const R ts = two_shock_guess(g, rl, ul, pl, cl, rr, ur, pr, cr);
  const double tol = kF32 ? fmax(P.exactTol, kExactTolF32) : P.exactTol;
  double pm;
  int it, pick;
  if constexpr (!SEEDED) {
    it = exact_newton_pick(g, rl, pl, cl, rr, pr, cr, du, double(ts), -1.0, tol, P.exactMaxIter,
                           pm, pick);
    if (it < 0) es.fallbacks += 1;
  } else {
    double pb = (kF32 && guess > 0.) ? fmax(guess, 1.e-30) : guess;
    for (;;) {
      it = exact_newton_pick(g, rl, pl, cl, rr, pr, cr, du, double(ts), pb, tol, P.exactMaxIter,
                             pm, pick);
      if (it < 0 && (pb > 0. || guess <= 0.)) es.fallbacks += 1;   // a restart counts once
      if (it >= 0 || pick == 0) break;
      es.iters += (unsigned long long)(-it);
      pb = -1.0;
    }
  }
  es.iters += (unsigned long long)(it < 0 ? -it : it);
  const R pmr = R(pm);
  R fl, fld, fr, frd;
  exact_fk(pmr, rl, pl, cl, g, fl, fld);
  exact_fk(pmr, rr, pr, cr, g, fr, frd);
  const R um = R(0.5) * (ul + ur) + R(0.5) * (fr - fl);
  exact_sample(g, rl, ul, pl, cl, rr, ur, pr, cr, pmr, um, r, u, pp, left);
  return 0;
}

// prim2flx with the direction chosen at run time (same operations and
// order as prim2flx<DIR>)
template <typename R>
__device__ __forceinline__ void prim2flx_dir(int d, R r, R vx, R vy, R vz, R p, R game, R fP,
                                             R F[6]) {
  const R u2 = vx * vx + vy * vy + vz * vz;
  const R E = R(0.5) * r * u2 + p / (game - R(1.));
  const R un = d == 0 ? vx : (d == 1 ? vy : vz);
  F[5] = p;
  F[0] = r * un;
  F[1] = F[0] * vx;
  F[2] = F[0] * vy;
  F[3] = F[0] * vz;
  const R fp = fP * p;
  if (d == 0) F[1] = F[1] + fp;
  else if (d == 1) F[2] = F[2] + fp;
  else F[3] = F[3] + fp;
  F[4] = (E + p) * un;
}

// Godunov flux F[0..7] across a face normal to direction d from the exact
// solution, with the aux internal-energy fluxes (hy_getFaceFlux.F90:428-442).
// The face states sit in global memory as components 0..5 = dens, velx,
// vely, velz, pres, eint at Lb[q * s] (left) and Rb[q * s] (right); reading
// them with the run-time offset d keeps them out of private arrays, and one
// copy of the solver serves all directions.  gamc is the left cell's, game
// the cells' on either side.  The states and the solve are in R; F is FP64
// for the update.  Returns exact_riemann's codes.
template <bool SEEDED, typename R>
__device__ __forceinline__ int exact_face_flux(const DevParams& P, int d, const R* Lb,
                                               const R* Rb, long s, double gamc,
                                               double gameL, double gameR, double guess,
                                               double F[8], ExactStats& es) {
  const R rl = Lb[0], pl = Lb[4 * s], rr = Rb[0], pr = Rb[4 * s];
  const R ul = Lb[(1 + d) * s], ur = Rb[(1 + d) * s];
  R r, u, pp;
  bool left;
  const int e = exact_riemann<SEEDED>(P, R(gamc), rl, ul, pl, rr, ur, pr, guess, es, r, u, pp,
                                      left);
  if (e) return e;
  // the sampled state takes its tangential velocities and game from the side
  // of the contact the face lies on
  const R* Sb = left ? Lb : Rb;
  R vx = Sb[s], vy = Sb[2 * s], vz = Sb[3 * s];
  if (d == 0) vx = u;
  else if (d == 1) vy = u;
  else vz = u;
  R Fr[6];
  prim2flx_dir(d, r, vx, vy, vz, pp, R(left ? gameL : gameR), R(P.fP), Fr);
#pragma unroll
  for (int q = 0; q < 6; ++q) F[q] = double(Fr[q]);
  F[6] = 0.;
  F[7] = 0.;
  if (P.aux) {
    if (Fr[0] > R(0.)) {
      F[7] = double(Fr[0] / rl);
      F[6] = double(Fr[0] * Lb[5 * s]);
    } else {
      F[7] = double(Fr[0] / rr);
      F[6] = double(Fr[0] * Rb[5 * s]);
    }
  }
  return 0;
}

// Gate of the learned seed: an interface is strong when its largest
// feature magnitude, max(|ln(rl/rr)|, |ln(pl/pr)|, |ur - ul| / c0) / 2 with
// c0^2 = gamma sqrt(pl pr) / sqrt(rl rr), reaches the gate.  The same test
// without logarithms or roots: with T = exp(2 gate) and K = 16 gate^4,
//   rl >= T rr  or  rr >= T rl  or  pl >= T pr  or  pr >= T pl
//   or  (ur - ul)^4 rl rr >= K gamma^2 pl pr.
struct SeedGate {
  double T, K;
};

__host__ __device__ inline SeedGate make_seed_gate(double gate) {
  const double g2 = gate * gate;
  return SeedGate{exp(2.0 * gate), 16.0 * g2 * g2};
}

__device__ __forceinline__ bool seed_gate_pass(const SeedGate& G, double g, double rl, double ul,
                                               double pl, double rr, double ur, double pr) {
  const double du = ur - ul, du2 = du * du;
  return rl >= G.T * rr || rr >= G.T * rl || pl >= G.T * pr || pr >= G.T * pl ||
         du2 * du2 * (rl * rr) >= G.K * (g * g) * (pl * pr);
}

// seed value of a face whose Newton iteration keeps the two-shock start
constexpr float kNoSeed = 3.0e38f;

// ---- gamma-law EOS, MODE_DENS_EI on one cell (Eos_getData.F90:189-215,
// eos_idealGamma.F90:323-333, Eos_putData.F90:197-210) ----
struct EosOut {
  double pres, temp, eint, ener, game;
};
__device__ __forceinline__ EosOut eos_dens_ei(const DevParams& P, double dens, double vx, double vy,
                                              double vz, double eintIn, double enerIn) {
  const double ekin = 0.5 * (vx * vx + vy * vy + vz * vz);
  double ie = eintIn;
  if (enerIn - ekin > fmax(P.smallE, P.eintSwitch * ekin)) ie = enerIn - ekin;
  ie = fmax(ie, P.smallE);
  EosOut o;
  o.eint = ie;
  o.pres = dens * ie * P.gam1inv;
  o.temp = ie * P.ggprodinv * P.abar;
  o.ener = ie + ekin;
  o.game = o.pres / (ie * dens) + 1;
  return o;
}

}  // namespace fx
