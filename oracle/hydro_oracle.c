/*
 * hydro_oracle.c -- TEST INFRASTRUCTURE: CPU FP64 restatement of FLASH5's
 * unsplit hydro block update, used only as the checker for the HIP product.
 *
 * Reference files (relative to source/physics/Hydro/HydroMain/unsplit):
 *   hy_computeFluxes.F90, hy_getRiemannState.F90, hy_dataReconstOneStep.F90,
 *   hy_DataReconstructNormalDir_PPM.F90, hy_eigenParameters.F90,
 *   hy_eigenValue.F90, OnlyHydro/hy_eigenVector.F90, hy_TVDslope.F90,
 *   hy_slopeLimiters.F90, hy_upwindTransverseFlux.F90, hy_getFaceFlux.F90,
 *   hy_HLLC.F90, hy_HLL.F90, hy_shockDetectBlk.F90, hy_prim2con.F90,
 *   hy_prim2flx.F90, hy_unsplitUpdate.F90,
 *   hy_energyFix.F90, hy_updateSolution.F90, Hydro_computeDt.F90;
 *   physics/Eos/EosMain/{Eos_getData,Eos_putData}.F90,
 *   physics/Eos/EosMain/Gamma/{eos_idealGamma,eos_initGamma}.F90.
 * Configuration restated: Cartesian geometry, no gravity, no species, no
 * MHD, no 3T, order=3 (PPM, NGUARD=4), no hybrid order, no upwind TVD,
 * transOrder=1, FLASH_EOS_GAMMA defined (gamma-law EOS).
 *
 * Each expression keeps the Fortran evaluation order.  Build with
 * -ffp-contract=off.
 */
#include "hydro_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* primitive slots inside a reconstruction vector (UHD.h:374-381, 0-based) */
enum { H_DENS = 0, H_VELX, H_VELY, H_VELZ, H_PRES, H_GAMC, H_GAME, H_EINT, H_GRAV };
/* waves (UHD.h:420-424, 0-based) */
enum { W_FL = 0, W_SL, W_EN, W_SR, W_FR };
#define NV9 9

void ox_default_params(ox_params* p, int ndim, int nb, double gamma) {
  memset(p, 0, sizeof(*p));
  p->ndim = ndim;
  p->nxb = nb;
  p->nyb = ndim > 1 ? nb : 1;
  p->nzb = ndim > 2 ? nb : 1;
  p->ng = 4;
  p->nunk = OX_NVAR_USED;
  for (int v = 0; v < OX_NVAR_USED; ++v) p->var[v] = v;
  p->gamma = gamma;
  p->limiter = OX_VANLEER;
  p->limitedSlopeBeta = 1.0;
  p->charLimiting = 1;
  p->flattening = 1;
  p->steepening = 0;
  p->use3dFullCTU = 1;
  p->useAuxEint = 1;
  p->riemann = OX_HLLC;
  p->fP = 1.0;
  p->smlrho = 1.e-10;
  p->smallp = 1.e-10;
  p->smallE = 1.e-10;
  p->eintSwitch = 0.0;
  p->tiny = 1.e-32;
  p->gasConstant = 8.3144598e7;
  p->singleSpeciesA = 1.0;
  p->exactMaxIter = 20;
  p->exactTol = 1.e-12;
  p->shockDetect = 0;
  p->shokVar = -1;
}

/* ------------------------------------------------------------------ */
/* block geometry helpers                                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int nx, ny, nz;        /* guard-inclusive extents */
  int lo[3], hi[3];      /* interior limits, 0-based inclusive */
  long ncell;
} geom_t;

static geom_t geom_of(const ox_params* p) {
  geom_t g;
  g.nx = p->nxb + 2 * p->ng;
  g.ny = p->ndim > 1 ? p->nyb + 2 * p->ng : 1;
  g.nz = p->ndim > 2 ? p->nzb + 2 * p->ng : 1;
  g.lo[0] = p->ng; g.hi[0] = p->ng + p->nxb - 1;
  g.lo[1] = p->ndim > 1 ? p->ng : 0; g.hi[1] = p->ndim > 1 ? p->ng + p->nyb - 1 : 0;
  g.lo[2] = p->ndim > 2 ? p->ng : 0; g.hi[2] = p->ndim > 2 ? p->ng + p->nzb - 1 : 0;
  g.ncell = (long)g.nx * g.ny * g.nz;
  return g;
}

long ox_block_cells(const ox_params* p) { return geom_of(p).ncell; }

long ox_flux_faces(const ox_params* p, int dir) {
  long n[3] = {p->nxb, p->ndim > 1 ? p->nyb : 1, p->ndim > 2 ? p->nzb : 1};
  if (dir < p->ndim) n[dir] += 1;
  return n[0] * n[1] * n[2];
}

#define CIDX(g, i, j, k) ((((long)(k)) * (g).ny + (j)) * (g).nx + (i))
#define UV(U, p, g, v, c) ((U)[(long)(p)->var[v] * (g).ncell + (c)])

/* flux index over interior faces: extents (nxb+[d==0]) x (nyb+[d==1]) x ... */
typedef struct { int fx, fy, fz; long nf; } fgeom_t;
static fgeom_t fgeom_of(const ox_params* p, int dir) {
  fgeom_t f;
  f.fx = p->nxb + (dir == 0);
  f.fy = (p->ndim > 1 ? p->nyb : 1) + (dir == 1 && p->ndim > 1);
  f.fz = (p->ndim > 2 ? p->nzb : 1) + (dir == 2 && p->ndim > 2);
  f.nf = (long)f.fx * f.fy * f.fz;
  return f;
}
/* (i,j,k) are guard-inclusive cell-style indices of the face (face i sits
 * on the low side of cell i) */
static long fidx(const ox_params* p, const geom_t* g, const fgeom_t* f, int i, int j, int k) {
  (void)p;
  return (((long)(k - g->lo[2])) * f->fy + (j - g->lo[1])) * f->fx + (i - g->lo[0]);
}

/* ------------------------------------------------------------------ */
/* slope limiters: hy_slopeLimiters.F90:38-60                           */
/* ------------------------------------------------------------------ */
static double f_sign1(double a) { return copysign(1.0, a); }
static double lim_minmod(double a, double b) {          /* :54-58 */
  return .5 * (f_sign1(a) + f_sign1(b)) * fmin(fabs(a), fabs(b));
}
static double lim_mc(double a, double b) {              /* :47-51 */
  return (f_sign1(a) + f_sign1(b)) * fmin(fmin(fabs(a), .25 * fabs(a + b)), fabs(b));
}
static double lim_vanleer(double a, double b) {         /* :38-45 */
  if (a * b <= 0.) return 0.;
  return 2. * a * b / (a + b);
}
static double lim_limited(double del1, double del2, double beta) {  /* hy_TVDslope.F90:220-229 */
  if (del2 > 0.)
    return fmax(fmax(0., fmin(beta * del1, del2)), fmin(del1, beta * del2));
  return fmin(fmin(0., fmax(beta * del1, del2)), fmax(del1, beta * del2));
}
double ox_limiter(int kind, double a, double b, double beta) {
  switch (kind) {
    case OX_MINMOD: return lim_minmod(a, b);
    case OX_MC: return lim_mc(a, b);
    case OX_VANLEER: return lim_vanleer(a, b);
    case OX_LIMITED: return lim_limited(a, b, beta);
    default: return lim_mc(a, b);
  }
}

/* ------------------------------------------------------------------ */
/* eigen system: hy_eigenParameters.F90:42-89, hy_eigenValue.F90:33-50,  */
/* OnlyHydro/hy_eigenVector.F90:46-151 (primitive form, cons=.false.)    */
/* leig[wave][var], reig[wave][var] (Fortran LeftEigvec(var,wave))       */
/* ------------------------------------------------------------------ */
int ox_eigen(const double V[9], int dir, double lambda[5], double leig[5][5],
             double reig[5][5]) {
  double uN = V[H_VELX + dir];
  double a2 = V[H_GAMC] * V[H_PRES] / V[H_DENS];          /* eigenParameters:81 */
  if (!(a2 > 0.)) return 1;                                /* :83-86 abort */
  double cf = sqrt(a2);
  lambda[W_FL] = uN - cf;                                  /* eigenValue:46-50 */
  lambda[W_SL] = uN;
  lambda[W_EN] = uN;
  lambda[W_SR] = uN;
  lambda[W_FR] = uN + cf;
  int v1, v2, v3;                                          /* eigenVector:75-89 */
  if (dir == 0) { v1 = H_VELX; v2 = H_VELY; v3 = H_VELZ; }
  else if (dir == 1) { v1 = H_VELY; v2 = H_VELZ; v3 = H_VELX; }
  else { v1 = H_VELZ; v2 = H_VELX; v3 = H_VELY; }
  memset(leig, 0, sizeof(double) * 25);
  memset(reig, 0, sizeof(double) * 25);
  double dinv = 1. / V[H_DENS];                            /* :102-108 */
  double a2e = cf * cf;
  double a2inv = 1. / a2e;
  double ahinv = 0.5 / cf;
  double Na = 0.5 * a2inv;
  leig[W_FL][v1] = -ahinv;                                 /* :115-133 */
  leig[W_FL][H_PRES] = dinv * Na;
  leig[W_SL][v2] = 1.;
  leig[W_EN][H_DENS] = 1.;
  leig[W_EN][H_PRES] = -a2inv;
  leig[W_SR][v3] = 1.;
  leig[W_FR][v1] = ahinv;
  leig[W_FR][H_PRES] = leig[W_FL][H_PRES];
  reig[W_FL][H_DENS] = V[H_DENS];                          /* :139-158 */
  reig[W_FL][v1] = -cf;
  reig[W_FL][H_PRES] = V[H_DENS] * a2e;
  reig[W_SL][v2] = 1.;
  reig[W_EN][H_DENS] = 1.;
  reig[W_SR][v3] = 1.;
  reig[W_FR][H_DENS] = V[H_DENS];
  reig[W_FR][v1] = cf;
  reig[W_FR][H_PRES] = reig[W_FL][H_PRES];
  return 0;
}

/* Fortran dot_product over vars 0..4, summed left to right */
static double dot5(const double* a, const double* b) {
  double s = 0.;
  for (int v = 0; v < 5; ++v) s = s + a[v] * b[v];
  return s;
}

/* ------------------------------------------------------------------ */
/* hy_TVDslope.F90:54-265 (hy_upwindTVD = .false.)                      */
/* ------------------------------------------------------------------ */
void ox_tvd_slope(const ox_params* p, const double VL[9], const double V0[9],
                  const double VR[9], const double lambda[5], double leig[5][5],
                  double delbar[9]) {
  (void)lambda;
  double dL[9], dR[9];
  for (int k = 0; k < NV9; ++k) { delbar[k] = 0.; dL[k] = V0[k] - VL[k]; dR[k] = VR[k] - V0[k]; }
  const int endVars = H_EINT;   /* HY_END_VARS = HY_EINT (no gravity, no 3T) */
  const double beta = p->limitedSlopeBeta;
  if (p->limiter == OX_HYBRID) {
    if (p->charLimiting) {                                  /* :145-175 */
      for (int w = 0; w < 5; ++w) {
        double a = dot5(leig[w], dL), b = dot5(leig[w], dR);
        delbar[w] = (w == W_EN) ? lim_mc(a, b) : lim_minmod(a, b);
      }
      for (int k = H_GAMC; k <= endVars; ++k) delbar[k] = lim_mc(dL[k], dR[k]);
    } else {                                                /* :177-196 */
      for (int k = 0; k <= endVars; ++k) {
        if (k == H_DENS || k == H_GAMC || k == H_GAME)
          delbar[k] = lim_mc(dL[k], dR[k]);
        else if (k >= H_VELX && k <= H_PRES)
          delbar[k] = lim_minmod(dL[k], dR[k]);
        /* HY_EINT has no case in the reference select: stays 0 */
      }
    }
    return;
  }
  if (p->charLimiting) {                                    /* :80-141 */
    for (int w = 0; w < 5; ++w)
      delbar[w] = ox_limiter(p->limiter, dot5(leig[w], dL), dot5(leig[w], dR), beta);
    for (int k = H_GAMC; k <= endVars; ++k)
      delbar[k] = ox_limiter(p->limiter, dL[k], dR[k], beta);
  } else {
    for (int k = 0; k <= endVars; ++k)
      delbar[k] = ox_limiter(p->limiter, dL[k], dR[k], beta);
  }
}

/* ------------------------------------------------------------------ */
/* hy_upwindTransverseFlux.F90:43-211, order=1, sigSize=HY_END_VARS      */
/* ------------------------------------------------------------------ */
void ox_upwind_trans_flux(const ox_params* p, const double vm1[9],
                          const double vc0[9], const double vp1[9],
                          const double lambda[5], double leig[5][5],
                          double reig[5][5], double sig[9]) {
  const int conEnd = H_PRES;                 /* hyConEnd = HY_ENER = 5 (1-based) */
  int varEnd = conEnd;                       /* FLASH_EOS_GAMMA: no gamc/game */
  if (p->useAuxEint) varEnd = H_EINT;        /* :101 */
  for (int v = 0; v < 8; ++v) sig[v] = 0.;   /* :122 sig(1:sigSize) = 0 */
  double tf[9];
  for (int n = 0; n < 5; ++n) {              /* :129-147 */
    if (lambda[n] > 0.) {
      for (int v = 0; v <= conEnd; ++v) tf[v] = vc0[v] - vm1[v];
    } else {
      for (int v = 0; v <= conEnd; ++v) tf[v] = vp1[v] - vc0[v];
    }
    double db = dot5(leig[n], tf);
    for (int v = 0; v <= conEnd; ++v) {
      tf[v] = lambda[n] * reig[n][v] * db;
      sig[v] = sig[v] + tf[v];
    }
  }
  int beg = p->useAuxEint ? H_EINT : H_EINT + 1;            /* :174-178 */
  if (varEnd > conEnd && beg >= varEnd) {                   /* :181-193 */
    for (int v = beg; v <= varEnd; ++v) {
      double d = (lambda[W_EN] > 0.) ? vc0[v] - vm1[v] : vp1[v] - vc0[v];
      sig[v] = lambda[W_EN] * d;
    }
  }
}

/* ------------------------------------------------------------------ */
/* hy_DataReconstructNormalDir_PPM.F90:73-958 (Cartesian hydro, NGUARD=4) */
/* ------------------------------------------------------------------ */
int ox_ppm(const ox_params* p, int dir, double dt, double delta,
           const double data[5][9], double flatCoeff, int trans,
           double lambda0[5], double leig0[5][5], double reig0[5][5],
           double Wp[9], double Wm[9], double sig[9]) {
  const int endPrim = H_PRES;                       /* hyEndPrimVar = HY_ENER(=5) */
  const int endVar = p->useAuxEint ? H_EINT : H_PRES; /* :168-169 */
  double Vmm[9], Vm[9], Vc[9], Vp[9], Vpp[9];
  memcpy(Vm, data[1], sizeof Vm);
  memcpy(Vc, data[2], sizeof Vc);
  memcpy(Vp, data[3], sizeof Vp);
  if (!trans) { memcpy(Vmm, data[0], sizeof Vmm); memcpy(Vpp, data[4], sizeof Vpp); }
  double dtn = dt / delta;                           /* :309-311 */
  double hdtn = 0.5 * dtn;
  for (int v = 0; v < NV9; ++v) { Wp[v] = 0.; Wm[v] = 0.; }   /* :314 */
  if (ox_eigen(Vc, dir, lambda0, leig0, reig0)) return 1;      /* :316-345 */
  if (p->ndim > 1) ox_upwind_trans_flux(p, Vm, Vc, Vp, lambda0, leig0, reig0, sig); /* :351-358 */
  if (trans) return 0;                                          /* :377 */

  double lP[5], lN[5], leP[5][5], reP[5][5], leN[5][5], reN[5][5];
  if (ox_eigen(Vp, dir, lP, leP, reP)) return 1;               /* :379-403 */
  if (ox_eigen(Vm, dir, lN, leN, reN)) return 1;               /* :406-430 */
  double dbN[9], db0[9], dbP[9];
  ox_tvd_slope(p, Vmm, Vm, Vc, lN, leN, dbN);                  /* :438-441 */
  ox_tvd_slope(p, Vm, Vc, Vp, lambda0, leig0, db0);
  ox_tvd_slope(p, Vc, Vp, Vpp, lP, leP, dbP);

  double Flat = p->flattening ? flatCoeff : 0.;                /* :514-519 */
  if (p->charLimiting) {                                       /* :523-570 */
    double t0[5], tP[5], tN[5];
    for (int v = 0; v <= endPrim; ++v) {
      t0[v] = reig0[W_FL][v] * db0[W_FL] + reig0[W_SL][v] * db0[W_SL] +
              reig0[W_EN][v] * db0[W_EN] + reig0[W_SR][v] * db0[W_SR] +
              reig0[W_FR][v] * db0[W_FR];
      tP[v] = reP[W_FL][v] * dbP[W_FL] + reP[W_SL][v] * dbP[W_SL] +
              reP[W_EN][v] * dbP[W_EN] + reP[W_SR][v] * dbP[W_SR] +
              reP[W_FR][v] * dbP[W_FR];
      tN[v] = reN[W_FL][v] * dbN[W_FL] + reN[W_SL][v] * dbN[W_SL] +
              reN[W_EN][v] * dbN[W_EN] + reN[W_SR][v] * dbN[W_SR] +
              reN[W_FR][v] * dbN[W_FR];
    }
    for (int v = 0; v <= endPrim; ++v) { db0[v] = t0[v]; dbP[v] = tP[v]; dbN[v] = tN[v]; }
  }

  double vecL[9], vecR[9];                                      /* :578-582 */
  for (int v = 0; v < NV9; ++v) { vecL[v] = 0.; vecR[v] = 0.; }
  for (int v = 0; v <= endVar; ++v) {
    vecL[v] = 0.5 * (Vc[v] + Vm[v]) - (db0[v] - dbN[v]) / 6.;
    vecR[v] = 0.5 * (Vc[v] + Vp[v]) - (dbP[v] - db0[v]) / 6.;
  }

  if (p->steepening) {                                          /* :639-671 */
    const double eta1 = 20.E0, eta2 = 0.05E0, epsln = 0.01E0, K0 = 0.1E0;
    double temp1 = Vp[H_DENS] - Vm[H_DENS];
    if (fabs(temp1) > p->tiny) {
      double del2rhoR = (Vpp[H_DENS] - 2. * Vp[H_DENS] + Vc[H_DENS]) / (6. * delta * delta);
      double del2rhoL = (Vc[H_DENS] - 2. * Vm[H_DENS] + Vmm[H_DENS]) / (6. * delta * delta);
      double eta = (del2rhoL - del2rhoR) * (delta * delta) / temp1;
      if (del2rhoR * del2rhoL >= 0.) eta = 0.;
      if (epsln * fmin(Vp[H_DENS], Vm[H_DENS]) - fabs(Vp[H_DENS] - Vm[H_DENS]) >= 0.) eta = 0.;
      eta = fmax(0., fmin(1., eta1 * (eta - eta2)));
      double temp2 = fabs(Vp[H_PRES] - Vm[H_PRES]) / fmin(Vp[H_PRES], Vm[H_PRES]);
      double temp3 = fabs(Vp[H_DENS] - Vm[H_DENS]) / fmin(Vp[H_DENS], Vm[H_DENS]);
      if (Vc[H_GAME] * K0 * temp3 - temp2 < 0.0) eta = 0.;
      vecL[H_DENS] = vecL[H_DENS] * (1. - eta) + (Vm[H_DENS] + 0.5 * dbN[H_DENS]) * eta;
      vecR[H_DENS] = vecR[H_DENS] * (1. - eta) + (Vp[H_DENS] - 0.5 * dbP[H_DENS]) * eta;
    }
  }

  if (p->flattening) {                                          /* :676-679 */
    for (int v = 0; v < NV9; ++v) {
      vecL[v] = Flat * Vc[v] + (1.0 - Flat) * vecL[v];
      vecR[v] = Flat * Vc[v] + (1.0 - Flat) * vecR[v];
    }
  }

  for (int n = 0; n <= endVar; ++n) {                           /* :686-697 */
    if ((vecR[n] - Vc[n]) * (Vc[n] - vecL[n]) <= 0.) { vecL[n] = Vc[n]; vecR[n] = Vc[n]; }
    double dd = vecR[n] - vecL[n];
    if (6. * (vecR[n] - vecL[n]) * (Vc[n] - 0.5 * (vecL[n] + vecR[n])) > dd * dd)
      vecL[n] = 3. * Vc[n] - 2. * vecR[n];
    dd = vecR[n] - vecL[n];
    if (6. * (vecR[n] - vecL[n]) * (Vc[n] - 0.5 * (vecL[n] + vecR[n])) < -(dd * dd))
      vecR[n] = 3. * Vc[n] - 2. * vecL[n];
  }

  double delW[9], W6[9];                                         /* :705-707 */
  for (int v = 0; v < NV9; ++v) { delW[v] = 0.; W6[v] = 0.; }
  for (int v = 0; v <= endVar; ++v) {
    delW[v] = vecR[v] - vecL[v];
    W6[v] = 6. * (Vc[v] - 0.5 * (vecR[v] + vecL[v]));
  }
  double lmax = fmax(lambda0[W_FR], 0.);                         /* :711-719 */
  for (int v = 0; v <= endPrim; ++v)
    Wp[v] = vecR[v] - lmax * hdtn * (delW[v] - (1.0 - lmax * hdtn * 4. / 3.) * W6[v]);
  lmax = fmax(lambda0[W_EN], 0.);
  for (int v = endPrim + 1; v <= endVar; ++v)
    Wp[v] = vecR[v] - lmax * hdtn * (delW[v] - (1.0 - lmax * hdtn * 4. / 3.) * W6[v]);
  double lmin = -fmin(lambda0[W_FL], 0.);                        /* :722-729 */
  for (int v = 0; v <= endPrim; ++v)
    Wm[v] = vecL[v] + lmin * hdtn * (delW[v] + (1.0 - lmin * hdtn * 4. / 3.) * W6[v]);
  lmin = -fmin(lambda0[W_EN], 0.);
  for (int v = endPrim + 1; v <= endVar; ++v)
    Wm[v] = vecL[v] + lmin * hdtn * (delW[v] + (1.0 - lmin * hdtn * 4. / 3.) * W6[v]);
  Wp[H_GAMC] = Vc[H_GAMC]; Wp[H_GAME] = Vc[H_GAME];              /* :733-736 */
  Wm[H_GAMC] = Vc[H_GAMC]; Wm[H_GAME] = Vc[H_GAME];

  double sigL[9], sigR[9];                                        /* :751-800 */
  for (int v = 0; v < NV9; ++v) { sigL[v] = 0.; sigR[v] = 0.; }
  double mW6[5];
  for (int v = 0; v < 5; ++v) mW6[v] = -W6[v];
  for (int n = 0; n < 5; ++n) {
    double cA = dot5(leig0[n], delW);
    double cB = dot5(leig0[n], mW6);
    double x = dtn * lambda0[n];
    if (p->riemann == OX_ROE) {
      if (lambda0[n] < 0.) {
        for (int v = 0; v <= endPrim; ++v)
          sigL[v] = sigL[v] + (.5 * (-1. - x) * reig0[n][v] * cA +
                               .25 * (1. + 2. * dtn * lambda0[n] + 4. / 3. * (x * x)) * reig0[n][v] * cB);
      } else if (lambda0[n] > 0.) {
        for (int v = 0; v <= endPrim; ++v)
          sigR[v] = sigR[v] + (.5 * (1. - x) * reig0[n][v] * cA +
                               .25 * (1. - 2. * dtn * lambda0[n] + 4. / 3. * (x * x)) * reig0[n][v] * cB);
      }
    } else {
      for (int v = 0; v <= endPrim; ++v) {
        double vl = .5 * (-1. - x) * reig0[n][v] * cA +
                    .25 * (1. + 2. * dtn * lambda0[n] + 4. / 3. * (x * x)) * reig0[n][v] * cB;
        sigL[v] = sigL[v] + vl;
        double vr = .5 * (1. - x) * reig0[n][v] * cA +
                    .25 * (1. - 2. * dtn * lambda0[n] + 4. / 3. * (x * x)) * reig0[n][v] * cB;
        sigR[v] = sigR[v] + vr;
      }
    }
  }
  for (int v = 0; v <= endPrim; ++v) {                              /* :841-842 */
    Wm[v] = Vc[v] + W6[v] / 12. + sigL[v];
    Wp[v] = Vc[v] + W6[v] / 12. + sigR[v];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* HLLC: hy_HLLC.F90:53-324 with hy_prim2con.F90:23-56,                */
/* hy_prim2flx.F90:25-140 (hydro)                                       */
/* ------------------------------------------------------------------ */
static void prim2con(const double V[9], double CU[5]) {
  double u2 = V[H_VELX] * V[H_VELX] + V[H_VELY] * V[H_VELY] + V[H_VELZ] * V[H_VELZ];
  CU[0] = V[H_DENS];
  CU[1] = V[H_DENS] * V[H_VELX];
  CU[2] = V[H_DENS] * V[H_VELY];
  CU[3] = V[H_DENS] * V[H_VELZ];
  CU[4] = 0.5 * V[H_DENS] * u2 + V[H_PRES] / (V[H_GAME] - 1.) + 0.5 * 0.;
}
static void prim2flx(int dir, const double V[9], double fP, double F[6]) {
  double u2 = V[H_VELX] * V[H_VELX] + V[H_VELY] * V[H_VELY] + V[H_VELZ] * V[H_VELZ];
  double E = 0.5 * V[H_DENS] * u2 + V[H_PRES] / (V[H_GAME] - 1.);
  F[5] = V[H_PRES];
  double un = V[H_VELX + dir];
  F[0] = V[H_DENS] * un;
  F[1] = F[0] * V[H_VELX];
  F[2] = F[0] * V[H_VELY];
  F[3] = F[0] * V[H_VELZ];
  F[1 + dir] = F[1 + dir] + fP * V[H_PRES];
  F[4] = (E + V[H_PRES]) * un;
}

int ox_hllc(const ox_params* p, int dir, const double Vm[9], const double Vp[9],
            double Fstar[6], double* speed) {
  double velNL = Vm[H_VELX + dir], velNR = Vp[H_VELX + dir];        /* :94-95 */
  double aL2 = Vm[H_GAMC] * Vm[H_PRES] / Vm[H_DENS];               /* :98-99 */
  double aR2 = Vp[H_GAMC] * Vp[H_PRES] / Vp[H_DENS];
  double magNL = 0., magNR = 0.;
  *speed = 0.;
  if ((Vm[H_DENS] < p->tiny && Vm[H_DENS] > 0.) || (Vp[H_DENS] < p->tiny && Vp[H_DENS] > 0.) ||
      (Vm[H_PRES] < p->tiny && Vm[H_PRES] > 0.) || (Vp[H_PRES] < p->tiny && Vp[H_PRES] > 0.)) {
    for (int f = 0; f < 6; ++f) Fstar[f] = 0.;                      /* :118-124 */
    return 0;
  } else if (aL2 < 0. || aR2 < 0.) {
    return 1;                                                       /* :125-127 */
  }
  double cfL = sqrt(aL2), cfR = sqrt(aR2);
  double SL = fmin(velNL - cfL, velNR - cfR);                        /* :141-142 */
  double SR = fmax(velNL + cfL, velNR + cfR);
  *speed = fmax(fabs(SL), fabs(SR));
  double tpL = Vm[H_PRES], tpR = Vp[H_PRES];
  double UL[5], UR[5], FL[6], FR[6];
  prim2con(Vm, UL); prim2con(Vp, UR);
  prim2flx(dir, Vm, p->fP, FL); prim2flx(dir, Vp, p->fP, FR);
  double qStar = (Vp[H_DENS] * velNR * (SR - velNR) - Vm[H_DENS] * velNL * (SL - velNL) +
                  tpL - tpR - magNL * magNL + magNR * magNR);        /* :189-193 */
  qStar = qStar / (Vp[H_DENS] * (SR - velNR) - Vm[H_DENS] * (SL - velNL));
  double numerL = SL - velNL, denomL = SL - qStar;
  double numerR = SR - velNR, denomR = SR - qStar;
  double pStar = Vm[H_DENS] * numerL * (qStar - velNL) + tpL;        /* :203 */
  double dStarL = UL[0] * numerL / denomL, dStarR = UR[0] * numerR / denomR;
  double UCL[5], UCR[5];
  UCL[0] = dStarL;
  UCL[4] = UL[4] * numerL / denomL + ((pStar * qStar - tpL * velNL)) / denomL;
  UCR[0] = dStarR;
  UCR[4] = UR[4] * numerR / denomR + ((pStar * qStar - tpR * velNR)) / denomR;
  for (int m = 0; m < 3; ++m) {                                      /* :233-283 */
    if (m == dir) { UCL[1 + m] = dStarL * qStar; UCR[1 + m] = dStarR * qStar; }
    else { UCL[1 + m] = UL[1 + m] * numerL / denomL; UCR[1 + m] = UR[1 + m] * numerR / denomR; }
  }
  if (SL >= 0.) {                                                    /* :295-311 */
    for (int f = 0; f < 6; ++f) Fstar[f] = FL[f];
  } else if (SL < 0. && qStar >= 0.) {
    for (int f = 0; f < 5; ++f) Fstar[f] = FL[f] + SL * (UCL[f] - UL[f]);
    Fstar[5] = FL[5];
  } else if (qStar < 0. && SR >= 0.) {
    for (int f = 0; f < 5; ++f) Fstar[f] = FR[f] + SR * (UCR[f] - UR[f]);
    Fstar[5] = FR[5];
  } else {
    for (int f = 0; f < 6; ++f) Fstar[f] = FR[f];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* HLL: hy_HLL.F90:82-165 (hydro; Davis speeds as in HLLC)              */
/* ------------------------------------------------------------------ */
int ox_hll(const ox_params* p, int dir, const double Vm[9], const double Vp[9],
           double Fstar[6], double* speed) {
  double velNL = Vm[H_VELX + dir], velNR = Vp[H_VELX + dir];        /* :97-98 */
  double aL2 = Vm[H_GAMC] * Vm[H_PRES] / Vm[H_DENS];               /* :101-102 */
  double aR2 = Vp[H_GAMC] * Vp[H_PRES] / Vp[H_DENS];
  *speed = 0.;
  if ((Vm[H_DENS] < p->tiny && Vm[H_DENS] > 0.) || (Vp[H_DENS] < p->tiny && Vp[H_DENS] > 0.) ||
      (Vm[H_PRES] < p->tiny && Vm[H_PRES] > 0.) || (Vp[H_PRES] < p->tiny && Vp[H_PRES] > 0.)) {
    for (int f = 0; f < 6; ++f) Fstar[f] = 0.;                      /* :112-118 */
    return 0;
  } else if (aL2 < 0. || aR2 < 0.) {
    return 1;                                                       /* :119-121 */
  }
  double cfL = sqrt(aL2), cfR = sqrt(aR2);
  double SL = fmin(velNL - cfL, velNR - cfR);                        /* :136-137 */
  double SR = fmax(velNL + cfL, velNR + cfR);
  *speed = fmax(fabs(SL), fabs(SR));                                 /* :141 */
  double UL[6], UR[6], FL[6], FR[6];
  prim2con(Vm, UL); prim2con(Vp, UR);                                /* :144-146 */
  UL[5] = 0.0; UR[5] = 0.0;
  prim2flx(dir, Vm, p->fP, FL); prim2flx(dir, Vp, p->fP, FR);
  if (SL > 0.) {                                                     /* :151-164 */
    for (int f = 0; f < 6; ++f) Fstar[f] = FL[f];
  } else if (SR < 0.) {
    for (int f = 0; f < 6; ++f) Fstar[f] = FR[f];
  } else {
    for (int f = 0; f < 6; ++f)
      Fstar[f] = (SR * FL[f] - SL * FR[f] + SR * SL * (UR[f] - UL[f])) / (SR - SL);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Shock flag: hy_shockDetectBlk.F90:93-229 (Balsara & Spicer, JCP      */
/* 149:270, 1999)                                                       */
/* ------------------------------------------------------------------ */
long ox_shock_detect(const ox_params* p, double* U) {
  if (p->shokVar < 0 || p->shokVar >= p->nunk) return -1;
  geom_t g = geom_of(p);
  double* S = U + (long)p->shokVar * g.ncell;
  for (long c = 0; c < g.ncell; ++c) S[c] = 0.;                      /* :117 */
  const int k2 = p->ndim > 1, k3 = p->ndim > 2;
  const long sx = 1, sy = g.nx, sz = (long)g.nx * g.ny;
  double* Vc = malloc(sizeof(double) * g.ncell);
  if (!Vc) return -1;
  for (int k = g.lo[2] - k3; k <= g.hi[2] + k3; ++k)                 /* :132-138 */
    for (int j = g.lo[1] - k2; j <= g.hi[1] + k2; ++j)
      for (int i = g.lo[0] - 1; i <= g.hi[0] + 1; ++i) {
        long c = CIDX(g, i, j, k);
        Vc[c] = sqrt(UV(U, p, g, OX_GAMC, c) * UV(U, p, g, OX_PRES, c) / UV(U, p, g, OX_DENS, c));
      }
  const double beta = 0.5, delta = 0.1;                              /* :94-95 */
  long n = 0;
  for (int k = g.lo[2]; k <= g.hi[2]; ++k)
    for (int j = g.lo[1]; j <= g.hi[1]; ++j)
      for (int i = g.lo[0]; i <= g.hi[0]; ++i) {
        long c = CIDX(g, i, j, k);
        double minP = UV(U, p, g, OX_PRES, c), minC = Vc[c];         /* :152-153 */
        for (int c3 = -k3; c3 <= k3; ++c3)
          for (int c2 = -k2; c2 <= k2; ++c2)
            for (int c1 = -1; c1 <= 1; ++c1) {
              long q = c + c1 * sx + c2 * sy + c3 * sz;
              minP = fmin(minP, UV(U, p, g, OX_PRES, q));
              minC = fmin(minC, Vc[q]);
            }
        double divv = UV(U, p, g, OX_VELX, c + sx) - UV(U, p, g, OX_VELX, c - sx);  /* :157-164 */
        if (k2) divv = divv + UV(U, p, g, OX_VELY, c + sy) - UV(U, p, g, OX_VELY, c - sy);
        if (k3) divv = divv + UV(U, p, g, OX_VELZ, c + sz) - UV(U, p, g, OX_VELZ, c - sz);
        divv = 0.5 * divv;
        double gradPx = 0.5 * (UV(U, p, g, OX_PRES, c + sx) - UV(U, p, g, OX_PRES, c - sx));
        double gradPy = 0., gradPz = 0.;                             /* :167-174 */
        if (k2) gradPy = 0.5 * (UV(U, p, g, OX_PRES, c + sy) - UV(U, p, g, OX_PRES, c - sy));
        if (k3) gradPz = 0.5 * (UV(U, p, g, OX_PRES, c + sz) - UV(U, p, g, OX_PRES, c - sz));
        int sw1 = fabs(gradPx) + fabs(gradPy) + fabs(gradPz) >= beta * minP;   /* :175-178 */
        int sw2 = -(delta * minC) >= divv;                                     /* :180-182 */
        if (sw1 && sw2) {                                                      /* :185-191 */
          S[c] = 1.;
          ++n;
        }
      }
  free(Vc);
  return n;
}

/* ------------------------------------------------------------------ */
/* Exact Riemann solver for the gamma-law gas.  Not in the reference:   */
/* Newton iteration on the exact pressure function (Toro, Riemann       */
/* Solvers and Numerical Methods, 3rd ed., sec. 4.3-4.5: f_K shock and   */
/* rarefaction branches, eq. 4.6-4.8, derivative eq. 4.37), seeded by    */
/* the two-shock guess (Toro eq. 9.42), sampled at x/t = 0 (sec. 4.5).   */
/* ------------------------------------------------------------------ */
static void exact_fk(double p, double rk, double pk, double ck, double g,
                     double* f, double* fd) {
  if (p <= pk) {                                     /* rarefaction, Toro 4.7 */
    double pratio = p / pk;
    *f = 2. / (g - 1.) * ck * (pow(pratio, (g - 1.) / (2. * g)) - 1.);
    *fd = (1. / (rk * ck)) * pow(pratio, -(g + 1.) / (2. * g));
  } else {                                           /* shock, Toro 4.6 */
    double ak = 2. / ((g + 1.) * rk);
    double bk = (g - 1.) / (g + 1.) * pk;
    double qrt = sqrt(ak / (bk + p));
    *f = (p - pk) * qrt;
    *fd = (1. - 0.5 * (p - pk) / (bk + p)) * qrt;
  }
}

static double two_shock_guess(double g, double rl, double ul, double pl,
                              double cl, double rr, double ur, double pr, double cr) {
  /* Toro eq. 9.42 with p0 = max(tol, p_PV) (eq. 9.32) */
  double ppv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr);
  double p0 = fmax(1.e-30, ppv);
  double gl = sqrt((2. / ((g + 1.) * rl)) / ((g - 1.) / (g + 1.) * pl + p0));
  double gr = sqrt((2. / ((g + 1.) * rr)) / ((g - 1.) / (g + 1.) * pr + p0));
  double pts = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
  return fmax(1.e-30, pts);
}

int ox_exact_pstar(double g, double rl, double ul, double pl, double rr,
                   double ur, double pr, double guess, double tol, int maxit,
                   double* pstar, double* ustar) {
  double cl = sqrt(g * pl / rl), cr = sqrt(g * pr / rr);
  double du = ur - ul;
  double p = guess > 0. ? guess : two_shock_guess(g, rl, ul, pl, cl, rr, ur, pr, cr);
  int it;
  int ok = 0;
  for (it = 1; it <= maxit; ++it) {
    double fl, fld, fr, frd;
    exact_fk(p, rl, pl, cl, g, &fl, &fld);
    exact_fk(p, rr, pr, cr, g, &fr, &frd);
    double pn = p - (fl + fr + du) / (fld + frd);
    if (!(pn > 0.)) pn = 1.e-30;   /* keep the iterate physical (Toro 4.3.2) */
    double change = 2. * fabs((pn - p) / (pn + p));
    p = pn;
    if (change <= tol) { ok = 1; break; }
  }
  double fl, fld, fr, frd;
  exact_fk(p, rl, pl, cl, g, &fl, &fld);
  exact_fk(p, rr, pr, cr, g, &fr, &frd);
  *pstar = p;
  *ustar = 0.5 * (ul + ur) + 0.5 * (fr - fl);        /* Toro eq. 4.9 */
  return ok ? it : -maxit;
}

/* sample the exact solution at S = x/t (Toro sec. 4.5); the Godunov flux
 * uses S = 0 */
static void exact_sample_at(double g, double rl, double ul, double pl, double rr,
                            double ur, double pr, double pm, double um, double S,
                            double* r, double* u, double* pp, int* leftside) {
  double cl = sqrt(g * pl / rl), cr = sqrt(g * pr / rr);
  if (S <= um) {                                     /* left of contact */
    *leftside = 1;
    if (pm <= pl) {                                  /* left rarefaction */
      double shl = ul - cl;
      if (S <= shl) { *r = rl; *u = ul; *pp = pl; return; }
      double cml = cl * pow(pm / pl, (g - 1.) / (2. * g));
      double stl = um - cml;
      if (S > stl) { *r = rl * pow(pm / pl, 1. / g); *u = um; *pp = pm; return; }
      double c = 2. / (g + 1.) * (cl + (g - 1.) / 2. * (ul - S));
      *u = 2. / (g + 1.) * (cl + (g - 1.) / 2. * ul + S);
      *r = rl * pow(c / cl, 2. / (g - 1.));
      *pp = pl * pow(c / cl, 2. * g / (g - 1.));
      return;
    }
    double pml = pm / pl;                            /* left shock */
    double sl = ul - cl * sqrt((g + 1.) / (2. * g) * pml + (g - 1.) / (2. * g));
    if (S <= sl) { *r = rl; *u = ul; *pp = pl; return; }
    *r = rl * (pml + (g - 1.) / (g + 1.)) / (pml * (g - 1.) / (g + 1.) + 1.);
    *u = um; *pp = pm;
    return;
  }
  *leftside = 0;
  if (pm > pr) {                                     /* right shock */
    double pmr = pm / pr;
    double sr = ur + cr * sqrt((g + 1.) / (2. * g) * pmr + (g - 1.) / (2. * g));
    if (S >= sr) { *r = rr; *u = ur; *pp = pr; return; }
    *r = rr * (pmr + (g - 1.) / (g + 1.)) / (pmr * (g - 1.) / (g + 1.) + 1.);
    *u = um; *pp = pm;
    return;
  }
  double shr = ur + cr;                              /* right rarefaction */
  if (S >= shr) { *r = rr; *u = ur; *pp = pr; return; }
  double cmr = cr * pow(pm / pr, (g - 1.) / (2. * g));
  double str = um + cmr;
  if (S <= str) { *r = rr * pow(pm / pr, 1. / g); *u = um; *pp = pm; return; }
  double c = 2. / (g + 1.) * (cr - (g - 1.) / 2. * (ur - S));
  *u = 2. / (g + 1.) * (-cr + (g - 1.) / 2. * ur + S);
  *r = rr * pow(c / cr, 2. / (g - 1.));
  *pp = pr * pow(c / cr, 2. * g / (g - 1.));
}

static void exact_sample(double g, double rl, double ul, double pl, double rr,
                         double ur, double pr, double pm, double um,
                         double* r, double* u, double* pp, int* leftside) {
  exact_sample_at(g, rl, ul, pl, rr, ur, pr, pm, um, 0., r, u, pp, leftside);
}

int ox_exact_sample(double g, double rl, double ul, double pl, double rr, double ur,
                    double pr, double S, double out[3]) {
  double pm, um;
  int it = ox_exact_pstar(g, rl, ul, pl, rr, ur, pr, -1., 1e-14, 100, &pm, &um);
  int left;
  exact_sample_at(g, rl, ul, pl, rr, ur, pr, pm, um, S, &out[0], &out[1], &out[2], &left);
  return it;
}

int ox_exact_riemann(const ox_params* p, int dir, const double VL[9],
                     const double VR[9], double Fstar[6], double* speed,
                     ox_stats* st) {
  double g = VL[H_GAMC];
  if (!(VL[H_DENS] > 0. && VR[H_DENS] > 0. && VL[H_PRES] > 0. && VR[H_PRES] > 0.)) return 1;
  double rl = VL[H_DENS], ul = VL[H_VELX + dir], pl = VL[H_PRES];
  double rr = VR[H_DENS], ur = VR[H_VELX + dir], pr = VR[H_PRES];
  double cl = sqrt(g * pl / rl), cr = sqrt(g * pr / rr);
  if (2. / (g - 1.) * (cl + cr) <= ur - ul) return 2;   /* vacuum (Toro 4.40) */
  double pm, um;
  int it = ox_exact_pstar(g, rl, ul, pl, rr, ur, pr, -1., p->exactTol, p->exactMaxIter, &pm, &um);
  if (st) { st->exactIters += it > 0 ? it : -it; if (it < 0) st->exactFallbacks++; }
  double r, u, pp;
  int left;
  exact_sample(g, rl, ul, pl, rr, ur, pr, pm, um, &r, &u, &pp, &left);
  double W[9];
  const double* side = left ? VL : VR;
  for (int v = 0; v < 9; ++v) W[v] = side[v];
  W[H_DENS] = r; W[H_VELX + dir] = u; W[H_PRES] = pp;
  prim2flx(dir, W, p->fP, Fstar);
  double SL = fmin(ul - cl, ur - cr), SR = fmax(ul + cl, ur + cr);
  *speed = fmax(fabs(SL), fabs(SR));
  return 0;
}

/* ------------------------------------------------------------------ */
/* ideal gamma-law EOS over a box                                       */
/* ------------------------------------------------------------------ */
void ox_eos_box(const ox_params* p, double* U, int mode, const int lo[3], const int hi[3]) {
  geom_t g = geom_of(p);
  double gammam1 = 1.0 / (p->gamma - 1.0);            /* eos_initGamma.F90:42 */
  double ggprod = gammam1 * p->gasConstant;           /* eos_idealGamma.F90:253 */
  double ggprodinv = 1. / ggprod;
  double gam1inv = 1. / gammam1;
  double abar = p->singleSpeciesA;
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        long c = CIDX(g, i, j, k);
        double vx = UV(U, p, g, OX_VELX, c), vy = UV(U, p, g, OX_VELY, c), vz = UV(U, p, g, OX_VELZ, c);
        double ekin = 0.5 * (vx * vx + vy * vy + vz * vz);   /* Eos_getData.F90:192-194 */
        double dens = UV(U, p, g, OX_DENS, c);
        double pres = UV(U, p, g, OX_PRES, c);
        double eint;
        if (mode == 1) {                                      /* MODE_DENS_PRES */
          double ie = UV(U, p, g, OX_EINT, c);                /* getData:201-207 */
          double ener = UV(U, p, g, OX_ENER, c);
          if (ener - ekin > fmax(p->smallE, p->eintSwitch * ekin)) ie = ener - ekin;
          ie = fmax(ie, p->smallE);
          (void)ie;                                           /* overwritten by the EOS */
          eint = pres * gammam1 / dens;                       /* eos_idealGamma:351-352 */
        } else {                                              /* MODE_DENS_EI */
          double ie = UV(U, p, g, OX_EINT, c);
          double ener = UV(U, p, g, OX_ENER, c);
          if (ener - ekin > fmax(p->smallE, p->eintSwitch * ekin)) ie = ener - ekin;
          ie = fmax(ie, p->smallE);
          eint = ie;
          pres = dens * eint * gam1inv;                       /* eos_idealGamma:327-328 */
        }
        double temp = eint * ggprodinv * abar;
        UV(U, p, g, OX_PRES, c) = pres;                       /* Eos_putData.F90:201-209 */
        UV(U, p, g, OX_TEMP, c) = temp;
        UV(U, p, g, OX_GAMC, c) = p->gamma;
        UV(U, p, g, OX_EINT, c) = eint;
        UV(U, p, g, OX_ENER, c) = eint + ekin;
        UV(U, p, g, OX_GAME, c) = pres / (eint * dens) + 1;
      }
}

/* ------------------------------------------------------------------ */
/* per-block scratch, laid out like the reference's local arrays         */
/* ------------------------------------------------------------------ */
typedef struct {
  double* flat;    /* FlatCoeff(dir, cell)          [c*3+d] */
  double* flatT;   /* FlatTilde                      [c*3+d] */
  double* sig;     /* sig(9, dir, cell)             [(c*3+d)*9+v] */
  double* lam;     /* lambda(5, dir, cell)          [(c*3+d)*5+w] */
  double* leig;    /* leftEig(var, wave, dir, cell) [((c*3+d)*5+w)*5+v] */
  double* reig;
  double* face;    /* scrchFace{X,Y,Z}: [(d*ncell + c)*16 + slot], P 0..7, N 8..15 */
  double* var1;    /* scrch_Ptr(HY_VAR1, cell) */
} scratch_t;

static int scratch_alloc(scratch_t* s, long ncell) {
  s->flat = calloc((size_t)ncell * 3, sizeof(double));
  s->flatT = calloc((size_t)ncell * 3, sizeof(double));
  s->sig = calloc((size_t)ncell * 27, sizeof(double));
  s->lam = calloc((size_t)ncell * 15, sizeof(double));
  s->leig = calloc((size_t)ncell * 75, sizeof(double));
  s->reig = calloc((size_t)ncell * 75, sizeof(double));
  s->face = calloc((size_t)ncell * 48, sizeof(double));
  s->var1 = calloc((size_t)ncell, sizeof(double));
  return (s->flat && s->flatT && s->sig && s->lam && s->leig && s->reig && s->face && s->var1) ? 0 : 1;
}
static void scratch_free(scratch_t* s) {
  free(s->flat); free(s->flatT); free(s->sig); free(s->lam);
  free(s->leig); free(s->reig); free(s->face); free(s->var1);
}

/* primitive reconstruction vector of one cell (hy_getRiemannState.F90:475-495) */
static void load_prim(const ox_params* p, const double* U, const geom_t* g, long c, double V[9]) {
  V[H_DENS] = UV(U, p, (*g), OX_DENS, c);
  V[H_VELX] = UV(U, p, (*g), OX_VELX, c);
  V[H_VELY] = UV(U, p, (*g), OX_VELY, c);
  V[H_VELZ] = UV(U, p, (*g), OX_VELZ, c);
  V[H_PRES] = UV(U, p, (*g), OX_PRES, c);
  V[H_GAMC] = UV(U, p, (*g), OX_GAMC, c);
  V[H_GAME] = UV(U, p, (*g), OX_GAME, c);
  V[H_EINT] = UV(U, p, (*g), OX_EINT, c);
  V[H_GRAV] = 0.;
}

/* ------------------------------------------------------------------ */
/* hy_dataReconstOneStep.F90:82-1211 for one cell                        */
/* ------------------------------------------------------------------ */
static int reconst_one_step(const ox_params* p, const double* U, const geom_t* g,
                            int ix, int iy, int iz, double dt, const double del[3],
                            const double* flat, const int transOnly[3],
                            double Wp[3][9], double Wn[3][9], double sigc[3][9],
                            double lamc[3][5], double lec[3][5][5], double rec[3][5][5],
                            long long* fallbacks) {
  double hdtx = 0.5 * dt / del[0];                                  /* :268-283 */
  double hdty = p->ndim > 1 ? 0.5 * dt / del[1] : 0.;
  double hdtz = p->ndim > 2 ? 0.5 * dt / del[2] : 0.;
  int ci[3] = {ix, iy, iz};
  for (int d = 0; d < p->ndim; ++d) {                               /* :499-846 */
    int kOrder = transOnly[d] ? 1 : 2;
    double data[5][9];
    memset(data, 0, sizeof data);
    for (int s = -kOrder; s <= kOrder; ++s) {
      int c3[3] = {ci[0], ci[1], ci[2]};
      c3[d] += s;
      int slot = transOnly[d] ? s + 2 : s + 2;
      load_prim(p, U, g, CIDX(*g, c3[0], c3[1], c3[2]), data[slot]);
    }
    long c = CIDX(*g, ix, iy, iz);
    if (ox_ppm(p, d, dt, del[d], (const double(*)[9])data, flat[c * 3 + d], transOnly[d],
               lamc[d], lec[d], rec[d], Wp[d], Wn[d], sigc[d]))
      return 1;
  }
  if (p->ndim > 1) {                                                  /* :906-914 */
    for (int v = 0; v <= H_EINT; ++v) {
      Wn[0][v] = Wn[0][v] - hdty * sigc[1][v];
      Wp[0][v] = Wp[0][v] - hdty * sigc[1][v];
      Wn[1][v] = Wn[1][v] - hdtx * sigc[0][v];
      Wp[1][v] = Wp[1][v] - hdtx * sigc[0][v];
    }
    if (p->ndim > 2) {                                                /* :940-951 */
      for (int v = 0; v <= H_EINT; ++v) {
        Wn[0][v] = Wn[0][v] - hdtz * sigc[2][v];
        Wp[0][v] = Wp[0][v] - hdtz * sigc[2][v];
        Wn[1][v] = Wn[1][v] - hdtz * sigc[2][v];
        Wp[1][v] = Wp[1][v] - hdtz * sigc[2][v];
        Wn[2][v] = Wn[2][v] - hdtx * sigc[0][v] - hdty * sigc[1][v];
        Wp[2][v] = Wp[2][v] - hdtx * sigc[0][v] - hdty * sigc[1][v];
      }
    }
  }
  long c = CIDX(*g, ix, iy, iz);                                      /* :1120-1154 */
  double Vc[9];
  load_prim(p, U, g, c, Vc);
  for (int d = 0; d < p->ndim; ++d) {
    if (Wn[d][H_DENS] < 0. || Wp[d][H_DENS] < 0. || Wn[d][H_PRES] < 0. || Wp[d][H_PRES] < 0.) {
      if (!transOnly[d]) (*fallbacks)++;
      for (int v = 0; v <= H_EINT; ++v) { Wn[d][v] = Vc[v]; Wp[d][v] = Vc[v]; }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* hy_getRiemannState.F90:68-1476                                       */
/* ------------------------------------------------------------------ */
static int get_riemann_state(const ox_params* p, const double* U, const geom_t* g,
                             double dt, const double del[3], scratch_t* s, ox_stats* st) {
  const int nd = p->ndim;
  const int* lo = g->lo;
  const int* hi = g->hi;
  int ext[3] = {1, nd > 1, nd > 2};          /* per-dim stencil switch (k2/k3) */
  if (p->flattening) {                                               /* :356-454 */
    for (int k = lo[2] - 2 * ext[2]; k <= hi[2] + 2 * ext[2]; ++k)
      for (int j = lo[1] - 2 * ext[1]; j <= hi[1] + 2 * ext[1]; ++j)
        for (int i = lo[0] - 2; i <= hi[0] + 2; ++i) {
          long c = CIDX(*g, i, j, k);
          for (int d = 0; d < nd; ++d) {
            int o[3] = {0, 0, 0};
            o[d] = 1;
            long cp1 = CIDX(*g, i + o[0], j + o[1], k + o[2]);
            long cm1 = CIDX(*g, i - o[0], j - o[1], k - o[2]);
            long cp2 = CIDX(*g, i + 2 * o[0], j + 2 * o[1], k + 2 * o[2]);
            long cm2 = CIDX(*g, i - 2 * o[0], j - 2 * o[1], k - 2 * o[2]);
            double dp1 = (UV(U, p, (*g), OX_PRES, cp1) - UV(U, p, (*g), OX_PRES, cm1));
            double dp2 = (UV(U, p, (*g), OX_PRES, cp2) - UV(U, p, (*g), OX_PRES, cm2));
            double dv1 = UV(U, p, (*g), OX_VELX + d, cp1) - UV(U, p, (*g), OX_VELX + d, cm1);
            double presL = UV(U, p, (*g), OX_PRES, cm1), presR = UV(U, p, (*g), OX_PRES, cp1);
            double Sp = (fabs(dp2) > 1.e-15) ? dp1 / dp2 - 0.75 : 0.;
            double ft = fmax(0.0, fmin(1.0, 10.0 * Sp));
            if ((fabs(dp1) / fmin(presL, presR) < 1. / 3.) || dv1 > 0.) ft = 0.;
            s->flatT[c * 3 + d] = ft;
          }
        }
    for (int k = lo[2] - ext[2]; k <= hi[2] + ext[2]; ++k)
      for (int j = lo[1] - ext[1]; j <= hi[1] + ext[1]; ++j)
        for (int i = lo[0] - 1; i <= hi[0] + 1; ++i) {
          long c = CIDX(*g, i, j, k);
          for (int d = 0; d < nd; ++d) {
            int o[3] = {0, 0, 0};
            o[d] = 1;
            long cp1 = CIDX(*g, i + o[0], j + o[1], k + o[2]);
            long cm1 = CIDX(*g, i - o[0], j - o[1], k - o[2]);
            double dp1 = (UV(U, p, (*g), OX_PRES, cp1) - UV(U, p, (*g), OX_PRES, cm1));
            double fc;
            if (dp1 < 0.0) fc = fmax(s->flatT[c * 3 + d], s->flatT[cp1 * 3 + d]);
            else if (dp1 == 0.) fc = s->flatT[c * 3 + d];
            else fc = fmax(s->flatT[c * 3 + d], s->flatT[cm1 * 3 + d]);
            s->flat[c * 3 + d] = fc;
          }
        }
  }

  long ncell = g->ncell;
  for (int k = lo[2] - ext[2]; k <= hi[2] + ext[2]; ++k)             /* :470-901 */
    for (int j = lo[1] - ext[1]; j <= hi[1] + ext[1]; ++j)
      for (int i = lo[0] - 1; i <= hi[0] + 1; ++i) {
        int tr[3] = {0, 0, 0};                                       /* :553-591 */
        if (i > hi[0] + 1 || i < lo[0] - 1) tr[0] = 1;
        if (i > hi[0] || i < lo[0]) { tr[1] = 1; tr[2] = 1; }
        if (nd > 1) {
          if (j > hi[1] + 1 || j < lo[1] - 1) tr[1] = 1;
          if (j > hi[1] || j < lo[1]) { tr[0] = 1; tr[2] = 1; }
          if (nd > 2) {
            if (k > hi[2] + 1 || k < lo[2] - 1) tr[2] = 1;
            if (k > hi[2] || k < lo[2]) { tr[0] = 1; tr[1] = 1; }
          }
        }
        double Wp[3][9], Wn[3][9], sigc[3][9], lamc[3][5], lec[3][5][5], rec[3][5][5];
        memset(sigc, 0, sizeof sigc);
        if (reconst_one_step(p, U, g, i, j, k, dt, del, s->flat, tr, Wp, Wn, sigc, lamc, lec,
                             rec, &st->fallbacks)) {
          st->eigenErrors++;
          return 1;
        }
        long c = CIDX(*g, i, j, k);
        for (int d = 0; d < nd; ++d) {
          memcpy(&s->sig[(c * 3 + d) * 9], sigc[d], sizeof(double) * 9);
          memcpy(&s->lam[(c * 3 + d) * 5], lamc[d], sizeof(double) * 5);
          memcpy(&s->leig[(c * 3 + d) * 25], lec[d], sizeof(double) * 25);
          memcpy(&s->reig[(c * 3 + d) * 25], rec[d], sizeof(double) * 25);
          if (!tr[d]) {                                              /* :843-881 */
            double* f = &s->face[((long)d * ncell + c) * 16];
            for (int v = 0; v < 8; ++v) { f[v] = Wp[d][v]; f[8 + v] = Wn[d][v]; }
          }
        }
      }

  if (nd == 3 && p->use3dFullCTU) {                                  /* :952-1327 */
    double dt2dxdy6 = dt * dt / (6. * del[0] * del[1]);
    double dt2dydz6 = dt * dt / (6. * del[1] * del[2]);
    double dt2dzdx6 = dt * dt / (6. * del[2] * del[0]);
    for (int k = lo[2] - 2; k <= hi[2] + 2; ++k)
      for (int j = lo[1] - 2; j <= hi[1] + 2; ++j)
        for (int i = lo[0] - 2; i <= hi[0] + 2; ++i) {
          long c = CIDX(*g, i, j, k);
          double Vc[9];
          load_prim(p, U, g, c, Vc);
          /* (transverse sig, eig direction, neighbour step direction) per
           * state direction: X: (Z over j, Y eig), (Y over k, Z eig);
           * Y: (X over k, Z eig), (Z over i, X eig); Z: (Y over i, X eig),
           * (X over j, Y eig) */
          static const int tab[3][2][3] = {{{2, 1, 1}, {1, 2, 2}},
                                           {{0, 2, 2}, {2, 0, 0}},
                                           {{1, 0, 0}, {0, 1, 1}}};
          const double* fac[3] = {&dt2dydz6, &dt2dzdx6, &dt2dxdy6};
          int ci[3] = {i, j, k};
          for (int d = 0; d < 3; ++d) {
            if (!(ci[d] >= lo[d] - 1 && ci[d] <= hi[d] + 1)) continue;
            int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
            if (!(ci[o1] >= lo[o1] && ci[o1] <= hi[o1] && ci[o2] >= lo[o2] && ci[o2] <= hi[o2])) continue;
            double tf[2][9];
            for (int t = 0; t < 2; ++t) {
              int sd = tab[d][t][0], ed = tab[d][t][1], nbr = tab[d][t][2];
              int om[3] = {i, j, k}, op[3] = {i, j, k};
              om[nbr] -= 1; op[nbr] += 1;
              const double* sm = &s->sig[(CIDX(*g, om[0], om[1], om[2]) * 3 + sd) * 9];
              const double* sc = &s->sig[(c * 3 + sd) * 9];
              const double* sp = &s->sig[(CIDX(*g, op[0], op[1], op[2]) * 3 + sd) * 9];
              double le[5][5], re[5][5];
              memcpy(le, &s->leig[(c * 3 + ed) * 25], sizeof le);
              memcpy(re, &s->reig[(c * 3 + ed) * 25], sizeof re);
              ox_upwind_trans_flux(p, sm, sc, sp, &s->lam[(c * 3 + ed) * 5], le, re, tf[t]);
            }
            double* f = &s->face[((long)d * ncell + c) * 16];
            double fc = *fac[d];
            for (int v = 0; v < 8; ++v) {
              f[v] = f[v] + (tf[0][v] + tf[1][v]) * fc;
              f[8 + v] = f[8 + v] + (tf[0][v] + tf[1][v]) * fc;
            }
            if (f[H_DENS] <= 0. || f[8 + H_DENS] <= 0. || f[H_PRES] <= 0. || f[8 + H_PRES] <= 0.) {
              st->fallbacks++;                                        /* :1037-1051 */
              for (int v = 0; v < 8; ++v) { f[v] = Vc[v]; f[8 + v] = Vc[v]; }
            }
          }
        }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* hy_getFaceFlux.F90:53-1275 (Cartesian, hydro)                        */
/* ------------------------------------------------------------------ */
static int get_face_flux(const ox_params* p, const double* U, const geom_t* g, scratch_t* s,
                         double* flux[3], ox_stats* st) {
  const int nd = p->ndim;
  const int* lo = g->lo;
  const int* hi = g->hi;
  long ncell = g->ncell;
  for (int d = 0; d < nd; ++d) {
    fgeom_t fg = fgeom_of(p, d);
    int o[3] = {0, 0, 0};
    o[d] = 1;
    for (int k = lo[2]; k <= hi[2] + o[2]; ++k)
      for (int j = lo[1]; j <= hi[1] + o[1]; ++j)
        for (int i = lo[0]; i <= hi[0] + o[0]; ++i) {
          long c = CIDX(*g, i, j, k);
          long cm = CIDX(*g, i - o[0], j - o[1], k - o[2]);
          double VL[9], VR[9];
          const double* fl = &s->face[((long)d * ncell + cm) * 16];
          const double* fr = &s->face[((long)d * ncell + c) * 16];
          for (int v = 0; v < 8; ++v) { VL[v] = fl[v]; VR[v] = fr[8 + v]; }
          VL[8] = 0.; VR[8] = 0.;
          double F[8], speed;
          int ierr;
          if (p->riemann == OX_EXACT) ierr = ox_exact_riemann(p, d, VL, VR, F, &speed, st);
          else if (p->riemann == OX_HLL) ierr = ox_hll(p, d, VL, VR, F, &speed);
          else if (p->riemann == OX_HYBR &&                            /* :353-358 */
                   U[(long)p->shokVar * ncell + cm] + U[(long)p->shokVar * ncell + c] > 0.)
            ierr = ox_hll(p, d, VL, VR, F, &speed);
          else ierr = ox_hllc(p, d, VL, VR, F, &speed);
          if (ierr) { st->riemannErrors++; return 1; }                /* :1217 do_error */
          F[OX_F_EINT] = 0.; F[OX_F_VOLU] = 0.;
          if (p->useAuxEint) {                                        /* :433-442 */
            if (F[OX_F_DENS] > 0.) {
              F[OX_F_VOLU] = F[OX_F_DENS] / VL[H_DENS];
              F[OX_F_EINT] = F[OX_F_DENS] * VL[H_EINT];
            } else {
              F[OX_F_VOLU] = F[OX_F_DENS] / VR[H_DENS];
              F[OX_F_EINT] = F[OX_F_DENS] * VR[H_EINT];
            }
          }
          if (flux[d]) {
            long f = fidx(p, g, &fg, i, j, k);
            for (int q = 0; q < OX_NFLUX; ++q) flux[d][(long)q * fg.nf + f] = F[q];
          }
        }
    if (p->useAuxEint) {                     /* :507-561 (x), y and z analogues */
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i) {
            long c = CIDX(*g, i, j, k);
            const double* f = &s->face[((long)d * ncell + c) * 16];
            if (d == 0) s->var1[c] = 0.5 * (f[8 + H_PRES] + f[H_PRES]);
            else if (d == 1) s->var1[c] = s->var1[c] * 0.5 + (f[8 + H_PRES] + f[H_PRES]) * 0.25;
            else s->var1[c] = (s->var1[c] * 4.0 + (f[8 + H_PRES] + f[H_PRES])) / 6.0;
          }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* hy_unsplitUpdate.F90:73-1182 (UPDATE_ALL, Cartesian) +                */
/* hy_energyFix.F90:188-251 + Eos_wrapped(MODE_DENS_EI) on the interior  */
/* ------------------------------------------------------------------ */
static void unsplit_update(const ox_params* p, double* U, const geom_t* g,
                           double* const flux[3], const double* var1, double dt,
                           const double del[3]) {
  const int nd = p->ndim;
  const int* lo = g->lo;
  const int* hi = g->hi;
  double dx = del[0], dy = nd > 1 ? del[1] : 1., dz = nd > 2 ? del[2] : 1.;
  fgeom_t fg[3];
  for (int d = 0; d < 3; ++d) fg[d] = fgeom_of(p, d);
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        long c = CIDX(*g, i, j, k);
        double FL[8], FR[8], GL[8], GR[8], HL[8], HR[8];
        long f0 = fidx(p, g, &fg[0], i, j, k), f1 = fidx(p, g, &fg[0], i + 1, j, k);
        for (int q = 0; q < 8; ++q) {
          FL[q] = flux[0][(long)q * fg[0].nf + f0];
          FR[q] = flux[0][(long)q * fg[0].nf + f1];
          GL[q] = GR[q] = HL[q] = HR[q] = 0.0;
        }
        if (nd >= 2) {
          long g0 = fidx(p, g, &fg[1], i, j, k), g1 = fidx(p, g, &fg[1], i, j + 1, k);
          for (int q = 0; q < 8; ++q) {
            GL[q] = flux[1][(long)q * fg[1].nf + g0] * 1;
            GR[q] = flux[1][(long)q * fg[1].nf + g1] * 1;
          }
        }
        if (nd == 3) {
          long h0 = fidx(p, g, &fg[2], i, j, k), h1 = fidx(p, g, &fg[2], i, j, k + 1);
          for (int q = 0; q < 8; ++q) {
            HL[q] = flux[2][(long)q * fg[2].nf + h0] * 1;
            HR[q] = flux[2][(long)q * fg[2].nf + h1] * 1;
          }
        }
        double Sphys[5] = {0., 0., 0., 0., 0.};                       /* :496 */
        Sphys[1] = (FL[OX_F_PRES] - FR[OX_F_PRES]) * (1.0 - p->fP) / dx;
        if (nd >= 2) Sphys[2] = (GL[OX_F_PRES] - GR[OX_F_PRES]) * (1.0 - p->fP) / dy;
        if (nd == 3) Sphys[3] = (HL[OX_F_PRES] - HR[OX_F_PRES]) * (1.0 - p->fP) / dz;
        double dens0 = UV(U, p, (*g), OX_DENS, c);
        double IntEner = 0.;
        if (p->useAuxEint) {                                         /* :642-655, 1279-1298 */
          IntEner = dens0 * UV(U, p, (*g), OX_EINT, c);
          double pres = var1[c];
          IntEner = IntEner + dt / dx * (FL[OX_F_EINT] - FR[OX_F_EINT] + pres * (FL[OX_F_VOLU] - FR[OX_F_VOLU]));
          if (nd > 1) {
            IntEner = IntEner + dt / dy * (GL[OX_F_EINT] - GR[OX_F_EINT] + pres * (GL[OX_F_VOLU] - GR[OX_F_VOLU]));
            if (nd > 2)
              IntEner = IntEner + dt / dz * (HL[OX_F_EINT] - HR[OX_F_EINT] + pres * (HL[OX_F_VOLU] - HR[OX_F_VOLU]));
          }
          IntEner = IntEner + dt * 0.0;
        }
        double U0[5];                                                 /* :706-707 */
        U0[0] = dens0;
        U0[1] = UV(U, p, (*g), OX_VELX, c) * dens0;
        U0[2] = UV(U, p, (*g), OX_VELY, c) * dens0;
        U0[3] = UV(U, p, (*g), OX_VELZ, c) * dens0;
        U0[4] = dens0 * UV(U, p, (*g), OX_ENER, c);
        for (int q = 0; q < 5; ++q) {                                 /* :1187-1220 */
          U0[q] = U0[q] - dt / dx * (FR[q] - FL[q]);
          if (nd > 1) {
            U0[q] = U0[q] - dt / dy * (GR[q] - GL[q]);
            if (nd > 2) U0[q] = U0[q] - dt / dz * (HR[q] - HL[q]);
          }
          U0[q] = U0[q] + dt * (0. + Sphys[q]);
        }
        double densNew = fmax(U0[0], p->smlrho);                       /* :811 */
        UV(U, p, (*g), OX_DENS, c) = densNew;
        UV(U, p, (*g), OX_VELX, c) = U0[1] / densNew;                  /* :878 */
        UV(U, p, (*g), OX_VELY, c) = U0[2] / densNew;
        UV(U, p, (*g), OX_VELZ, c) = U0[3] / densNew;
        UV(U, p, (*g), OX_ENER, c) = U0[4] / densNew;
        if (p->useAuxEint) UV(U, p, (*g), OX_EINT, c) = IntEner / densNew;
      }
  /* hy_energyFix.F90:188-251 */
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        long c = CIDX(*g, i, j, k);
        double vx = UV(U, p, (*g), OX_VELX, c), vy = UV(U, p, (*g), OX_VELY, c), vz = UV(U, p, (*g), OX_VELZ, c);
        double ekin = 0.5 * (vx * vx + vy * vy + vz * vz);
        double eint = UV(U, p, (*g), OX_ENER, c) - ekin - 0.0;
        double aux = UV(U, p, (*g), OX_EINT, c);
        double newEint;
        if (!p->useAuxEint || eint > p->eintSwitch * (ekin + 0.0)) newEint = fmax(p->smallE, eint);
        else newEint = fmax(p->smallE, aux);
        UV(U, p, (*g), OX_ENER, c) = newEint + ekin;
        UV(U, p, (*g), OX_EINT, c) = newEint;
      }
  int elo[3] = {lo[0], lo[1], lo[2]}, ehi[3] = {hi[0], hi[1], hi[2]};
  ox_eos_box(p, U, 2, elo, ehi);
}

int ox_compute_fluxes(const ox_params* p, double* U, double* flux[3], double* pavg,
                      double dt, const double del[3], int update, ox_stats* st) {
  if (p->riemann == OX_HYBR && (p->shokVar < 0 || p->shokVar >= p->nunk))
    return -1;  /* hy_shockDetectBlk.F90:119-122: HYBR needs SHOK_VAR */
  geom_t g = geom_of(p);
  scratch_t s;
  ox_stats local;
  memset(&local, 0, sizeof local);
  if (!st) st = &local;
  if (scratch_alloc(&s, g.ncell)) return -1;
  double* fl[3] = {flux[0], flux[1], flux[2]};
  double* own[3] = {NULL, NULL, NULL};
  for (int d = 0; d < p->ndim; ++d)
    if (!fl[d] && update) { own[d] = malloc(sizeof(double) * OX_NFLUX * fgeom_of(p, d).nf); fl[d] = own[d]; }
  int rc = get_riemann_state(p, U, &g, dt, del, &s, st);
  if (!rc) rc = get_face_flux(p, U, &g, &s, fl, st);
  if (!rc && pavg) {
    long n = 0;
    for (int k = g.lo[2]; k <= g.hi[2]; ++k)
      for (int j = g.lo[1]; j <= g.hi[1]; ++j)
        for (int i = g.lo[0]; i <= g.hi[0]; ++i) pavg[n++] = s.var1[CIDX(g, i, j, k)];
  }
  if (!rc && update) unsplit_update(p, U, &g, fl, s.var1, dt, del);
  for (int d = 0; d < 3; ++d) free(own[d]);
  scratch_free(&s);
  return rc;
}

int ox_update_solution(const ox_params* p, double* U, double* const flux[3],
                       const double* pavg, double dt, const double del[3], ox_stats* st) {
  (void)st;
  geom_t g = geom_of(p);
  double* var1 = calloc((size_t)g.ncell, sizeof(double));
  if (!var1) return -1;
  long n = 0;
  for (int k = g.lo[2]; k <= g.hi[2]; ++k)
    for (int j = g.lo[1]; j <= g.hi[1]; ++j)
      for (int i = g.lo[0]; i <= g.hi[0]; ++i) var1[CIDX(g, i, j, k)] = pavg[n++];
  unsplit_update(p, U, &g, flux, var1, dt, del);
  free(var1);
  return 0;
}

int ox_compute_fluxes_batch(const ox_params* p, double* U, long nblocks, double dt,
                            const double del[3], int nthreads, ox_stats* st) {
  long ncell = ox_block_cells(p);
  int rc = 0;
  long long fb = 0, re = 0, ee = 0, xi = 0, xf = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fb, re, ee, xi, xf) reduction(| : rc)
#else
  (void)nthreads;
#endif
  for (long b = 0; b < nblocks; ++b) {
    ox_stats s;
    memset(&s, 0, sizeof s);
    double* fl[3] = {NULL, NULL, NULL};
    double* Ub = U + b * ncell * p->nunk;
    /* Hydro_init.F90:387-395 turns detection on for HYBR */
    if ((p->shockDetect || p->riemann == OX_HYBR) && ox_shock_detect(p, Ub) < 0) rc |= 1;
    rc |= ox_compute_fluxes(p, Ub, fl, NULL, dt, del, 1, &s);
    fb += s.fallbacks; re += s.riemannErrors; ee += s.eigenErrors;
    xi += s.exactIters; xf += s.exactFallbacks;
  }
  if (st) {
    st->fallbacks += fb; st->riemannErrors += re; st->eigenErrors += ee;
    st->exactIters += xi; st->exactFallbacks += xf;
  }
  return rc;
}

double ox_block_dt(const ox_params* p, const double* U, const double del[3], double cfl) {
  geom_t g = geom_of(p);
  double dt_temp = 0., tempCfl = cfl;
  for (int k = g.lo[2]; k <= g.hi[2]; ++k)
    for (int j = g.lo[1]; j <= g.hi[1]; ++j)
      for (int i = g.lo[0]; i <= g.hi[0]; ++i) {
        long c = CIDX(g, i, j, k);
        double snd2 = UV(U, p, g, OX_GAMC, c) * UV(U, p, g, OX_PRES, c) / UV(U, p, g, OX_DENS, c);
        double dtl = (fabs(UV(U, p, g, OX_VELX, c) - 0.) + sqrt(snd2)) * (1. / del[0]);
        if (p->ndim > 1) dtl = fmax(dtl, (fabs(UV(U, p, g, OX_VELY, c) - 0.) + sqrt(snd2)) * (1. / del[1]));
        if (p->ndim > 2) dtl = fmax(dtl, (fabs(UV(U, p, g, OX_VELZ, c) - 0.) + sqrt(snd2)) * (1. / del[2]));
        if (dtl * tempCfl > dt_temp * cfl) { dt_temp = dtl; tempCfl = cfl; }
      }
  return dt_temp != 0.0 ? tempCfl / dt_temp : 1.e300;
}

/* ------------------------------------------------------------------ */
/* flux register: Grid_addFineToFluxRegister.F90:128-171 (coef 1, 1/2,   */
/* 1/4 by NDIM; ratio 2), Grid_conserveFluxes.F90:95-108; arithmetic of  */
/* AMReX fluxregister fineadd / overwrite as restated in the header.     */
/* fineFlux/coarseFlux: 3 pointers per tile (tile*3 + dir).              */
/* ------------------------------------------------------------------ */
static long reg_face_index(const ox_params* p, int d, int a, int b, int c) {
  fgeom_t f = fgeom_of(p, d);
  return ((long)c * f.fy + b) * f.fx + a;
}

void ox_fluxreg_put(const ox_params* p, double* const* fineFlux, const ox_fluxreg_face* faces,
                    int nFaces, double* reg) {
  const double coef = p->ndim == 3 ? 0.25 : (p->ndim == 2 ? 0.5 : 1.0);
  for (int e = 0; e < nFaces; ++e) {
    const ox_fluxreg_face* fc = &faces[e];
    const int d = fc->dir;
    fgeom_t fg = fgeom_of(p, d);
    const double* F = fineFlux[fc->fine_tile * 3 + d];
    /* transverse offset axes: outer = higher axis, inner = lower axis */
    int inner = d == 0 ? 1 : 0, outer = d == 2 ? 1 : 2;
    int nIn = p->ndim > 1 ? 2 : 1, nOut = p->ndim > 2 ? 2 : 1;
    for (int q = 0; q < OX_NFLUX; ++q) {
      double acc = 0.0;
      for (int o = 0; o < nOut; ++o)
        for (int in = 0; in < nIn; ++in) {
          int idx[3] = {fc->fi, fc->fj, fc->fk};
          idx[inner] += in;
          idx[outer] += o;
          acc = acc + coef * F[(long)q * fg.nf + reg_face_index(p, d, idx[0], idx[1], idx[2])];
        }
      reg[(long)q * nFaces + e] = acc;
    }
  }
}

void ox_fluxreg_conserve(const ox_params* p, double* const* coarseFlux,
                         const ox_fluxreg_face* faces, int nFaces, const double* reg) {
  for (int e = 0; e < nFaces; ++e) {
    const ox_fluxreg_face* fc = &faces[e];
    const int d = fc->dir;
    fgeom_t fg = fgeom_of(p, d);
    double* F = coarseFlux[fc->coarse_tile * 3 + d];
    long f = reg_face_index(p, d, fc->ci, fc->cj, fc->ck);
    for (int q = 0; q < OX_NFLUX; ++q) F[(long)q * fg.nf + f] = reg[(long)q * nFaces + e] * 1.0;
  }
}
