/*
 * hydro_oracle.h -- CPU FP64 restatement of FLASH5's unsplit hydro block
 * update (PPM + characteristic limiting + flattening + CTU + HLLC + aux
 * internal-energy update + energy fix + ideal gamma-law EOS).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in flash5_amd/; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product never links or calls it.
 *
 * Every routine follows the reference file:line cited in hydro_oracle.c
 * (paths relative to source/physics/Hydro/HydroMain/unsplit unless stated).
 * The operation order of each expression follows the Fortran source
 * (left-to-right evaluation, no contraction) so that the restatement is
 * deterministic and reproducible; build with -ffp-contract=off.
 */
#ifndef HYDRO_ORACLE_H
#define HYDRO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Cell-centred variable slots used by the path (a block stores them
 * var-major: U[var][k][j][i], i fastest, guard cells included).  The slot
 * numbers below are the oracle's canonical map; ox_params.var lets a caller
 * use any other map (e.g. the one FLASH's generated Flash.h defines). */
enum { OX_DENS = 0, OX_VELX, OX_VELY, OX_VELZ, OX_PRES, OX_ENER, OX_GAMC,
       OX_GAME, OX_TEMP, OX_EINT, OX_NVAR_USED };

/* Face flux slots (OnlyHydro/Config:6-13, UHD.h:405-412). */
enum { OX_F_DENS = 0, OX_F_XMOM, OX_F_YMOM, OX_F_ZMOM, OX_F_ENER, OX_F_PRES,
       OX_F_EINT, OX_F_VOLU, OX_NFLUX };

/* Slope limiters, numbered as UHD.h:15-19. */
enum { OX_MINMOD = 1, OX_MC = 2, OX_HYBRID = 3, OX_VANLEER = 4, OX_LIMITED = 5 };

/* Riemann solvers, numbered as UHD.h:59-66 (HLL, HLLC, HYBR = HLL at
 * flagged shocks and HLLC elsewhere); OX_EXACT is the exact two-shock
 * iterative solver, which the reference does not contain. */
enum { OX_ROE = 1, OX_HLL = 3, OX_HLLC = 4, OX_HYBR = 8, OX_EXACT = 20 };

typedef struct {
  int ndim;             /* 1, 2 or 3 */
  int nxb, nyb, nzb;    /* interior cells per block (1 for absent dims) */
  int ng;               /* guard cells per side (NGUARD = 4) */
  int nunk;             /* number of variables stored per cell */
  int var[OX_NVAR_USED];/* slot of each used variable in U's var axis */
  double gamma;         /* eos_gamma */
  int limiter;          /* OX_MINMOD / OX_MC / OX_VANLEER / OX_HYBRID / OX_LIMITED */
  double limitedSlopeBeta;
  int charLimiting;     /* 1 = characteristic limiting */
  int flattening;       /* use_flattening */
  int steepening;       /* use_steepening */
  int use3dFullCTU;     /* use_3dFullCTU */
  int useAuxEint;       /* use_auxEintEqn */
  int riemann;          /* OX_HLL, OX_HLLC, OX_HYBR or OX_EXACT */
  double fP;            /* hy_fPresInMomFlux */
  double smlrho, smallp, smallE, eintSwitch;
  double tiny;          /* hy_tiny (Hydro_data.F90:65) */
  double gasConstant;   /* "ideal gas constant" (PhysicalConstants_init.F90:175) */
  double singleSpeciesA;/* eos_singleSpeciesA */
  /* exact solver controls */
  int exactMaxIter;
  double exactTol;
  /* shock detection (hy_shockDetectBlk.F90): shockDetect as the runtime
   * parameter; shokVar = slot of SHOK_VAR in U's var axis, -1 if absent */
  int shockDetect;
  int shokVar;
} ox_params;

typedef struct {
  long long fallbacks;      /* cells reset to first order (positivity) */
  long long riemannErrors;  /* imaginary sound speed in the Riemann solver */
  long long eigenErrors;    /* imaginary sound speed in the eigen system */
  long long exactIters;     /* total Newton iterations of the exact solver */
  long long exactFallbacks; /* exact-solver interfaces that did not converge */
} ox_stats;

/* Default knobs: unsplit/Config defaults with order=3 (PPM), vanLeer,
 * charLimiting on, flattening on, steepening off, 3D full CTU on,
 * aux eint on, HLLC, fP=1, small*=1e-10, eintSwitch=0, tiny=1e-32. */
void ox_default_params(ox_params* p, int ndim, int nb, double gamma);

/* Cell count of one block including guard cells. */
long ox_block_cells(const ox_params* p);
/* Face count of one direction's flux array (interior faces only). */
long ox_flux_faces(const ox_params* p, int dir);

/* Ideal gamma-law EOS over a box of cells (0-based, guard-inclusive index
 * ranges lo..hi inclusive).  mode 1 = MODE_DENS_PRES, 2 = MODE_DENS_EI.
 * Follows physics/Eos/EosMain/Eos_getData.F90, Gamma/eos_idealGamma.F90,
 * Eos_putData.F90. */
void ox_eos_box(const ox_params* p, double* U, int mode, const int lo[3], const int hi[3]);

/* hy_computeFluxes (hy_computeFluxes.F90:76-406) on one block.
 *   U      : block, updated in place when update != 0 (Uout => Uin)
 *   flux[d]: face flux densities per direction d (may be NULL to skip
 *            storing); layout F[f][k][j][i] over interior faces
 *   pavg   : per-interior-cell face-averaged pressure for the aux eint
 *            equation (scrch_Ptr HY_VAR1), written when non-NULL
 * Returns 0 on success, nonzero on an error the reference would abort on. */
int ox_compute_fluxes(const ox_params* p, double* U, double* flux[3],
                      double* pavg, double dt, const double del[3], int update,
                      ox_stats* st);

/* hy_updateSolution (hy_updateSolution.F90:104-369): conservative update
 * from stored (possibly corrected) fluxes plus the saved pavg, then
 * energy fix and EOS. */
int ox_update_solution(const ox_params* p, double* U, double* const flux[3],
                       const double* pavg, double dt, const double del[3],
                       ox_stats* st);

/* Batched driver over independent blocks, OpenMP over blocks
 * (nthreads <= 0 uses the OpenMP default).  With p->shockDetect each block
 * first runs ox_shock_detect (Hydro.F90:94-110; the blocks have no guard
 * fill, so their SHOK guard cells stay 0). */
int ox_compute_fluxes_batch(const ox_params* p, double* U, long nblocks,
                            double dt, const double del[3], int nthreads,
                            ox_stats* st);

/* CFL time step of one block (unsplit/Hydro_computeDt.F90:186-282). */
double ox_block_dt(const ox_params* p, const double* U, const double del[3], double cfl);

/* ---- per-stage entry points (used by the pinning tests) ---- */
/* V = (dens, vx, vy, vz, pres, gamc, game, eint, grav) */
int ox_eigen(const double V[9], int dir, double lambda[5], double leig[5][5],
             double reig[5][5]);
double ox_limiter(int kind, double a, double b, double beta);
void ox_tvd_slope(const ox_params* p, const double VL[9], const double V0[9],
                  const double VR[9], const double lambda[5],
                  double leig[5][5], double delbar[9]);
void ox_upwind_trans_flux(const ox_params* p, const double vm1[9],
                          const double vc0[9], const double vp1[9],
                          const double lambda[5], double leig[5][5],
                          double reig[5][5], double sig[9]);
/* One PPM reconstruction (hy_DataReconstructNormalDir_PPM.F90:73-958).
 * data: 5 cells x 9 prims (Vmm, Vm, Vc, Vp, Vpp); trans != 0 uses only
 * data[1..3].  Returns nonzero on an eigen-system error. */
int ox_ppm(const ox_params* p, int dir, double dt, double delta,
           const double data[5][9], double flatCoeff, int trans,
           double lambda[5], double leig[5][5], double reig[5][5],
           double Wp[9], double Wm[9], double sig[9]);
/* HLLC flux (hy_HLLC.F90:53-324): Fstar[0..5]; returns ierr. */
int ox_hllc(const ox_params* p, int dir, const double VL[9], const double VR[9],
            double Fstar[6], double* speed);
/* HLL flux (hy_HLL.F90:82-165): Fstar[0..5]; returns ierr. */
int ox_hll(const ox_params* p, int dir, const double VL[9], const double VR[9],
           double Fstar[6], double* speed);
/* Shock flag of one block (hy_shockDetectBlk.F90:93-229, Balsara & Spicer
 * 1999): SHOK_VAR := 0 over the guard-inclusive box, then 1 on interior
 * cells where the undivided pressure jump reaches 0.5 of the local minimum
 * pressure and the undivided velocity divergence is at most -0.1 of the
 * local minimum sound speed.  Reads dens, pres, gamc and velocities on the
 * interior and one guard layer.  Returns the number of flagged cells, or
 * -1 when p->shokVar < 0. */
long ox_shock_detect(const ox_params* p, double* U);
/* Exact two-shock Riemann solver with the same contract. */
int ox_exact_riemann(const ox_params* p, int dir, const double VL[9],
                     const double VR[9], double Fstar[6], double* speed,
                     ox_stats* st);
/* Exact-solver star pressure for one interface (gamma-law), returns iters
 * (negative when not converged); pstar/ustar out.  guess <= 0 uses the
 * two-shock default guess. */
int ox_exact_pstar(double gamma, double rl, double ul, double pl, double rr,
                   double ur, double pr, double guess, double tol, int maxit,
                   double* pstar, double* ustar);

/* Exact solution (rho, u, p) at x/t = S of a Riemann problem (Toro sec.
 * 4.5); returns the Newton iterations spent on p*. */
int ox_exact_sample(double gamma, double rl, double ul, double pl, double rr,
                    double ur, double pr, double S, double out[3]);

/* ---- flux register (Cartesian, refinement ratio 2) ----
 * Grid_putFluxData -> Grid_addFineToFluxRegister(zeroFullRegister)
 * (Grid/GridMain/AMR/Amrex/Grid_addFineToFluxRegister.F90:128-171) and
 * Grid_conserveFluxes (Grid_conserveFluxes.F90:95-108).  The register
 * arithmetic lives in AMReX (amrex_fluxregister fineadd / overwrite, not
 * vendored): the register is zeroed, each covering fine face flux is added
 * times coef = 1, 1/2, 1/4 (1-D, 2-D, 3-D), the fine faces taken in order of
 * ascending transverse offset with the higher axis outermost; overwrite
 * copies the register into the coarse flux array.  One entry per copy of a
 * coarse face (a face shared by two coarse blocks is listed for each). */
typedef struct {
  int coarse_tile, fine_tile, dir;
  int ci, cj, ck;   /* face in the coarse tile's flux[dir] array */
  int fi, fj, fk;   /* first covering face in the fine tile's flux[dir] array */
} ox_fluxreg_face;
/* reg[q * nFaces + e]; fine[t][d] is tile t's flux array for direction d */
void ox_fluxreg_put(const ox_params* p, double* const* fineFlux, const ox_fluxreg_face* faces,
                    int nFaces, double* reg);
void ox_fluxreg_conserve(const ox_params* p, double* const* coarseFlux,
                         const ox_fluxreg_face* faces, int nFaces, const double* reg);

#ifdef __cplusplus
}
#endif
#endif
