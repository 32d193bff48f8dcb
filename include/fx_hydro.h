/*
 * fx_hydro.h -- C ABI of the MI355X-native unsplit hydro block update.
 *
 * Drop-in boundary for FLASH5's unsplit Hydro on the ideal gamma-law EOS.
 * The reference entry points this ABI replaces (paths relative to
 * source/physics/Hydro):
 *   Hydro(simTime, dt, dtOld)         Hydro_interface.F90:52-57,
 *                                     HydroMain/unsplit/Hydro.F90:6-343
 *   hy_computeFluxes(tile, Uin, Uout, del, timeEndAdv, dt, dtOld, sweep)
 *                                     HydroMain/unsplit/hy_computeFluxes.F90:76
 *   hy_updateSolution(tile, Uin, Uout, del, timeEndAdv, dt, dtOld, sweep)
 *                                     HydroMain/unsplit/hy_updateSolution.F90:104
 *   Hydro_init / Hydro_finalize       HydroMain/unsplit/Hydro_init.F90:24
 *   Hydro_computeDt                   HydroMain/unsplit/Hydro_computeDt.F90:186
 *   Grid_putFluxData / Grid_conserveFluxes (flux-register correction,
 *   Cartesian, refinement ratio 2)    Grid/GridMain/AMR/Amrex/
 *                                     Grid_addFineToFluxRegister.F90:144-171,
 *                                     Grid_conserveFluxes.F90:108
 *
 * Layout (the index-reordered AMReX layout, Grid_tile.F90:285): a tile's
 * cell data U(i,j,k,var) in Fortran order over the guard-inclusive box,
 * i.e. U[var][k][j][i] with i fastest.  Face flux densities F(i,j,k,flux)
 * per direction over the interior faces only (extent nxb+1 along the face
 * normal), NFLUXES = 8 in the order dens, xmom, ymom, zmom, ener, pres,
 * eint, volu (OnlyHydro/Config:6-13).  All pointers passed to compute
 * entry points are device pointers (hipMalloc'd or managed memory); the
 * library never frees caller memory.
 *
 * Errors: every entry returns 0 on success.  On failure the fx_status
 * carries the first failing (tile, i, j, k, dir, code), where the Fortran
 * host calls Driver_abortFlash as the reference would
 * (hy_getFaceFlux.F90:1217, hy_eigenParameters.F90:84).
 */
#ifndef FX_HYDRO_H
#define FX_HYDRO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* variable slots inside fx_params.var (FLASH's *_VAR indices, 0-based) */
enum { FX_DENS = 0, FX_VELX, FX_VELY, FX_VELZ, FX_PRES, FX_ENER, FX_GAMC,
       FX_GAME, FX_TEMP, FX_EINT, FX_NVAR_USED };
enum { FX_NFLUX = 8 };

/* slope limiters (UHD.h:15-19) */
enum { FX_MINMOD = 1, FX_MC = 2, FX_HYBRID = 3, FX_VANLEER = 4, FX_LIMITED = 5 };
/* Riemann solvers numbered as UHD.h:59-66: FX_HLL, FX_HLLC and FX_HYBR (HLL
 * at faces next to a flagged shock cell, HLLC elsewhere,
 * hy_getFaceFlux.F90:353-367); FX_EXACT = exact iterative solver */
enum { FX_ROE = 1, FX_HLL = 3, FX_HLLC = 4, FX_HYBR = 8, FX_EXACT = 20 };
/* precision modes (value 1 is unused; fx_hydro_init rejects it) */
enum { FX_PREC_FP64 = 0,      /* plain FP64, IEEE ops without contraction */
       FX_PREC_MIXED = 2 };   /* FP32 reconstruction and fluxes (HLLC, HLL
                                 and the exact solver, whose face states are
                                 stored in FP32); FP64 update, e_int
                                 recovery, EOS and the exact solver's Newton
                                 iterate and convergence test (tolerance
                                 max(exactTol, 1e-5)); 1-D blocks run FP64 */

/* Runtime knobs read by Hydro_init (Hydro_init.F90:60-420) and the EOS
 * (Eos_init.F90:66-73, eos_initGamma.F90:42). */
typedef struct {
  int ndim;                 /* NDIM: 1, 2, 3 */
  int nxb, nyb, nzb;        /* NXB, NYB, NZB (1 for absent dims) */
  int ng;                   /* NGUARD; must be 4 */
  int nunk;                 /* NUNK_VARS: variables stored per cell */
  int var[FX_NVAR_USED];    /* 0-based slot of DENS, VELX, ... in U's var axis */
  double gamma;             /* gamma */
  int limiter;              /* slopeLimiter */
  double limitedSlopeBeta;  /* LimitedSlopeBeta */
  int charLimiting;         /* charLimiting */
  int flattening;           /* use_flattening */
  int steepening;           /* use_steepening */
  int use3dFullCTU;         /* use_3dFullCTU */
  int useAuxEint;           /* use_auxEintEqn */
  int riemann;              /* RiemannSolver: FX_HLL, FX_HLLC, FX_HYBR, FX_EXACT */
  double fP;                /* hy_fPresInMomFlux */
  double smlrho, smallp, smallE, eintSwitch;
  double tiny;              /* hy_tiny */
  double gasConstant;       /* ideal gas constant */
  double singleSpeciesA;    /* eos_singleSpeciesA */
  int exactMaxIter;         /* Newton iteration cap of the exact solver */
  double exactTol;          /* relative p* convergence tolerance */
  int precision;            /* FX_PREC_* */
  int learnedSeed;          /* 1: seed exact-solver Newton from the MLP */
  /* shock detection (Hydro_init.F90:117-124, 387-395) */
  int shockDetect;          /* shockDetect; on whenever riemann = FX_HYBR */
  int shockLowerCFL;        /* shockLowerCFL; turns shockDetect on */
  int shokVar;              /* 0-based slot of SHOK_VAR in U's var axis (the
                               unsplit Config's "shok"); read only while
                               shock detection is on */
  /* CFL after the positivity fallback (unsplit/Config:137-138) */
  int fallbackLowerCFL;     /* hy_fallbackLowerCFL */
  double cflFallbackFactor; /* hy_cflFallbackFactor (0.9 in the Config);
                               must be > 0 while fallbackLowerCFL is on */
} fx_params;

/* One tile as the reference's tile iterator yields it with tiling off
 * (Grid_tile.F90; gr_useTiling defaults to .FALSE., Grid/GridMain/Config:183):
 * a whole block.  There are no tile limits: the library computes every
 * interior cell of the block, so a caller iterating with tiling on must still
 * pass each block once. */
typedef struct {
  double* U;                /* U[var][k][j][i], guard-inclusive */
  double* flux[3];          /* flux densities per direction, or NULL */
  double* pavg;             /* per-interior-cell face-averaged pressure for
                               the aux internal-energy update (the
                               reference's scrch_Ptr HY_VAR1), or NULL */
  double del[3];            /* cell widths (tileDesc%deltas) */
  int level;                /* refinement level, 1-based */
  int block_class;          /* index into fx_hydro_set_block_classes's list;
                               ignored (0) while no classes are set */
} fx_tile;

/* Status of one call.  The graph-captured runs (fx_grid_run_graph,
 * fx_amr_run_graph) report one status for the whole run: the first error of
 * any step with its location (the captured steps still run to the end of the
 * last replay; the host then aborts, as the reference would), and the
 * fallback and Newton counters summed over every step of every replay. */
typedef struct {
  int code;                 /* 0 ok, 1 eigen (imaginary c), 2 Riemann error,
                               3 bad argument, 4 HIP error */
  int first_tile, first_i, first_j, first_k, first_dir;
  long long fallbacks;      /* cells reset to first-order states */
  long long newton_iters;   /* exact solver: total Newton iterations */
  long long newton_fallbacks; /* exact solver: learned seeds that failed */
  double min_cell_cfl;      /* calls that run the detector
                               (fx_hydro_shock_detect*, fx_amr_step, the
                               graph-captured runs' last step): the CFL cap
                               0.60, 0.45, 0.25 (NDIM 1, 2, 3) when
                               shockLowerCFL is on and a cell was flagged
                               (the reference sets hy_cfl = min(hy_cfl, cap),
                               hy_shockDetectBlk.F90:211-227); else 0 */
  double fallback_cfl;      /* calls that compute fluxes, NDIM 2 and 3, with
                               fallbackLowerCFL on: cflFallbackFactor / NDIM
                               when a state fell back to first order
                               (fallbacks > 0; graph-captured runs: in any
                               step), which the reference applies after the
                               fluxes as hy_cfl = min(hy_cfl, that)
                               (hy_dataReconstOneStep.F90:1125-1129,
                               hy_getRiemannState.F90:934-941, 1303-1319,
                               1473); else 0 */
} fx_status;

/* Hydro_init: set knobs, allocate nothing (scratch grows on first use). */
int fx_hydro_init(const fx_params* p);
/* Hydro_finalize: free library-owned scratch. */
void fx_hydro_finalize(void);

/* hy_computeFluxes over nTiles tiles of the configured block shape.
 * updateInPlace[t] != 0 (or updateInPlace == NULL) updates U of tile t in
 * place (the reference's Uout => Uin case, hy_computeFluxes.F90:334-378):
 * conservative update, energy fix and EOS; otherwise only fluxes and pavg
 * are produced, for hy_updateSolution after flux correction. */
int fx_hydro_compute_fluxes(const fx_tile* tiles, int nTiles, double dt,
                            double dtOld, const unsigned char* updateInPlace,
                            fx_status* st);

/* hy_updateSolution: update from each tile's (corrected) flux arrays and
 * pavg, then energy fix and EOS. */
int fx_hydro_update_solution(const fx_tile* tiles, int nTiles, double dt,
                             double dtOld, fx_status* st);

/* hy_shockDetect (hy_shockDetect.F90, hy_shockDetectBlk.F90:93-229) over
 * the tiles: SHOK_VAR := 0 over each guard-inclusive box, then 1 on the
 * interior cells where the undivided pressure jump reaches 0.5 of the 3^d
 * minimum pressure and the undivided velocity divergence is at most -0.1 of
 * the 3^d minimum sound speed.  Reads dens, pres, gamc and the velocities on
 * the interior and the first guard layer, which the host has filled (the
 * reference's first guard fill, Hydro.F90:102-104).  The host then fills the
 * guard cells (SHOK_VAR included) before fx_hydro_compute_fluxes.  Returns 3
 * when shock detection is off.  Shock detection needs
 * 2 (nxb+2)(nyb+2) + 4 nxb nyb doubles of LDS in 3-D (2 (nxb+2)(nyb+2) in
 * 1-D and 2-D) within 64 KiB, i.e. x-y block extents up to 36 x 36;
 * fx_hydro_init and fx_hydro_set_block_classes refuse larger ones. */
int fx_hydro_shock_detect(const fx_tile* tiles, int nTiles, fx_status* st);

/* fx_hydro_shock_detect over a strided batch (block b's U starts at
 * U + b*blockStride doubles), no host-side tile list. */
int fx_hydro_shock_detect_batch(double* U, long nBlocks, long blockStride,
                                const double del[3], fx_status* st);

/* Strided batch of independent blocks: block b's U starts at
 * U + b*blockStride (doubles).  Equivalent to fx_hydro_compute_fluxes with
 * updateInPlace all 1 and no flux output; no host-side tile list. */
int fx_hydro_advance_batch(double* U, long nBlocks, long blockStride,
                           const double del[3], double dt, fx_status* st);

/* fx_hydro_advance_batch without any allocation.  No counterpart in the
 * reference: it serves callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and bits as fx_hydro_advance_batch,
 * for the HLLC, HLL and hybrid solvers on 2-D and 3-D blocks whose interior
 * x and y extents are multiples of 8; other configurations return 3.  Its
 * scratch (status record, tile table, face fluxes and face-averaged
 * pressure: 848 KiB per 16^3 block, 112 KiB per 8^3 block) lives in a fixed
 * 64 MiB device array that each GPU holds until the process ends, so one
 * call takes at most 77 blocks of 16^3 or 584 of 8^3; a larger call
 * returns 3 with a message on stderr.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other host threads or
 * streams) race.  fx_hydro_init must have run. */
int fx_hydro_advance_batch_noalloc(double* U, long nBlocks, long blockStride,
                                   const double del[3], double dt, fx_status* st);

/* Hydro_computeDt over a strided batch (device-side min reduction):
 * dt = cfl / max over interior cells of max_d (|v_d| + c) / del_d. */
int fx_hydro_compute_dt(const double* U, long nBlocks, long blockStride,
                        const double del[3], double cfl, double* dtOut);

/* ---- flux-register correction (Cartesian, refinement ratio 2) ----
 * One entry per copy of a coarse face on a fine-coarse boundary: a face
 * shared by two coarse blocks is listed once for each block's flux array.
 * Indices are 0-based positions in the tiles' interior-face flux arrays
 * (extent nxb+1 along the face normal).  The 2^(NDIM-1) fine faces covering
 * the coarse face start at (fi, fj, fk) and step +1 along the transverse axes
 * of the same fine tile. */
typedef struct {
  int coarse_tile, fine_tile;  /* indices into the coarse / fine tile lists */
  int dir;                     /* face normal, 0 .. ndim-1 */
  int ci, cj, ck;              /* face in the coarse tile's flux[dir] array */
  int fi, fj, fk;              /* first covering face in the fine tile's flux[dir] */
} fx_fluxreg_face;

/* Grid_putFluxData(level) for the fine level: reg := 0, then
 * reg += coef * each covering fine flux, coef = 1, 1/2, 1/4 in 1-, 2-, 3-D
 * (Grid_putFluxData.F90:75 -> Grid_addFineToFluxRegister.F90:128-171,
 * AMReX fluxregister fineadd).  reg is a device array reg[q*nFaces + e] of
 * FX_NFLUX*nFaces doubles; the fine tiles' flux arrays are read. */
int fx_fluxreg_put(const fx_tile* fine, int nFine, const fx_fluxreg_face* faces,
                   int nFaces, double* reg, fx_status* st);
/* Grid_conserveFluxes(ALLDIR, level) for the coarse level: each listed
 * coarse face flux := reg (Grid_conserveFluxes.F90:95-108, AMReX
 * fluxregister overwrite).  Writes the coarse tiles' flux arrays. */
int fx_fluxreg_conserve(const fx_tile* coarse, int nCoarse, const fx_fluxreg_face* faces,
                        int nFaces, const double* reg, fx_status* st);

/* ---- learned Newton seed of the exact solver (no counterpart in the
 * reference; for callers that train or use the seed network) ----
 * Network: 6 -> 64 -> 64 -> 1, tanh, BF16 weights on MFMA with FP32
 * accumulation.  params: FX_MLP_NPARAM FP32 values on the device, in the
 * order W1[64][6], b1[64], W2[64][64] (out x in), b2[64], w3[64], b3.
 * Features of an interface with P0 = sqrt(pL pR), R0 = sqrt(rhoL rhoR),
 * C0 = sqrt(gamma P0/R0), ubar = (uL+uR)/2: (ln rhoL/R0, ln pL/P0,
 * (uL-ubar)/C0, ln rhoR/R0, ln pR/P0, (uR-ubar)/C0); target ln(pstar/P0).
 * The seed P0 exp(y) is considered where some feature has magnitude at or
 * above the gate (fx_mlp_set_gate, default 0.2); Newton then starts from
 * whichever of the seed and the two-shock guess has the smaller residual of
 * the pressure function.  Weaker interfaces start from the two-shock guess. 
 * states: device SoA [6][n] = rhoL, uL, pL, rhoR, uR, pR. */
enum { FX_MLP_NPARAM = 4673 };
int fx_mlp_features(const double* states, long n, double gamma, float* x /* [n][6] */);
int fx_mlp_forward(const float* params, const float* x, long n, float* y);
/* grad += d/dparams of scale*sum_n (y_n - t_n)^2; *loss += the same sum */
int fx_mlp_grad(const float* params, const float* x, const float* t, long n, float scale,
                float* grad, float* loss);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on FP32 master parameters */
int fx_mlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr);
/* Batched exact Riemann problems: pstar by FP64 Newton from the two-shock guess
 * (params == NULL) or from the network's seed; a seeded start that does not
 * converge restarts from the two-shock guess (counted in
 * st->newton_fallbacks).  iters[i]: Newton iterations spent on interface i,
 * negative when even the restart did not converge. */
int fx_riemann_batch(const double* states, long n, double gamma, const float* params,
                     int maxit, double tol, double* pstar, int* iters, fx_status* st);
int fx_mlp_set_gate(double gate);
/* Network used when fx_params.learnedSeed = 1 and riemann = FX_EXACT
 * (NULL clears it).  The seeded step needs the fused kernel path (2-D, 3-D). */
int fx_hydro_set_seed_params(const float* params);
/* The face states of the last exact-solver step as Riemann problems in the
 * fx_riemann_batch layout (u = face-normal velocity), for training the seed
 * on states the solver itself produced (FP32 values after an FX_PREC_MIXED
 * step, whose face states are FP32).  *n receives the face count; returns
 * nonzero when cap < *n or no such step ran. */
int fx_hydro_face_states(double* states, long cap, long* n);

/* ---- guard cells on one level of equal blocks ----
 * Grid_fillGuardCells(CENTER, ALLDIR) for a uniform grid of independent
 * block storage (Grid/GridMain/UG/Grid_fillGuardCells.F90): every variable
 * of every guard cell is copied from the neighbour block's interior, or set
 * by the boundary condition.  nbr: host array [nBlocks][6] in the order
 * -x, +x, -y, +y, -z, +z holding a block index or one of the codes below;
 * it is checked, then kept on the device until it changes.  Blocks need at
 * least NGUARD interior cells per axis. */
enum { FX_BC_OUTFLOW = -1,   /* zero gradient */
       FX_BC_REFLECT = -2 }; /* mirror, normal velocity negated */
int fx_grid_fill_guards(double* U, long nBlocks, long blockStride, const int* nbr,
                        fx_status* st);

/* Whole steps of a uniform grid captured in a HIP graph: each captured step
 * fills the guard cells (as fx_grid_fill_guards), computes the CFL dt on
 * the device (as fx_hydro_compute_dt, kept in device memory), and advances
 * every block (as fx_hydro_advance_batch with that dt).  stepsPerGraph steps
 * are captured once and the graph is replayed nReplays times; *tOut gets the
 * elapsed simulation time (the clock starts at 0).  With shock detection on,
 * each step fills, flags and fills SHOK_VAR's guard cells (Hydro.F90:94-123);
 * under shockLowerCFL a step after one with a flagged cell takes CFL
 * min(cfl, 1.2 * cap) (Hydro.F90:104-141).  Under fallbackLowerCFL a step
 * with a fallback lowers the CFL to at most cflFallbackFactor / NDIM, and
 * each later step raises it by 1.2 up to cfl (Hydro.F90:128-141; reset to
 * cfl before detection when the detector runs).  The CFL is carried on the
 * device across steps and replays, starting at cfl.  Not with learnedSeed. */
int fx_grid_run_graph(double* U, long nBlocks, long blockStride, const int* nbr,
                      const double del[3], double cfl, int stepsPerGraph, int nReplays,
                      double* tOut, fx_status* st);

/* Work on the given HIP stream (hipStream_t as void*); NULL = default. */
void fx_hydro_set_stream(void* stream);
/* 1: entry points return without synchronising (needed for graph capture). */
void fx_hydro_set_async(int async_);
/* Per-stage timing: when on, each compute call records HIP events on the
 * library's stream around its kernels; fx_hydro_stage_times copies the last
 * call's durations in ms (states, fluxes, update) and returns the count. */
void fx_hydro_set_timing(int on);
int fx_hydro_stage_times(float* ms, int n);
/* Kernel path: 0 = fused column-march kernel where the block shape allows
 * (default; batch steps of 3-D blocks up to 16 x 16 across with the exact
 * solver also run its face solve and the conservative update as one
 * kernel), 1 = three-kernel path (states, fluxes, update), 2 = as 0 but the
 * exact solver's batch steps keep the update as a kernel of its own. */
void fx_hydro_set_kernel_path(int path);

/* Block classes: tiles of several interior shapes in one call, e.g. levels
 * whose blocks differ in size.  No counterpart in the reference, where every
 * block has the compile-time shape NXB x NYB x NZB; it serves callers that
 * batch levels of different block sizes.  nb[3*c + a] is class c's interior
 * extent along axis a (1 beyond ndim); every physics knob stays that of
 * fx_hydro_init.  Once set, fx_hydro_compute_fluxes, fx_hydro_update_solution
 * and the flux-register calls read each tile's shape from
 * fx_tile.block_class; tiles of several classes go through one persistent
 * launch per stage.  The shapes must admit the fused column-march kernel
 * (interior x and y extents multiples of 8, 2-D or 3-D).  nClasses = 0
 * returns to the fx_hydro_init shape.  The batch entry points
 * (fx_hydro_advance_batch, fx_grid_*) keep the fx_hydro_init shape.
 * fx_hydro_init with different knobs clears the classes. 
 * Returns 0, or 3 for a bad shape or count (at most FX_MAX_CLASSES). */
#define FX_MAX_CLASSES 8
int fx_hydro_set_block_classes(const int* nb, int nClasses);

/* ---- autotuner (no counterpart in the reference) ----
 * Times every candidate launch shape of the flux stages (column march, and
 * the exact solver's face stage) on the caller's own batch and keeps the
 * fastest for later calls.  Candidates: the column-tile width (16 or 8
 * columns across x, where every tile shape divides it, or per class for
 * tiles of several classes); in the mixed mode in
 * 3-D, a register budget for 4 waves per SIMD (occupancy against spills: two
 * 512-thread workgroups share a CU);
 * and for tiles of several classes the persistent workgroups per CU (1, 2
 * or 4 times what stays resident) and the order of the column work items
 * (longest columns first, or tile by tile).  Each candidate runs reps times after one
 * warm-up, timed with HIP events; the minimum counts.  The flux stages read
 * U and write only the tiles' flux arrays and library scratch, so U is left
 * as it was.  out[0 .. *n-1] receives every candidate's time (up to cap);
 * the chosen one has chosen = 1. */
typedef struct {
  int tile_x;               /* columns across x per workgroup (8 or 16; 1 = per
                               class: 16 where the class allows it, else 8) */
  int persist;              /* workgroups per CU multiple (multi-class only; 0 = n/a) */
  int order;                /* multi-class items: 0 longest first, 1 tile by tile */
  int occupancy;            /* waves-per-SIMD hint of the 3-D march: 4 (FP32),
                               0 = none */
  int chosen;
  float ms;                 /* fastest of reps */
} fx_tune_result;
int fx_hydro_autotune(const fx_tile* tiles, int nTiles, double dt, int reps, fx_tune_result* out,
                      int cap, int* n);
/* The same on the multi-level grid of fx_amr_setup. */
int fx_amr_autotune(double dt, int reps, fx_tune_result* out, int cap, int* n);
/* Forget the tuned choice (back to the built-in one). */
void fx_hydro_autotune_reset(void);

/* ---- multi-level block grids held on the device ----
 * The Grid operations a multi-level hydro step needs between the Hydro
 * calls, for grids whose levels differ in block size: guard fill
 * (Grid_fillGuardCells(CENTER, ALLDIR), Amrex/Grid_fillGuardCells.F90:287-369:
 * same-level copies, conservative linear prolongation from the next coarser
 * level, EOS on the prolonged cells), the flux registers of every level pair
 * (Grid_putFluxData / Grid_conserveFluxes, finest pair first) and restriction
 * (Grid_restrictAllLevels -> Amrex/gr_restrictAllLevels.F90:86-130).  The
 * refinement ratio is 2; each level has one block class; no subcycling (one
 * dt for all levels).  Tiles carry block_class (fx_hydro_set_block_classes
 * first) and the level's cell widths; null flux / pavg arrays are supplied
 * by the library.  fx_hydro_set_block_classes with other shapes, or
 * fx_hydro_init with other knobs, drops the grid (set it up again). */
typedef struct {
  int level;                /* 1-based refinement level */
  int pos[3];               /* block position on its level, in blocks */
} fx_amr_block;
/* levelExtent[3*L + a]: blocks along axis a that would cover the domain at
 * level L+1 (each level doubles the cells across the domain).  periodic[a]:
 * 1 wraps axis a, 0 is outflow (zero gradient).  Returns 3 for an
 * inconsistent grid (levels not nested by ratio 2, two blocks at one
 * position, a level with two block shapes). */
int fx_amr_setup(const fx_tile* tiles, const fx_amr_block* blocks, int nBlocks, int nLevels,
                 const int* levelExtent, const int* periodic, fx_status* st);
/* Guard cells of every block (code 3 if a fine block's coarse neighbourhood
 * is missing: levels not properly nested). */
int fx_amr_fill_guards(fx_status* st);
/* Every covered cell from its children, finest level first. */
int fx_amr_restrict(fx_status* st);
/* dt = cfl / max over all levels' leaf cells (interior cells no finer block
 * covers, as Driver_computeDt.F90:219 iterates LEAF blocks) of
 * (|v_d| + c) / del_d (Hydro_computeDt.F90:186-270 without the reductions
 * over ranks).  Under shockLowerCFL or fallbackLowerCFL the caller passes
 * the lowered CFL: the last step's fx_status.min_cell_cfl and fallback_cfl
 * give the caps. */
int fx_amr_compute_dt(double cfl, double* dtOut);
/* One step of every level: guard fill, fluxes of all blocks (one persistent
 * launch per stage), flux correction at fine-coarse faces, update, then
 * restriction. */
int fx_amr_step(double dt, fx_status* st);
/* nReplays of a HIP graph of stepsPerGraph steps (dt on the device each
 * step, as fx_amr_compute_dt; under shockLowerCFL or fallbackLowerCFL the
 * lowered CFL is carried on the device, as in fx_grid_run_graph); *tOut =
 * simulated time. */
int fx_amr_run_graph(double cfl, int stepsPerGraph, int nReplays, double* tOut, fx_status* st);
/* Library version string. */
const char* fx_hydro_version(void);

#ifdef __cplusplus
}
#endif
#endif
