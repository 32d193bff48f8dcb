// fx_amr.hpp -- device side of multi-level block grids: guard fill with
// coarse-to-fine prolongation, restriction, and the CFL reduction over all
// levels.  Included by fx_hydro.hip after the hydro kernels.
//
// Reference semantics (paths relative to source/Grid/GridMain/AMR/Amrex):
//   guard fill       Grid_fillGuardCells.F90:287-369: same-level copies, then
//                    amrex_fillpatch from the next coarser level with the
//                    interpolator amrexInterpolator = CELL_CONSERVATIVE_LINEAR
//                    (Config:39-40), on data in conserved form
//                    (gr_primitiveToConserve.F90:107-116) and back, then EOS
//                    on the guard cells (Hydro.F90:121).
//   restriction      gr_restrictAllLevels.F90:86-130: PER_MASS variables
//                    times density on the leaves, amrex_average_down finest to
//                    coarsest, then back to per-mass form on every block.
// AMReX itself is not part of the reference tree.  FLASH's
// CELL_CONSERVATIVE_LINEAR is amrex_interp_cell_cons
// (RuntimeParameters_mapStrToInt.F90:482-485): AMReX's cell-conservative
// linear interpolator without its cross-component limiting.  Its published
// algorithm (mf_cell_cons_lin_interp_mcslope / mf_cell_cons_lin_interp):
// per component and axis the MC slope sign(dc) min(|dc|, |df|, |db|) of the
// central and doubled one-sided differences (0 unless df db >= 0), all
// scaled by one factor alpha <= 1 that keeps every fine value inside the
// min / max of the coarse cell's 3^ndim neighbourhood, evaluated at the
// fine-cell offsets +-1/4 of a coarse cell.  PER_MASS variables (velx, vely,
// velz, ener, eint; hydro unsplit Config:44-53) are interpolated and averaged
// as density-weighted quantities.
#pragma once

struct AmrBlock {
  double* U;      // guard-inclusive storage of the block
  int level;      // 0-based
  int cls;        // block class
  int pos[3];     // block position at its level, in blocks
};

struct AmrLevelDev {
  int nb[3];           // interior cells per block along each axis (1 beyond ndim)
  int ext[3];          // blocks per axis that would cover the domain at this level
  const int* index;    // dense [(z * ext1 + y) * ext0 + x] -> block, or -1
  double del[3];       // cell widths
};

struct AmrLaunch {
  const AmrBlock* blocks;
  int nBlocks;
  const AmrLevelDev* levels;
  int nLevels;
  const DevParams* cls;
  const int* list;          // block subset of the launch (restriction: one level), or null
  int nList;                // blocks in the launch
  const long* start;        // prefix over the launch's blocks of its work items
  const int* wtab;          // block of every 64th work item (find_tile_w), or null
  int periodic[3];
  unsigned long long perMass;  // bit v: variable slot v is PER_MASS
  DevStatus* status;
  int onlyVar;              // k_amr_fill: >= 0 fills that slot alone, without the EOS
  const int* guard;         // k_amr_fill: guard-cell box indices of every class, class by class
  const int* guardOff;      // class c's guard cells are guard[guardOff[c] .. guardOff[c+1])
};

// global cell g (level lev) wrapped or clamped into the domain: periodic
// axes wrap, the others clamp (zero-gradient outflow)
__device__ __forceinline__ void amr_domain(const AmrLaunch& A, int lev, int ndim, int g[3]) {
  const AmrLevelDev& L = A.levels[lev];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a >= ndim) break;
    const int n = L.ext[a] * L.nb[a];
    if (A.periodic[a]) g[a] = ((g[a] % n) + n) % n;
    else g[a] = g[a] < 0 ? 0 : (g[a] >= n ? n - 1 : g[a]);
  }
}

// block holding global cell g of level lev and the cell's guard-inclusive
// index in it; false when the level has no block there
__device__ __forceinline__ bool amr_find(const AmrLaunch& A, int lev, const int g[3], int& b,
                                         long& c) {
  const AmrLevelDev& L = A.levels[lev];
  const int bx = g[0] / L.nb[0], by = g[1] / L.nb[1], bz = g[2] / L.nb[2];
  const int id = L.index[((long)bz * L.ext[1] + by) * L.ext[0] + bx];
  if (id < 0) return false;
  const DevParams& P = A.cls[A.blocks[id].cls];
  c = cidx(P, g[0] - bx * L.nb[0] + P.lo[0], g[1] - by * L.nb[1] + P.lo[1],
           g[2] - bz * L.nb[2] + P.lo[2]);
  b = id;
  return true;
}

// conserved-form value of variable v at cell c of U
__device__ __forceinline__ double amr_q(const AmrLaunch& A, const DevParams& P, const double* U,
                                        long c, int v) {
  const gdouble* G = as_global(U);
  const double x = G[(long)v * P.ncell + c];
  return (A.perMass >> v) & 1ULL ? G[P.voff[FX_DENS] + c] * x : x;
}

// AMReX's MC slope of one axis from the coarse values u-, u0, u+
__device__ __forceinline__ double amr_mcslope(double u0, double um, double up) {
  const double dc = 0.5 * (up - um);
  const double df = 2.0 * (up - u0);
  const double db = 2.0 * (u0 - um);
  const double s = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
  return copysign(1.0, dc) * fmin(s, fabs(dc));
}

// Source of one guard-fill work slot (chunk w, lane t: the guard cell
// guard[guardOff[cls] + chunks[w].y + t] of block chunks[w].x), fixed by
// the grid's layout: x >= 0: the same-level block x holds the cell, at cell
// index y of its box; x <= kPlanCoarse: prolongation from the coarse cell y
// of block kPlanCoarse - x of the next coarser level; kPlanNone: an idle
// lane; kPlanBad: neither exists (the levels are not properly nested, or a
// level-0 cell has no block).
constexpr int kPlanNone = -1, kPlanBad = -2, kPlanCoarse = -3;

// the guard cell of a work slot: its block, cell index in that block's box
// and global cell at the block's level (wrapped or clamped into the domain)
struct AmrGuardCell {
  int b, r, loc[3], g[3];
};
__device__ __forceinline__ bool amr_guard_cell(const AmrLaunch& A, const int2* chunks, int w,
                                               int t, AmrGuardCell& gc) {
  const int2 ch = chunks[w];
  gc.b = ch.x;
  const AmrBlock& B = A.blocks[gc.b];
  const int o = A.guardOff[B.cls] + ch.y + t;
  if (o >= A.guardOff[B.cls + 1]) return false;
  gc.r = A.guard[o];
  const DevParams& P = A.cls[B.cls];
  gc.loc[0] = gc.r % P.nx;
  gc.loc[1] = (gc.r / P.nx) % P.ny;
  gc.loc[2] = gc.r / (P.nx * P.ny);
  const AmrLevelDev& Lf = A.levels[B.level];
#pragma unroll
  for (int a = 0; a < 3; ++a) gc.g[a] = a < P.ndim ? B.pos[a] * Lf.nb[a] + gc.loc[a] - P.lo[a] : 0;
  amr_domain(A, B.level, P.ndim, gc.g);
  return true;
}

// The guard-fill plan of every work slot (run once, when the grid is set
// up): the source lookups that the fill would otherwise repeat every time,
// each a chain of dependent loads through the level tables.  Slots that are
// not same-level copies (prolongations, and lookups that fail) are also
// listed in rest (count: *nRest; order free), for k_amr_prolong.
__global__ void __launch_bounds__(kThreads) k_amr_plan(AmrLaunch A,
                                                        const int2* __restrict__ chunks,
                                                        int2* __restrict__ plan,
                                                        const double** __restrict__ src,
                                                        int* __restrict__ rest,
                                                        unsigned* __restrict__ nRest) {
  AmrGuardCell gc;
  int2 e = make_int2(kPlanNone, 0);
  if (amr_guard_cell(A, chunks, blockIdx.x, threadIdx.x, gc)) {
    const int lev = A.blocks[gc.b].level;
    int sb;
    long sc;
    if (amr_find(A, lev, gc.g, sb, sc)) {
      e = make_int2(sb, (int)sc);
    } else {
      e = make_int2(kPlanBad, 0);
      const int c3[3] = {gc.g[0] / 2, gc.g[1] / 2, gc.g[2] / 2};
      if (lev > 0 && amr_find(A, lev - 1, c3, sb, sc)) e = make_int2(kPlanCoarse - sb, (int)sc);
    }
  }
  const int slot = (int)blockIdx.x * kThreads + (int)threadIdx.x;
  plan[slot] = e;
  // a copy's source (variable-0 slot of the source cell), else null
  src[slot] = e.x >= 0 ? A.blocks[e.x].U + e.y : nullptr;
  const bool listed = e.x <= kPlanBad;
  const unsigned long long mask = __ballot(listed);
  if (mask) {
    const int lane = (int)(threadIdx.x & 63), leader = __builtin_ctzll(mask);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(nRest, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    if (listed) rest[base + (unsigned)__popcll(mask & ((1ULL << lane) - 1ULL))] = slot;
  }
}

// Guard cells of every block, in two kernels over one plan (k_amr_plan):
// k_amr_fill copies the same-level guard cells (from src, the plan's source
// addresses), k_amr_prolong fills the
// others by conservative linear prolongation from the next coarser level
// followed by the EOS (MODE_DENS_EI).  The copy kernel's workgroup w takes
// kThreads guard cells of one block, chunks[w] = (block, first guard cell
// in its class's list), so the block and its parameters are uniform over
// the workgroup and every lane has a guard cell (the lists hold guard cells
// only, in box order).  Apart, the copy kernel keeps few registers.
__global__ void __launch_bounds__(kThreads) k_amr_fill(AmrLaunch A,
                                                        const int2* __restrict__ chunks,
                                                        const double* const* __restrict__ src) {
  const double* const sp = src[(long)blockIdx.x * kThreads + threadIdx.x];
  if (!sp) return;  // idle lane, or a k_amr_prolong slot
  const int2 ch = chunks[blockIdx.x];
  const AmrBlock& B = A.blocks[ch.x];
  const DevParams& P = A.cls[B.cls];
  const long cd = A.guard[A.guardOff[B.cls] + ch.y + (int)threadIdx.x];
  gdouble* Ud = as_global(B.U);
  // blocks of one level have one block shape, hence one box size
  const gdouble* Us = as_global(sp);
  const int v0 = A.onlyVar >= 0 ? A.onlyVar : 0, v1 = A.onlyVar >= 0 ? A.onlyVar + 1 : P.nunk;
  // up to sixteen variables' loads in flight at a time
  for (int v = v0; v < v1; v += 16) {
    double t[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (v + q < v1) t[q] = Us[(long)(v + q) * P.ncell];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (v + q < v1) Ud[(long)(v + q) * P.ncell + cd] = t[q];
  }
}

// Prolongation in two passes.  Every fine guard cell under one coarse cell
// takes the same coarse value q0 and limited slopes s_a * alpha of each
// variable (the fine cells differ only in their offsets +-1/4), so
// k_amr_coarse computes them once per distinct coarse cell (a record,
// listed when the grid is set up) and k_amr_prolong combines them per fine
// cell: q = q0 + sum_a (+-0.25) (s_a alpha), in the order a one-pass
// interpolation would take.  Record r of n holds, per variable v,
// cv[(v * 4 + e) * n + r]: e = 0 the conserved value q0, 1..3 s_a * alpha;
// recOk[r] = 0 when the coarse cell's neighbourhood is incomplete.
struct AmrCoarseRec {
  int block;  // coarse block
  int cell;   // the coarse cell's index in its box
};

__global__ void __launch_bounds__(kThreads) k_amr_coarse(AmrLaunch A,
                                                          const AmrCoarseRec* __restrict__ recs,
                                                          int nRec, double* __restrict__ cv,
                                                          int* __restrict__ recOk) {
  auto one = [&](int rr) {
    const AmrCoarseRec rc = recs[rr];
    const AmrBlock& Cb = A.blocks[rc.block];
    const DevParams& PC = A.cls[Cb.cls];
    const int lev = Cb.level;
    const AmrLevelDev& Lc = A.levels[lev];
    const int ndim = PC.ndim;
    // the coarse cell's global position at its level
    const int loc[3] = {rc.cell % PC.nx, (rc.cell / PC.nx) % PC.ny, rc.cell / (PC.nx * PC.ny)};
    int gc[3] = {0, 0, 0};
    bool inner = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a >= ndim) break;
      const int lc = loc[a] - PC.lo[a];
      gc[a] = Cb.pos[a] * Lc.nb[a] + lc;
      if (lc < 1 || lc > Lc.nb[a] - 2) inner = false;
    }
    // the 3^ndim neighbourhood (centre included), as addresses of each
    // cell's variable-0 slot; hood[(dz+1)*9 + (dy+1)*3 + (dx+1)].  A cell
    // one or more cells inside its block's interior on every axis has the
    // whole neighbourhood in that block: no level lookups.
    const gdouble* hood[27];
    bool ok = true;
    const gdouble* const cU = as_global(Cb.U) + rc.cell;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
      const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      hood[o] = nullptr;
      if ((ndim < 2 && dy != 0) || (ndim < 3 && dz != 0)) continue;
      if (inner || o == 13) {
        hood[o] = cU + ((long)dz * PC.ny + dy) * PC.nx + dx;
        continue;
      }
      int gn[3] = {gc[0] + dx, gc[1] + dy, gc[2] + dz};
      amr_domain(A, lev, ndim, gn);
      int hb;
      long hc;
      if (amr_find(A, lev, gn, hb, hc)) hood[o] = as_global(A.blocks[hb].U) + hc;
      else ok = false;
    }
    recOk[rr] = ok ? 1 : 0;
    if (!ok) return;
    const long dOff = PC.voff[FX_DENS];
    // conserved-form value of variable v at neighbourhood cell o
    auto qv = [&](int o, int v) {
      const double x = hood[o][(long)v * PC.ncell];
      return (A.perMass >> v) & 1ULL ? hood[o][dOff] * x : x;
    };
    const int densSlot = (int)(dOff / PC.ncell);
    for (int v = 0; v < PC.nunk; ++v) {
      if (A.onlyVar >= 0 && v != A.onlyVar && v != densSlot) continue;
      const double q0 = qv(13, v);
      double sl[3] = {0.0, 0.0, 0.0};
      sl[0] = amr_mcslope(q0, qv(12, v), qv(14, v));
      if (ndim > 1) sl[1] = amr_mcslope(q0, qv(10, v), qv(16, v));
      if (ndim > 2) sl[2] = amr_mcslope(q0, qv(4, v), qv(22, v));
      double alpha = 1.0;
      if (sl[0] != 0.0 || sl[1] != 0.0 || sl[2] != 0.0) {
        double dumax = 0.0;
        for (int a = 0; a < ndim; ++a) dumax = dumax + fabs(sl[a]) * 0.25;
        double umin = q0, umax = q0;
#pragma unroll
        for (int o = 0; o < 27; ++o) {
          if (!hood[o]) continue;
          const double x = qv(o, v);
          umin = fmin(umin, x);
          umax = fmax(umax, x);
        }
        if (dumax * alpha > (umax - q0)) alpha = (umax - q0) / dumax;
        if (dumax * alpha > (q0 - umin)) alpha = (q0 - umin) / dumax;
      }
      cv[((long)v * 4 + 0) * nRec + rr] = q0;
#pragma unroll
      for (int a = 0; a < 3; ++a) cv[((long)v * 4 + 1 + a) * nRec + rr] = sl[a] * alpha;
    }
  };
  for (int rr = blockIdx.x * blockDim.x + threadIdx.x; rr < nRec; rr += gridDim.x * blockDim.x)
    one(rr);
}

// The listed slots of the plan (rest[0 .. n-1]): the fine guard cell from
// its coarse record recOf[q] (k_amr_coarse), then the EOS; or the error of
// a guard cell that no level holds (recOf[q] < 0, or an incomplete record).
__global__ void __launch_bounds__(kThreads) k_amr_prolong(AmrLaunch A,
                                                           const int2* __restrict__ chunks,
                                                           const int* __restrict__ rest,
                                                           const int* __restrict__ recOf,
                                                           int nRest, const double* __restrict__ cv,
                                                           const int* __restrict__ recOk,
                                                           int nRec) {
  auto one = [&](int q) {
    const int slot = rest[q];
    const int w = slot / kThreads, t = slot - w * kThreads;
    AmrGuardCell gcl;
    amr_guard_cell(A, chunks, w, t, gcl);
    const AmrBlock& B = A.blocks[gcl.b];
    const DevParams& P = A.cls[B.cls];
    const int rr = recOf[q];
    if (rr < 0 || !recOk[rr]) {
      report_error(A.status, 3, gcl.b, gcl.loc[0], gcl.loc[1], gcl.loc[2], -1);
      return;
    }
    gdouble* Ud = as_global(B.U);
    const long cd = (long)gcl.r;
    double h[3] = {0.0, 0.0, 0.0};  // the fine cell's offsets in its coarse cell
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a < P.ndim) h[a] = (gcl.g[a] & 1) ? 0.25 : -0.25;
    auto interp = [&](int v) {
      double qq = cv[((long)v * 4) * nRec + rr];
      for (int a = 0; a < P.ndim; ++a) qq = qq + h[a] * cv[((long)v * 4 + 1 + a) * nRec + rr];
      return qq;
    };
    const double dens = interp((int)(P.voff[FX_DENS] / P.ncell));
    const int v0 = A.onlyVar >= 0 ? A.onlyVar : 0, v1 = A.onlyVar >= 0 ? A.onlyVar + 1 : P.nunk;
    for (int v = v0; v < v1; ++v) {
      const double qq = interp(v);
      Ud[(long)v * P.ncell + cd] = (A.perMass >> v) & 1ULL ? qq / dens : qq;
    }
    if (A.onlyVar >= 0) return;
    const EosOut eo = eos_dens_ei(P, dens, Ud[P.voff[FX_VELX] + cd], Ud[P.voff[FX_VELY] + cd],
                                  Ud[P.voff[FX_VELZ] + cd], Ud[P.voff[FX_EINT] + cd],
                                  Ud[P.voff[FX_ENER] + cd]);
    Ud[P.voff[FX_PRES] + cd] = eo.pres;
    Ud[P.voff[FX_TEMP] + cd] = eo.temp;
    Ud[P.voff[FX_ENER] + cd] = eo.ener;
    Ud[P.voff[FX_EINT] + cd] = eo.eint;
    Ud[P.voff[FX_GAMC] + cd] = P.gamma;
    Ud[P.voff[FX_GAME] + cd] = eo.game;
  };
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nRest; q += gridDim.x * blockDim.x)
    one(q);
}

// Restriction of one level onto its parent level: every interior cell of
// the parent covered by the listed fine blocks becomes the average of its
// 2^ndim children (z, then y, then x outermost to innermost), PER_MASS
// variables density-weighted.  One thread per covered parent cell.
__global__ void __launch_bounds__(kThreads) k_amr_restrict(AmrLaunch A) {
  const long total = A.start[A.nList];
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    const int q = find_tile(A.start, A.nList, gid);
    const int r = (int)(gid - A.start[q]);
    const int fb = A.list[q];
    const AmrBlock& F = A.blocks[fb];
    const DevParams& P = A.cls[F.cls];
    const int hx = P.nxb / 2, hy = P.ndim > 1 ? P.nyb / 2 : 1;
    const int c[3] = {r % hx, (r / hx) % hy, r / (hx * hy)};
    int G[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a < P.ndim) G[a] = F.pos[a] * (A.levels[F.level].nb[a] / 2) + c[a];
    int pb;
    long pc;
    if (!amr_find(A, F.level - 1, G, pb, pc)) {
      report_error(A.status, 3, fb, c[0], c[1], c[2], -1);
      continue;
    }
    const AmrBlock& Pb = A.blocks[pb];
    const DevParams& PP = A.cls[Pb.cls];
    const int nz = P.ndim > 2 ? 2 : 1, ny = P.ndim > 1 ? 2 : 1;
    const double frac = 1.0 / (double)(2 * ny * nz);
    // children sum of variable v in conserved form
    auto child_sum = [&](int v) {
      double s = 0.;
      for (int oz = 0; oz < nz; ++oz)
        for (int oy = 0; oy < ny; ++oy)
          for (int ox = 0; ox < 2; ++ox) {
            const long fc = cidx(P, 2 * c[0] + ox + P.lo[0], 2 * c[1] + oy + P.lo[1],
                                 2 * c[2] + oz + P.lo[2]);
            s = s + amr_q(A, P, F.U, fc, v);
          }
      return s;
    };
    const double dens = child_sum((int)(P.voff[FX_DENS] / P.ncell)) * frac;
    for (int v = 0; v < P.nunk; ++v) {
      const double avg = child_sum(v) * frac;
      as_global(Pb.U)[(long)v * PP.ncell + pc] = (A.perMass >> v) & 1ULL ? avg / dens : avg;
    }
  }
}

// the CFL dt of each block (one workgroup per block, as dt_pick_block
// describes) over its leaf cells: interior cells that no block of the next
// finer level covers.  A block whose cells are all covered gets 1e300.
// The CFL is cflDev[0] when cflDev is set, else cfl.
__global__ void __launch_bounds__(kThreads) k_amr_dt(AmrLaunch A, double cfl, const double* cflDev,
                                                     double* __restrict__ dtBlock) {
  const double c = cflDev ? cflDev[0] : cfl;
  const int b = blockIdx.x;
  const AmrBlock& B = A.blocks[b];
  const DevParams& P = A.cls[B.cls];
  const AmrLevelDev& L = A.levels[B.level];
  const int nxi = P.nxb, nyi = P.ndim > 1 ? P.nyb : 1, n = (int)P.icells;
  const gdouble* U = as_global(B.U);
  DtPick best = dt_pick_none();
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const int ii = r % nxi, jj = (r / nxi) % nyi, kk = r / (nxi * nyi);
    if (B.level + 1 < A.nLevels) {
      // Driver_computeDt.F90:219 takes leaf blocks only: skip a cell that
      // the next finer level covers (its first fine cell's fine block exists)
      const AmrLevelDev& F = A.levels[B.level + 1];
      const int fx = 2 * (B.pos[0] * nxi + ii) / F.nb[0];
      const int fy = P.ndim > 1 ? 2 * (B.pos[1] * nyi + jj) / F.nb[1] : 0;
      const int fz = P.ndim > 2 ? 2 * (B.pos[2] * P.nzb + kk) / F.nb[2] : 0;
      if (F.index[((long)fz * F.ext[1] + fy) * F.ext[0] + fx] >= 0) continue;
    }
    const long cc = cidx(P, ii + P.lo[0], jj + P.lo[1], kk + P.lo[2]);
    const double cs = sqrt(U[P.voff[FX_GAMC] + cc] * U[P.voff[FX_PRES] + cc] / U[P.voff[FX_DENS] + cc]);
    double dtl = (fabs(U[P.voff[FX_VELX] + cc] - 0.) + cs) * (1. / L.del[0]);
    if (P.ndim > 1) dtl = fmax(dtl, (fabs(U[P.voff[FX_VELY] + cc] - 0.) + cs) * (1. / L.del[1]));
    if (P.ndim > 2) dtl = fmax(dtl, (fabs(U[P.voff[FX_VELZ] + cc] - 0.) + cs) * (1. / L.del[2]));
    dt_pick_cell(best, dtl, c, r);
  }
  const double dt = dt_pick_block(best, c);
  if (threadIdx.x == 0) dtBlock[b] = dt;
}
