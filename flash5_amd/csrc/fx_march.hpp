// fx_march.hpp -- fused face-state + flux kernel: one workgroup per
// (block, TX x TY column tile), marching along z through the block.
//
// The kernel templates are instantiated in fx_march_inst.hip, one unit per
// (solver family, real type), behind the launchers of fx_march_launch.hpp.
//
// Per z-plane k the workgroup computes, each value once (the work of
// hy_getRiemannState.F90:356-1327 regrouped by quantity instead of by cell):
//   stage A  transverse fluxes sigma_x, sigma_y of plane k+1 into 3-plane LDS
//            rings (they feed the 3-D CTU terms of planes k and k+2) and,
//            from the same stencils, the limited characteristic slopes
//            along x and y of plane k+1 into 2-plane rings; sigma_z of
//            plane k; slopes along z (with the z flattening tilde) of plane
//            k+1 into 3-plane rings
//   stage B1 x-, y- and z-face states of plane k: PPM with characteristic
//            tracing, 2-D CTU transverse corrections, positivity fallback
//   stage B2 3-D full-CTU corner terms added to those states
//   stage C  x-, y- and z-face fluxes (HLLC, HLL or the hybrid switch, and
//            the aux internal-energy fluxes, hy_getFaceFlux.F90) to the
//            tile's flux arrays and the
//            face-averaged pressure of each interior cell to pavg.
// The conservative update then runs as k_update from those arrays.
// x/y states exist for interior planes only; z states for planes lo-1..hi+1.

#pragma once
#include "fx_common.hpp"

namespace fx {

// LDS index helpers, tile-local coordinates a = i - ox, b = j - oy
template <int TX, int TY> struct TileIdx {
  // R_x: a in [0,TX), b in [-1,TY]   (sigma_x, y-face states)
  static __device__ __forceinline__ int rx(int a, int b) { return (b + 1) * TX + a; }
  // R_y: a in [-1,TX], b in [0,TY)   (sigma_y, x-face states)
  static __device__ __forceinline__ int ry(int a, int b) { return b * (TX + 2) + (a + 1); }
  // C: a in [-1,TX], b in [-1,TY]    (sigma_z)
  static __device__ __forceinline__ int cc(int a, int b) { return (b + 1) * (TX + 2) + (a + 1); }
  // slopes along x: a in [-2,TX+1], b in [0,TY)
  static __device__ __forceinline__ int lx(int a, int b) { return b * (TX + 4) + (a + 2); }
  // slopes along y: a in [0,TX), b in [-2,TY+1]
  static __device__ __forceinline__ int ly(int a, int b) { return (b + 2) * TX + a; }
  static constexpr int NRX = TX * (TY + 2), NRY = (TX + 2) * TY, NC = (TX + 2) * (TY + 2);
  static constexpr int NLX = (TX + 4) * TY, NLY = TX * (TY + 4);
  // stage-A items along x: the cells of R_x and of the x slopes, rows
  // b = -1, [0,TY), TY; along y: the cells of R_y and of the y slopes,
  // rows b = -2, -1, [0,TY), TY, TY+1
  static constexpr int NUX = 2 * TX + (TX + 4) * TY, NUY = 4 * TX + (TX + 2) * TY;
};

// LDS of one column tile (march_tile), as offsets into one array of R
template <int NDIM, int TX, int TY> struct MarchLds {
  using I = TileIdx<TX, TY>;
  static constexpr int NR = NDIM == 3 ? 3 : 1, NCOL = TX * TY, Z = NDIM == 3 ? 1 : 0;
  static constexpr int SX = 0;                                // sigma_x ring
  static constexpr int SY = SX + NR * 6 * I::NRX;             // sigma_y ring
  static constexpr int SZ = SY + NR * 6 * I::NRY;             // sigma_z of plane k
  static constexpr int NRL = NDIM == 3 ? 2 : 1;               // slope ring depth
  static constexpr int LX = SZ + (Z ? 6 * I::NC : 1);         // slopes along x, ring
  static constexpr int LY = LX + NRL * 6 * I::NLX;            // slopes along y, ring
  static constexpr int LZ = LY + NRL * 6 * I::NLY;            // slopes along z, ring
  static constexpr int FZ = LZ + (Z ? 3 * 6 * NCOL : 1);      // flattening tilde along z, ring
  static constexpr int WX = FZ + (Z ? 3 * NCOL : 1);          // x-face states (Wp, Wn)
  static constexpr int WY = WX + 12 * I::NRY;                 // y-face states
  static constexpr int WZ = WY + 12 * I::NRX;                 // z-face states of plane k
  static constexpr int WZP = WZ + (Z ? 12 * NCOL : 1);        // z-face Wp of plane k-1
  static constexpr int CST = WZP + (Z ? 6 * NCOL : 1);        // the step's ratios, widths (FP32)
  static constexpr int total = CST + 12;
};

// components stored for sigma and slopes: dens, vx, vy, vz, pres, eint
__device__ __forceinline__ int comp_slot(int q) { return q < 5 ? q : P_EINT; }

struct MarchLaunch {
  const fx_tile* tiles;
  int nTiles;
  int tilesX, tilesY;  // column tiles per block
  DevStatus* status;
  double dt;
  void* fstates;       // RS_STATES: per-face states for the separate Riemann stage, in R
  const double* dtp;   // device-resident dt (graph-captured steps), or null
};

// Riemann-stage variant of the march: stage C stores each face's left and
// right states (dens, vx, vy, vz, pres, eint) instead of solving there.
constexpr int RS_STATES = 99;
constexpr bool is_states(int rs) { return rs == RS_STATES; }
constexpr int kFaceComps = 12;

// face-state scratch of one tile: [dir][comp 0..11][face], in the march's
// real type R (FP32 in the mixed mode); offsets below count elements
__device__ __forceinline__ long fstate_tile_stride(const DevParams& P) {
  return (long)kFaceComps * (P.nface[0] + P.nface[1] + P.nface[2]);
}
__device__ __forceinline__ long fstate_dir_offset(const DevParams& P, int d) {
  long o = 0;
  for (int e = 0; e < d; ++e) o += (long)kFaceComps * P.nface[e];
  return o;
}

template <int D, typename R>
__device__ __forceinline__ void write_face_states(const DevParams& P, R* base, long r,
                                                  const R VL[8], const R VR[8]) {
  const long nf = P.nface[D];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    base[q * nf + r] = VL[comp_slot(q)];
    base[(6 + q) * nf + r] = VR[comp_slot(q)];
  }
}

template <int SD, typename R>
__device__ __forceinline__ bool sigma_cell(const DevParams& P, const gdouble* __restrict__ U, long c,
                                           R sig[8]) {
  return sigma_at<SD>(P, U, c, sig, nullptr);
}

template <int DIR, typename R>
__device__ __forceinline__ bool slope_cell(const DevParams& P, const gdouble* __restrict__ U, long c,
                                           R db[8]) {
  const long s = axis_stride(P, DIR);
  R Vm[8], Vc[8], Vp[8];
  load_prim(P, U, c - s, Vm);
  load_prim(P, U, c, Vc);
  load_prim(P, U, c + s, Vp);
  EigT<R> e;
  if (!eig_make<DIR>(Vc, e)) return false;
  tvd_slope<DIR>(P, Vm, Vc, Vp, e, db);
  return true;
}

template <typename R>
__device__ __forceinline__ void lds_put6(R* base, int n, int idx, const R v[8]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) base[q * n + idx] = v[comp_slot(q)];
}
template <typename R>
__device__ __forceinline__ void lds_get6(const R* base, int n, int idx, R v[8]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) v[comp_slot(q)] = base[q * n + idx];
  v[P_GAMC] = R(0.);
  v[P_GAME] = R(0.);
}

// First-order reset of one state cell's face states.  A state cell on a
// column-tile seam is computed by both tiles; only its owner (own = true)
// counts the reset, so the count is one per (state cell, direction), as the
// reference's (hy_dataReconstOneStep.F90:1124-1129).
template <typename R>
__device__ __forceinline__ void fallback_first_order(const R* Vc, R Wp[8], R Wn[8],
                                                     DevStatus* st, bool own) {
  if (own) atomicAdd(&st->fallbacks, 1ULL);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    Wn[v] = Vc[v];
    Wp[v] = Vc[v];
  }
}

template <typename R>
__device__ __forceinline__ void sub_sigma(R Wp[8], R Wn[8], const R* sg, R hdt) {
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    Wn[v] = Wn[v] - hdt * sg[v];
    Wp[v] = Wp[v] - hdt * sg[v];
  }
}

template <typename R>
__device__ __forceinline__ void add_ctu(R Wp[8], R Wn[8], const R* tA, const R* tB, R fac) {
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    Wp[v] = Wp[v] + (tA[v] + tB[v]) * fac;
    Wn[v] = Wn[v] + (tA[v] + tB[v]) * fac;
  }
}

template <typename R>
__device__ __forceinline__ bool nonpositive(const R* Wp, const R* Wn) {
  return Wp[P_DENS] <= R(0.) || Wn[P_DENS] <= R(0.) || Wp[P_PRES] <= R(0.) || Wn[P_PRES] <= R(0.);
}
template <typename R>
__device__ __forceinline__ bool negative(const R* Wp, const R* Wn) {
  return Wn[P_DENS] < R(0.) || Wp[P_DENS] < R(0.) || Wn[P_PRES] < R(0.) || Wp[P_PRES] < R(0.);
}

// store/load a face-state pair (Wp, Wn) of 6 components each
template <typename R>
__device__ __forceinline__ void put_state(R* base, int n, int idx, const R Wp[8], const R Wn[8]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    base[q * n + idx] = Wp[comp_slot(q)];
    base[(6 + q) * n + idx] = Wn[comp_slot(q)];
  }
}

template <int D>
__device__ __forceinline__ void write_flux(const DevParams& P, gdouble* Fo, long r, const double F[8]) {
  const long nf = P.nface[D];
#pragma unroll
  for (int q = 0; q < FX_NFLUX; ++q) Fo[q * nf + r] = F[q];
}

template <typename R>
__device__ __forceinline__ void get_state(const R* base, int n, int idx, R Wp[8], R Wn[8]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    Wp[comp_slot(q)] = base[q * n + idx];
    Wn[comp_slot(q)] = base[(6 + q) * n + idx];
  }
  Wp[P_GAMC] = Wp[P_GAME] = Wn[P_GAMC] = Wn[P_GAME] = R(0.);
}

template <typename R>
__device__ __forceinline__ R flat_tilde_vals(R pm2, R pm1, R pp1, R pp2, R vm1, R vp1) {
  const R dp1 = (pp1 - pm1);
  const R dp2 = (pp2 - pm2);
  const R dv1 = vp1 - vm1;
  const R Sp = (fabs(dp2) > R(1.e-15)) ? dp1 / dp2 - R(0.75) : R(0.);
  R ft = fmax(R(0.0), fmin(R(1.0), R(10.0) * Sp));
  if ((fabs(dp1) / fmin(pm1, pp1) < R(1.) / R(3.)) || dv1 > R(0.)) ft = R(0.);
  return ft;
}

// Item counts rounded up to whole wavefronts: each stage lays its item kinds
// out in wave-aligned segments, so that no wavefront runs two kinds of item
// (two code paths) one after the other.
constexpr int wave_up(int n) { return (n + 63) / 64 * 64; }

// ---- a tile's cell data as the march reads it ----
// 32-bit byte offsets of cells (a tile holds far fewer than 2^29 cells),
// which serve every variable.  Outputs are stored through a uniform base
// address plus one vector offset (st_at).
typedef __attribute__((address_space(1))) char gchar;
template <typename T>
__device__ __forceinline__ void st_at(__attribute__((address_space(1))) T* base, unsigned boff,
                                      T v) {
  *reinterpret_cast<__attribute__((address_space(1))) T*>(reinterpret_cast<gchar*>(base) + boff) = v;
}

// The tile's primitives are read as one buffer (a scalar resource) with the
// variable's byte offset in a scalar operand and the cell's in a vector
// one, so no 64-bit address is formed per lane.  With 64-bit addresses the
// compiler kept several per stencil point live across the item loops: the
// FP32 march at 4 waves per SIMD spilled 36 registers instead of 21, and
// the 8-wide march tiles ran at 7.6 instead of 10.1 ms (FP32) and 11.6
// instead of 20.1 ms (FP64) on 4096 x 16^3.
struct TileView {
  __amdgpu_buffer_rsrc_t rs;  // the tile's guard-inclusive array
  unsigned ob[8];             // byte offsets of the primitive slots P_DENS .. P_EINT
  unsigned obShok;            // SHOK_VAR's (hybrid solver)
};

// The workgroup's tile is uniform; saying so (readfirstlane) keeps the
// compiler from making each buffer load uniform on its own where the tile
// and its class come from a work list.
__device__ __forceinline__ unsigned fx_uniform(unsigned x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ TileView tile_view(const DevParams& P, const gdouble* U) {
  TileView tv;
  const unsigned long long u = reinterpret_cast<unsigned long long>(U);
  void* const base = reinterpret_cast<void*>(
      (static_cast<unsigned long long>(fx_uniform((unsigned)(u >> 32))) << 32) |
      fx_uniform((unsigned)u));
  // raw buffer: no stride, a byte range past any tile's array (offsets stay
  // below 2^31), and the resource word of untyped dword accesses
  tv.rs = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, 0x7fffffff, 0x00020000);
  tv.ob[P_DENS] = fx_uniform((unsigned)P.voff[FX_DENS] * 8u);
  tv.ob[P_VELX] = fx_uniform((unsigned)P.voff[FX_VELX] * 8u);
  tv.ob[P_VELY] = fx_uniform((unsigned)P.voff[FX_VELY] * 8u);
  tv.ob[P_VELZ] = fx_uniform((unsigned)P.voff[FX_VELZ] * 8u);
  tv.ob[P_PRES] = fx_uniform((unsigned)P.voff[FX_PRES] * 8u);
  tv.ob[P_GAMC] = fx_uniform((unsigned)P.voff[FX_GAMC] * 8u);
  tv.ob[P_GAME] = fx_uniform((unsigned)P.voff[FX_GAME] * 8u);
  tv.ob[P_EINT] = fx_uniform((unsigned)P.voff[FX_EINT] * 8u);
  tv.obShok = fx_uniform(P.shokOff >= 0 ? (unsigned)P.shokOff * 8u : 0u);
  return tv;
}
// variable slot q (byte offset ob) of the cell at byte offset cb
__device__ __forceinline__ double tv_ld(const TileView& tv, unsigned ob, unsigned cb) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(tv.rs, (int)cb, (int)ob, 0));
}

// A stage's item number.  With OPAQUE it passes through an empty asm
// statement, which the compiler cannot see through: the index arithmetic of
// each stage's items (item kind, column, cell offsets) is then redone per
// plane instead of being hoisted out of the plane loop and kept live.  That
// took the FP32 16x8 march from 160 registers (128 with 21 spilled at 4 waves
// per SIMD) to 104, and 4096 x 16^3 from 5.95 to 4.81 ms (HLLC) and 6.82 to
// 5.97 ms (exact-solver face states); the FP64 multi-class kernels from 256
// registers with 51 spilled to 195.  The single-class FP64 march, which LDS
// holds at 2 waves per SIMD whatever its registers, keeps the hoisted form
// (1 % faster).
template <bool OPAQUE>
__device__ __forceinline__ int item_index(int w) {
  if constexpr (OPAQUE) asm volatile("" : "+v"(w));
  return w;
}

// byte offset of cell (i, j, k) in its tile
__device__ __forceinline__ unsigned cell_b(const DevParams& P, int i, int j, int k) {
  return (unsigned)((k * P.ny + j) * P.nx + i) * 8u;
}

template <typename R>
__device__ __forceinline__ void tv_prim(const TileView& tv, unsigned cb, R V[8]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) V[q] = R(tv_ld(tv, tv.ob[q], cb));
}

// sigma of the cell at cb along SD (the transverse-only PPM call; sb: the
// axis stride in bytes)
template <int SD, typename R>
__device__ __forceinline__ bool tv_sigma(const DevParams& P, const TileView& tv, unsigned cb,
                                         unsigned sb, R sig[8]) {
  R Vm[8], Vc[8], Vp[8];
  tv_prim(tv, cb - sb, Vm);
  tv_prim(tv, cb, Vc);
  tv_prim(tv, cb + sb, Vp);
  EigT<R> e;
  if (!eig_make<SD>(Vc, e)) return false;
  trans_sigma<SD>(P, Vm, Vc, Vp, e, sig);
  return true;
}

template <int DIR, typename R>
__device__ __forceinline__ bool tv_slope(const DevParams& P, const TileView& tv, unsigned cb,
                                         unsigned sb, R db[8]) {
  R Vm[8], Vc[8], Vp[8];
  tv_prim(tv, cb - sb, Vm);
  tv_prim(tv, cb, Vc);
  tv_prim(tv, cb + sb, Vp);
  EigT<R> e;
  if (!eig_make<DIR>(Vc, e)) return false;
  tvd_slope<DIR>(P, Vm, Vc, Vp, e, db);
  return true;
}

// FlatTilde and FlatCoeff along DIR (hy_getRiemannState.F90:365-453)
template <int DIR, typename R>
__device__ __forceinline__ R tv_flat_tilde(const TileView& tv, unsigned cb, unsigned sb) {
  const unsigned pr = tv.ob[P_PRES], vn = tv.ob[P_VELX + DIR];
  return flat_tilde_vals(R(tv_ld(tv, pr, cb - 2 * sb)), R(tv_ld(tv, pr, cb - sb)),
                         R(tv_ld(tv, pr, cb + sb)), R(tv_ld(tv, pr, cb + 2 * sb)),
                         R(tv_ld(tv, vn, cb - sb)), R(tv_ld(tv, vn, cb + sb)));
}
template <int DIR, typename R>
__device__ __forceinline__ R tv_flat_coeff(const TileView& tv, unsigned cb, unsigned sb) {
  const unsigned pr = tv.ob[P_PRES];
  const R dp1 = (R(tv_ld(tv, pr, cb + sb)) - R(tv_ld(tv, pr, cb - sb)));
  const R f0 = tv_flat_tilde<DIR, R>(tv, cb, sb);
  if (dp1 < R(0.0)) return fmax(f0, tv_flat_tilde<DIR, R>(tv, cb + sb, sb));
  if (dp1 == R(0.)) return f0;
  return fmax(f0, tv_flat_tilde<DIR, R>(tv, cb - sb, sb));
}

// Whether the face between the cells at cbl and cbr takes HLL (HLL, or the
// hybrid switch with a flagged cell on either side, hy_getFaceFlux.F90:353-367)
template <int RS>
__device__ __forceinline__ bool tv_use_hll(const DevParams& P, const TileView& tv, unsigned cbl,
                                           unsigned cbr) {
  if constexpr (RS == FX_HYBR) {
    if (P.riemann == FX_HLL) return true;
    if (P.riemann != FX_HYBR) return false;
    return tv_ld(tv, tv.obShok, cbl) + tv_ld(tv, tv.obShok, cbr) > 0.;
  } else {
    return false;
  }
}

// face flux densities F[0..7] to the tile's flux array of direction D at the
// face's byte offset rb (q-major, nface[D] faces per flux)
template <int D>
__device__ __forceinline__ void write_flux_b(const DevParams& P, gdouble* Fo, unsigned rb,
                                             const double F[8]) {
  const long nf = P.nface[D];
#pragma unroll
  for (int q = 0; q < FX_NFLUX; ++q) st_at(Fo + q * nf, rb, F[q]);
}

// One column tile (TX x TY columns at ox, oy, all planes) of tile t.  Every
// thread of the workgroup calls it with the same arguments; it ends on a
// barrier, so a workgroup may call it again for another column tile.  fsb:
// tile t's face-state scratch (RS_STATES).
//
// Work items are laid out over fixed extents so that item numbers split by
// compile-time divisors: x-face states over a in [-1, TX], x faces over
// a in [0, TX] (y likewise); the high row (a = TX) is skipped unless the
// tile sits at the block's high side, where it owns that row.
template <int NDIM, int TX, int TY, int NT, int RS, typename R, bool OPAQUE = (sizeof(R) == 4)>
__device__ __forceinline__ void march_tile(const DevParams& P, const MarchLaunch& L, int t, int ox,
                                           int oy, R* fsb, ExactStats& es,
                                           R* smem) {
  using I = TileIdx<TX, TY>;
  using M = MarchLds<NDIM, TX, TY>;
  constexpr int NR = NDIM == 3 ? 3 : 1;        // ring depth
  constexpr int NCOL = TX * TY;
  constexpr int Z = NDIM == 3 ? 1 : 0;
  R* const SX = smem + M::SX;
  R* const SY = smem + M::SY;
  R* const SZ = smem + M::SZ;
  R* const LX = smem + M::LX;
  R* const LY = smem + M::LY;
  R* const LZ = smem + M::LZ;
  R* const FZ = smem + M::FZ;
  R* const WX = smem + M::WX;
  R* const WY = smem + M::WY;
  R* const WZ = smem + M::WZ;
  R* const WZP = smem + M::WZP;

  const int tid = threadIdx.x;
  // A tile owns its high x (y) face row only at the block's high side;
  // otherwise the neighbouring tile computes that row (and the states
  // behind it) as its low row.
  const bool ex = (ox + TX > P.hi[0]);
  const bool ey = (oy + TY > P.hi[1]);
  const fx_tile& T = L.tiles[t];
  const TileView tv = tile_view(P, as_global(T.U));
  const double dt = step_dt(L.dt, L.dtp);
  const double dx = T.del[0], dy = T.del[1], dz = T.del[2];
  const unsigned sxb = 8u, syb = (unsigned)P.nx * 8u, szb = (unsigned)(P.nx * P.ny) * 8u;
  const int klo = P.lo[2], khi = P.hi[2];
  DevStatus* st = L.status;
  bool bad = false;  // eigen failure seen by this thread
  // The step's ratios of dt and the cell widths, and the widths, uniform
  // over the tile.  The FP64 march keeps them in registers (the compiler
  // hoists them); the FP32 march reads them from LDS, written once per
  // column tile by one thread (stage B1 first reads them after stage A's
  // barrier, and the previous column tile ended on a barrier): 12 fewer live
  // registers, mixed-mode march 6.29 -> 6.08 ms (HLLC) and 7.85 -> 7.08 ms
  // (exact-solver face states) on 4096 x 16^3 at 4 waves per SIMD, bit for
  // bit the same.
  const R rat[12] = {R(dt / dx), R(dt / dy), R(dt / dz),
                     R(0.5 * dt / dx), R(0.5 * dt / dy), R(0.5 * dt / dz),
                     R(dt * dt / (6. * dy * dz)), R(dt * dt / (6. * dz * dx)),
                     R(dt * dt / (6. * dx * dy)), R(dx), R(dy), R(dz)};
  R* const CST = smem + M::CST;
  if constexpr (sizeof(R) == 4) {
    if (tid == 0)
      for (int i = 0; i < 12; ++i) CST[i] = rat[i];
  }
  auto ratio = [&](int i) -> R {
    if constexpr (sizeof(R) == 4) return CST[i];
    else return rat[i];
  };

  auto ring = [](int k) { return NR == 3 ? (k % 3) : 0; };
  auto ringL = [](int k) { return M::NRL == 2 ? (k & 1) : 0; };

  // stage-A items: sigma_x (w < NRX) and sigma_y (w < NRY) of plane kk
  auto sig_x_item = [&](int w, int kk) {
    R sg[8];
    const int a = w % TX, b = w / TX - 1;
    if (!tv_sigma<0>(P, tv, cell_b(P, ox + a, oy + b, kk), sxb, sg)) { bad = true; return; }
    lds_put6(SX + ring(kk) * 6 * I::NRX, I::NRX, w, sg);
  };
  auto sig_y_item = [&](int w, int kk) {
    R sg[8];
    const int a = w % (TX + 2) - 1, b = w / (TX + 2);
    if (!tv_sigma<1>(P, tv, cell_b(P, ox + a, oy + b, kk), syb, sg)) { bad = true; return; }
    lds_put6(SY + ring(kk) * 6 * I::NRY, I::NRY, w, sg);
  };
  // stage-A items along x (w < NUX) and y (w < NUY) at plane p: one cell's
  // three-cell stencil and eigen system serve its sigma (sig: p is a plane
  // the sigma ring takes, and the cell lies in R_x / R_y) and its limited
  // slope (sl: p is interior, and the cell lies in the slope rows)
  auto x_item = [&](int w, int p, bool sig, bool sl) {
    int a, b;
    if (w < TX) {
      a = w;
      b = -1;
    } else if (w < TX + (TX + 4) * TY) {
      const int u = w - TX;
      a = u % (TX + 4) - 2;
      b = u / (TX + 4);
    } else {
      a = w - TX - (TX + 4) * TY;
      b = TY;
    }
    const bool doS = sig && a >= 0 && a < TX;
    const bool doL = sl && b >= 0 && b < TY;
    if (!doS && !doL) return;
    const unsigned cb = cell_b(P, ox + a, oy + b, p);
    R Vm[8], Vc[8], Vp[8];
    tv_prim(tv, cb - sxb, Vm);
    tv_prim(tv, cb, Vc);
    tv_prim(tv, cb + sxb, Vp);
    EigT<R> e;
    if (!eig_make<0>(Vc, e)) { bad = true; return; }
    if (doS) {
      R sg[8];
      trans_sigma<0>(P, Vm, Vc, Vp, e, sg);
      lds_put6(SX + ring(p) * 6 * I::NRX, I::NRX, I::rx(a, b), sg);
    }
    if (doL) {
      R db[8];
      tvd_slope<0>(P, Vm, Vc, Vp, e, db);
      lds_put6(LX + ringL(p) * 6 * I::NLX, I::NLX, I::lx(a, b), db);
    }
  };
  auto y_item = [&](int w, int p, bool sig, bool sl) {
    int a, b;
    if (w < 2 * TX) {
      a = w % TX;
      b = w / TX - 2;
    } else if (w < 2 * TX + (TX + 2) * TY) {
      const int u = w - 2 * TX;
      a = u % (TX + 2) - 1;
      b = u / (TX + 2);
    } else {
      const int u = w - 2 * TX - (TX + 2) * TY;
      a = u % TX;
      b = TY + u / TX;
    }
    const bool doS = sig && b >= 0 && b < TY;
    const bool doL = sl && a >= 0 && a < TX;
    if (!doS && !doL) return;
    const unsigned cb = cell_b(P, ox + a, oy + b, p);
    R Vm[8], Vc[8], Vp[8];
    tv_prim(tv, cb - syb, Vm);
    tv_prim(tv, cb, Vc);
    tv_prim(tv, cb + syb, Vp);
    EigT<R> e;
    if (!eig_make<1>(Vc, e)) { bad = true; return; }
    if (doS) {
      R sg[8];
      trans_sigma<1>(P, Vm, Vc, Vp, e, sg);
      lds_put6(SY + ring(p) * 6 * I::NRY, I::NRY, I::ry(a, b), sg);
    }
    if (doL) {
      R db[8];
      tvd_slope<1>(P, Vm, Vc, Vp, e, db);
      lds_put6(LY + ringL(p) * 6 * I::NLY, I::NLY, I::ly(a, b), db);
    }
  };
  // stage-A item: z-slope and z flattening tilde of column w at plane m
  auto zslope_item = [&](int w, int m) {
    const int a = w % TX, b = w / TX;
    const unsigned cb = cell_b(P, ox + a, oy + b, m);
    R db[8];
    if (!tv_slope<2>(P, tv, cb, szb, db)) { bad = true; return; }
    lds_put6(LZ + ring(m) * 6 * NCOL, NCOL, w, db);
    FZ[ring(m) * NCOL + w] = tv_flat_tilde<2, R>(tv, cb, szb);
  };

  const int kBeg = Z ? klo - 1 : 0;
  const int kEnd = Z ? khi + 1 : 0;
  if (Z) {
    // prologue: sigma_x/y of plane kBeg, z slopes of kBeg-1 and kBeg
    constexpr int Q1 = wave_up(I::NRX), Q2 = Q1 + wave_up(I::NRY), Q3 = Q2 + wave_up(NCOL);
    for (int w = tid; w < Q3 + NCOL; w += NT) {
      if (w < Q1) { if (w < I::NRX) sig_x_item(w, kBeg); }
      else if (w < Q2) { if (w - Q1 < I::NRY) sig_y_item(w - Q1, kBeg); }
      else if (w < Q3) { if (w - Q2 < NCOL) zslope_item(w - Q2, kBeg - 1); }
      else zslope_item(w - Q3, kBeg);
    }
  }

  // fixed item extents of stages B and C (see above)
  constexpr int NBX = (TX + 2) * TY, NBY = TX * (TY + 2);
  constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1);

  for (int k = kBeg; k <= kEnd; ++k) {
    const bool interior = !Z || (k >= klo && k <= khi);
    // ------------------------------ stage A ------------------------------
    // Plane k+1's sigma_x/y (the ring's next plane) and x/y slopes (the
    // slope ring's next plane), plane k's sigma_z, plane k+1's z slopes.
    // (2-D: plane 0's sigmas and slopes.)
    {
      const int pA = Z ? k + 1 : 0;
      const bool wSig = !Z || k + 1 <= kEnd;
      const bool wL = !Z || (k + 1 >= klo && k + 1 <= khi);
      const bool uXY = wSig || wL;
      const int nSz = (Z && interior) ? I::NC : 0;
      const int nZs = Z ? NCOL : 0;
      constexpr int PUX = wave_up(I::NUX), PUY = wave_up(I::NUY), PZ = wave_up(I::NC);
      // Segments in wave-aligned order (x items, y items, sigma_z, z
      // slopes): one item kind per wavefront.  Each stage's work item is a
      // lambda whose early exits are returns, not loop continues: the
      // compiler then allocates registers per item kind.  With loop-level
      // continues the FP64 16x8 march spilled 272 B per lane at 256 VGPRs;
      // as lambdas it needs 251 and spills nothing (FP64 march 11.7 -> 9.3
      // ms on 4096 x 16^3, FP32 10.4 -> 7.7 ms).  One x item serves a
      // cell's sigma_x and x slope (one stencil, one eigen system), which
      // keeps the slopes of two planes in LDS.
      const int s1 = uXY ? PUX + PUY : 0, s2 = s1 + (nSz ? PZ : 0), sEnd = s2 + nZs;
      auto itemA = [&](int v) {
        if (v < s1) {
          if (v < PUX) {
            if (v < I::NUX) x_item(v, pA, wSig, wL);
          } else if (v - PUX < I::NUY) {
            y_item(v - PUX, pA, wSig, wL);
          }
          return;
        }
        if (v < s2) {
          const int r = v - s1;
          if (r >= I::NC) return;
          R v8[8];
          const int a = r % (TX + 2) - 1, b = r / (TX + 2) - 1;
          if (!tv_sigma<2>(P, tv, cell_b(P, ox + a, oy + b, k), szb, v8)) { bad = true; return; }
          lds_put6(SZ, I::NC, r, v8);
          return;
        }
        zslope_item(v - s2, k + 1);
      };
      for (int v = tid; v < sEnd; v += NT) itemA(item_index<OPAQUE>(v));
    }
    __syncthreads();

    const R* sxk = SX + ring(k) * 6 * I::NRX;
    const R* syk = SY + ring(k) * 6 * I::NRY;
    // ------------------------------ stage B1 -----------------------------
    {
      // x-, y- and z-face states in wave-aligned segments
      constexpr int B1 = wave_up(NBX), B2 = B1 + wave_up(NBY), B3 = B2 + (Z ? NCOL : 0);
      auto itemB1 = [&](int w) {
        R Wp[8], Wn[8], Vm[8], Vc[8], Vp[8], dbN[8], db0[8], dbP[8];
        EigT<R> e0;
        if (w < B1) {
          if (!interior || w >= NBX) return;
          const int a = w % (TX + 2) - 1, b = w / (TX + 2);
          if (a == TX && !ex) return;
          const unsigned cb = cell_b(P, ox + a, oy + b, k);
          tv_prim(tv, cb - sxb, Vm);
          tv_prim(tv, cb, Vc);
          tv_prim(tv, cb + sxb, Vp);
          const R flat = P.flat ? tv_flat_coeff<0, R>(tv, cb, sxb) : R(0.);
          const R* lxk = LX + ringL(k) * 6 * I::NLX;
          lds_get6(lxk, I::NLX, I::lx(a - 1, b), dbN);
          lds_get6(lxk, I::NLX, I::lx(a, b), db0);
          lds_get6(lxk, I::NLX, I::lx(a + 1, b), dbP);
          if (!eig_make<0>(Vc, e0)) { bad = true; return; }
          const R dmm = P.steep ? R(tv_ld(tv, tv.ob[P_DENS], cb - 2 * sxb)) : R(0.);
          const R dpp = P.steep ? R(tv_ld(tv, tv.ob[P_DENS], cb + 2 * sxb)) : R(0.);
          ppm_trace<0>(P, ratio(0), ratio(9), Vm, Vc, Vp, dmm, dpp, dbN, db0, dbP, flat, e0, Wp, Wn);
          R sg[8];
          lds_get6(syk, I::NRY, I::ry(a, b), sg);
          sub_sigma(Wp, Wn, sg, ratio(4));
          if (Z) {
            lds_get6(SZ, I::NC, I::cc(a, b), sg);
            sub_sigma(Wp, Wn, sg, ratio(5));
          }
          if (negative(Wp, Wn)) fallback_first_order(Vc, Wp, Wn, st, a >= 0 || ox == P.lo[0]);
          put_state(WX, I::NRY, I::ry(a, b), Wp, Wn);
        } else if (w < B2) {
          const int w2 = w - B1;
          if (!interior || w2 >= NBY) return;
          const int a = w2 % TX, b = w2 / TX - 1;
          if (b == TY && !ey) return;
          const unsigned cb = cell_b(P, ox + a, oy + b, k);
          tv_prim(tv, cb - syb, Vm);
          tv_prim(tv, cb, Vc);
          tv_prim(tv, cb + syb, Vp);
          const R flat = P.flat ? tv_flat_coeff<1, R>(tv, cb, syb) : R(0.);
          const R* lyk = LY + ringL(k) * 6 * I::NLY;
          lds_get6(lyk, I::NLY, I::ly(a, b - 1), dbN);
          lds_get6(lyk, I::NLY, I::ly(a, b), db0);
          lds_get6(lyk, I::NLY, I::ly(a, b + 1), dbP);
          if (!eig_make<1>(Vc, e0)) { bad = true; return; }
          const R dmm = P.steep ? R(tv_ld(tv, tv.ob[P_DENS], cb - 2 * syb)) : R(0.);
          const R dpp = P.steep ? R(tv_ld(tv, tv.ob[P_DENS], cb + 2 * syb)) : R(0.);
          ppm_trace<1>(P, ratio(1), ratio(10), Vm, Vc, Vp, dmm, dpp, dbN, db0, dbP, flat, e0, Wp, Wn);
          R sg[8];
          lds_get6(sxk, I::NRX, I::rx(a, b), sg);
          sub_sigma(Wp, Wn, sg, ratio(3));
          if (Z) {
            lds_get6(SZ, I::NC, I::cc(a, b), sg);
            sub_sigma(Wp, Wn, sg, ratio(5));
          }
          if (negative(Wp, Wn)) fallback_first_order(Vc, Wp, Wn, st, b >= 0 || oy == P.lo[1]);
          put_state(WY, I::NRX, I::rx(a, b), Wp, Wn);
        } else {
          const int w2 = w - B2;
          const int a = w2 % TX, b = w2 / TX;
          const unsigned cb = cell_b(P, ox + a, oy + b, k);
          tv_prim(tv, cb - szb, Vm);
          tv_prim(tv, cb, Vc);
          tv_prim(tv, cb + szb, Vp);
          R flat = R(0.);
          if (P.flat) {
            const R dp1 = Vp[P_PRES] - Vm[P_PRES];
            const R f0 = FZ[ring(k) * NCOL + w2];
            flat = dp1 < R(0.0) ? fmax(f0, FZ[ring(k + 1) * NCOL + w2])
                                : (dp1 == R(0.) ? f0 : fmax(f0, FZ[ring(k - 1) * NCOL + w2]));
          }
          lds_get6(LZ + ring(k - 1) * 6 * NCOL, NCOL, w2, dbN);
          lds_get6(LZ + ring(k) * 6 * NCOL, NCOL, w2, db0);
          lds_get6(LZ + ring(k + 1) * 6 * NCOL, NCOL, w2, dbP);
          if (!eig_make<2>(Vc, e0)) { bad = true; return; }
          const R dmm = P.steep ? R(tv_ld(tv, tv.ob[P_DENS], cb - 2 * szb)) : R(0.);
          const R dpp = P.steep ? R(tv_ld(tv, tv.ob[P_DENS], cb + 2 * szb)) : R(0.);
          ppm_trace<2>(P, ratio(2), ratio(11), Vm, Vc, Vp, dmm, dpp, dbN, db0, dbP, flat, e0, Wp, Wn);
          R sg[8];
          lds_get6(sxk, I::NRX, I::rx(a, b), sg);
          sub_sigma(Wp, Wn, sg, ratio(3));
          lds_get6(syk, I::NRY, I::ry(a, b), sg);
          sub_sigma(Wp, Wn, sg, ratio(4));
          if (negative(Wp, Wn)) fallback_first_order(Vc, Wp, Wn, st, true);
          put_state(WZ, NCOL, w2, Wp, Wn);
        }
      };
      for (int w = tid; w < B3; w += NT) itemB1(item_index<OPAQUE>(w));
    }
    __syncthreads();

    // ------------------------------ stage B2 -----------------------------
    if (Z && P.ctu3d) {
      constexpr int C1 = wave_up(NBX), C2 = C1 + wave_up(NBY), C3 = C2 + NCOL;
      auto itemB2 = [&](int w) {
        R Wp[8], Wn[8], Vc[8], sm[8], sc[8], sp[8], tA[8], tB[8];
        EigT<R> e1, e2;
        R fac;
        R* base;
        int n, idx;
        bool own = true;  // the tile owns this state cell (counts its reset)
        if (w < C1) {
          // X: sigma_z over j with the y eigen system, sigma_y over k with z's
          if (!interior || w >= NBX) return;
          const int a = w % (TX + 2) - 1, b = w / (TX + 2);
          if (a == TX && !ex) return;
          own = a >= 0 || ox == P.lo[0];
          tv_prim(tv, cell_b(P, ox + a, oy + b, k), Vc);
          if (!(eig_make<1>(Vc, e1) &&
                eig_make<2>(Vc, e2))) { bad = true; return; }
          lds_get6(SZ, I::NC, I::cc(a, b - 1), sm);
          lds_get6(SZ, I::NC, I::cc(a, b), sc);
          lds_get6(SZ, I::NC, I::cc(a, b + 1), sp);
          trans_sigma<1>(P, sm, sc, sp, e1, tA);
          lds_get6(SY + ring(k - 1) * 6 * I::NRY, I::NRY, I::ry(a, b), sm);
          lds_get6(syk, I::NRY, I::ry(a, b), sc);
          lds_get6(SY + ring(k + 1) * 6 * I::NRY, I::NRY, I::ry(a, b), sp);
          trans_sigma<2>(P, sm, sc, sp, e2, tB);
          fac = ratio(6);
          base = WX; n = I::NRY; idx = I::ry(a, b);
        } else if (w < C2) {
          // Y: sigma_x over k with z's eigen system, sigma_z over i with x's
          const int w2 = w - C1;
          if (!interior || w2 >= NBY) return;
          const int a = w2 % TX, b = w2 / TX - 1;
          if (b == TY && !ey) return;
          own = b >= 0 || oy == P.lo[1];
          tv_prim(tv, cell_b(P, ox + a, oy + b, k), Vc);
          if (!(eig_make<2>(Vc, e1) &&
                eig_make<0>(Vc, e2))) { bad = true; return; }
          lds_get6(SX + ring(k - 1) * 6 * I::NRX, I::NRX, I::rx(a, b), sm);
          lds_get6(sxk, I::NRX, I::rx(a, b), sc);
          lds_get6(SX + ring(k + 1) * 6 * I::NRX, I::NRX, I::rx(a, b), sp);
          trans_sigma<2>(P, sm, sc, sp, e1, tA);
          lds_get6(SZ, I::NC, I::cc(a - 1, b), sm);
          lds_get6(SZ, I::NC, I::cc(a, b), sc);
          lds_get6(SZ, I::NC, I::cc(a + 1, b), sp);
          trans_sigma<0>(P, sm, sc, sp, e2, tB);
          fac = ratio(7);
          base = WY; n = I::NRX; idx = I::rx(a, b);
        } else {
          // Z: sigma_y over i with x's eigen system, sigma_x over j with y's
          const int w2 = w - C2;
          const int a = w2 % TX, b = w2 / TX;
          tv_prim(tv, cell_b(P, ox + a, oy + b, k), Vc);
          if (!(eig_make<0>(Vc, e1) &&
                eig_make<1>(Vc, e2))) { bad = true; return; }
          lds_get6(syk, I::NRY, I::ry(a - 1, b), sm);
          lds_get6(syk, I::NRY, I::ry(a, b), sc);
          lds_get6(syk, I::NRY, I::ry(a + 1, b), sp);
          trans_sigma<0>(P, sm, sc, sp, e1, tA);
          lds_get6(sxk, I::NRX, I::rx(a, b - 1), sm);
          lds_get6(sxk, I::NRX, I::rx(a, b), sc);
          lds_get6(sxk, I::NRX, I::rx(a, b + 1), sp);
          trans_sigma<1>(P, sm, sc, sp, e2, tB);
          fac = ratio(8);
          base = WZ; n = NCOL; idx = w2;
        }
        get_state(base, n, idx, Wp, Wn);
        add_ctu(Wp, Wn, tA, tB, fac);
        if (nonpositive(Wp, Wn)) fallback_first_order(Vc, Wp, Wn, st, own);
        put_state(base, n, idx, Wp, Wn);
      };
      for (int w = tid; w < C3; w += NT) itemB2(item_index<OPAQUE>(w));
      __syncthreads();
    }

    // ------------------------------ stage C ------------------------------
    {
      // x faces, y faces, column items (z face, pavg) in wave-aligned segments
      constexpr int D1 = wave_up(NFX), D2 = D1 + wave_up(NFY), D3 = D2 + NCOL;
      const bool cols = Z || interior;
      auto itemC = [&](int w) {
        R VL[8], VR[8];
        double F[8];
        if (w < D1) {
          if (!interior || w >= NFX) return;
          const int a = w % (TX + 1), b = w / (TX + 1);
          if (a == TX && !ex) return;
          lds_get6(WX, I::NRY, I::ry(a - 1, b), VL);
          lds_get6(WX + 6 * I::NRY, I::NRY, I::ry(a, b), VR);
          const unsigned cb = cell_b(P, ox + a, oy + b, k);
          VL[P_GAMC] = R(tv_ld(tv, tv.ob[P_GAMC], cb - sxb));
          VL[P_GAME] = R(tv_ld(tv, tv.ob[P_GAME], cb - sxb));
          VR[P_GAMC] = R(tv_ld(tv, tv.ob[P_GAMC], cb));
          VR[P_GAME] = R(tv_ld(tv, tv.ob[P_GAME], cb));
          const unsigned fr = (unsigned)face_index(P, 0, ox + a, oy + b, k);
          if constexpr (is_states(RS)) {
            write_face_states<0>(P, fsb + fstate_dir_offset(P, 0), fr, VL, VR);
            return;
          } else {
            if (face_flux<0, RS>(P, VL, VR, tv_use_hll<RS>(P, tv, cb - sxb, cb), F, es)) { report_error(st, 2, t, ox + a, oy + b, k, 0); return; }
            write_flux_b<0>(P, as_global(T.flux[0]), fr * 8u, F);
          }
        } else if (w < D2) {
          const int w2 = w - D1;
          if (!interior || w2 >= NFY) return;
          const int a = w2 % TX, b = w2 / TX;
          if (b == TY && !ey) return;
          lds_get6(WY, I::NRX, I::rx(a, b - 1), VL);
          lds_get6(WY + 6 * I::NRX, I::NRX, I::rx(a, b), VR);
          const unsigned cb = cell_b(P, ox + a, oy + b, k);
          VL[P_GAMC] = R(tv_ld(tv, tv.ob[P_GAMC], cb - syb));
          VL[P_GAME] = R(tv_ld(tv, tv.ob[P_GAME], cb - syb));
          VR[P_GAMC] = R(tv_ld(tv, tv.ob[P_GAMC], cb));
          VR[P_GAME] = R(tv_ld(tv, tv.ob[P_GAME], cb));
          const unsigned fr = (unsigned)face_index(P, 1, ox + a, oy + b, k);
          if constexpr (is_states(RS)) {
            write_face_states<1>(P, fsb + fstate_dir_offset(P, 1), fr, VL, VR);
            return;
          } else {
            if (face_flux<1, RS>(P, VL, VR, tv_use_hll<RS>(P, tv, cb - syb, cb), F, es)) { report_error(st, 2, t, ox + a, oy + b, k, 1); return; }
            write_flux_b<1>(P, as_global(T.flux[1]), fr * 8u, F);
          }
        } else {
          // column item: z-face flux at face k, pavg of the cell, carry Wp
          if (!cols) return;
          const int w2 = w - D2;
          const int a = w2 % TX, b = w2 / TX;
          const unsigned cb = cell_b(P, ox + a, oy + b, k);
          if (Z && k >= klo) {
            lds_get6(WZP, NCOL, w2, VL);
            lds_get6(WZ + 6 * NCOL, NCOL, w2, VR);
            VL[P_GAMC] = R(tv_ld(tv, tv.ob[P_GAMC], cb - szb));
            VL[P_GAME] = R(tv_ld(tv, tv.ob[P_GAME], cb - szb));
            VR[P_GAMC] = R(tv_ld(tv, tv.ob[P_GAMC], cb));
            VR[P_GAME] = R(tv_ld(tv, tv.ob[P_GAME], cb));
            const unsigned fr = (unsigned)face_index(P, 2, ox + a, oy + b, k);
            if constexpr (is_states(RS)) {
              write_face_states<2>(P, fsb + fstate_dir_offset(P, 2), fr, VL, VR);
            } else {
              if (face_flux<2, RS>(P, VL, VR, tv_use_hll<RS>(P, tv, cb - szb, cb), F, es)) report_error(st, 2, t, ox + a, oy + b, k, 2);
              else write_flux_b<2>(P, as_global(T.flux[2]), fr * 8u, F);
            }
          }
          if (interior && P.aux) {
            // face-averaged pressure (hy_getFaceFlux.F90:507-561 and the
            // y, z analogues)
            R pres = R(0.5) * (WX[10 * I::NRY + I::ry(a, b)] + WX[4 * I::NRY + I::ry(a, b)]);
            pres = pres * R(0.5) + (WY[10 * I::NRX + I::rx(a, b)] + WY[4 * I::NRX + I::rx(a, b)]) * R(0.25);
            if (Z) pres = div_const<6>(pres * R(4.0) + (WZ[10 * NCOL + w2] + WZ[4 * NCOL + w2]));
            const unsigned r = (unsigned)(((k - P.lo[2]) * P.nyb + (oy + b - P.lo[1])) * P.nxb + (ox + a - P.lo[0]));
            st_at(as_global(T.pavg), r * 8u, double(pres));
          }
          if (Z) {
#pragma unroll
            for (int q = 0; q < 6; ++q) WZP[q * NCOL + w2] = WZ[q * NCOL + w2];
          }
        }
      };
      for (int w = tid; w < D3; w += NT) itemC(item_index<OPAQUE>(w));
    }
    __syncthreads();
  }
  if (bad) report_error(st, 1, t, ox, oy, -1, -1);
}

// One workgroup per column tile of a batch of equal-shape tiles.  OCC > 0
// asks the compiler for OCC waves per SIMD (fewer registers, maybe spills);
// the autotuner decides whether that pays.  The FP64 16x8 tile runs 768
// threads, 3 waves per SIMD: its LDS allows one workgroup per CU, and with
// the item numbers opaque the march fits the 168 registers of 3 waves
// without spilling.  Stage A's 708 items then take one pass instead of two
// (B1, B2 and C keep 8 busy waves of 12): 8.83 -> 8.41 ms on 4096 x 16^3.
template <int NDIM, int TX, int TY, int NT, int RS, typename R, int OCC = 0>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1)))
k_march(DevParams P, MarchLaunch L) {
  const int tilesPerBlock = L.tilesX * L.tilesY;
  const int t = blockIdx.x / tilesPerBlock;
  const int tr = blockIdx.x - t * tilesPerBlock;
  const int ox = P.lo[0] + (tr % L.tilesX) * TX;
  const int oy = P.lo[1] + (tr / L.tilesX) * TY;
  ExactStats es{0ULL, 0ULL};
  R* fsb = is_states(RS) ? static_cast<R*>(L.fstates) + t * fstate_tile_stride(P) : nullptr;
  __shared__ R smem[MarchLds<NDIM, TX, TY>::total];
  march_tile<NDIM, TX, TY, NT, RS, R, (sizeof(R) == 4 || NT > 512)>(P, L, t, ox, oy, fsb, es,
                                                                    smem);
  flush_exact(L.status, es);
}

// Tiles of several block classes in one persistent launch: workgroup b takes
// the column tiles b, b + G, b + 2G, ... of the item list (G = grid size).
// The host orders the items by falling column height, so the round-robin
// deals the long columns first and the short ones fill in behind them.
struct MarchItem {
  int t, cls, ox, oy;  // tile, its class, column-tile origin (guard-inclusive)
  int tx;              // column-tile width of the item (mixed-width launches)
};

template <int NDIM, int TX, int TY, int NT, int RS, typename R, int OCC = 0>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1)))
k_march_multi(const DevParams* __restrict__ cls,
                                                    MarchLaunch L,
                                                    const MarchItem* __restrict__ items,
                                                    int nItems, const long* __restrict__ faceStart) {
  ExactStats es{0ULL, 0ULL};
  __shared__ R smem[MarchLds<NDIM, TX, TY>::total];
  for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
    const MarchItem w = items[it];
    const DevParams& P = cls[w.cls];
    const long faceBase = is_states(RS) ? faceStart[w.t] : 0;
    R* fsb = is_states(RS) ? static_cast<R*>(L.fstates) + kFaceComps * faceBase : nullptr;
    march_tile<NDIM, TX, TY, NT, RS, R, true>(P, L, w.t, w.ox, w.oy, fsb, es, smem);
  }
  flush_exact(L.status, es);
}

// The same with the column-tile width chosen per item: 16 columns across x
// for classes whose x extent allows it, 8 for the others.  One LDS array of
// the larger layout serves both; NT is the 16-wide tile's thread count.
template <int NDIM, int NT, int RS, typename R, int OCC = 0>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1)))
k_march_mixw(const DevParams* __restrict__ cls, MarchLaunch L,
             const MarchItem* __restrict__ items, int nItems, const long* __restrict__ faceStart) {
  ExactStats es{0ULL, 0ULL};
  constexpr int n16 = MarchLds<NDIM, 16, 8>::total, n8 = MarchLds<NDIM, 8, 8>::total;
  __shared__ R smem[n16 > n8 ? n16 : n8];
  for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
    const MarchItem w = items[it];
    const DevParams& P = cls[w.cls];
    const long faceBase = is_states(RS) ? faceStart[w.t] : 0;
    R* fsb = is_states(RS) ? static_cast<R*>(L.fstates) + kFaceComps * faceBase : nullptr;
    if (w.tx == 16) march_tile<NDIM, 16, 8, NT, RS, R, true>(P, L, w.t, w.ox, w.oy, fsb, es, smem);
    else march_tile<NDIM, 8, 8, NT, RS, R, true>(P, L, w.t, w.ox, w.oy, fsb, es, smem);
  }
  flush_exact(L.status, es);
}

}  // namespace fx
