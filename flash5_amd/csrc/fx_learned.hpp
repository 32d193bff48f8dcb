// fx_learned.hpp -- kernels of the learned Newton seed (no counterpart in
// the reference): interface features, MLP inference and training on BF16
// MFMA (fx_mlp.hpp), Adam, and the batched exact Riemann solver whose Newton
// iteration starts from the two-shock guess or from the network.
//
// Included inside fx_hydro.hip's anonymous namespace (uses its helpers).
// Every loop that calls a wavefront-wide routine (MFMA, shuffles) is
// uniform per wavefront: a wavefront takes 64 consecutive interfaces per
// trip and lanes past the end carry a harmless dummy interface.

constexpr int kLearnThreads = 256;          // 4 wavefronts per workgroup
constexpr int kLearnWaves = kLearnThreads / 64;

// states: SoA [6][n] = rhoL, uL, pL, rhoR, uR, pR
struct IfaceState {
  double rl, ul, pl, rr, ur, pr;
};
__device__ __forceinline__ IfaceState load_iface(const double* __restrict__ S, long n, long i,
                                                 bool act) {
  if (!act) return IfaceState{1.0, 0.0, 1.0, 1.0, 0.0, 1.0};
  return IfaceState{S[i], S[n + i], S[2 * n + i], S[3 * n + i], S[4 * n + i], S[5 * n + i]};
}

__global__ void __launch_bounds__(kThreads) k_mlp_features(const double* __restrict__ S, long n,
                                                           double gamma, float* __restrict__ x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const IfaceState s = load_iface(S, n, i, true);
    float f[6];
    double P0;
    mlp_features(gamma, s.rl, s.ul, s.pl, s.rr, s.ur, s.pr, f, P0);
#pragma unroll
    for (int k = 0; k < 6; ++k) x[i * 6 + k] = f[k];
  }
}

// stage this wavefront's 64 feature rows (x[n][6]) into LDS rows of 8
__device__ __forceinline__ void stage_x(const float* __restrict__ x, long n, long i, bool act,
                                        float* xs, int lane) {
#pragma unroll
  for (int k = 0; k < 6; ++k) xs[lane * 8 + k] = act ? x[i * 6 + k] : 0.f;
  xs[lane * 8 + 6] = 0.f;
  xs[lane * 8 + 7] = 0.f;
  wave_lds_sync();
}

__global__ void __launch_bounds__(kLearnThreads) k_mlp_forward(const MlpPacked* __restrict__ W,
                                                               const float* __restrict__ x, long n,
                                                               float* __restrict__ y) {
  __shared__ float xs[kLearnWaves][64 * 8];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const long nw = (long)gridDim.x * kLearnWaves;
  for (long base = ((long)blockIdx.x * kLearnWaves + wv) * 64; base < n; base += nw * 64) {
    const long i = base + lane;
    const bool act = i < n;
    stage_x(x, n, i, act, xs[wv], lane);
    MlpActs A;
    float yo;
    mlp_forward_wave(*W, xs[wv], lane, A, yo);
    if (act) y[i] = yo;
    wave_lds_sync();
  }
}

// grad += d/dparams of scale * sum (y - t)^2 ; loss += scale * sum (y - t)^2
__global__ void __launch_bounds__(kLearnThreads) k_mlp_grad(const MlpPacked* __restrict__ W,
                                                            const float* __restrict__ x,
                                                            const float* __restrict__ t, long n,
                                                            float scale, float* grad, float* loss) {
  __shared__ float xs[kLearnWaves][64 * 8];
  // 16-byte aligned: the weight-gradient products read 8 bf16 at a time
  __shared__ __attribute__((aligned(16))) unsigned short sZ[kLearnWaves][64 * 64];
  __shared__ __attribute__((aligned(16))) unsigned short sH[kLearnWaves][64 * 64];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const long nw = (long)gridDim.x * kLearnWaves;
  float lsum = 0.f;
  for (long base = ((long)blockIdx.x * kLearnWaves + wv) * 64; base < n; base += nw * 64) {
    const long i = base + lane;
    const bool act = i < n;
    stage_x(x, n, i, act, xs[wv], lane);
    MlpActs A;
    float yo;
    mlp_forward_wave(*W, xs[wv], lane, A, yo);
    const float r = act ? yo - t[i] : 0.f;
    lsum += scale * r * r;
    mlp_backward_wave(*W, xs[wv], lane, A, 2.f * scale * r, sZ[wv], sH[wv], grad);
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) lsum += __shfl_xor(lsum, o);
  if (lane == 0) atomicAdd(loss, lsum);
}

__global__ void __launch_bounds__(kThreads) k_mlp_adam(float* __restrict__ prm,
                                                       const float* __restrict__ grad,
                                                       float* __restrict__ m, float* __restrict__ v,
                                                       float lr, float b1, float b2, float c1,
                                                       float c2, float eps) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= MLP_NPARAM) return;
  const float gk = grad[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  prm[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// The network predicts y = ln(p* / P0), P0 = sqrt(pL pR).  Weak interfaces
// (every |feature| below the gate, fx_mlp_set_gate; seed_gate_pass) keep the two-shock guess, which is
// already accurate to high order there; for the others Newton starts from
// whichever of P0 * exp(y) and the two-shock guess has the smaller residual
// of the pressure function.
__device__ __forceinline__ double seed_from_y(float y, double pl, double pr) {
  return sqrt(pl * pr) * exp((double)fminf(fmaxf(y, -60.f), 60.f));
}

// Network seed y = ln(p*/P0) for every interface of a batch (inference only;
// the Newton kernel below reads it, so each keeps its own register budget).
__global__ void __launch_bounds__(kLearnThreads) k_mlp_seed_states(
    const double* __restrict__ S, long n, double gamma, const MlpPacked* __restrict__ W,
    float* __restrict__ y, SeedGate gate) {
  __shared__ float xs[kLearnWaves][64 * 8];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const long nw = (long)gridDim.x * kLearnWaves;
  for (long base = ((long)blockIdx.x * kLearnWaves + wv) * 64; base < n; base += nw * 64) {
    const long i = base + lane;
    const bool act = i < n;
    const IfaceState s = load_iface(S, n, i, act);
    // wavefronts whose interfaces are all weak skip the network
    const bool need = act && seed_gate_pass(gate, gamma, s.rl, s.ul, s.pl, s.rr, s.ur, s.pr);
    if (__ballot(need) == 0) continue;
    float f[6];
    double P0;
    mlp_features(gamma, s.rl, s.ul, s.pl, s.rr, s.ur, s.pr, f, P0);
#pragma unroll
    for (int k = 0; k < 6; ++k) xs[wv][lane * 8 + k] = f[k];
    xs[wv][lane * 8 + 6] = 0.f;
    xs[wv][lane * 8 + 7] = 0.f;
    wave_lds_sync();
    MlpActs A;
    float yo;
    mlp_forward_wave(*W, xs[wv], lane, A, yo);
    if (act) y[i] = yo;
    wave_lds_sync();
  }
}

// Exact Riemann problems in a batch: p* by Newton (FP64 iterate and
// convergence test) from the two-shock guess, or from the network's seed
// (SEEDED); a seeded start that does not converge restarts from the
// two-shock guess.  iters: Newton iterations spent, negative when even the
// two-shock start did not converge.
template <bool SEEDED>
__global__ void __launch_bounds__(kThreads) k_riemann_batch(
    const double* __restrict__ S, long n, double gamma, const float* __restrict__ yseed,
    double tol, int maxit, double* __restrict__ pstar, int* __restrict__ iters, DevStatus* st,
    SeedGate gate) {
  ExactStats es{0ULL, 0ULL};
  const double g = gamma;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const IfaceState s = load_iface(S, n, i, true);
    const double cl = sqrt(g * s.pl / s.rl), cr = sqrt(g * s.pr / s.rr);
    if (!(s.rl > 0. && s.rr > 0. && s.pl > 0. && s.pr > 0.) ||
        2. / (g - 1.) * (cl + cr) <= s.ur - s.ul) {
      pstar[i] = 0.0;
      iters[i] = 0;
      report_error(st, 2, -1, (int)(i & 0x7fffffff), 0, 0, 0);
      continue;
    }
    const double du = s.ur - s.ul;
    const double ts = two_shock_guess(g, s.rl, s.ul, s.pl, cl, s.rr, s.ur, s.pr, cr);
    const bool use = SEEDED && seed_gate_pass(gate, g, s.rl, s.ul, s.pl, s.rr, s.ur, s.pr);
    double pm, pb = use ? seed_from_y(yseed[i], s.pl, s.pr) : -1.0;
    int it, total = 0;
    // one Newton loop for the seeded start and its two-shock restart
    for (;;) {
      int pick;
      it = exact_newton_pick(g, s.rl, s.pl, cl, s.rr, s.pr, cr, du, ts, pb, tol, maxit, pm, pick);
      total += it < 0 ? -it : it;
      if (it < 0 && (pb > 0. || !use)) es.fallbacks += 1;   // a restart counts once
      if (it >= 0 || pick == 0) break;
      pb = -1.0;
    }
    es.iters += (unsigned long long)total;
    pstar[i] = pm;
    iters[i] = it < 0 ? -total : total;
  }
  flush_exact(st, es);
}

// ---- separate Riemann stage of the hydro step (exact solver) ----
// The march kernel stored each face's left/right states; these kernels solve
// one face per thread at full occupancy, optionally from the network seed.
struct FaceLaunch {
  const fx_tile* tiles;
  int nTiles;
  const void* fstates;     // [tile][dir][comp][face], from the march's RS_STATES stage;
                           // double, or float in the mixed-precision mode
  const float* y;          // network seed per face (flattened like gid), or null
  DevStatus* status;
  // tiles of several block classes (the MULTI kernels): parameters per class,
  // class of each tile, faces before each tile (faceStart[nTiles] = all)
  const DevParams* cls;
  const unsigned char* tcls;
  const long* faceStart;
  const int* wtab;          // tile of every 64th face (find_tile_w)
};

struct FaceRef {
  int t, d;
  long r;          // face index within the direction's flux array
  long cR, cL;     // right / left cells (guard-inclusive index)
};

// all faces of the launch
template <bool MULTI>
__device__ __forceinline__ long face_total(const DevParams& P0, const FaceLaunch& F) {
  if constexpr (MULTI) return F.faceStart[F.nTiles];
  else return (P0.nface[0] + P0.nface[1] + P0.nface[2]) * F.nTiles;
}

// Tile t holding flat face gid, gid's index r within it, and the tile's
// parameters: the launch's own, or its class's (MULTI).
template <bool MULTI>
__device__ __forceinline__ const DevParams& face_tile(const DevParams& P0, const FaceLaunch& F,
                                                      long gid, int& t, long& r) {
  if constexpr (MULTI) {
    t = find_tile_w(F.faceStart, F.wtab, F.nTiles, gid);
    r = gid - F.faceStart[t];
    return F.cls[F.tcls[t]];
  } else {
    const long perTile = P0.nface[0] + P0.nface[1] + P0.nface[2];
    t = (int)(gid / perTile);
    r = gid - (long)t * perTile;
    return P0;
  }
}

// offset (in elements) of tile t's face states in F.fstates
template <bool MULTI>
__device__ __forceinline__ long face_state_offset(const DevParams& P, const FaceLaunch& F, int t) {
  if constexpr (MULTI) return kFaceComps * F.faceStart[t];
  else return t * fstate_tile_stride(P);
}

// tile t's face states of direction d, stored as S
template <typename S, bool MULTI>
__device__ __forceinline__ const S* face_states_at(const DevParams& P, const FaceLaunch& F, int t,
                                                   int d) {
  return static_cast<const S*>(F.fstates) + face_state_offset<MULTI>(P, F, t) +
         fstate_dir_offset(P, d);
}

// face rt of tile t (index arithmetic in 32 bits: a tile holds < 2^31 faces)
__device__ __forceinline__ FaceRef face_ref(const DevParams& P, int t, long rt) {
  FaceRef f;
  f.t = t;
  int r = (int)rt;
  int d = 0;
  while (d < P.ndim - 1 && r >= (int)P.nface[d]) {
    r -= (int)P.nface[d];
    ++d;
  }
  f.d = d;
  f.r = r;
  const int f0 = P.fext[d][0], f1 = P.fext[d][1];
  const int a0 = r % f0, a1 = (r / f0) % f1, a2 = r / (f0 * f1);
  f.cR = cidx(P, a0 + P.lo[0], a1 + P.lo[1], a2 + P.lo[2]);
  f.cL = f.cR - axis_stride(P, d);
  return f;
}

// One code path serves the three directions: the face-normal velocity and
// the flux component that takes the pressure are picked at run time.  R is
// the solver's arithmetic: double, or float in the mixed-precision mode
// (FP64 Newton iterate and convergence test, FP64 fluxes out).
// (3 waves per SIMD fit the FP64 solver's registers without spilling, 4 the
// FP32 one's)
template <bool SEEDED, bool MULTI, typename R>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == sizeof(float) ? 4 : 3)))
k_face_solve(DevParams P0, FaceLaunch F) {
  const long total = face_total<MULTI>(P0, F);
  ExactStats es{0ULL, 0ULL};
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (long)gridDim.x * blockDim.x) {
    int t;
    long rt;
    const DevParams& P = face_tile<MULTI>(P0, F, gid, t, rt);
    const FaceRef f = face_ref(P, t, rt);
    const int d = f.d;
    const R* b = face_states_at<R, MULTI>(P, F, t, d);
    const long nf = P.nface[d];
    // face states: comps 0..5 dens, velx, vely, velz, pres, eint; left, then right
    const R* Lb = b + f.r;
    const R* Rb = b + 6 * nf + f.r;
    // a face without a seed takes the unseeded solver, as the batch step's
    // solve+update kernel gives it (the same FP32 code, the same bits)
    const float y = SEEDED ? F.y[gid] : kNoSeed;
    const gdouble* U = as_global(F.tiles[f.t].U);
    const double gamc = U[P.voff[FX_GAMC] + f.cL], gL = U[P.voff[FX_GAME] + f.cL],
                 gR = U[P.voff[FX_GAME] + f.cR];
    double Fo[8];
    const int ierr =
        y != kNoSeed
            ? exact_face_flux<true, R>(P, d, Lb, Rb, nf, gamc, gL, gR,
                                       seed_from_y(y, Lb[4 * nf], Rb[4 * nf]), Fo, es)
            : exact_face_flux<false, R>(P, d, Lb, Rb, nf, gamc, gL, gR, -1.0, Fo, es);
    if (ierr) {
      report_error(F.status, 2, f.t, (int)(f.cR % P.nx), (int)((f.cR / P.nx) % P.ny),
                   (int)(f.cR / ((long)P.nx * P.ny)), d);
      continue;
    }
    gdouble* out = as_global(F.tiles[f.t].flux[d]);
#pragma unroll
    for (int q = 0; q < FX_NFLUX; ++q) out[q * nf + f.r] = Fo[q];
  }
  flush_exact(F.status, es);
}

// Bits of the seeded faces of a batch step: tile t owns words
// [t * W, (t + 1) * W) with W = seed_bit_words(P), bit rt of them for its
// face rt (directions in order), so that no word spans two tiles.
__host__ __device__ inline long seed_bit_words(const DevParams& P) {
  return (P.nface[0] + P.nface[1] + P.nface[2] + 31) / 32;
}

// The learned seed's strong faces of a batch step, solved ahead of
// k_solve_update: each listed face (its network seed already in F.y) is
// solved seeded, its flux goes to over[q * faces + gid] in R and its bit in
// bits is set; k_solve_update takes that flux and clears the bit, so the
// unseeded solver there carries no seeded code.  A face the seeded solve
// fails on gets a zero flux, as k_solve_update gives its own failures.
template <typename R>
__global__ void __launch_bounds__(kThreads) k_seeded_solve(DevParams P, FaceLaunch F,
                                                           const unsigned* __restrict__ list,
                                                           const unsigned* __restrict__ count,
                                                           long cap, long faces, R* over,
                                                           unsigned* bits) {
  const long total = (long)*count < cap ? (long)*count : cap;
  ExactStats es{0ULL, 0ULL};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long gid = (long)list[i];
    int t;
    long rt;
    face_tile<false>(P, F, gid, t, rt);
    const FaceRef f = face_ref(P, t, rt);
    const int d = f.d;
    const R* b = face_states_at<R, false>(P, F, t, d);
    const long nf = P.nface[d];
    const R* Lb = b + f.r;
    const R* Rb = b + 6 * nf + f.r;
    const float y = F.y[gid];
    const double guess = y != kNoSeed ? seed_from_y(y, Lb[4 * nf], Rb[4 * nf]) : -1.0;
    const gdouble* U = as_global(F.tiles[t].U);
    double Fo[8];
    const int ierr = exact_face_flux<true, R>(P, d, Lb, Rb, nf, U[P.voff[FX_GAMC] + f.cL],
                                              U[P.voff[FX_GAME] + f.cL],
                                              U[P.voff[FX_GAME] + f.cR], guess, Fo, es);
    if (ierr)
      report_error(F.status, 2, t, (int)(f.cR % P.nx), (int)((f.cR / P.nx) % P.ny),
                   (int)(f.cR / ((long)P.nx * P.ny)), d);
#pragma unroll
    for (int q = 0; q < 8; ++q) over[q * faces + gid] = ierr ? R(0.) : R(Fo[q]);
    atomicOr(&bits[t * seed_bit_words(P) + (rt >> 5)], 1u << (rt & 31));
  }
  flush_exact(F.status, es);
}

// The exact solver's face stage and the conservative update in one kernel,
// for batches of 3-D blocks (no caller tile list, every block in place):
// one workgroup per block sweeps its planes upward.  Per plane k it solves
// the x and y faces of plane k and the z face below it (face k), keeping
// the fluxes in LDS; then each column finishes its cell of plane k-1 (the z
// term with faces k-1 and k, the energy fix, the EOS) and starts its cell of
// plane k (the x and y terms), as upd_start / upd_finish, so the bits are
// those of k_face_solve followed by k_update.  Every face reads a cell's
// gamc and game before that cell is updated: the faces of plane k are
// solved before plane k-1 is written.  Blocks of at most 16 x 16 cells
// across (the LDS holds one plane of fluxes).  SEEDED: faces whose bit is
// set in bits take k_seeded_solve's flux from over; the block's bit words
// are copied to LDS at the start and cleared in memory (at most
// kSolveMaxBitWords of them).
constexpr int kSolveThreads = 256;
constexpr int kSolveMaxCols = kSolveThreads;   // nxb * nyb: one column per thread
constexpr int kSolveMaxFaces = 17 * 16 * 2 + 16 * 16;
constexpr int kSolveMaxBitWords = 2048;

template <bool SEEDED, typename R>
__global__ void __launch_bounds__(kSolveThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_solve_update(DevParams P, FaceLaunch F, double dt0, const double* dtp, const R* over,
               long faces, unsigned* bits) {
  // this plane's face fluxes (x, y, then z), in the solver's type R, which
  // holds them exactly (the solver forms them in R)
  __shared__ R flx[8 * kSolveMaxFaces];
  const int t = blockIdx.x;
  const int c = threadIdx.x;                     // this thread's column
  const double dt = step_dt(dt0, dtp);
  const fx_tile& T = F.tiles[t];
  gdouble* const U = as_global(T.U);
  const int nx = P.nxb, ny = P.nyb, ncol = nx * ny;
  const int nfx = (nx + 1) * ny, nfy = nx * (ny + 1);   // faces of one plane
  const long perTile = P.nface[0] + P.nface[1] + P.nface[2];
  const double dx = T.del[0], dy = T.del[1], dz = T.del[2];
  const int a = c % nx, bb = c / nx;
  UpdPart part{};                                // the column's cell in flight
  double zprev[8];                               // the z flux below it
  ExactStats es{0ULL, 0ULL};
  __shared__ unsigned sbits[SEEDED ? kSolveMaxBitWords : 1];
  if (SEEDED) {
    const long nw = seed_bit_words(P);
    unsigned* const tb = bits + (long)t * nw;
    for (int i = c; i < nw; i += kSolveThreads) {
      const unsigned v = tb[i];
      sbits[i] = v;
      if (v) tb[i] = 0u;
    }
    __syncthreads();
  }
  for (int k = P.lo[2]; k <= P.hi[2] + 1; ++k) {
    const bool inPlane = k <= P.hi[2];
    const int zBase = inPlane ? nfx + nfy : 0;
    const int nItems = zBase + ncol;
    for (int w = c; w < nItems; w += kSolveThreads) {
      int d, a0, a1;
      if (w < zBase && w < nfx) {
        d = 0; a0 = w % (nx + 1); a1 = w / (nx + 1);
      } else if (w < zBase) {
        d = 1; const int r = w - nfx; a0 = r % nx; a1 = r / nx;
      } else {
        d = 2; const int r = w - zBase; a0 = r % nx; a1 = r / nx;
      }
      const int i = a0 + P.lo[0], j = a1 + P.lo[1];
      const long r = face_index(P, d, i, j, k);
      const long cR = cidx(P, i, j, k), cL = cR - axis_stride(P, d);
      const int slot = d == 2 ? nfx + nfy + (w - zBase) : w;
      if (SEEDED) {
        long rt = r;
        for (int e = 0; e < d; ++e) rt += P.nface[e];
        if ((sbits[rt >> 5] >> (rt & 31)) & 1u) {
          const long gid = (long)t * perTile + rt;
#pragma unroll
          for (int q = 0; q < 8; ++q) flx[q * kSolveMaxFaces + slot] = over[q * faces + gid];
          continue;
        }
      }
      const R* b = face_states_at<R, false>(P, F, t, d);
      const long nf = P.nface[d];
      const R* Lb = b + r;
      const R* Rb = b + 6 * nf + r;
      double Fo[8];
      const int ierr = exact_face_flux<false, R>(P, d, Lb, Rb, nf, U[P.voff[FX_GAMC] + cL],
                                                 U[P.voff[FX_GAME] + cL], U[P.voff[FX_GAME] + cR],
                                                 -1.0, Fo, es);
      if (ierr) report_error(F.status, 2, t, i, j, k, d);
#pragma unroll
      for (int q = 0; q < 8; ++q) flx[q * kSolveMaxFaces + slot] = ierr ? R(0.) : R(Fo[q]);
    }
    __syncthreads();
    if (c < ncol) {
      const int i = a + P.lo[0], j = bb + P.lo[1];
      double FR[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) FR[q] = double(flx[q * kSolveMaxFaces + nfx + nfy + c]);
      if (k - 1 >= P.lo[2]) {
        double out[10];
        const long cb = cidx(P, i, j, k - 1);
        upd_finish(P, dt, dz, part, zprev, FR, P.aux ? 0. : U[P.voff[FX_EINT] + cb], out);
#pragma unroll
        for (int v = 0; v < 10; ++v) U[P.voff[upd_var(v)] + cb] = out[v];
      }
      if (inPlane) {
        const long cb = cidx(P, i, j, k);
        double FLx[8], FRx[8], FLy[8], FRy[8];
        const int fx0 = bb * (nx + 1) + a, fy0 = nfx + bb * nx + a;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          FLx[q] = double(flx[q * kSolveMaxFaces + fx0]);
          FRx[q] = double(flx[q * kSolveMaxFaces + fx0 + 1]);
          FLy[q] = double(flx[q * kSolveMaxFaces + fy0]);
          FRy[q] = double(flx[q * kSolveMaxFaces + fy0 + nx]);
        }
        const long rp = ((long)(k - P.lo[2]) * ny + bb) * nx + a;
        part = upd_start(P, dt, dx, dy, U[P.voff[FX_DENS] + cb], U[P.voff[FX_VELX] + cb],
                         U[P.voff[FX_VELY] + cb], U[P.voff[FX_VELZ] + cb],
                         U[P.voff[FX_ENER] + cb], U[P.voff[FX_EINT] + cb],
                         P.aux ? as_global(T.pavg)[rp] : 0., FLx, FRx, FLy, FRy);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) zprev[q] = FR[q];
    }
    __syncthreads();
  }
  flush_exact(F.status, es);
}

// Network seed ln(p*/P0) of the faces the march listed as strong (at or
// above the gate); every other face already holds "no seed".  The list
// length is read on the device, so the launch needs no host sync.
template <bool MULTI, typename S>
__global__ void __launch_bounds__(kLearnThreads) k_face_seed(DevParams P0, FaceLaunch F,
                                                             const MlpPacked* __restrict__ W,
                                                             float* __restrict__ y,
                                                             const unsigned* __restrict__ list,
                                                             const unsigned* __restrict__ count,
                                                             long cap) {
  __shared__ float xs[kLearnWaves][64 * 8];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const long total = (long)*count < cap ? (long)*count : cap;
  const long nw = (long)gridDim.x * kLearnWaves;
  for (long base = ((long)blockIdx.x * kLearnWaves + wv) * 64; base < total; base += nw * 64) {
    const long i = base + lane;
    const bool act = i < total;
    const long gid = act ? (long)list[i] : 0;
    double rl = 1.0, ul = 0.0, pl = 1.0, rr = 1.0, ur = 0.0, pr = 1.0;
    if (act) {
      int t;
      long rt;
      const DevParams& P = face_tile<MULTI>(P0, F, gid, t, rt);
      const FaceRef f = face_ref(P, t, rt);
      const S* b = face_states_at<S, MULTI>(P, F, t, f.d);
      const long nf = P.nface[f.d];
      rl = b[0 * nf + f.r];
      ul = b[(1 + f.d) * nf + f.r];
      pl = b[4 * nf + f.r];
      rr = b[6 * nf + f.r];
      ur = b[(7 + f.d) * nf + f.r];
      pr = b[10 * nf + f.r];
    }
    float x[6];
    double pref;
    mlp_features(P0.gamma, rl, ul, pl, rr, ur, pr, x, pref);
#pragma unroll
    for (int k = 0; k < 6; ++k) xs[wv][lane * 8 + k] = x[k];
    xs[wv][lane * 8 + 6] = 0.f;
    xs[wv][lane * 8 + 7] = 0.f;
    wave_lds_sync();
    MlpActs A;
    float yo;
    mlp_forward_wave(*W, xs[wv], lane, A, yo);
    if (act) y[gid] = yo;
    wave_lds_sync();
  }
}

__global__ void k_zero_count(unsigned* c) { c[threadIdx.x] = 0u; }

// The learned seed's gate over every face of the launch, from the face
// states the march stored: a face at or above the gate (seed_gate_pass, in
// FP64 from the stored values) joins the strong list (one atomic per
// wavefront, ranks by ballot) for k_face_seed.  With y, every face's seed
// slot also becomes "no seed" first (k_face_seed then fills the listed
// ones): the face stage that reads a seed per face needs it, the batch
// step's list-driven solve does not.  One pass over the states, apart from
// the march, so that neither march carries the gate.
template <bool MULTI, typename S>
__global__ void __launch_bounds__(kThreads) k_face_gate(DevParams P0, FaceLaunch F, SeedGate G,
                                                        float* __restrict__ y,
                                                        unsigned* __restrict__ list,
                                                        unsigned* __restrict__ count,
                                                        unsigned cap,
                                                        const int2* __restrict__ chunks) {
  const int lane = (int)(threadIdx.x & 63);
  // one face: its gate test, and its seed slot reset
  auto gate_one = [&](const DevParams& P, int t, long gid, int r) {
    int d = 0;
    while (d < P.ndim - 1 && r >= (int)P.nface[d]) {
      r -= (int)P.nface[d];
      ++d;
    }
    const S* b = face_states_at<S, MULTI>(P, F, t, d);
    const long nf = P.nface[d];
    const double rl = b[r], ul = b[(1 + d) * nf + r], pl = b[4 * nf + r];
    const double rr = b[6 * nf + r], ur = b[(7 + d) * nf + r], pr = b[10 * nf + r];
    if (y) y[gid] = kNoSeed;
    return rl > 0. && rr > 0. && pl > 0. && pr > 0. &&
           seed_gate_pass(G, P.gamma, rl, ul, pl, rr, ur, pr);
  };
  auto append = [&](bool strong, long gid) {
    const unsigned long long mask = __ballot(strong);
    if (mask) {
      const int leader = __builtin_ctzll(mask);
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(mask));
      base = __shfl(base, leader);
      const unsigned slot = base + (unsigned)__popcll(mask & ((1ULL << lane) - 1ULL));
      if (strong && slot < cap) list[slot] = (unsigned)gid;
    }
  };
  if constexpr (MULTI) {
    // workgroup w takes kThreads faces of one tile, chunks[w] = (tile, first
    // face of the tile): the tile's class is uniform, no per-face search
    const int2 ch = chunks[blockIdx.x];
    const int t = ch.x;
    const DevParams& P = F.cls[F.tcls[t]];
    const long perTile = P.nface[0] + P.nface[1] + P.nface[2];
    const int rt = ch.y + (int)threadIdx.x;
    const long gid = F.faceStart[t] + rt;
    const bool strong = rt < perTile && gate_one(P, t, gid, rt);
    append(strong, gid);
  } else {
    // face numbers fit 32 bits (the strong list holds them); the loop bound
    // is the same for every lane of a workgroup, so whole wavefronts reach
    // the ballot together
    const unsigned pt = (unsigned)(P0.nface[0] + P0.nface[1] + P0.nface[2]);
    const long total = (long)pt * F.nTiles;
    for (long g0 = (long)blockIdx.x * blockDim.x; g0 < total;
         g0 += (long)gridDim.x * blockDim.x) {
      const long gid = g0 + threadIdx.x;
      bool strong = false;
      if (gid < total) {
        const unsigned ut = (unsigned)gid / pt;
        strong = gate_one(P0, (int)ut, gid, (int)((unsigned)gid - ut * pt));
      }
      append(strong, gid);
    }
  }
}

// face states of the last exact-solver step as Riemann problems, SoA [6][n]
// (rhoL, uL, pL, rhoR, uR, pR with u the face-normal velocity); S is the
// type they were stored in
template <bool MULTI, typename S>
__global__ void __launch_bounds__(kThreads) k_collect_faces(DevParams P0, FaceLaunch F, long n,
                                                            double* __restrict__ out) {
  for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < n;
       gid += (long)gridDim.x * blockDim.x) {
    int t;
    long rt;
    const DevParams& P = face_tile<MULTI>(P0, F, gid, t, rt);
    const FaceRef f = face_ref(P, t, rt);
    const S* b = face_states_at<S, MULTI>(P, F, t, f.d);
    const long nf = P.nface[f.d];
    out[gid] = b[0 * nf + f.r];
    out[n + gid] = b[(1 + f.d) * nf + f.r];
    out[2 * n + gid] = b[4 * nf + f.r];
    out[3 * n + gid] = b[6 * nf + f.r];
    out[4 * n + gid] = b[(7 + f.d) * nf + f.r];
    out[5 * n + gid] = b[10 * nf + f.r];
  }
}
