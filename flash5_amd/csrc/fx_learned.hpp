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

// Network seed y = ln(p*/P0) for every interface of a batch (inference only;
// the Newton kernel below reads it, so each keeps its own register budget).
__global__ void __launch_bounds__(kLearnThreads) k_mlp_seed_states(
    const double* __restrict__ S, long n, double gamma, const MlpPacked* __restrict__ W,
    float* __restrict__ y) {
  __shared__ float xs[kLearnWaves][64 * 8];
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const long nw = (long)gridDim.x * kLearnWaves;
  for (long base = ((long)blockIdx.x * kLearnWaves + wv) * 64; base < n; base += nw * 64) {
    const long i = base + lane;
    const bool act = i < n;
    const IfaceState s = load_iface(S, n, i, act);
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

// start of Newton from a network output y = ln(p*/P0)
__device__ __forceinline__ double seed_from_y(float y, double pl, double pr) {
  return sqrt(pl * pr) * exp((double)fminf(fmaxf(y, -60.f), 60.f));
}

// Exact Riemann problems in a batch: p* by Newton (FP64 iterate and
// convergence test) from the two-shock guess, or from the network's seed
// (SEEDED); a seeded start that does not converge restarts from the
// two-shock guess.  iters: Newton iterations spent, negative when even the
// two-shock start did not converge.
template <bool SEEDED>
__global__ void __launch_bounds__(kThreads) k_riemann_batch(
    const double* __restrict__ S, long n, double gamma, const float* __restrict__ yseed,
    double tol, int maxit, double* __restrict__ pstar, int* __restrict__ iters, DevStatus* st) {
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
    double pm;
    int it = exact_newton(g, s.rl, s.pl, cl, s.rr, s.pr, cr, du,
                          SEEDED ? seed_from_y(yseed[i], s.pl, s.pr) : ts, tol, maxit, pm);
    int total = it < 0 ? -it : it;
    if (SEEDED && it < 0) {
      es.fallbacks += 1;
      it = exact_newton(g, s.rl, s.pl, cl, s.rr, s.pr, cr, du, ts, tol, maxit, pm);
      total += it < 0 ? -it : it;
    } else if (!SEEDED && it < 0) {
      es.fallbacks += 1;
    }
    es.iters += (unsigned long long)total;
    pstar[i] = pm;
    iters[i] = it < 0 ? -total : total;
  }
  flush_exact(st, es);
}
