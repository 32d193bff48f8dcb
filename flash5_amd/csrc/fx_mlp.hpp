// fx_mlp.hpp -- the learned Newton seed of the exact Riemann solver: a
// 6 -> 64 -> 64 -> 1 tanh MLP evaluated and trained one wavefront (64
// interfaces) at a time on BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation.  No counterpart in the reference.
//
// Features of an interface (Galilean and scale invariant): with
// P0 = sqrt(pL pR), R0 = sqrt(rhoL rhoR), C0 = sqrt(gamma P0 / R0),
//   x = (ln rhoL/R0, ln pL/P0, (uL - ubar)/C0, ln rhoR/R0, ln pR/P0, (uR - ubar)/C0)
// and the network predicts y = ln(p*/P0); the seed is P0 exp(y) (used where
// the interface is not weak, see fx_learned.hpp).
//
// Layout.  Layers compute H^T = W X^T (units as MFMA rows, interfaces as
// columns).  A 16x16 MFMA result tile sits in lane l as column l%16 and rows
// 4(l/16)+i, i<4, which is exactly the B-operand fragment of a K=16 slice;
// two result tiles form one K=32 operand when the K slots of lane group
// g = l/16 are taken as units {4g..4g+3} of the first tile then of the
// second.  The packed weight fragments use the same slot permutation, so
// layers chain in registers without shuffles.
#pragma once
#include <hip/hip_runtime.h>

namespace fx {

constexpr int MLP_IN = 6, MLP_H = 64;
constexpr int OFF_W1 = 0;                          // W1[64][6]
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;    // b1[64]
constexpr int OFF_W2 = OFF_B1 + MLP_H;             // W2[64 out][64 in]
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;     // b2[64]
constexpr int OFF_W3 = OFF_B2 + MLP_H;             // w3[64]
constexpr int OFF_B3 = OFF_W3 + MLP_H;             // b3
constexpr int MLP_NPARAM = OFF_B3 + 1;             // 4673

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ bf16x8 to_bf16x8(const float v[8]) {
  u16x8 h;
#pragma unroll
  for (int j = 0; j < 8; ++j) h[j] = f2bf(v[j]);
  return __builtin_bit_cast(bf16x8, h);
}

// tanh from the hardware exponential: 1 - 2 / (exp(2x) + 1), saturating to
// +-1 at both ends; a few ulp in FP32, ample for a Newton seed
__device__ __forceinline__ float act_tanh(float x) {
  return 1.f - __fdividef(2.f, __expf(2.f * x) + 1.f);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// unit held in K slot j of lane group g for K step ks (see the header note)
__device__ __forceinline__ int kslot_unit(int ks, int g, int j) {
  return 32 * ks + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// BF16 weight fragments in lane order plus FP32 biases.
struct MlpPacked {
  bf16x8 w1[4 * 64];     // [mt][lane]: rows 16mt + l%16, K slots = features 8g+j (< 6)
  bf16x8 w2[8 * 64];     // [mt][ks][lane]: W2[16mt + l%16][kslot_unit]
  bf16x8 w2t[8 * 64];    // [mt][ks][lane]: W2[kslot_unit][16mt + l%16]  (for W2^T)
  float b1[64], b2[64], w3[64], b3;
};

__global__ void k_mlp_pack(const float* __restrict__ prm, MlpPacked* out) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < 4 * 64) {
    const int mt = tid / 64, l = tid % 64, g = l / 16, c = l % 16;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 8 * g + j;
      v[j] = f < MLP_IN ? prm[OFF_W1 + (16 * mt + c) * MLP_IN + f] : 0.f;
    }
    out->w1[tid] = to_bf16x8(v);
  }
  if (tid < 8 * 64) {
    const int mt = tid / 128, ks = (tid / 64) % 2, l = tid % 64, g = l / 16, c = l % 16;
    float v[8], vt[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int u = kslot_unit(ks, g, j);
      v[j] = prm[OFF_W2 + (16 * mt + c) * MLP_H + u];
      vt[j] = prm[OFF_W2 + u * MLP_H + (16 * mt + c)];
    }
    out->w2[tid] = to_bf16x8(v);
    out->w2t[tid] = to_bf16x8(vt);
  }
  if (tid < 64) {
    out->b1[tid] = prm[OFF_B1 + tid];
    out->b2[tid] = prm[OFF_B2 + tid];
    out->w3[tid] = prm[OFF_W3 + tid];
  }
  if (tid == 0) out->b3 = prm[OFF_B3];
}

__device__ __forceinline__ void mlp_features(double g, double rl, double ul, double pl, double rr,
                                             double ur, double pr, float x[6], double& P0) {
  P0 = sqrt(pl * pr);
  const double R0 = sqrt(rl * rr);
  const double C0 = sqrt(g * P0 / R0);
  const float hr = 0.5f * logf((float)(rl / rr));
  const float hp = 0.5f * logf((float)(pl / pr));
  const float hu = (float)(0.5 * (ur - ul) / C0);
  x[0] = hr; x[1] = hp; x[2] = -hu;
  x[3] = -hr; x[4] = -hp; x[5] = hu;
}

// Activations of one wavefront (64 interfaces): result tile [mt][nt] of
// each layer, 4 floats per lane (column 16nt + l%16, rows 16mt + 4g + i).
struct MlpActs {
  f32x4 h1[4][4];
  f32x4 h2[4][4];
};

// xs: this wavefront's LDS block of 64 x 8 floats (features 0..5 used)
__device__ __forceinline__ void mlp_forward_wave(const MlpPacked& W, const float* xs, int lane,
                                                 MlpActs& A, float& yOut) {
  const int g = lane / 16, c = lane % 16;
  bf16x8 bx[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (g == 0 && j < MLP_IN) ? xs[(16 * nt + c) * 8 + j] : 0.f;
    bx[nt] = to_bf16x8(v);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const bf16x8 a = W.w1[mt * 64 + lane];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 d = mfma(a, bx[nt], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] = act_tanh(d[i] + W.b1[16 * mt + 4 * g + i]);
      A.h1[mt][nt] = d;
    }
  }
  // layer 2: K = 64 as two K=32 steps over result tiles (2ks, 2ks+1)
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    bf16x8 b[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = A.h1[2 * ks][nt][j];
        v[4 + j] = A.h1[2 * ks + 1][nt][j];
      }
      b[ks] = to_bf16x8(v);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 d = mfma(W.w2[(mt * 2 + 0) * 64 + lane], b[0], f32x4{0.f, 0.f, 0.f, 0.f});
      d = mfma(W.w2[(mt * 2 + 1) * 64 + lane], b[1], d);
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] = act_tanh(d[i] + W.b2[16 * mt + 4 * g + i]);
      A.h2[mt][nt] = d;
    }
  }
  // layer 3 (64 -> 1) on the VALU: partial dot products, then across groups
  float part[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float s = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) s += W.w3[16 * mt + 4 * g + i] * A.h2[mt][nt][i];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    part[nt] = s;
  }
  // interface n = lane: its sum sits in part[n/16] of lane n%16
  float y = 0.f;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float v = __shfl(part[nt], c);
    if (g == nt) y = v;
  }
  yOut = y + W.b3;
}

// LDS ordering between lanes of one wavefront (no cross-wave barrier: the
// callers' loops are wavefront-uniform, not workgroup-uniform)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum over the 16 lanes of a lane group (same g)
__device__ __forceinline__ float group_sum(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}

// Gradient of sum_n dy[n] * y[n] with respect to every parameter, added to
// grad (global FP32) with atomics.  dy: this lane's interface (0 for lanes
// without one).  sZ, sH: this wavefront's LDS scratch, 64*64 bf16 each.
__device__ __forceinline__ void mlp_backward_wave(const MlpPacked& W, const float* xs, int lane,
                                                  const MlpActs& A, float dy, unsigned short* sZ,
                                                  unsigned short* sH, float* grad) {
  const int g = lane / 16, c = lane % 16;
  float dyc[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) dyc[nt] = __shfl(dy, 16 * nt + c);
  float b3 = dy;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) b3 += __shfl_xor(b3, o);
  if (lane == 0) atomicAdd(&grad[OFF_B3], b3);
  // output layer and layer-2 pre-activation gradient
  f32x4 dz2[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = 16 * mt + 4 * g + i;
      float s3 = 0.f, s2 = 0.f;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const float h = A.h2[mt][nt][i];
        const float z = W.w3[u] * dyc[nt] * (1.f - h * h);
        dz2[mt][nt][i] = z;
        s3 += dyc[nt] * h;
        s2 += z;
      }
      s3 = group_sum(s3);
      s2 = group_sum(s2);
      if (c == 0) {
        atomicAdd(&grad[OFF_W3 + u], s3);
        atomicAdd(&grad[OFF_B2 + u], s2);
      }
    }
  // dH1 = W2^T dZ2, chained from the dz2 tiles
  f32x4 dh1[4][4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    bf16x8 b[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = dz2[2 * ks][nt][j];
        v[4 + j] = dz2[2 * ks + 1][nt][j];
      }
      b[ks] = to_bf16x8(v);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 d = mfma(W.w2t[(mt * 2 + 0) * 64 + lane], b[0], f32x4{0.f, 0.f, 0.f, 0.f});
      dh1[mt][nt] = mfma(W.w2t[(mt * 2 + 1) * 64 + lane], b[1], d);
    }
  }
  // layer 1: dZ1, db1, dW1 (K = 6 features, on the VALU)
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = 16 * mt + 4 * g + i;
      float sb = 0.f, sw[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const float h = A.h1[mt][nt][i];
        const float z = dh1[mt][nt][i] * (1.f - h * h);
        sb += z;
#pragma unroll
        for (int f = 0; f < MLP_IN; ++f) sw[f] += z * xs[(16 * nt + c) * 8 + f];
      }
      sb = group_sum(sb);
      if (c == 0) atomicAdd(&grad[OFF_B1 + u], sb);
#pragma unroll
      for (int f = 0; f < MLP_IN; ++f) {
        const float v = group_sum(sw[f]);
        if (c == 0) atomicAdd(&grad[OFF_W1 + u * MLP_IN + f], v);
      }
    }
  // dW2[o][in] = sum_n dZ2[o][n] H1[in][n]: transpose both through LDS
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * mt + 4 * g + i, n = 16 * nt + c;
        sZ[u * 64 + n] = f2bf(dz2[mt][nt][i]);
        sH[u * 64 + n] = f2bf(A.h1[mt][nt][i]);
      }
  wave_lds_sync();
  bf16x8 bh[2][4];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      bh[ks][nt] = *reinterpret_cast<const bf16x8*>(&sH[(16 * nt + c) * 64 + 32 * ks + 8 * g]);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    bf16x8 az[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
      az[ks] = *reinterpret_cast<const bf16x8*>(&sZ[(16 * mt + c) * 64 + 32 * ks + 8 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 d = mfma(az[0], bh[0][nt], f32x4{0.f, 0.f, 0.f, 0.f});
      d = mfma(az[1], bh[1][nt], d);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        atomicAdd(&grad[OFF_W2 + (16 * mt + 4 * g + i) * MLP_H + 16 * nt + c], d[i]);
    }
  }
  wave_lds_sync();
}

}  // namespace fx
