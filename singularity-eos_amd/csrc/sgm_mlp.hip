// Learned initial guess for the mixed-cell closure: a small MLP predicts the
// equilibrium volume fractions and log temperature of a cell from its inputs;
// the FP64 Newton solver starts from that guess when it lowers the initial
// residual and from its own guess otherwise, so results keep the solver's
// 1e-10 accuracy either way.
//
// Network: 64 inputs (per material slot s < 6: one-hot of the material id,
// alpha, ln rho, ln T; per cell: asinh(sie_tot / 1e10), ln rho_total; padded
// from 56) -> 128 -> 128 -> 128 (GELU, tanh form) -> 16 outputs (6 slot
// logits, masked softmax over the cell's materials; standardized ln T; 9
// padding). BF16 weights and activations, FP32 accumulation on the matrix
// cores (gfx950's v_mfma_f32_16x16x32_bf16), FP32 master weights, Adam.
//
// Work split: a block of 4 waves owns 64 rows (cells); each wave computes 16
// rows x all columns of a layer with 16x16 MFMA tiles; weights of the layer
// being applied are staged in LDS (as [out][in] for the forward, [in][out]
// for the input-gradient product), activations stay in LDS between layers.
#include "sgm_internal.hpp"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace sgm {
namespace mlp {

constexpr int KIN = 64, H = 128, KOUT = 16, SLOTS = 6, NFEAT = 9 * SLOTS + 2;
constexpr int ROWS = 64;  // rows per block
// flat parameter layout (FP32 master, gradients, Adam moments)
constexpr int OW1 = 0, OB1 = OW1 + H * KIN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H * H, OW4 = OB3 + H, OB4 = OW4 + KOUT * H, NPARAM = OB4 + KOUT;
// bf16 weight layout (no biases): W1 [128x64], W2, W3 [128x128], W4 [16x128]
constexpr int QW1 = 0, QW2 = QW1 + H * KIN, QW3 = QW2 + H * H, QW4 = QW3 + H * H,
              NQ = QW4 + KOUT * H;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

__host__ __device__ inline bf16 f2bf(float x) {  // round to nearest even
  unsigned u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  const unsigned u = (unsigned)h << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
// GELU, tanh form, and its derivative
__device__ inline float gelu(float z) {
  const float k = 0.7978845608028654f, c = 0.044715f;
  return 0.5f * z * (1.0f + tanhf(k * (z + c * z * z * z)));
}
__device__ inline float gelu_d(float z) {
  const float k = 0.7978845608028654f, c = 0.044715f;
  const float t = tanhf(k * (z + c * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * k * (1.0f + 3.0f * c * z * z);
}

// One wave: acc[j] (+)= A[16 rows x K] * Bt[16 cols of tile j x K]^T, bf16 in
// LDS. A row-major with stride lda; Bt row-major ([N][K]) with stride ldb
// (16-byte aligned rows). gfx950's 16x16x32 step: lane l holds
// A[l%16][8(l/16)..+7] and Bt[l%16][8(l/16)..+7]; a K that leaves 16 over
// (the 16-wide output layer's gradient product) ends with one 16x16x16 step,
// lane l holding 4(l/16)..+3 of those 16. D[4(l/16)+i][l%16] comes back in
// acc[j][i].
template <int NT>
__device__ inline void wave_gemm(const bf16 *A, int lda, const bf16 *Bt, int ldb, int K,
                                 f4 (&acc)[NT]) {
  const int l = threadIdx.x & 63;
  const int r = l & 15;
  int k0 = 0;
  for (; k0 + 32 <= K; k0 += 32) {
    const int q = (l >> 4) * 8;
    const b8 a = *reinterpret_cast<const b8 *>(A + r * lda + k0 + q);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const b8 b = *reinterpret_cast<const b8 *>(Bt + (j * 16 + r) * ldb + k0 + q);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
  }
  if (k0 < K) {
    const int q = (l >> 4) * 4;
    const s4 a = *reinterpret_cast<const s4 *>(A + r * lda + k0 + q);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const s4 b = *reinterpret_cast<const s4 *>(Bt + (j * 16 + r) * ldb + k0 + q);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
    }
  }
}

constexpr int LDK = KIN + 8, LDH = H + 8, LDR = ROWS + 8;  // padded LDS strides

// stage a [rows x cols] bf16 matrix from global (row-major, stride cols) into LDS
__device__ inline void stage(bf16 *dst, int ld, const bf16 *src, int rows, int cols) {
  for (int i = threadIdx.x; i < rows * cols; i += blockDim.x) {
    const int r = i / cols, c = i - r * cols;
    dst[r * ld + c] = src[i];
  }
}

// Applies layers 1..4 to the 64-row tile in X (LDS, [64][LDK]). Training mode
// writes A1..A3 (bf16) and Z1..Z3, Y (fp32) for rows [row0, row0 + nrow).
struct FwdOut {
  bf16 *A[4];   // A[0] = X, A[1..3] = GELU(Z) ([B][KIN] / [B][H])
  float *Z[4];  // Z[1..3] pre-activations [B][H]
  float *Y;     // [B][KOUT]
};

__device__ void forward_tile(const bf16 *__restrict__ wq, const float *__restrict__ P, bf16 *X,
                             bf16 *act0, bf16 *act1, bf16 *wl, float *ylds, long row0,
                             int nrow, const FwdOut *o) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r16 = w * 16;
  const bf16 *in = X;
  int ldin = LDK, kin = KIN;
  bf16 *outs[2] = {act0, act1};
  for (int layer = 1; layer <= 4; ++layer) {
    const int nout = layer == 4 ? KOUT : H;
    const bf16 *wsrc = wq + (layer == 1 ? QW1 : layer == 2 ? QW2 : layer == 3 ? QW3 : QW4);
    const float *bias = P + (layer == 1 ? OB1 : layer == 2 ? OB2 : layer == 3 ? OB3 : OB4);
    __syncthreads();
    stage(wl, kin == KIN ? LDK : LDH, wsrc, nout, kin);
    __syncthreads();
    const int ldw = kin == KIN ? LDK : LDH;
    if (layer < 4) {
      f4 acc[H / 16];
#pragma unroll
      for (int j = 0; j < H / 16; ++j) acc[j] = f4{0, 0, 0, 0};
      wave_gemm<H / 16>(in + r16 * ldin, ldin, wl, ldw, kin, acc);
      bf16 *out = outs[(layer - 1) & 1];
#pragma unroll
      for (int j = 0; j < H / 16; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rr = r16 + (l >> 4) * 4 + i, cc = j * 16 + (l & 15);
          const float z = acc[j][i] + bias[cc];
          const bf16 a = f2bf(gelu(z));
          out[rr * LDH + cc] = a;
          if (o && rr < nrow) {
            o->Z[layer][(row0 + rr) * H + cc] = z;
            o->A[layer][(row0 + rr) * H + cc] = a;
          }
        }
      in = out;
      ldin = LDH;
      kin = H;
    } else {
      f4 acc[1] = {f4{0, 0, 0, 0}};
      wave_gemm<1>(in + r16 * ldin, ldin, wl, ldw, kin, acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rr = r16 + (l >> 4) * 4 + i, cc = l & 15;
        const float y = acc[0][i] + bias[cc];
        ylds[rr * KOUT + cc] = y;
        if (o && rr < nrow) o->Y[(row0 + rr) * KOUT + cc] = y;
      }
    }
  }
  __syncthreads();
}

// training forward over rows idx[0..B) of the feature matrix Xall [N][KIN]
__global__ void __launch_bounds__(256) fwd_kernel(const bf16 *__restrict__ wq,
                                                  const float *__restrict__ P,
                                                  const bf16 *__restrict__ Xall,
                                                  const int *__restrict__ idx, long B, FwdOut o) {
  __shared__ __attribute__((aligned(16))) bf16 X[ROWS * LDK];
  __shared__ __attribute__((aligned(16))) bf16 act0[ROWS * LDH], act1[ROWS * LDH];
  __shared__ __attribute__((aligned(16))) bf16 wl[H * LDH];
  __shared__ float ylds[ROWS * KOUT];
  const long row0 = (long)blockIdx.x * ROWS;
  const int nrow = (int)((B - row0) < ROWS ? (B - row0) : ROWS);
  for (int i = threadIdx.x; i < ROWS * KIN; i += blockDim.x) {
    const int r = i / KIN, c = i - r * KIN;
    bf16 v = 0;
    if (r < nrow) v = Xall[(long)idx[row0 + r] * KIN + c];
    X[r * LDK + c] = v;
    if (r < nrow) o.A[0][(row0 + r) * KIN + c] = v;
  }
  forward_tile(wq, P, X, act0, act1, wl, ylds, row0, nrow, &o);
}

// Loss: squared error of the log volume fractions (log-softmax of the slot
// logits over the cell's materials, targets floored at 1e-8) + wT times the
// squared error of the standardized ln T; mean over the batch.
// targets row: [alpha0..alpha5, lnT_std, nmat (0 = row excluded)]
__global__ void loss_kernel(const float *__restrict__ Y, const float *__restrict__ tgt,
                            const int *__restrict__ idx, long B, float wT, float inv_b,
                            float *__restrict__ dY, float *__restrict__ loss) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const float *t = tgt + (long)idx[r] * 8;
  const float *y = Y + r * KOUT;
  float *d = dY + r * KOUT;
  for (int c = 0; c < KOUT; ++c) d[c] = 0.f;
  const int nm = (int)t[7];
  if (nm < 2 || nm > SLOTS) {
    loss[r] = 0.f;
    return;
  }
  float mx = -1e30f;
  for (int s = 0; s < nm; ++s) mx = fmaxf(mx, y[s]);
  float p[SLOTS], sum = 0.f;
  for (int s = 0; s < nm; ++s) {
    p[s] = expf(y[s] - mx);
    sum += p[s];
  }
  const float lse = mx + logf(sum);
  float L = 0.f, e[SLOTS], esum = 0.f;
  for (int s = 0; s < nm; ++s) {
    p[s] /= sum;
    e[s] = (y[s] - lse) - logf(fmaxf(t[s], 1e-8f));
    L += e[s] * e[s];
    esum += e[s];
  }
  // d/dy_j sum_s e_s^2 = 2 (e_j - p_j sum_s e_s)
  for (int s = 0; s < nm; ++s) d[s] = inv_b * 2.f * (e[s] - p[s] * esum);
  const float et = y[6] - t[6];
  L += wT * et * et;
  d[6] = inv_b * 2.f * wT * et;
  loss[r] = L;
}

// Backward through layer `layer` for one 64-row tile: partial dW, db of the
// tile, and dZ of the layer below (through GELU') when layer > 1.
struct BwdArgs {
  const bf16 *wq_io;   // bf16 weights as [in][out] (the transpose of the forward layout)
  const float *dZ;     // [B][nout] gradient of this layer's pre-activation
  const bf16 *Aprev;   // [B][nin] input of this layer
  const float *Zprev;  // [B][nin] pre-activation of the layer below (layer > 1)
  float *dZprev;       // [B][nin]
  float *part;         // [nblocks][NPARAM]
  long B;
  int layer;
};

__global__ void __launch_bounds__(256) bwd_kernel(BwdArgs a) {
  __shared__ __attribute__((aligned(16))) bf16 dZr[ROWS * LDH];   // [r][n]
  __shared__ __attribute__((aligned(16))) bf16 dZt[H * LDR];      // [n][r]
  __shared__ __attribute__((aligned(16))) bf16 At[H * LDR];       // [k][r]
  __shared__ __attribute__((aligned(16))) bf16 Wio[H * LDH];      // [k][n] : Bt for dA = dZ W
  const int layer = a.layer;
  const int nout = layer == 4 ? KOUT : H, nin = layer == 1 ? KIN : H;
  const long row0 = (long)blockIdx.x * ROWS;
  const int nrow = (int)((a.B - row0) < ROWS ? (a.B - row0) : ROWS);
  for (int i = threadIdx.x; i < ROWS * nout; i += blockDim.x) {
    const int r = i / nout, n = i - r * nout;
    const bf16 v = r < nrow ? f2bf(a.dZ[(row0 + r) * nout + n]) : (bf16)0;
    dZr[r * LDH + n] = v;
    dZt[n * LDR + r] = v;
  }
  for (int i = threadIdx.x; i < ROWS * nin; i += blockDim.x) {
    const int r = i / nin, k = i - r * nin;
    At[k * LDR + r] = r < nrow ? a.Aprev[(row0 + r) * nin + k] : (bf16)0;
  }
  const int woff = layer == 1 ? QW1 : layer == 2 ? QW2 : layer == 3 ? QW3 : QW4;
  if (layer > 1) {
    // Wio[k][n] for n < nout, k < nin; zero-pad n to 16 columns for layer 4
    for (int i = threadIdx.x; i < nin * nout; i += blockDim.x) {
      const int k = i / nout, n = i - k * nout;
      Wio[k * LDH + n] = a.wq_io[woff + i];
    }
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  float *part = a.part + (long)blockIdx.x * NPARAM;
  const int pw = layer == 1 ? OW1 : layer == 2 ? OW2 : layer == 3 ? OW3 : OW4;
  const int pb = layer == 1 ? OB1 : layer == 2 ? OB2 : layer == 3 ? OB3 : OB4;
  // dWt[n][k] = sum_r dZ[r][n] A[r][k]: M = nout (tiles of 16 over waves), N = nin, K = 64
  for (int mt = w; mt < nout / 16; mt += 4) {
    if (nin == KIN) {
      f4 acc[KIN / 16];
#pragma unroll
      for (int j = 0; j < KIN / 16; ++j) acc[j] = f4{0, 0, 0, 0};
      wave_gemm<KIN / 16>(dZt + mt * 16 * LDR, LDR, At, LDR, ROWS, acc);
#pragma unroll
      for (int j = 0; j < KIN / 16; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int n = mt * 16 + (l >> 4) * 4 + i, k = j * 16 + (l & 15);
          part[pw + n * nin + k] = acc[j][i];
        }
    } else {
      f4 acc[H / 16];
#pragma unroll
      for (int j = 0; j < H / 16; ++j) acc[j] = f4{0, 0, 0, 0};
      wave_gemm<H / 16>(dZt + mt * 16 * LDR, LDR, At, LDR, ROWS, acc);
#pragma unroll
      for (int j = 0; j < H / 16; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int n = mt * 16 + (l >> 4) * 4 + i, k = j * 16 + (l & 15);
          part[pw + n * nin + k] = acc[j][i];
        }
    }
  }
  // db[n] = sum_r dZ[r][n] (FP32 from global)
  for (int n = threadIdx.x; n < nout; n += blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < nrow; ++r) s += a.dZ[(row0 + r) * nout + n];
    part[pb + n] = s;
  }
  if (layer == 1) return;
  // dA[r][k] = sum_n dZ[r][n] W[n][k]: each wave 16 rows, N = nin = 128, K = nout
  f4 acc[H / 16];
#pragma unroll
  for (int j = 0; j < H / 16; ++j) acc[j] = f4{0, 0, 0, 0};
  wave_gemm<H / 16>(dZr + w * 16 * LDH, LDH, Wio, LDH, nout, acc);
#pragma unroll
  for (int j = 0; j < H / 16; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = w * 16 + (l >> 4) * 4 + i, k = j * 16 + (l & 15);
      if (r < nrow) {
        const long g = (row0 + r) * H + k;
        a.dZprev[g] = acc[j][i] * gelu_d(a.Zprev[g]);
      }
    }
}

// grad[p] = sum over blocks of the partials
__global__ void reduce_kernel(const float *__restrict__ part, int nblocks,
                              float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += part[(long)b * NPARAM + p];
  grad[p] = s;
}

// bf16 copies of the weights in both layouts
__device__ inline void write_bf(bf16 *wq, bf16 *wq_io, int p, float v) {
  int q, nout, nin, local;
  if (p < OB1) { q = QW1; nout = H; nin = KIN; local = p - OW1; }
  else if (p >= OW2 && p < OB2) { q = QW2; nout = H; nin = H; local = p - OW2; }
  else if (p >= OW3 && p < OB3) { q = QW3; nout = H; nin = H; local = p - OW3; }
  else if (p >= OW4 && p < OB4) { q = QW4; nout = KOUT; nin = H; local = p - OW4; }
  else return;
  const int n = local / nin, k = local - n * nin;
  const bf16 h = f2bf(v);
  wq[q + local] = h;
  wq_io[q + k * nout + n] = h;
}

__global__ void adam_kernel(float *P, float *m, float *v, const float *g, bf16 *wq, bf16 *wq_io,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.f - b1) * gp;
  const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  const float x = P[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
  P[p] = x;
  write_bf(wq, wq_io, p, x);
}

__global__ void sync_bf_kernel(const float *P, bf16 *wq, bf16 *wq_io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NPARAM) write_bf(wq, wq_io, p, P[p]);
}

// Features of cell c (CSR inputs) -> raw[64]; returns nmat if the cell can be
// described (2..6 materials, registry ids < 6), else 0.
__device__ inline int cell_features(const PteArgs &a, long c, float *raw) {
  for (int i = 0; i < KIN; ++i) raw[i] = 0.f;
  const long o0 = a.offsets[c], o1 = a.offsets[c + 1];
  if (!(o0 >= 0 && o1 >= o0 && o1 <= a.nnz)) return 0;
  const int nm = (int)(o1 - o0);
  if (nm < 2 || nm > SLOTS) return 0;
  double rt = 0;
  for (int s = 0; s < nm; ++s) {
    const int m = a.mat_ids[o0 + s];
    if (m >= SLOTS) return 0;
    const double r = a.rho[o0 + s], al = a.vfrac[o0 + s], T = a.temp[o0 + s];
    if (!(r > 0) || !(al > 0) || !(T > 0)) return 0;
    raw[9 * s + m] = 1.f;
    raw[9 * s + 6] = (float)al;
    raw[9 * s + 7] = (float)log(r);
    raw[9 * s + 8] = (float)log(T);
    rt += r * al;
  }
  raw[54] = (float)asinh(a.sie_tot[c] * 1e-10);
  raw[55] = (float)log(rt);
  return nm;
}

// normalized feature row: active slots and the cell features standardized,
// inactive slots exactly zero
__device__ inline void normalize(const float *raw, int nm, const float *mu, const float *sg,
                                 bf16 *out) {
  for (int i = 0; i < KIN; ++i) {
    const bool on = i < 9 * nm || i == 54 || i == 55;
    out[i] = on && i < NFEAT ? f2bf((raw[i] - mu[i]) / sg[i]) : (bf16)0;
  }
}

__global__ void features_kernel(PteArgs a, const float *mu, const float *sg, float *raw_out,
                                bf16 *X, int *nmat) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  float raw[KIN];
  const int nm = cell_features(a, c, raw);
  nmat[c] = nm;
  if (raw_out)
    for (int i = 0; i < KIN; ++i) raw_out[c * KIN + i] = raw[i];
  if (X) {
    bf16 row[KIN];
    normalize(raw, nm, mu, sg, row);
    for (int i = 0; i < KIN; ++i) X[c * KIN + i] = row[i];
  }
}

// training targets from a solved closure: converged alphas per slot (over
// vfrac_tot) and the standardized ln T; nmat 0 for cells left out
__global__ void targets_kernel(PteArgs a, const int *nmat_feat, const int *conv, float t_mu,
                               float t_sg, float *tgt) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  float *t = tgt + c * 8;
  for (int i = 0; i < 8; ++i) t[i] = 0.f;
  const int nm = nmat_feat[c];
  if (nm == 0 || !conv[c]) return;
  const long o0 = a.offsets[c];
  const double vt = a.vfrac_tot[c];
  for (int s = 0; s < nm; ++s) t[s] = (float)(a.vfrac[o0 + s] / vt);
  t[6] = ((float)log(a.temp[o0]) - t_mu) / t_sg;
  t[7] = (float)nm;
}

// inference: features, the four layers, softmax -> guesses for the closure
__global__ void __launch_bounds__(256) guess_kernel(PteArgs a, const bf16 *__restrict__ wq,
                                                    const float *__restrict__ P,
                                                    const float *mu, const float *sg, float t_mu,
                                                    float t_sg, double *gv, double *gT) {
  __shared__ __attribute__((aligned(16))) bf16 X[ROWS * LDK];
  __shared__ __attribute__((aligned(16))) bf16 act0[ROWS * LDH], act1[ROWS * LDH];
  __shared__ __attribute__((aligned(16))) bf16 wl[H * LDH];
  __shared__ float ylds[ROWS * KOUT];
  __shared__ int snm[ROWS];
  const long c0 = (long)blockIdx.x * ROWS;
  if (threadIdx.x < ROWS) {
    const long c = c0 + threadIdx.x;
    float raw[KIN];
    bf16 row[KIN];
    int nm = 0;
    if (c < a.ncell) nm = cell_features(a, c, raw);
    else
      for (int i = 0; i < KIN; ++i) raw[i] = 0.f;
    normalize(raw, nm, mu, sg, row);
    for (int i = 0; i < KIN; ++i) X[threadIdx.x * LDK + i] = row[i];
    snm[threadIdx.x] = nm;
  }
  forward_tile(wq, P, X, act0, act1, wl, ylds, 0, 0, nullptr);
  if (threadIdx.x < ROWS) {
    const long c = c0 + threadIdx.x;
    if (c >= a.ncell) return;
    const int nm = snm[threadIdx.x];
    const float *y = ylds + threadIdx.x * KOUT;
    if (nm == 0) {
      gT[c] = __builtin_nan("");
      return;
    }
    float mx = -1e30f;
    for (int s = 0; s < nm; ++s) mx = fmaxf(mx, y[s]);
    float p[SLOTS], sum = 0.f;
    for (int s = 0; s < nm; ++s) {
      p[s] = expf(y[s] - mx);
      sum += p[s];
    }
    const long o0 = a.offsets[c];
    const double vt = a.vfrac_tot[c];
    for (int s = 0; s < nm; ++s) gv[o0 + s] = (double)(p[s] / sum) * vt;
    gT[c] = exp((double)(t_mu + t_sg * y[6]));
  }
}

} // namespace mlp
} // namespace sgm

using namespace sgm;
using namespace sgm::mlp;

struct sgm_mlp {
  float *P = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  bf16 *wq = nullptr, *wq_io = nullptr;
  float *mu = nullptr, *sg = nullptr;  // [KIN]
  float t_mu = 0.f, t_sg = 1.f;
  int step = 0;
  // training workspace
  long cap = 0;
  bf16 *A[4] = {nullptr, nullptr, nullptr, nullptr};
  float *Z[4] = {nullptr, nullptr, nullptr, nullptr};
  float *Y = nullptr, *dY = nullptr, *dZa = nullptr, *dZb = nullptr, *loss = nullptr,
        *part = nullptr;
};

namespace {
template <class T>
bool dalloc(T *&p, size_t n) {
  return hipMalloc((void **)&p, sizeof(T) * n) == hipSuccess;
}
void free_ws(sgm_mlp *M) {
  for (int i = 0; i < 4; ++i) {
    if (M->A[i]) (void)hipFree(M->A[i]);
    if (M->Z[i]) (void)hipFree(M->Z[i]);
    M->A[i] = nullptr;
    M->Z[i] = nullptr;
  }
  for (float **p : {&M->Y, &M->dY, &M->dZa, &M->dZb, &M->loss, &M->part}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  M->cap = 0;
}
bool ensure_ws(sgm_mlp *M, long B) {
  if (B <= M->cap) return true;
  free_ws(M);
  const long nb = (B + ROWS - 1) / ROWS;
  bool ok = dalloc(M->A[0], B * KIN);
  for (int i = 1; i < 4; ++i) ok = ok && dalloc(M->A[i], B * H) && dalloc(M->Z[i], B * H);
  ok = ok && dalloc(M->Y, B * KOUT) && dalloc(M->dY, B * KOUT) && dalloc(M->dZa, B * H) &&
       dalloc(M->dZb, B * H) && dalloc(M->loss, B) && dalloc(M->part, nb * (long)NPARAM);
  if (!ok) {
    free_ws(M);
    return false;
  }
  M->cap = B;
  return true;
}
PteArgs csr_args(sgm_ctx *ctx, long ncell, long nnz, const int *off, const uint8_t *mat,
                 const double *rho, const double *vfrac, const double *temp,
                 const double *vfrac_tot, const double *sie_tot) {
  PteArgs a{};
  a.d_eos = ctx ? ctx->d_eos : nullptr;
  a.nreg = ctx ? ctx->nmat : 0;
  a.ncell = ncell;
  a.nnz = nnz;
  a.offsets = off;
  a.mat_ids = mat;
  a.rho = const_cast<double *>(rho);
  a.vfrac = const_cast<double *>(vfrac);
  a.temp = const_cast<double *>(temp);
  a.vfrac_tot = vfrac_tot;
  a.sie_tot = sie_tot;
  return a;
}
int grid1(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// forward + loss + backward over rows idx[0..B) into M->g (no update)
int gradients(sgm_mlp *M, const bf16 *X, const float *tgt, const int *idx, long B, float wT,
              hipStream_t s) {
  if (!ensure_ws(M, B)) return SGM_E_HIP;
  const int nb = (int)((B + ROWS - 1) / ROWS);
  FwdOut o;
  for (int i = 0; i < 4; ++i) {
    o.A[i] = M->A[i];
    o.Z[i] = M->Z[i];
  }
  o.Y = M->Y;
  fwd_kernel<<<nb, 256, 0, s>>>(M->wq, M->P, X, idx, B, o);
  loss_kernel<<<grid1(B), 256, 0, s>>>(M->Y, tgt, idx, B, wT, 1.0f / (float)B, M->dY, M->loss);
  const float *dz = M->dY;
  float *bufs[2] = {M->dZa, M->dZb};
  for (int layer = 4; layer >= 1; --layer) {
    BwdArgs a;
    a.wq_io = M->wq_io;
    a.dZ = dz;
    a.Aprev = M->A[layer - 1];
    a.Zprev = layer > 1 ? M->Z[layer - 1] : nullptr;
    a.dZprev = bufs[layer & 1];
    a.part = M->part;
    a.B = B;
    a.layer = layer;
    bwd_kernel<<<nb, 256, 0, s>>>(a);
    dz = a.dZprev;
  }
  reduce_kernel<<<grid1(NPARAM), 256, 0, s>>>(M->part, nb, M->g);
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}
} // namespace

extern "C" {

int sgm_mlp_create(uint64_t seed, sgm_mlp **out) {
  if (!out) return SGM_E_ARG;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return SGM_E_NOGPU;
  auto *M = new sgm_mlp();
  bool ok = dalloc(M->P, NPARAM) && dalloc(M->m, NPARAM) && dalloc(M->v, NPARAM) &&
            dalloc(M->g, NPARAM) && dalloc(M->wq, NQ) && dalloc(M->wq_io, NQ) &&
            dalloc(M->mu, KIN) && dalloc(M->sg, KIN);
  // He-uniform weights, zero biases
  std::vector<float> h(NPARAM, 0.f), zero(NPARAM, 0.f), ones(KIN, 1.f), zk(KIN, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int nout, int nin) {
    const double b = std::sqrt(6.0 / nin);
    for (int i = 0; i < nout * nin; ++i) {
      const double u = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);
      h[off + i] = (float)((2 * u - 1) * b);
    }
  };
  fill(OW1, H, KIN);
  fill(OW2, H, H);
  fill(OW3, H, H);
  fill(OW4, KOUT, H);
  ok = ok && hipMemcpy(M->P, h.data(), 4 * NPARAM, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(M->m, zero.data(), 4 * NPARAM, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(M->v, zero.data(), 4 * NPARAM, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(M->mu, zk.data(), 4 * KIN, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(M->sg, ones.data(), 4 * KIN, hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    sync_bf_kernel<<<grid1(NPARAM), 256>>>(M->P, M->wq, M->wq_io);
    ok = hipDeviceSynchronize() == hipSuccess;
  }
  if (!ok) {
    sgm_mlp_destroy(M);
    return SGM_E_HIP;
  }
  *out = M;
  return SGM_OK;
}

int sgm_mlp_destroy(sgm_mlp *M) {
  if (!M) return SGM_E_ARG;
  (void)hipDeviceSynchronize();
  free_ws(M);
  for (float *p : {M->P, M->m, M->v, M->g, M->mu, M->sg})
    if (p) (void)hipFree(p);
  if (M->wq) (void)hipFree(M->wq);
  if (M->wq_io) (void)hipFree(M->wq_io);
  delete M;
  return SGM_OK;
}

int sgm_mlp_nparam(void) { return NPARAM; }

int sgm_mlp_params(sgm_mlp *M, float *host_out, const float *host_in) {
  if (!M) return SGM_E_ARG;
  if (host_out && hipMemcpy(host_out, M->P, 4 * NPARAM, hipMemcpyDeviceToHost) != hipSuccess)
    return SGM_E_HIP;
  if (host_in) {
    if (hipMemcpy(M->P, host_in, 4 * NPARAM, hipMemcpyHostToDevice) != hipSuccess)
      return SGM_E_HIP;
    sync_bf_kernel<<<grid1(NPARAM), 256>>>(M->P, M->wq, M->wq_io);
    if (hipDeviceSynchronize() != hipSuccess) return SGM_E_HIP;
  }
  return SGM_OK;
}

int sgm_mlp_set_norm(sgm_mlp *M, const float *mu, const float *sg, float t_mu, float t_sg) {
  if (!M || !mu || !sg || !(t_sg > 0)) return SGM_E_ARG;
  for (int i = 0; i < KIN; ++i)
    if (!(sg[i] > 0)) return SGM_E_ARG;
  if (hipMemcpy(M->mu, mu, 4 * KIN, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M->sg, sg, 4 * KIN, hipMemcpyHostToDevice) != hipSuccess)
    return SGM_E_HIP;
  M->t_mu = t_mu;
  M->t_sg = t_sg;
  return SGM_OK;
}

int sgm_mlp_features(sgm_mlp *M, long ncell, long nnz, const int *offsets,
                     const uint8_t *mat_ids, const double *rho, const double *vfrac,
                     const double *temp, const double *sie_tot, float *raw_out, void *X_out,
                     int *nmat_out, sgm_stream_t stream) {
  if (!M || ncell < 0 || !offsets || !mat_ids || !rho || !vfrac || !temp || !sie_tot ||
      !nmat_out)
    return SGM_E_ARG;
  if (ncell == 0) return SGM_OK;
  PteArgs a = csr_args(nullptr, ncell, nnz, offsets, mat_ids, rho, vfrac, temp, nullptr,
                       sie_tot);
  features_kernel<<<grid1(ncell), 256, 0, (hipStream_t)stream>>>(a, M->mu, M->sg, raw_out,
                                                                  (bf16 *)X_out, nmat_out);
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

int sgm_mlp_targets(sgm_mlp *M, long ncell, long nnz, const int *offsets, const double *vfrac,
                    const double *temp, const double *vfrac_tot, const int *nmat,
                    const int *converged, float *targets, sgm_stream_t stream) {
  if (!M || ncell < 0 || !offsets || !vfrac || !temp || !vfrac_tot || !nmat || !converged ||
      !targets)
    return SGM_E_ARG;
  if (ncell == 0) return SGM_OK;
  PteArgs a = csr_args(nullptr, ncell, nnz, offsets, nullptr, nullptr, vfrac, temp, vfrac_tot,
                       nullptr);
  targets_kernel<<<grid1(ncell), 256, 0, (hipStream_t)stream>>>(a, nmat, converged, M->t_mu,
                                                                 M->t_sg, targets);
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

int sgm_mlp_gradients(sgm_mlp *M, const void *X, const float *targets, const int *idx,
                      long batch, float wT, float *grad_out, float *loss_out,
                      sgm_stream_t stream) {
  if (!M || !X || !targets || !idx || batch < 1 || batch > (1L << 24)) return SGM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  int rc = gradients(M, (const bf16 *)X, targets, idx, batch, wT, s);
  if (rc != SGM_OK) return rc;
  if (grad_out &&
      hipMemcpyAsync(grad_out, M->g, 4 * NPARAM, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SGM_E_HIP;
  if (loss_out &&
      hipMemcpyAsync(loss_out, M->loss, 4 * batch, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SGM_E_HIP;
  return SGM_OK;
}

int sgm_mlp_train_step(sgm_mlp *M, const void *X, const float *targets, const int *idx,
                       long batch, float wT, float lr, float *loss_out, sgm_stream_t stream) {
  if (!M || !X || !targets || !idx || batch < 1 || batch > (1L << 24) || !(lr >= 0))
    return SGM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  int rc = gradients(M, (const bf16 *)X, targets, idx, batch, wT, s);
  if (rc != SGM_OK) return rc;
  ++M->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
  adam_kernel<<<grid1(NPARAM), 256, 0, s>>>(M->P, M->m, M->v, M->g, M->wq, M->wq_io, lr, b1, b2,
                                            1e-8f, c1, c2);
  if (loss_out &&
      hipMemcpyAsync(loss_out, M->loss, 4 * batch, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SGM_E_HIP;
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

int sgm_mlp_forward(sgm_mlp *M, const void *X, const int *idx, long batch, float *Y_out,
                    sgm_stream_t stream) {
  if (!M || !X || !idx || batch < 1 || batch > (1L << 24) || !Y_out) return SGM_E_ARG;
  if (!ensure_ws(M, batch)) return SGM_E_HIP;
  hipStream_t s = (hipStream_t)stream;
  FwdOut o;
  for (int i = 0; i < 4; ++i) {
    o.A[i] = M->A[i];
    o.Z[i] = M->Z[i];
  }
  o.Y = M->Y;
  fwd_kernel<<<(int)((batch + ROWS - 1) / ROWS), 256, 0, s>>>(M->wq, M->P, (const bf16 *)X, idx,
                                                              batch, o);
  if (hipMemcpyAsync(Y_out, M->Y, 4 * KOUT * batch, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SGM_E_HIP;
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

int sgm_mlp_guess(sgm_mlp *M, sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                  const uint8_t *mat_ids, const double *rho, const double *vfrac,
                  const double *temp, const double *vfrac_tot, const double *sie_tot,
                  double *guess_vfrac, double *guess_T, sgm_stream_t stream) {
  if (!M || !ctx || ncell < 0 || !offsets || !mat_ids || !rho || !vfrac || !temp ||
      !vfrac_tot || !sie_tot || !guess_vfrac || !guess_T)
    return SGM_E_ARG;
  if (ctx->host_only) return SGM_E_NOGPU;
  if (ncell == 0) return SGM_OK;
  PteArgs a = csr_args(ctx, ncell, nnz, offsets, mat_ids, rho, vfrac, temp, vfrac_tot, sie_tot);
  guess_kernel<<<(int)((ncell + ROWS - 1) / ROWS), 256, 0, (hipStream_t)stream>>>(
      a, M->wq, M->P, M->mu, M->sg, M->t_mu, M->t_sg, guess_vfrac, guess_T);
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

} // extern "C"
