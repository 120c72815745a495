// Multilayer perceptron on MFMA: BF16 operands, FP32 accumulation.
//
// The network maps a chemistry state and time step to the increment of the
// state over that step; it is trained on the GPU with the kernels below
// (forward, mean-squared-error loss, backward, Adam) and evaluated by the
// surrogate step.  Layers: nin -> hidden (GELU) x nhl -> nout, all linear
// layers y = x W^T + b with W[out][in] row-major.  Widths are padded to
// multiples of 64 (the GEMM tile); padded weights, biases and activations
// stay exactly zero, so the padded network computes the unpadded one.
//
// GEMM: C[M][N] = sum_k A(m,k) B(k,n) on 64x64 output tiles, 4 waves per
// workgroup, each wave 32x32 as 2x2 v_mfma_f32_16x16x32_bf16 blocks, K in
// steps of 32 staged through LDS.  Operands are addressed by element strides
// so the three products of a layer (x W^T, dY W, dY^T x) use the same kernel
// with no transposed copies; the load path is chosen by which stride is 1.
// Entry points: include/tiger_mlp.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tiger_mlp.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(x)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(TGR_MLP_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;   // output tile (M and N)
constexpr int kKStep = 32;  // K per LDS stage
constexpr int kLdsRow = kKStep + 8;  // bf16 elements per LDS row (80 B: 16-byte aligned, staggered banks)

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even; NaN stays a quiet NaN
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU(x) = x Phi(x) with the normal CDF through erf; derivative Phi(x) + x phi(x)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum Act { kNone = 0, kGelu = 1, kGeluGrad = 2 };

struct GemmArgs {
  const uint16_t* A;
  int64_t sam, sak;  // A(m,k) = A[m*sam + k*sak]
  const uint16_t* B;
  int64_t sbk, sbn;  // B(k,n) = B[k*sbk + n*sbn]
  int M, N, K;       // multiples of 64, 64, 32
  const float* bias;  // [N] or null
  int act;
  const float* aux;  // kGeluGrad: pre-activation, aux[m*ldaux + n]
  int64_t ldaux;
  float* pre;  // kGelu: pre-activation out (or null)
  int64_t ldpre;
  float* out_f;  // fp32 out (or null)
  int64_t ldf;
  uint16_t* out_h;  // bf16 out (or null)
  int64_t ldh;
  int n_valid;  // columns n >= n_valid are written as 0
  int ksplit;   // > 1: workgroup z sums K range z of ksplit into out_f + z * M * ldf (no epilogue)
};

// AK: A's k stride is 1 (else its m stride is 1); BK: B's k stride is 1 (else its n stride is 1).
template <bool AK, bool BK>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kTile * kLdsRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile * kLdsRow];
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int kchunk = g.K / g.ksplit;
  const int kbeg = blockIdx.z * kchunk, kend = kbeg + kchunk;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kbeg; k0 < kend; k0 += kKStep) {
    // global -> registers: 8 bf16 of A and 8 of B per thread (16-byte loads)
    uint4 ra, rb;
    int ar, ac, br, bc;
    if (AK) {  // row m = t/4, k chunk (t%4)*8
      ar = t >> 2;
      ac = (t & 3) * 8;
      ra = *reinterpret_cast<const uint4*>(g.A + (int64_t)(m0 + ar) * g.sam + (k0 + ac));
    } else {  // k = t/8, m chunk (t%8)*8
      ar = (t & 7) * 8;
      ac = t >> 3;
      ra = *reinterpret_cast<const uint4*>(g.A + (int64_t)(k0 + ac) * g.sak + (m0 + ar));
    }
    if (BK) {
      br = t >> 2;
      bc = (t & 3) * 8;
      rb = *reinterpret_cast<const uint4*>(g.B + (int64_t)(n0 + br) * g.sbn + (k0 + bc));
    } else {
      br = (t & 7) * 8;
      bc = t >> 3;
      rb = *reinterpret_cast<const uint4*>(g.B + (int64_t)(k0 + bc) * g.sbk + (n0 + br));
    }
    __syncthreads();  // previous stage's reads are done
    if (AK) {
      *reinterpret_cast<uint4*>(&As[ar * kLdsRow + ac]) = ra;
    } else {
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&ra);
      for (int i = 0; i < 8; ++i) As[(ar + i) * kLdsRow + ac] = e[i];
    }
    if (BK) {
      *reinterpret_cast<uint4*>(&Bs[br * kLdsRow + bc]) = rb;
    } else {
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&rb);
      for (int i = 0; i < 8; ++i) Bs[(br + i) * kLdsRow + bc] = e[i];
    }
    __syncthreads();
    // MFMA 16x16x32: lane l holds row (l%16) of A and column (l%16) of B,
    // k = (l/16)*8 .. +8 of this stage (the same k split for A and B)
    const int r = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; ++i)
      a[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&As[(wm + i * 16 + r) * kLdsRow + kq]));
    for (int j = 0; j < 2; ++j)
      b[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[(wn + j * 16 + r) * kLdsRow + kq]));
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }

  // epilogue: lane l holds rows (l/16)*4 + 0..3, column l%16 of each 16x16 block
  if (g.ksplit > 1) {  // partial sums of this K range
    float* part = g.out_f + (int64_t)blockIdx.z * g.M * g.ldf;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn + j * 16 + (lane & 15);
        for (int q = 0; q < 4; ++q) part[(int64_t)(m0 + wm + i * 16 + (lane >> 4) * 4 + q) * g.ldf + col] = acc[i][j][q];
      }
    return;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
      const float bias = g.bias ? g.bias[col] : 0.0f;
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
        float v = acc[i][j][q] + bias;
        if (g.act == kGelu) {
          if (g.pre) g.pre[(int64_t)row * g.ldpre + col] = v;
          v = gelu(v);
        } else if (g.act == kGeluGrad) {
          v *= gelu_grad(g.aux[(int64_t)row * g.ldaux + col]);
        }
        if (col >= g.n_valid) v = 0.0f;
        if (g.out_f) g.out_f[(int64_t)row * g.ldf + col] = v;
        if (g.out_h) g.out_h[(int64_t)row * g.ldh + col] = f2bf(v);
      }
    }
}

// Mean-squared error over the first nout columns of out[B][ld] against
// tgt[B][nout]: dY = 2 (out - tgt) / (B nout) (0 in the padded columns) as
// fp32 and bf16, and per-block partial sums of the squared error.
__global__ __launch_bounds__(256) void mse_grad_kernel(const float* out, int64_t ld, const float* tgt, int nout,
                                                      int64_t B, float* dyf, uint16_t* dyh, float* part) {
  __shared__ float red[256];
  const float scale = 2.0f / ((float)B * (float)nout);
  float s = 0.0f;
  const int64_t total = B * ld;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / ld;
    const int col = (int)(e - row * ld);
    float d = 0.0f;
    if (col < nout) d = out[e] - tgt[row * nout + col];
    s += d * d;
    dyf[e] = scale * d;
    dyh[e] = f2bf(scale * d);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// loss = sum(part[0:n]) / count, one workgroup
__global__ __launch_bounds__(256) void finish_loss_kernel(const float* part, int n, float inv_count, float* loss) {
  __shared__ float red[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] * inv_count;
}

// part[chunk][n] = sum over the rows of chunk `blockIdx.y` of y[row][n]; 64
// columns per workgroup, 4 row groups
__global__ __launch_bounds__(256) void colsum_kernel(const float* y, int64_t ld, int64_t rows, int64_t rows_per_chunk,
                                                    float* part, int n) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), rg = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  const int64_t r1 = r0 + rows_per_chunk < rows ? r0 + rows_per_chunk : rows;
  float s = 0.0f;
  for (int64_t r = r0 + rg; r < r1; r += 4) s += y[r * ld + c];
  red[rg][threadIdx.x & 63] = s;
  __syncthreads();
  if (rg == 0) part[(int64_t)blockIdx.y * n + c] = (red[0][c & 63] + red[1][c & 63]) + (red[2][c & 63] + red[3][c & 63]);
}

// out[i] = sum_z part[z][i], i < n (split-K and chunked column sums)
__global__ __launch_bounds__(256) void sum_parts_kernel(const float* part, int nparts, int64_t n, float* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int z = 0; z < nparts; ++z) s += part[(int64_t)z * n + i];
    out[i] = s;
  }
}

// Adam with bias correction folded into lr_t and eps_t; keeps the bf16 copy in step
__global__ __launch_bounds__(256) void adam_kernel(float* p, const float* gr, float* m, float* v, uint16_t* pb,
                                                  int64_t n, float lr_t, float b1, float b2, float eps_t) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr_t * mi / (sqrtf(vi) + eps_t);
    p[i] = pi;
    pb[i] = f2bf(pi);
  }
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float* p, uint16_t* pb, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    pb[i] = f2bf(p[i]);
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

}  // namespace

struct tgr_mlp {
  int nin = 0, nout = 0, hidden = 0, nhl = 0, nlin = 0;
  std::vector<int> fin, fout;        // padded fan-in / fan-out per linear layer
  std::vector<int64_t> offw, offb;   // offsets in the flat parameter vector
  int64_t nparam = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* pb = nullptr;
  int64_t step = 0;
  // batch workspace
  int64_t cap = 0;
  std::vector<float*> pre;      // [nlin-1] hidden pre-activations, fp32 [cap][hidden]
  std::vector<uint16_t*> act;   // [nlin] act[0] unused (input), act[l] bf16 [cap][fin[l]]
  float* out = nullptr;         // [cap][fout.back()]
  float* dyf[2] = {nullptr, nullptr};
  uint16_t* dyh[2] = {nullptr, nullptr};
  float* part = nullptr;
  float* loss = nullptr;
  float* kpart = nullptr;  // split-K / column-sum partials
  int64_t kpart_n = 0;
  static constexpr int kLossBlocks = 256;
  static constexpr int kMaxSplit = 32;

  void free_ws() {
    for (float* q : pre) (void)hipFree(q);
    for (size_t l = 1; l < act.size(); ++l) (void)hipFree(act[l]);
    pre.clear();
    act.clear();
    for (void* q : {(void*)out, (void*)dyf[0], (void*)dyf[1], (void*)dyh[0], (void*)dyh[1], (void*)part,
                    (void*)loss, (void*)kpart})
      if (q) (void)hipFree(q);
    kpart = nullptr;
    kpart_n = 0;
    out = nullptr;
    dyf[0] = dyf[1] = nullptr;
    dyh[0] = dyh[1] = nullptr;
    part = loss = nullptr;
    cap = 0;
  }

  int ensure(int64_t rows) {
    if (rows <= cap) return 0;
    free_ws();
    const int wmax = std::max(hidden, std::max(fin[0], fout.back()));
    pre.assign(nlin - 1, nullptr);
    act.assign(nlin, nullptr);
    for (int l = 0; l < nlin - 1; ++l) HIP_TRY(hipMalloc(&pre[l], sizeof(float) * rows * hidden));
    for (int l = 1; l < nlin; ++l) HIP_TRY(hipMalloc(&act[l], sizeof(uint16_t) * rows * fin[l]));
    HIP_TRY(hipMalloc(&out, sizeof(float) * rows * fout.back()));
    for (int i = 0; i < 2; ++i) {
      HIP_TRY(hipMalloc(&dyf[i], sizeof(float) * rows * wmax));
      HIP_TRY(hipMalloc(&dyh[i], sizeof(uint16_t) * rows * wmax));
    }
    HIP_TRY(hipMalloc(&part, sizeof(float) * kLossBlocks));
    HIP_TRY(hipMalloc(&loss, sizeof(float)));
    // partials of the weight-gradient products (at most kMaxSplit slices of
    // the largest weight matrix) followed by those of the bias column sums
    // (at most kMaxSplit rows of the widest layer)
    kpart_n = (int64_t)kMaxSplit * ((int64_t)wmax * wmax + wmax);
    HIP_TRY(hipMalloc(&kpart, sizeof(float) * kpart_n));
    cap = rows;
    return 0;
  }
};

namespace {

int launch_gemm(const GemmArgs& a, hipStream_t s) {
  if (a.M % kTile || a.N % kTile || a.K % kKStep || a.M <= 0 || a.N <= 0 || a.K <= 0)
    return fail(TGR_MLP_E_ARG, "gemm: M, N must be positive multiples of 64 and K of 32");
  const bool ak = a.sak == 1, bk = a.sbk == 1;
  if (!ak && a.sam != 1) return fail(TGR_MLP_E_ARG, "gemm: A needs a unit stride in m or k");
  if (!bk && a.sbn != 1) return fail(TGR_MLP_E_ARG, "gemm: B needs a unit stride in k or n");
  // 16-byte loads: the other stride and the base must keep 8-element alignment
  if ((ak ? a.sam : a.sak) % 8 || (bk ? a.sbn : a.sbk) % 8 || ((uintptr_t)a.A & 15) || ((uintptr_t)a.B & 15))
    return fail(TGR_MLP_E_ARG, "gemm: operands must be 16-byte aligned with strides that are multiples of 8");
  const int ks = a.ksplit > 1 ? a.ksplit : 1;
  if (a.K % (ks * kKStep)) return fail(TGR_MLP_E_ARG, "gemm: K must split into multiples of 32");
  dim3 grid(a.N / kTile, a.M / kTile, ks), block(256);
  if (ak && bk) gemm_bf16_kernel<true, true><<<grid, block, 0, s>>>(a);
  else if (ak) gemm_bf16_kernel<true, false><<<grid, block, 0, s>>>(a);
  else if (bk) gemm_bf16_kernel<false, true><<<grid, block, 0, s>>>(a);
  else gemm_bf16_kernel<false, false><<<grid, block, 0, s>>>(a);
  HIP_TRY(hipGetLastError());
  return 0;
}

GemmArgs plain_args() {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.act = kNone;
  a.ksplit = 1;
  return a;
}

// forward over B rows of x (bf16 [B][fin[0]]); keep_pre: store the hidden
// pre-activations that the backward needs
int forward(tgr_mlp* h, const uint16_t* x, int64_t B, float* out, int64_t ldo, bool keep_pre, hipStream_t s) {
  const uint16_t* in = x;
  for (int l = 0; l < h->nlin; ++l) {
    GemmArgs a = plain_args();
    a.A = in;
    a.sam = h->fin[l];
    a.sak = 1;
    a.B = h->pb + h->offw[l];  // W[out][in]: B(k=in, n=out) = W[n*fin + k]
    a.sbk = 1;
    a.sbn = h->fin[l];
    a.M = (int)B;
    a.N = h->fout[l];
    a.K = h->fin[l];
    a.bias = h->p + h->offb[l];
    a.n_valid = h->fout[l];
    if (l < h->nlin - 1) {
      a.act = kGelu;
      a.pre = keep_pre ? h->pre[l] : nullptr;
      a.ldpre = h->hidden;
      a.out_h = h->act[l + 1];
      a.ldh = h->fin[l + 1];
    } else {
      a.out_f = out;
      a.ldf = ldo;
      a.n_valid = h->nout;
    }
    int rc = launch_gemm(a, s);
    if (rc) return rc;
    in = h->act[l + 1];
  }
  return 0;
}

// loss and gradient into h->g for targets tgt[B][nout]
int loss_and_grad(tgr_mlp* h, const uint16_t* x, const float* tgt, int64_t B, hipStream_t s) {
  int rc = h->ensure(B);
  if (rc) return rc;
  const int ldo = h->fout.back();
  if ((rc = forward(h, x, B, h->out, ldo, true, s))) return rc;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(tgr_mlp::kLossBlocks), dim3(256), 0, s, h->out, (int64_t)ldo, tgt,
                     h->nout, B, h->dyf[0], h->dyh[0], h->part);
  hipLaunchKernelGGL(finish_loss_kernel, dim3(1), dim3(256), 0, s, h->part, tgr_mlp::kLossBlocks,
                     1.0f / ((float)B * (float)h->nout), h->loss);
  HIP_TRY(hipGetLastError());
  int cur = 0;
  for (int l = h->nlin - 1; l >= 0; --l) {
    const int no = h->fout[l], ni = h->fin[l];
    const uint16_t* xin = l == 0 ? x : h->act[l];
    // dW[o][i] = sum_b dY[b][o] x[b][i]
    GemmArgs a = plain_args();
    a.A = h->dyh[cur];
    a.sam = 1;
    a.sak = no;
    a.B = xin;
    a.sbk = ni;
    a.sbn = 1;
    a.M = no;
    a.N = ni;
    a.K = (int)B;
    a.ldf = ni;
    a.n_valid = ni;
    // split K (the batch) until the grid has >= 512 workgroups: the weight
    // matrices alone give only (no / 64) x (ni / 64)
    int ks = 1;
    while (ks < tgr_mlp::kMaxSplit && (int64_t)(no / 64) * (ni / 64) * ks < 512 && B % (2 * ks * kKStep) == 0 &&
           B / (2 * ks) >= 128)
      ks *= 2;
    a.ksplit = ks;
    a.out_f = ks > 1 ? h->kpart : h->g + h->offw[l];
    if ((rc = launch_gemm(a, s))) return rc;
    if (ks > 1)
      hipLaunchKernelGGL(sum_parts_kernel, dim3(256), dim3(256), 0, s, h->kpart, ks, (int64_t)no * ni,
                         h->g + h->offw[l]);
    // bias gradient: column sums of dY in row chunks, then the chunk sums
    const int64_t nch = std::min<int64_t>(tgr_mlp::kMaxSplit, (B + 255) / 256);
    const int64_t rpc = (B + nch - 1) / nch;
    hipLaunchKernelGGL(colsum_kernel, dim3(no / 64, (unsigned)nch), dim3(256), 0, s, h->dyf[cur], (int64_t)no, B,
                       rpc, h->kpart + (int64_t)tgr_mlp::kMaxSplit * no * ni, no);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, s, h->kpart + (int64_t)tgr_mlp::kMaxSplit * no * ni,
                       (int)nch, (int64_t)no, h->g + h->offb[l]);
    HIP_TRY(hipGetLastError());
    if (l == 0) break;
    // dY_prev = (dY W) * gelu'(pre_prev)
    GemmArgs d = plain_args();
    d.A = h->dyh[cur];
    d.sam = no;
    d.sak = 1;
    d.B = h->pb + h->offw[l];  // B(k=o, n=i) = W[k*ni + n]
    d.sbk = ni;
    d.sbn = 1;
    d.M = (int)B;
    d.N = ni;
    d.K = no;
    d.act = kGeluGrad;
    d.aux = h->pre[l - 1];
    d.ldaux = h->hidden;
    d.out_f = h->dyf[cur ^ 1];
    d.ldf = ni;
    d.out_h = h->dyh[cur ^ 1];
    d.ldh = ni;
    d.n_valid = ni;
    if ((rc = launch_gemm(d, s))) return rc;
    cur ^= 1;
  }
  return 0;
}

constexpr int64_t kInferChunk = 262144;  // rows per inference pass (a multiple of 64)

int check_batch(const tgr_mlp* h, int64_t B) {
  if (!h) return fail(TGR_MLP_E_ARG, "null network");
  if (B <= 0 || B % 64) return fail(TGR_MLP_E_ARG, "batch rows must be a positive multiple of 64");
  if (B > (int64_t)1 << 26) return fail(TGR_MLP_E_ARG, "batch too large");
  return 0;
}

}  // namespace

extern "C" {

int tgr_mlp_create(int nin, int nout, int hidden, int n_hidden_layers, tgr_mlp** out) {
  if (!out) return fail(TGR_MLP_E_ARG, "null output pointer");
  *out = nullptr;
  if (nin < 1 || nout < 1 || hidden < 64 || hidden % 64 || n_hidden_layers < 1 || n_hidden_layers > 16 ||
      nin > 4096 || nout > 4096 || hidden > 4096)
    return fail(TGR_MLP_E_ARG, "bad network shape (hidden must be a multiple of 64)");
  tgr_mlp* h = new tgr_mlp();
  h->nin = nin;
  h->nout = nout;
  h->hidden = hidden;
  h->nhl = n_hidden_layers;
  h->nlin = n_hidden_layers + 1;
  for (int l = 0; l < h->nlin; ++l) {
    h->fin.push_back(l == 0 ? round_up(nin, 64) : hidden);
    h->fout.push_back(l == h->nlin - 1 ? round_up(nout, 64) : hidden);
    h->offw.push_back(h->nparam);
    h->nparam += (int64_t)h->fin[l] * h->fout[l];
    h->offb.push_back(h->nparam);
    h->nparam += h->fout[l];
  }
  const size_t n = h->nparam;
  if (hipMalloc(&h->p, n * 4) != hipSuccess || hipMalloc(&h->g, n * 4) != hipSuccess ||
      hipMalloc(&h->m, n * 4) != hipSuccess || hipMalloc(&h->v, n * 4) != hipSuccess ||
      hipMalloc(&h->pb, n * 2) != hipSuccess) {
    tgr_mlp_destroy(h);
    return fail(TGR_MLP_E_NOMEM, "hipMalloc failed");
  }
  if (hipMemset(h->p, 0, n * 4) != hipSuccess || hipMemset(h->g, 0, n * 4) != hipSuccess ||
      hipMemset(h->m, 0, n * 4) != hipSuccess || hipMemset(h->v, 0, n * 4) != hipSuccess ||
      hipMemset(h->pb, 0, n * 2) != hipSuccess) {
    tgr_mlp_destroy(h);
    return fail(TGR_MLP_E_HIP, "hipMemset failed");
  }
  *out = h;
  return 0;
}

int tgr_mlp_destroy(tgr_mlp* h) {
  if (!h) return 0;
  h->free_ws();
  for (void* q : {(void*)h->p, (void*)h->g, (void*)h->m, (void*)h->v, (void*)h->pb})
    if (q) (void)hipFree(q);
  delete h;
  return 0;
}

int tgr_mlp_layout(const tgr_mlp* h, int64_t* nparam, int* in_pad, int* out_pad) {
  if (!h) return fail(TGR_MLP_E_ARG, "null network");
  if (nparam) *nparam = h->nparam;
  if (in_pad) *in_pad = h->fin[0];
  if (out_pad) *out_pad = h->fout.back();
  return 0;
}

int tgr_mlp_shape(const tgr_mlp* h, int* nin, int* nout, int* hidden, int* n_hidden_layers) {
  if (!h) return fail(TGR_MLP_E_ARG, "null network");
  if (nin) *nin = h->nin;
  if (nout) *nout = h->nout;
  if (hidden) *hidden = h->hidden;
  if (n_hidden_layers) *n_hidden_layers = h->nhl;
  return 0;
}

int tgr_mlp_set_params(tgr_mlp* h, const float* host) {
  if (!h || !host) return fail(TGR_MLP_E_ARG, "null argument");
  // padded entries must be zero for the padded network to equal the unpadded one
  for (int l = 0; l < h->nlin; ++l) {
    const int ni = l == 0 ? h->nin : h->fin[l], no = l == h->nlin - 1 ? h->nout : h->fout[l];
    for (int o = 0; o < h->fout[l]; ++o) {
      for (int i = 0; i < h->fin[l]; ++i)
        if ((o >= no || i >= ni) && host[h->offw[l] + (int64_t)o * h->fin[l] + i] != 0.0f)
          return fail(TGR_MLP_E_ARG, "nonzero weight in the padding of layer " + std::to_string(l));
      if (o >= no && host[h->offb[l] + o] != 0.0f)
        return fail(TGR_MLP_E_ARG, "nonzero bias in the padding of layer " + std::to_string(l));
    }
  }
  HIP_TRY(hipMemcpy(h->p, host, h->nparam * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(256), dim3(256), 0, 0, h->p, h->pb, h->nparam);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemset(h->m, 0, h->nparam * 4));
  HIP_TRY(hipMemset(h->v, 0, h->nparam * 4));
  HIP_TRY(hipDeviceSynchronize());
  h->step = 0;
  return 0;
}

int tgr_mlp_get_params(const tgr_mlp* h, float* host) {
  if (!h || !host) return fail(TGR_MLP_E_ARG, "null argument");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(host, h->p, h->nparam * 4, hipMemcpyDeviceToHost));
  return 0;
}

int tgr_mlp_forward(tgr_mlp* h, const uint16_t* x, int64_t B, float* out, void* stream) {
  int rc = check_batch(h, B);
  if (rc) return rc;
  if (!x || !out) return fail(TGR_MLP_E_ARG, "null argument");
  // inference in row chunks: the workspace stays at most kInferChunk rows
  const int64_t chunk = std::min<int64_t>(B, kInferChunk);
  if ((rc = h->ensure(chunk))) return rc;
  const int ldo = h->fout.back();
  for (int64_t r0 = 0; r0 < B; r0 += chunk) {
    const int64_t rows = std::min<int64_t>(chunk, B - r0);
    if ((rc = forward(h, x + r0 * h->fin[0], rows, out + r0 * ldo, ldo, false, (hipStream_t)stream))) return rc;
  }
  return 0;
}

int tgr_mlp_gradients(tgr_mlp* h, const uint16_t* x, const float* tgt, int64_t B, float* grad_host,
                      float* loss_host, void* stream) {
  int rc = check_batch(h, B);
  if (rc) return rc;
  if (!x || !tgt) return fail(TGR_MLP_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if ((rc = loss_and_grad(h, x, tgt, B, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  if (grad_host) HIP_TRY(hipMemcpy(grad_host, h->g, h->nparam * 4, hipMemcpyDeviceToHost));
  if (loss_host) HIP_TRY(hipMemcpy(loss_host, h->loss, 4, hipMemcpyDeviceToHost));
  return 0;
}

int tgr_mlp_train_step(tgr_mlp* h, const uint16_t* x, const float* tgt, int64_t B, float lr, float beta1,
                       float beta2, float eps, float* loss_host, void* stream) {
  int rc = check_batch(h, B);
  if (rc) return rc;
  if (!x || !tgt) return fail(TGR_MLP_E_ARG, "null argument");
  if (!(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
    return fail(TGR_MLP_E_ARG, "bad optimiser settings");
  hipStream_t s = (hipStream_t)stream;
  if ((rc = loss_and_grad(h, x, tgt, B, s))) return rc;
  h->step += 1;
  const double c1 = 1.0 - std::pow((double)beta1, (double)h->step);
  const double c2 = 1.0 - std::pow((double)beta2, (double)h->step);
  const float lr_t = (float)(lr * std::sqrt(c2) / c1);
  const float eps_t = (float)(eps * std::sqrt(c2));
  hipLaunchKernelGGL(adam_kernel, dim3(512), dim3(256), 0, s, h->p, h->g, h->m, h->v, h->pb, h->nparam, lr_t, beta1,
                     beta2, eps_t);
  HIP_TRY(hipGetLastError());
  if (loss_host) {
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(loss_host, h->loss, 4, hipMemcpyDeviceToHost));
  }
  return 0;
}

int tgr_mlp_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                      int M, int N, int K, float* C, int64_t ldc, void* stream) {
  if (!A || !B || !C) return fail(TGR_MLP_E_ARG, "null argument");
  GemmArgs a = plain_args();
  a.A = A;
  a.sam = sam;
  a.sak = sak;
  a.B = B;
  a.sbk = sbk;
  a.sbn = sbn;
  a.M = M;
  a.N = N;
  a.K = K;
  a.out_f = C;
  a.ldf = ldc;
  a.n_valid = N;
  return launch_gemm(a, (hipStream_t)stream);
}

const char* tgr_mlp_last_error(void) { return g_err.c_str(); }

}  // extern "C"
