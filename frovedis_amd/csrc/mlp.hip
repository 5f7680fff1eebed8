// FP32 MLP 1 -> 64 -> 64 -> 1 (ReLU) for the learned CDF model, with
// hand-written forward, backward and Adam kernels for gfx950.  No
// counterpart in the reference: it drives the learned partition sort.
//
// Parameter layout (MLP_NPARAM floats): W1[64] b1[64] W2[64][64] b2[64]
// W3[64] b3.  Input: x in [0,1) (a normalised log-key); target: the
// normalised rank of the key.  Loss: mean squared error over the batch.
//
// Training kernel: a workgroup walks its share of the batch in tiles of 64
// samples; the 64x64 products are spread over the 256 threads through LDS
// and every thread keeps 16 of the 4096 W2 gradients in registers; the
// workgroup's gradients are added to the global gradient with FP32 atomics.
#include "fhip_common.hpp"

namespace fhip {

constexpr int MLP_H = 64;
constexpr int MLP_NPARAM = MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;
constexpr int OFF_W1 = 0, OFF_B1 = 64, OFF_W2 = 128, OFF_B2 = 128 + 4096, OFF_W3 = OFF_B2 + 64,
              OFF_B3 = OFF_W3 + 64;
static_assert(OFF_B3 + 1 == MLP_NPARAM, "layout");
extern const int kMlpParams = MLP_NPARAM;

namespace {

constexpr int MT = 256;  // threads
constexpr int TS = 64;   // samples per tile

struct MlpLds {
  float p[MLP_NPARAM];
  float h1[TS][MLP_H + 1];   // +1 pads LDS banks
  float h2[TS][MLP_H + 1];
  float d2[TS][MLP_H + 1];   // dL/dz2
  float d1[TS][MLP_H + 1];   // dL/dz1
  float x[TS], t[TS], dy[TS];
  float g_small[4 * MLP_H + 1];  // W1 b1 b2 W3 grads (+ b3)
  float loss;
};

__global__ __launch_bounds__(MT) void mlp_grad_kernel(const float* __restrict__ params,
                                                      const float* __restrict__ xs,
                                                      const float* __restrict__ ts, size_t S,
                                                      size_t per_block, float inv_total,
                                                      float* __restrict__ grad,
                                                      float* __restrict__ loss_out) {
  __shared__ MlpLds s;
  const int tid = threadIdx.x;
  for (int i = tid; i < MLP_NPARAM; i += MT) s.p[i] = params[i];
  for (int i = tid; i < 4 * MLP_H + 1; i += MT) s.g_small[i] = 0.f;
  if (tid == 0) s.loss = 0.f;
  float gw2[16];
#pragma unroll
  for (int e = 0; e < 16; e++) gw2[e] = 0.f;
  // thread owns W2 gradient entries (j, i) with j = tid / 4, i = (tid % 4) * 16 + e
  const int gj = tid >> 2, gi0 = (tid & 3) * 16;
  __syncthreads();

  const size_t beg = size_t(blockIdx.x) * per_block;
  const size_t end = beg + per_block < S ? beg + per_block : S;
  for (size_t base = beg; base < end; base += TS) {
    const int cnt = int(end - base < size_t(TS) ? end - base : size_t(TS));
    if (tid < TS) {
      s.x[tid] = tid < cnt ? xs[base + tid] : 0.f;
      s.t[tid] = tid < cnt ? ts[base + tid] : 0.f;
    }
    __syncthreads();
    // layer 1: 4096 activations, 16 per thread
    for (int k = tid; k < TS * MLP_H; k += MT) {
      const int sm = k / MLP_H, j = k % MLP_H;
      float z = s.p[OFF_W1 + j] * s.x[sm] + s.p[OFF_B1 + j];
      s.h1[sm][j] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    // layer 2
    for (int k = tid; k < TS * MLP_H; k += MT) {
      const int sm = k / MLP_H, j = k % MLP_H;
      float z = s.p[OFF_B2 + j];
      const float* w = s.p + OFF_W2 + j * MLP_H;
#pragma unroll 8
      for (int i = 0; i < MLP_H; i++) z += w[i] * s.h1[sm][i];
      s.h2[sm][j] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    // output and dL/dy
    if (tid < TS) {
      float y = s.p[OFF_B3];
#pragma unroll 8
      for (int j = 0; j < MLP_H; j++) y += s.p[OFF_W3 + j] * s.h2[tid][j];
      const float e = tid < cnt ? (y - s.t[tid]) : 0.f;
      s.dy[tid] = 2.f * e * inv_total;
      atomicAdd(&s.loss, e * e * inv_total);
    }
    __syncthreads();
    for (int k = tid; k < TS * MLP_H; k += MT) {
      const int sm = k / MLP_H, j = k % MLP_H;
      s.d2[sm][j] = s.h2[sm][j] > 0.f ? s.dy[sm] * s.p[OFF_W3 + j] : 0.f;
    }
    __syncthreads();
    // W2 gradient: sum over samples of d2[s][j] * h1[s][i]
    for (int sm = 0; sm < TS; sm++) {
      const float dj = s.d2[sm][gj];
#pragma unroll
      for (int e = 0; e < 16; e++) gw2[e] += dj * s.h1[sm][gi0 + e];
    }
    // dL/dz1 = (W2^T d2) * relu'(z1)
    for (int k = tid; k < TS * MLP_H; k += MT) {
      const int sm = k / MLP_H, i = k % MLP_H;
      float a = 0.f;
#pragma unroll 8
      for (int j = 0; j < MLP_H; j++) a += s.p[OFF_W2 + j * MLP_H + i] * s.d2[sm][j];
      s.d1[sm][i] = s.h1[sm][i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // small gradients: W1, b1, b2, W3 (one thread per entry), b3
    if (tid < MLP_H) {
      float gw1 = 0.f, gb1 = 0.f, gb2 = 0.f, gw3 = 0.f;
      for (int sm = 0; sm < TS; sm++) {
        gw1 += s.d1[sm][tid] * s.x[sm];
        gb1 += s.d1[sm][tid];
        gb2 += s.d2[sm][tid];
        gw3 += s.dy[sm] * s.h2[sm][tid];
      }
      s.g_small[tid] += gw1;
      s.g_small[MLP_H + tid] += gb1;
      s.g_small[2 * MLP_H + tid] += gb2;
      s.g_small[3 * MLP_H + tid] += gw3;
    } else if (tid == MLP_H) {
      float gb3 = 0.f;
      for (int sm = 0; sm < TS; sm++) gb3 += s.dy[sm];
      s.g_small[4 * MLP_H] += gb3;
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < 16; e++) unsafeAtomicAdd(&grad[OFF_W2 + gj * MLP_H + gi0 + e], gw2[e]);
  if (tid < MLP_H) {
    unsafeAtomicAdd(&grad[OFF_W1 + tid], s.g_small[tid]);
    unsafeAtomicAdd(&grad[OFF_B1 + tid], s.g_small[MLP_H + tid]);
    unsafeAtomicAdd(&grad[OFF_B2 + tid], s.g_small[2 * MLP_H + tid]);
    unsafeAtomicAdd(&grad[OFF_W3 + tid], s.g_small[3 * MLP_H + tid]);
  } else if (tid == MLP_H) {
    unsafeAtomicAdd(&grad[OFF_B3], s.g_small[4 * MLP_H]);
    if (loss_out) unsafeAtomicAdd(loss_out, s.loss);
  }
}

__global__ void mlp_forward_kernel(const float* __restrict__ params, const float* __restrict__ xs,
                                   size_t S, float* __restrict__ ys) {
  __shared__ float p[MLP_NPARAM];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) p[i] = params[i];
  __syncthreads();
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < S; k += stride) {
    const float x = xs[k];
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
      float z = p[OFF_W1 + j] * x + p[OFF_B1 + j];
      h1[j] = z > 0.f ? z : 0.f;
    }
    float y = p[OFF_B3];
    for (int j = 0; j < MLP_H; j++) {
      float z = p[OFF_B2 + j];
#pragma unroll
      for (int i = 0; i < MLP_H; i++) z += p[OFF_W2 + j * MLP_H + i] * h1[i];
      y += p[OFF_W3 + j] * (z > 0.f ? z : 0.f);
    }
    ys[k] = y;
  }
}

// Adam (Kingma & Ba 2015) with bias correction; zeroes the gradient.
__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    g[i] = 0.f;
  }
}

}  // namespace

void mlp_grad(const float* params, const float* xs, const float* ts, size_t S, float* grad,
              float* loss, hipStream_t st) {
  if (S == 0) return;
  const int blocks = int(std::min<size_t>(ceil_div(S, TS), size_t(fill_blocks(MT, 1))));
  const size_t per_block = ceil_div(ceil_div(S, size_t(blocks)), TS) * TS;
  const int grid = int(ceil_div(S, per_block));
  mlp_grad_kernel<<<grid, MT, 0, st>>>(params, xs, ts, S, per_block, 1.f / float(S), grad, loss);
  FHIP_CHECK(hipGetLastError());
}

void mlp_forward(const float* params, const float* xs, size_t S, float* ys, hipStream_t st) {
  if (S == 0) return;
  const int grid = int(std::min<size_t>(ceil_div(S, 256), size_t(fill_blocks(256, 4))));
  mlp_forward_kernel<<<grid, 256, 0, st>>>(params, xs, S, ys);
  FHIP_CHECK(hipGetLastError());
}

void mlp_adam(float* p, float* g, float* m, float* v, int step, float lr, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, float(step)), c2 = 1.f - std::pow(b2, float(step));
  adam_kernel<<<int(ceil_div(MLP_NPARAM, 256)), 256, 0, st>>>(p, g, m, v, MLP_NPARAM, lr, b1, b2,
                                                               1e-8f, c1, c2);
  FHIP_CHECK(hipGetLastError());
}

}  // namespace fhip

extern "C" {

int fhip_mlp_nparam(void) { return fhip::MLP_NPARAM; }

int fhip_dev_mlp_grad(const float* params, const float* x, const float* t, size_t S, float* grad,
                      float* loss, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(params && grad && (S == 0 || (x && t)), FHIP_ERR_ARG, "mlp_grad: null pointer");
    fhip::mlp_grad(params, x, t, S, grad, loss, fhip::as_stream(stream));
  });
}

int fhip_dev_mlp_forward(const float* params, const float* x, size_t S, float* y, void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(params && (S == 0 || (x && y)), FHIP_ERR_ARG, "mlp_forward: null pointer");
    fhip::mlp_forward(params, x, S, y, fhip::as_stream(stream));
  });
}

int fhip_dev_mlp_adam(float* params, float* grad, float* m, float* v, int step, float lr,
                      void* stream) {
  return fhip::guarded([&] {
    FHIP_REQUIRE(params && grad && m && v && step >= 1, FHIP_ERR_ARG, "mlp_adam: bad argument");
    fhip::mlp_adam(params, grad, m, v, step, lr, fhip::as_stream(stream));
  });
}

}  // extern "C"
