// mlp.hip -- training and inference of the learned con2prim pressure guess
// (3 -> 64 -> 64 -> 1, ReLU, bf16 weights with FP32 masters and FP32
// accumulation, Adam) on gfx950.  No reference counterpart.
//
// Training step: one wavefront per 64 samples (mlp_train_kernel) runs the
// forward pass of mlp_device.h, the MSE loss and the backward pass:
//   dW3, db2, db3: per-feature sums over the wave's points (LDS);
//   dW2 = dZ2 . H1 and dH1^T = W2^T . dZ2 on the matrix cores (bf16 operands,
//   FP32 accumulate), the same 32x32x16 layout as the forward pass;
//   dW1, db1: per-unit sums over the wave's points (LDS);
// and writes its partial gradient.  mlp_update_kernel sums the partials in a
// fixed order (deterministic), applies Adam to the FP32 masters and refreshes
// the bf16 working copies.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "bamhip_internal.h"
#include "mlp_device.h"

struct bamhip_mlp {
  float *p = nullptr;       // FP32 masters
  float *mom = nullptr;     // Adam first moment
  float *vel = nullptr;     // Adam second moment
  float *wq = nullptr;      // bf16-rounded W1, W3 as float (256)
  __bf16 *w2 = nullptr;     // [f][k]
  __bf16 *w2t = nullptr;    // [k][f]
  float norm[8] = {0, 0, 0, 1, 1, 1, 0, 1};  // xs[3], xk[3], ys, yk
  long step = 0;
  float *part = nullptr;    // per-wave partial gradients
  long part_blocks = 0;
  float *lossp = nullptr;   // per-wave loss partials
  float *grad = nullptr;    // reduced gradient (for bamhip_mlp_grad)
};

namespace bamhip {

MlpView mlp_view(const bamhip_mlp *m) {
  MlpView v;
  v.p = m->p;
  v.wq = m->wq;
  v.w2 = m->w2;
  v.w2t = m->w2t;
  for (int i = 0; i < 3; i++) {
    v.xs[i] = m->norm[i];
    v.xk[i] = m->norm[3 + i];
  }
  v.ys = m->norm[6];
  v.yk = m->norm[7];
  return v;
}

namespace {

constexpr int kWave = 64;

__global__ void mlp_train_kernel(MlpView M, const float *__restrict__ feat, const float *__restrict__ target, long n,
                                 long q0, long batch, unsigned long long pa, unsigned long long pb,
                                 float *__restrict__ part, float *__restrict__ lossp, float inv_batch) {
  __shared__ __attribute__((aligned(16))) __bf16 h1[kWave * kHStr];    // [pt][k]
  __shared__ __attribute__((aligned(16))) __bf16 h1t[kMlpH * kHStr];   // [k][pt]
  __shared__ __attribute__((aligned(16))) __bf16 dz2[kMlpH * kHStr];   // [f][pt]
  __shared__ __attribute__((aligned(16))) __bf16 dz2t[kWave * kHStr];  // [pt][f]
  __shared__ float hs[kMlpH * (kWave + 1)];                            // h2 [f][pt], then dz1 [k][pt]
  __shared__ float xsh[kWave * 3], dys[kWave];
  const int l = threadIdx.x, lr = l % 32, lh = l / 32;
  const long slot = (long)blockIdx.x * kWave + l;
  // ---- sample: a fixed affine permutation of [0, n) per epoch
  bool valid = slot < batch && q0 + slot < n;
  float x[3] = {0.f, 0.f, 0.f}, t = 0.f;
  if (valid) {
    const long idx = (long)((pa * (unsigned long long)(q0 + slot) + pb) % (unsigned long long)n);
    for (int j = 0; j < 3; j++) x[j] = (feat[3 * idx + j] - M.xs[j]) * M.xk[j];
    t = (target[idx] - M.ys) / M.yk;
    valid = isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]) && isfinite(t);
    if (!valid) x[0] = x[1] = x[2] = t = 0.f;
  }
  for (int j = 0; j < 3; j++) xsh[3 * l + j] = x[j];
  // ---- forward
  mlp_layer1(M, x, h1 + l * kHStr);
  for (int k = 0; k < kMlpH; k++) h1t[k * kHStr + l] = h1[l * kHStr + k];
  __syncthreads();
  f32x16 c[2][2];
  const float y = mlp_layer23(M, h1, c, M.p + kOffB2, M.wq + 192);
  const float r = valid ? y - t : 0.f;
  const float dy = 2.f * r * inv_batch;
  float lsum = r * r;
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if (l == 0) lossp[blockIdx.x] = lsum;
  dys[l] = dy;
  // ---- dZ2 = relu'(z2) * W3 * dy, kept as bf16 in both layouts; h2 kept in FP32
#pragma unroll
  for (int n2 = 0; n2 < 2; n2++) {
    const float dyp = __shfl(dy, 32 * n2 + lr);
    const int pt = 32 * n2 + lr;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int f = 32 * m + mfma_row(v, l);
        const float h = c[m][n2][v];
        const float dz = h > 0.f ? M.wq[192 + f] * dyp : 0.f;
        hs[f * (kWave + 1) + pt] = h;
        dz2[f * kHStr + pt] = (__bf16)dz;
        dz2t[pt * kHStr + f] = (__bf16)dz;
      }
  }
  __syncthreads();
  float *g = part + (long)blockIdx.x * kMlpNParam;
  {
    // lane f: dW3[f], db2[f]; lane 0 also db3
    const int f = l;
    const float w3 = M.wq[192 + f];
    float gw3 = 0.f, gb2 = 0.f;
    for (int pt = 0; pt < kWave; pt++) {
      const float h = hs[f * (kWave + 1) + pt];
      gw3 = fmaf(h, dys[pt], gw3);
      gb2 += h > 0.f ? w3 * dys[pt] : 0.f;
    }
    g[kOffW3 + f] = gw3;
    g[kOffB2 + f] = gb2;
    if (l == 0) {
      float gb3 = 0.f;
      for (int pt = 0; pt < kWave; pt++) gb3 += dys[pt];
      g[kOffB3] = gb3;
    }
  }
  // ---- dW2[f][k] = sum_pt dZ2[f][pt] H1[pt][k]   (A = dZ2, B = H1 with K = points)
  {
    f32x16 acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n2 = 0; n2 < 2; n2++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[m][n2][v] = 0.f;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int m = 0; m < 2; m++) a[m] = *(const bf16x8 *)(dz2 + (32 * m + lr) * kHStr + 16 * s + 8 * lh);
#pragma unroll
      for (int n2 = 0; n2 < 2; n2++) b[n2] = *(const bf16x8 *)(h1t + (32 * n2 + lr) * kHStr + 16 * s + 8 * lh);
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n2 = 0; n2 < 2; n2++)
          acc[m][n2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[m], b[n2], acc[m][n2], 0, 0, 0);
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n2 = 0; n2 < 2; n2++)
#pragma unroll
        for (int v = 0; v < 16; v++) g[kOffW2 + (32 * m + mfma_row(v, l)) * kMlpH + 32 * n2 + lr] = acc[m][n2][v];
  }
  __syncthreads();   // hs is rewritten below
  // ---- dH1^T[k][pt] = sum_f W2[f][k] dZ2[f][pt]   (A = W2^T, B = dZ2 with K = features)
  {
    f32x16 acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n2 = 0; n2 < 2; n2++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[m][n2][v] = 0.f;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int m = 0; m < 2; m++) a[m] = *(const bf16x8 *)(M.w2t + (32 * m + lr) * kMlpH + 16 * s + 8 * lh);
#pragma unroll
      for (int n2 = 0; n2 < 2; n2++) b[n2] = *(const bf16x8 *)(dz2t + (32 * n2 + lr) * kHStr + 16 * s + 8 * lh);
#pragma unroll
      for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n2 = 0; n2 < 2; n2++)
          acc[m][n2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[m], b[n2], acc[m][n2], 0, 0, 0);
    }
    // dZ1 = relu'(z1) * dH1, with the forward pass's mask (h1 > 0)
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n2 = 0; n2 < 2; n2++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const int k = 32 * m + mfma_row(v, l), pt = 32 * n2 + lr;
          const bool on = (float)h1[pt * kHStr + k] > 0.f;
          hs[k * (kWave + 1) + pt] = on ? acc[m][n2][v] : 0.f;
        }
  }
  __syncthreads();
  {
    // lane k: dW1[k][0..2], db1[k]
    const int k = l;
    float gw[3] = {0.f, 0.f, 0.f}, gb = 0.f;
    for (int pt = 0; pt < kWave; pt++) {
      const float d = hs[k * (kWave + 1) + pt];
      gb += d;
      for (int j = 0; j < 3; j++) gw[j] = fmaf(d, xsh[3 * pt + j], gw[j]);
    }
    for (int j = 0; j < 3; j++) g[kOffW1 + 3 * k + j] = gw[j];
    g[kOffB1 + k] = gb;
  }
}

// gradient reduction in block order, Adam on the masters (skipped when
// lr == 0), bf16 working copies
__global__ void mlp_update_kernel(float *__restrict__ p, float *__restrict__ mom, float *__restrict__ vel,
                                  float *__restrict__ wq, __bf16 *__restrict__ w2, __bf16 *__restrict__ w2t,
                                  const float *__restrict__ part, int nblk, float *__restrict__ grad, float lr,
                                  float beta1, float beta2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpNParam) return;
  if (nblk > 0) {
    float gs = 0.f;
    for (int b = 0; b < nblk; b++) gs += part[(long)b * kMlpNParam + i];
    if (grad) grad[i] = gs;
    if (lr > 0.f) {
      const float m1 = beta1 * mom[i] + (1.f - beta1) * gs;
      const float v1 = beta2 * vel[i] + (1.f - beta2) * gs * gs;
      mom[i] = m1;
      vel[i] = v1;
      p[i] -= lr * (m1 / bc1) / (sqrtf(v1 / bc2) + eps);
    }
  }
  const float w = p[i];
  if (i < kOffB1) wq[i] = (float)(__bf16)w;
  else if (i >= kOffW3 && i < kOffB3) wq[192 + (i - kOffW3)] = (float)(__bf16)w;
  else if (i >= kOffW2 && i < kOffB2) {
    const int f = (i - kOffW2) / kMlpH, k = (i - kOffW2) % kMlpH;
    w2[f * kMlpH + k] = (__bf16)w;
    w2t[k * kMlpH + f] = (__bf16)w;
  }
}

// network output (log(p / p_hi)) for n raw feature triples
__global__ void __launch_bounds__(256) mlp_predict_kernel(MlpView M, const float *__restrict__ feat, long n,
                                                          float *__restrict__ y) {
  __shared__ __attribute__((aligned(16))) __bf16 h1[4][kWave * kHStr];
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  for (long base = (long)blockIdx.x * 256; base < n; base += (long)gridDim.x * 256) {
    const long q = base + threadIdx.x;
    float x[3] = {0.f, 0.f, 0.f};
    if (q < n)
      for (int j = 0; j < 3; j++) x[j] = (feat[3 * q + j] - M.xs[j]) * M.xk[j];
    mlp_layer1(M, x, h1[w] + l * kHStr);
    __syncthreads();
    f32x16 c[2][2];
    const float yn = mlp_layer23(M, h1[w], c, M.p + kOffB2, M.wq + 192);
    if (q < n) y[q] = M.ys + M.yk * yn;
    __syncthreads();
  }
}

int ensure_part(bamhip_mlp *m, long nblk) {
  if (nblk <= m->part_blocks) return BAMHIP_OK;
  if (m->part) (void)hipFree(m->part);
  if (m->lossp) (void)hipFree(m->lossp);
  m->part = nullptr;
  m->lossp = nullptr;
  m->part_blocks = 0;
  BH_HIP(hipMalloc(&m->part, sizeof(float) * kMlpNParam * nblk));
  BH_HIP(hipMalloc(&m->lossp, sizeof(float) * nblk));
  m->part_blocks = nblk;
  return BAMHIP_OK;
}

int refresh(bamhip_mlp *m) {
  mlp_update_kernel<<<(kMlpNParam + 255) / 256, 256, 0, stream()>>>(m->p, m->mom, m->vel, m->wq, m->w2, m->w2t,
                                                                    nullptr, 0, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f,
                                                                    1.f);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

// one optimisation step over `batch` samples starting at permuted position q0
int train_step(bamhip_mlp *m, const float *feat, const float *target, long n, long q0, long batch,
               unsigned long long pa, unsigned long long pb, float lr, float *grad_out) {
  const long nblk = (batch + kWave - 1) / kWave;
  if (int e = ensure_part(m, nblk)) return e;
  mlp_train_kernel<<<(unsigned)nblk, kWave, 0, stream()>>>(mlp_view(m), feat, target, n, q0, batch, pa, pb, m->part,
                                                           m->lossp, 1.f / (float)batch);
  BH_HIP(hipGetLastError());
  const float b1 = 0.9f, b2 = 0.999f;
  if (lr > 0.f) m->step++;
  const float bc1 = 1.f - powf(b1, (float)std::max<long>(m->step, 1));
  const float bc2 = 1.f - powf(b2, (float)std::max<long>(m->step, 1));
  mlp_update_kernel<<<(kMlpNParam + 255) / 256, 256, 0, stream()>>>(m->p, m->mom, m->vel, m->wq, m->w2, m->w2t,
                                                                    m->part, (int)nblk, grad_out, lr, b1, b2,
                                                                    1e-8f, bc1, bc2);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

unsigned long long gcd64(unsigned long long a, unsigned long long b) {
  while (b) {
    const unsigned long long t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace
}  // namespace bamhip

using namespace bamhip;

extern "C" {

int bamhip_mlp_create(bamhip_mlp **out, unsigned long long seed) {
  if (!out) return set_error(BAMHIP_EINVAL, "bamhip_mlp_create: NULL");
  bamhip_mlp *m = new bamhip_mlp();
  const size_t np = sizeof(float) * kMlpNParam;
  if (hipMalloc(&m->p, np) != hipSuccess || hipMalloc(&m->mom, np) != hipSuccess ||
      hipMalloc(&m->vel, np) != hipSuccess || hipMalloc(&m->wq, sizeof(float) * 256) != hipSuccess ||
      hipMalloc(&m->w2, sizeof(__bf16) * kMlpH * kMlpH) != hipSuccess ||
      hipMalloc(&m->w2t, sizeof(__bf16) * kMlpH * kMlpH) != hipSuccess ||
      hipMalloc(&m->grad, np) != hipSuccess) {
    bamhip_mlp_free(m);
    return set_error(BAMHIP_ENOMEM, "bamhip_mlp_create: out of device memory");
  }
  // He-normal initialisation for the ReLU layers, small output layer, zero biases
  std::vector<float> h(kMlpNParam, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < 192; i++) h[kOffW1 + i] = (float)(nd(rng) * std::sqrt(2.0 / 3.0));
  for (int i = 0; i < kMlpH * kMlpH; i++) h[kOffW2 + i] = (float)(nd(rng) * std::sqrt(2.0 / kMlpH));
  for (int i = 0; i < kMlpH; i++) h[kOffW3 + i] = (float)(nd(rng) * std::sqrt(1.0 / kMlpH));
  hipStream_t st = stream();
  if (hipMemcpyAsync(m->p, h.data(), np, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(m->mom, 0, np, st) != hipSuccess || hipMemsetAsync(m->vel, 0, np, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    bamhip_mlp_free(m);
    return set_error(BAMHIP_EDEVICE, "bamhip_mlp_create: initialisation failed");
  }
  if (int e = refresh(m)) {
    bamhip_mlp_free(m);
    return e;
  }
  *out = m;
  return BAMHIP_OK;
}

int bamhip_mlp_free(bamhip_mlp *m) {
  if (!m) return BAMHIP_OK;
  (void)hipStreamSynchronize(stream());
  for (void *p : {(void *)m->p, (void *)m->mom, (void *)m->vel, (void *)m->wq, (void *)m->w2, (void *)m->w2t,
                  (void *)m->part, (void *)m->lossp, (void *)m->grad})
    if (p) (void)hipFree(p);
  delete m;
  return BAMHIP_OK;
}

int bamhip_mlp_get_params(const bamhip_mlp *m, float *params, float *norm) {
  if (!m) return set_error(BAMHIP_EINVAL, "bamhip_mlp_get_params: NULL model");
  if (params) {
    BH_HIP(hipMemcpyAsync(params, m->p, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
  }
  if (norm) std::copy(m->norm, m->norm + 8, norm);
  return BAMHIP_OK;
}

int bamhip_mlp_set_params(bamhip_mlp *m, const float *params, const float *norm) {
  if (!m) return set_error(BAMHIP_EINVAL, "bamhip_mlp_set_params: NULL model");
  if (norm) {
    for (int i = 3; i < 6; i++)
      if (!(norm[i] != 0.f) || !std::isfinite(norm[i]))
        return set_error(BAMHIP_EINVAL, "bamhip_mlp_set_params: input scale %d is %g", i - 3, (double)norm[i]);
    if (!(norm[7] != 0.f)) return set_error(BAMHIP_EINVAL, "bamhip_mlp_set_params: output scale is 0");
    std::copy(norm, norm + 8, m->norm);
  }
  if (params) {
    const size_t np = sizeof(float) * kMlpNParam;
    BH_HIP(hipMemcpyAsync(m->p, params, np, hipMemcpyHostToDevice, stream()));
    BH_HIP(hipMemsetAsync(m->mom, 0, np, stream()));
    BH_HIP(hipMemsetAsync(m->vel, 0, np, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    m->step = 0;
    if (int e = refresh(m)) return e;
  }
  return BAMHIP_OK;
}

int bamhip_mlp_grad(bamhip_mlp *m, long n, const float *feat, const float *target, float *grad, float *loss) {
  if (!m || !feat || !target || n < 1) return set_error(BAMHIP_EINVAL, "bamhip_mlp_grad: bad arguments");
  if (int e = train_step(m, feat, target, n, 0, n, 1ULL, 0ULL, 0.f, m->grad)) return e;
  const long nblk = (n + kWave - 1) / kWave;
  std::vector<float> lp(nblk);
  if (grad) BH_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost, stream()));
  BH_HIP(hipMemcpyAsync(lp.data(), m->lossp, sizeof(float) * nblk, hipMemcpyDeviceToHost, stream()));
  BH_HIP(hipStreamSynchronize(stream()));
  if (loss) {
    double s = 0;
    for (float v : lp) s += v;
    *loss = (float)(s / n);
  }
  return BAMHIP_OK;
}

int bamhip_mlp_train(bamhip_mlp *m, long n, const float *feat, const float *target, int epochs, long batch,
                     float lr, unsigned long long seed, float *epoch_loss) {
  if (!m || !feat || !target || n < 1 || epochs < 1 || batch < kWave || lr <= 0.f)
    return set_error(BAMHIP_EINVAL, "bamhip_mlp_train: bad arguments");
  std::mt19937_64 rng(seed);
  const long steps_per_epoch = (n + batch - 1) / batch;
  const long total = steps_per_epoch * epochs;
  const long nblk = (batch + kWave - 1) / kWave;
  long t = 0;
  std::vector<float> lp(nblk);
  TimerScope ts("mlp_train");
  for (int ep = 0; ep < epochs; ep++) {
    // affine permutation q -> (a q + b) mod n with gcd(a, n) = 1
    unsigned long long a;
    do {
      a = (rng() % (unsigned long long)n) | 1ULL;
    } while (gcd64(a, (unsigned long long)n) != 1ULL);
    const unsigned long long b = rng() % (unsigned long long)n;
    double lsum = 0;
    for (long s = 0; s < steps_per_epoch; s++, t++) {
      const long q0 = s * batch;
      const long bs = std::min(batch, n - q0);
      const float lrt = lr * (0.02f + 0.98f * 0.5f * (1.f + cosf(3.14159265f * (float)t / (float)total)));
      if (int e = train_step(m, feat, target, n, q0, bs, a, b, lrt, nullptr)) return e;
      if (epoch_loss && s == steps_per_epoch - 1) {
        const long nb = (bs + kWave - 1) / kWave;
        BH_HIP(hipMemcpyAsync(lp.data(), m->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, stream()));
        BH_HIP(hipStreamSynchronize(stream()));
        for (long i = 0; i < nb; i++) lsum += lp[i];
        epoch_loss[ep] = (float)(lsum / bs);
      }
    }
  }
  return BAMHIP_OK;
}

int bamhip_mlp_predict(const bamhip_mlp *m, long n, const float *feat, float *y) {
  if (!m || !feat || !y || n < 0) return set_error(BAMHIP_EINVAL, "bamhip_mlp_predict: bad arguments");
  if (n == 0) return BAMHIP_OK;
  long blocks = std::min<long>((n + 255) / 256, 4096);
  TimerScope ts("mlp");
  mlp_predict_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(mlp_view(m), feat, n, y);
  BH_HIP(hipGetLastError());
  return BAMHIP_OK;
}

}  // extern "C"
