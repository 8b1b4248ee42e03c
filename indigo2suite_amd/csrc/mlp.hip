// The style selector's network: a multilayer perceptron in -> H -> H -> out
// (ReLU, FP32) with hand-written forward, backward and Adam kernels.  (No
// reference counterpart.)
//
// Mapping: the hidden width is one wavefront (H = 64), so one wave evaluates
// one sample with lane j owning neuron j; layer inputs are exchanged through
// LDS.  Weight gradients are computed weight-stationary (one thread per
// weight, summing its batch contributions in sample order), so training is
// deterministic.  Loss: masked mean squared error between the outputs and the
// targets (the selector regresses each style's log slowdown vs the best).
//
// Parameter layout (float): W1[in][H], b1[H], W2[H][H], b2[H], W3[H][out], b3[out].
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "igx_internal.hpp"

namespace igx {
namespace {

constexpr int kH = 64;
constexpr int kMaxIn = 64;
constexpr int kMaxOut = 64;

struct Mlp {
  int in = 0, out = 0;
  size_t nparams = 0;
  float* params = nullptr;  // device
  float* grads = nullptr;   // device
  float* m1 = nullptr;      // Adam moments
  float* m2 = nullptr;
  int step = 0;
  // activations / deltas for up to `cap` samples
  int cap = 0;
  float *x = nullptr, *t = nullptr, *mask = nullptr, *h1 = nullptr, *h2 = nullptr, *y = nullptr;
  float *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *loss = nullptr;
};

inline size_t param_count(int in, int out) { return (size_t)in * kH + kH + kH * kH + kH + (size_t)kH * out + out; }

struct Offs {
  size_t w1, b1, w2, b2, w3, b3;
};
__host__ __device__ inline Offs offsets(int in, int out)
{
  Offs o;
  o.w1 = 0;
  o.b1 = o.w1 + (size_t)in * kH;
  o.w2 = o.b1 + kH;
  o.b2 = o.w2 + (size_t)kH * kH;
  o.w3 = o.b2 + kH;
  o.b3 = o.w3 + (size_t)kH * out;
  (void)out;
  return o;
}

// one wave per sample; lane j = neuron j
__global__ void __launch_bounds__(kBlock) mlp_forward_kernel(const float* __restrict__ P, int in, int out,
                                                              const float* __restrict__ X, int batch,
                                                              float* __restrict__ H1, float* __restrict__ H2,
                                                              float* __restrict__ Y)
{
  __shared__ float s_x[kWavesPerBlock][kMaxIn];
  __shared__ float s_h[kWavesPerBlock][kH];
  const Offs o = offsets(in, out);
  const int w = threadIdx.x / kWave;
  const int j = lane_id();
  const int b = blockIdx.x * kWavesPerBlock + w;
  if (b >= batch) return;  // whole wave leaves together
  if (j < in) s_x[w][j] = X[(size_t)b * in + j];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  float a = P[o.b1 + j];
  for (int i = 0; i < in; i++) a += s_x[w][i] * P[o.w1 + (size_t)i * kH + j];
  const float h1 = a > 0.0f ? a : 0.0f;
  H1[(size_t)b * kH + j] = h1;
  s_h[w][j] = h1;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  a = P[o.b2 + j];
  for (int k = 0; k < kH; k++) a += s_h[w][k] * P[o.w2 + (size_t)k * kH + j];
  const float h2 = a > 0.0f ? a : 0.0f;
  H2[(size_t)b * kH + j] = h2;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  s_h[w][j] = h2;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  if (j < out) {
    a = P[o.b3 + j];
    for (int k = 0; k < kH; k++) a += s_h[w][k] * P[o.w3 + (size_t)k * out + j];
    Y[(size_t)b * out + j] = a;
  }
}

// output deltas and per-sample loss terms: D3 = 2 mask (Y - T) / norm
__global__ void __launch_bounds__(kBlock) mlp_out_delta(const float* __restrict__ Y, const float* __restrict__ T,
                                                         const float* __restrict__ M, int count, float inv_norm,
                                                         float* __restrict__ D3, float* __restrict__ L)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < count) {
    const float d = (Y[i] - T[i]) * M[i];
    D3[i] = 2.0f * d * inv_norm;
    L[i] = d * d * inv_norm;
  }
}

// hidden deltas, one wave per sample: D2 = (W3 D3) * [H2 > 0]; D1 = (W2 D2) * [H1 > 0]
__global__ void __launch_bounds__(kBlock) mlp_backward_deltas(const float* __restrict__ P, int in, int out,
                                                               int batch, const float* __restrict__ H1,
                                                               const float* __restrict__ H2,
                                                               const float* __restrict__ D3, float* __restrict__ D1,
                                                               float* __restrict__ D2)
{
  __shared__ float s_d[kWavesPerBlock][kMaxOut > kH ? kMaxOut : kH];
  const Offs o = offsets(in, out);
  const int w = threadIdx.x / kWave;
  const int j = lane_id();
  const int b = blockIdx.x * kWavesPerBlock + w;
  if (b >= batch) return;
  if (j < out) s_d[w][j] = D3[(size_t)b * out + j];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  float a = 0.0f;
  for (int c = 0; c < out; c++) a += P[o.w3 + (size_t)j * out + c] * s_d[w][c];
  const float d2 = H2[(size_t)b * kH + j] > 0.0f ? a : 0.0f;
  D2[(size_t)b * kH + j] = d2;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  s_d[w][j] = d2;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  a = 0.0f;
  for (int k = 0; k < kH; k++) a += P[o.w2 + (size_t)j * kH + k] * s_d[w][k];
  D1[(size_t)b * kH + j] = H1[(size_t)b * kH + j] > 0.0f ? a : 0.0f;
}

// weight-stationary gradients: dW[r][c] = sum_b A[b][r] D[b][c], db[c] = sum_b D[b][c]
__global__ void __launch_bounds__(kBlock) mlp_weight_grads(int rows, int cols, int batch, const float* __restrict__ A,
                                                            const float* __restrict__ D, float* __restrict__ dW,
                                                            float* __restrict__ db)
{
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx < rows * cols) {
    const int r = idx / cols, c = idx % cols;
    float s = 0.0f;
    for (int b = 0; b < batch; b++) s += A[(size_t)b * rows + r] * D[(size_t)b * cols + c];
    dW[idx] = s;
  } else if (idx < rows * cols + cols) {
    const int c = idx - rows * cols;
    float s = 0.0f;
    for (int b = 0; b < batch; b++) s += D[(size_t)b * cols + c];
    db[c] = s;
  }
}

__global__ void __launch_bounds__(kBlock) mlp_adam(float* __restrict__ P, const float* __restrict__ G,
                                                    float* __restrict__ m1, float* __restrict__ m2, int count,
                                                    float lr, float beta1, float beta2, float eps, float c1, float c2)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < count) {
    const float g = G[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * g;
    const float v = beta2 * m2[i] + (1.0f - beta2) * g * g;
    m1[i] = a;
    m2[i] = v;
    P[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
  }
}

__global__ void mlp_sum(const float* __restrict__ L, int count, float* __restrict__ out)
{
  // single workgroup, fixed order per lane then a fixed tree: deterministic
  __shared__ float s[kBlock];
  float acc = 0.0f;
  for (int i = threadIdx.x; i < count; i += kBlock) acc += L[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

int ensure_cap(Mlp* m, int batch)
{
  if (batch <= m->cap) return IGX_OK;
  float** bufs[] = {&m->x, &m->t, &m->mask, &m->h1, &m->h2, &m->y, &m->d1, &m->d2, &m->d3, &m->loss};
  for (float** b : bufs) {
    if (*b) IGX_HIP(hipFree(*b));
    *b = nullptr;
  }
  const size_t B = (size_t)batch;
  IGX_HIP(hipMalloc((void**)&m->x, sizeof(float) * B * m->in));
  IGX_HIP(hipMalloc((void**)&m->t, sizeof(float) * B * m->out));
  IGX_HIP(hipMalloc((void**)&m->mask, sizeof(float) * B * m->out));
  IGX_HIP(hipMalloc((void**)&m->h1, sizeof(float) * B * kH));
  IGX_HIP(hipMalloc((void**)&m->h2, sizeof(float) * B * kH));
  IGX_HIP(hipMalloc((void**)&m->y, sizeof(float) * B * m->out));
  IGX_HIP(hipMalloc((void**)&m->d1, sizeof(float) * B * kH));
  IGX_HIP(hipMalloc((void**)&m->d2, sizeof(float) * B * kH));
  IGX_HIP(hipMalloc((void**)&m->d3, sizeof(float) * B * m->out));
  IGX_HIP(hipMalloc((void**)&m->loss, sizeof(float) * (B * m->out + 1)));
  m->cap = batch;
  return IGX_OK;
}

void enqueue_forward(Mlp* m, const float* x, int batch, hipStream_t s)
{
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks_for(batch, kWavesPerBlock)), dim3(kBlock), 0, s, m->params,
                     m->in, m->out, x, batch, m->h1, m->h2, m->y);
}

// gradients of the masked MSE for samples [x, t, mask] already on the device
void enqueue_grads(Mlp* m, const float* x, const float* t, const float* mask, int batch, float inv_norm,
                   hipStream_t s)
{
  enqueue_forward(m, x, batch, s);
  const int cnt = batch * m->out;
  hipLaunchKernelGGL(mlp_out_delta, dim3(blocks_for(cnt, kBlock)), dim3(kBlock), 0, s, m->y, t, mask, cnt, inv_norm,
                     m->d3, m->loss);
  hipLaunchKernelGGL(mlp_backward_deltas, dim3(blocks_for(batch, kWavesPerBlock)), dim3(kBlock), 0, s, m->params,
                     m->in, m->out, batch, m->h1, m->h2, m->d3, m->d1, m->d2);
  const Offs o = offsets(m->in, m->out);
  hipLaunchKernelGGL(mlp_weight_grads, dim3(blocks_for(m->in * kH + kH, kBlock)), dim3(kBlock), 0, s, m->in, kH,
                     batch, x, m->d1, m->grads + o.w1, m->grads + o.b1);
  hipLaunchKernelGGL(mlp_weight_grads, dim3(blocks_for(kH * kH + kH, kBlock)), dim3(kBlock), 0, s, kH, kH, batch,
                     m->h1, m->d2, m->grads + o.w2, m->grads + o.b2);
  hipLaunchKernelGGL(mlp_weight_grads, dim3(blocks_for(kH * m->out + m->out, kBlock)), dim3(kBlock), 0, s, kH, m->out,
                     batch, m->h2, m->d3, m->grads + o.w3, m->grads + o.b3);
  hipLaunchKernelGGL(mlp_sum, dim3(1), dim3(kBlock), 0, s, m->loss, cnt, m->loss + cnt);
}

}  // namespace
}  // namespace igx

using namespace igx;

extern "C" {

int igx_mlp_create(int in, int out, uint64_t seed, void** handle)
{
  IGX_TRY(ensure_init());
  if (handle == nullptr || in < 1 || in > kMaxIn || out < 1 || out > kMaxOut) return IGX_ERR_ARG;
  Mlp* m = new Mlp();
  m->in = in;
  m->out = out;
  m->nparams = param_count(in, out);
  IGX_HIP(hipMalloc((void**)&m->params, sizeof(float) * m->nparams));
  IGX_HIP(hipMalloc((void**)&m->grads, sizeof(float) * m->nparams));
  IGX_HIP(hipMalloc((void**)&m->m1, sizeof(float) * m->nparams));
  IGX_HIP(hipMalloc((void**)&m->m2, sizeof(float) * m->nparams));
  // He-uniform initial weights from a counter-based hash, zero biases
  std::vector<float> p(m->nparams, 0.0f);
  const Offs o = offsets(in, out);
  auto fill = [&](size_t off, int fan_in, size_t count) {
    const float lim = sqrtf(6.0f / (float)fan_in);
    for (size_t i = 0; i < count; i++) {
      uint64_t z = seed * 0x9E3779B97F4A7C15ull + off + i + 1;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const float u = (float)((z >> 40) * (1.0 / 16777216.0));
      p[off + i] = (2.0f * u - 1.0f) * lim;
    }
  };
  fill(o.w1, in, (size_t)in * kH);
  fill(o.w2, kH, (size_t)kH * kH);
  fill(o.w3, kH, (size_t)kH * out);
  IGX_HIP(hipMemcpy(m->params, p.data(), sizeof(float) * m->nparams, hipMemcpyHostToDevice));
  std::vector<float> zero(m->nparams, 0.0f);
  IGX_HIP(hipMemcpy(m->m1, zero.data(), sizeof(float) * m->nparams, hipMemcpyHostToDevice));
  IGX_HIP(hipMemcpy(m->m2, zero.data(), sizeof(float) * m->nparams, hipMemcpyHostToDevice));
  *handle = m;
  return IGX_OK;
}

int igx_mlp_destroy(void* handle)
{
  Mlp* m = (Mlp*)handle;
  if (m == nullptr) return IGX_ERR_ARG;
  hipStreamSynchronize(rt().stream);
  float* bufs[] = {m->params, m->grads, m->m1, m->m2, m->x, m->t, m->mask, m->h1, m->h2, m->y, m->d1, m->d2, m->d3,
                   m->loss};
  for (float* b : bufs) {
    if (b) hipFree(b);
  }
  delete m;
  return IGX_OK;
}

int igx_mlp_param_count(void* handle) { return handle ? (int)((Mlp*)handle)->nparams : -1; }

int igx_mlp_get_params(void* handle, float* host, size_t count)
{
  Mlp* m = (Mlp*)handle;
  if (m == nullptr || host == nullptr || count != m->nparams) return IGX_ERR_ARG;
  return igx_memcpy_d2h(host, m->params, sizeof(float) * count);
}

int igx_mlp_set_params(void* handle, const float* host, size_t count)
{
  Mlp* m = (Mlp*)handle;
  if (m == nullptr || host == nullptr || count != m->nparams) return IGX_ERR_ARG;
  m->step = 0;
  return igx_memcpy_h2d(m->params, host, sizeof(float) * count);
}

int igx_mlp_forward(void* handle, const float* host_x, int batch, float* host_y)
{
  Mlp* m = (Mlp*)handle;
  if (m == nullptr || host_x == nullptr || host_y == nullptr || batch < 1) return IGX_ERR_ARG;
  IGX_TRY(ensure_cap(m, batch));
  IGX_TRY(igx_memcpy_h2d(m->x, host_x, sizeof(float) * (size_t)batch * m->in));
  enqueue_forward(m, m->x, batch, rt().stream);
  IGX_HIP(hipGetLastError());
  return igx_memcpy_d2h(host_y, m->y, sizeof(float) * (size_t)batch * m->out);
}

// gradient of the masked MSE (sum over entries / norm) for one batch, for checking
int igx_mlp_gradients(void* handle, const float* x, const float* t, const float* mask, int batch, float norm,
                      float* host_grads, float* loss)
{
  Mlp* m = (Mlp*)handle;
  if (m == nullptr || !x || !t || !mask || batch < 1 || !(norm > 0.0f)) return IGX_ERR_ARG;
  IGX_TRY(ensure_cap(m, batch));
  IGX_TRY(igx_memcpy_h2d(m->x, x, sizeof(float) * (size_t)batch * m->in));
  IGX_TRY(igx_memcpy_h2d(m->t, t, sizeof(float) * (size_t)batch * m->out));
  IGX_TRY(igx_memcpy_h2d(m->mask, mask, sizeof(float) * (size_t)batch * m->out));
  enqueue_grads(m, m->x, m->t, m->mask, batch, 1.0f / norm, rt().stream);
  IGX_HIP(hipGetLastError());
  if (host_grads) IGX_TRY(igx_memcpy_d2h(host_grads, m->grads, sizeof(float) * m->nparams));
  if (loss) IGX_TRY(igx_memcpy_d2h(loss, m->loss + (size_t)batch * m->out, sizeof(float)));
  return IGX_OK;
}

// Adam on minibatches drawn in a seeded order from n samples resident on the
// device; returns the mean loss of the last epoch.
int igx_mlp_train(void* handle, const float* x, const float* t, const float* mask, int n, int epochs, int batch,
                  float lr, uint64_t seed, float* last_loss)
{
  Mlp* m = (Mlp*)handle;
  if (m == nullptr || !x || !t || !mask || n < 1 || epochs < 1 || batch < 1) return IGX_ERR_ARG;
  Runtime& R = rt();
  const int in = m->in, out = m->out;
  float *dx = nullptr, *dt = nullptr, *dm = nullptr;
  IGX_HIP(hipMalloc((void**)&dx, sizeof(float) * (size_t)n * in));
  IGX_HIP(hipMalloc((void**)&dt, sizeof(float) * (size_t)n * out));
  IGX_HIP(hipMalloc((void**)&dm, sizeof(float) * (size_t)n * out));
  // shuffled copies per epoch are staged on the host, one contiguous upload each
  std::vector<int> perm(n);
  std::vector<float> sx((size_t)n * in), st((size_t)n * out), sm((size_t)n * out);
  IGX_TRY(ensure_cap(m, batch));
  double epoch_loss = 0.0;
  for (int ep = 0; ep < epochs; ep++) {
    for (int i = 0; i < n; i++) perm[i] = i;
    uint64_t s = seed * 0x2545F4914F6CDD1Dull + (uint64_t)ep + 1;
    for (int i = n - 1; i > 0; i--) {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      const int j = (int)(s % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < n; i++) {
      memcpy(&sx[(size_t)i * in], &x[(size_t)perm[i] * in], sizeof(float) * in);
      memcpy(&st[(size_t)i * out], &t[(size_t)perm[i] * out], sizeof(float) * out);
      memcpy(&sm[(size_t)i * out], &mask[(size_t)perm[i] * out], sizeof(float) * out);
    }
    IGX_HIP(hipMemcpyAsync(dx, sx.data(), sizeof(float) * sx.size(), hipMemcpyHostToDevice, R.stream));
    IGX_HIP(hipMemcpyAsync(dt, st.data(), sizeof(float) * st.size(), hipMemcpyHostToDevice, R.stream));
    IGX_HIP(hipMemcpyAsync(dm, sm.data(), sizeof(float) * sm.size(), hipMemcpyHostToDevice, R.stream));
    epoch_loss = 0.0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int bs = std::min(batch, n - b0);
      float msum = 0.0f;
      for (int i = b0 * out; i < (b0 + bs) * out; i++) msum += sm[i];
      const float norm = msum > 0.0f ? msum : 1.0f;
      enqueue_grads(m, dx + (size_t)b0 * in, dt + (size_t)b0 * out, dm + (size_t)b0 * out, bs, 1.0f / norm, R.stream);
      m->step++;
      const float c1 = 1.0f - powf(0.9f, (float)m->step);
      const float c2 = 1.0f - powf(0.999f, (float)m->step);
      hipLaunchKernelGGL(mlp_adam, dim3(blocks_for((long long)m->nparams, kBlock)), dim3(kBlock), 0, R.stream,
                         m->params, m->grads, m->m1, m->m2, (int)m->nparams, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
      float l = 0.0f;
      IGX_HIP(hipMemcpyAsync(&l, m->loss + (size_t)bs * out, sizeof(float), hipMemcpyDeviceToHost, R.stream));
      IGX_HIP(hipStreamSynchronize(R.stream));
      epoch_loss += (double)l * bs;
    }
    epoch_loss /= n;
  }
  IGX_HIP(hipGetLastError());
  hipFree(dx);
  hipFree(dt);
  hipFree(dm);
  if (last_loss) *last_loss = (float)epoch_loss;
  return IGX_OK;
}

}  // extern "C"
