// convnet.hip -- the learned early stop: a causal dilated 1-D convolution
// network trained and run on the device with hand-written kernels.
//
// No reference counterpart (the reference always runs to the end
// criterion). The network maps port traces observed over the first part
// of a run, plus the known source, to the whole trace; see
// include/ems_learn.h for its shape and the C ABI.
//
// Layout: activations [batch][channel][time] f32; weights
// [out][in][tap] with an f32 master copy (Adam) and a bf16 copy that the
// forward and backward passes read (f32 accumulation).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/ems_gpu.h"
#include "../../include/ems_learn.h"

// The network tolerates reassociation (its parity is a tolerance, not bit
// identity): let the compiler contract multiply-adds here.
#pragma clang fp contract(fast)

namespace {

constexpr int CN_K = 9;  // taps

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// Round to nearest even (finite inputs), as the host does.
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t u = __float_as_uint(f);
#else
  uint32_t u;
  std::memcpy(&u, &f, 4);
#endif
  const uint32_t r = ((u >> 16) & 1u) + 0x7FFFu;
  return (uint16_t)((u + r) >> 16);
}

// z[b][co][t] = bias[co] + sum_ci sum_k w[co][ci][k] x[b][ci][t - dil (K-1-k)]
// (x = 0 before t = 0). A thread per output.
__global__ void __launch_bounds__(256) cn_conv_fwd(const float *__restrict__ x, int cin,
                                                   const uint16_t *__restrict__ w,
                                                   const float *__restrict__ bias,
                                                   float *__restrict__ z, int cout, int B, int T,
                                                   int dil) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)B * cout * T) return;
  const int t = (int)(id % T);
  const int co = (int)((id / T) % cout);
  const int b = (int)(id / ((size_t)T * cout));
  float acc = bias[co];
  const float *xb = x + (size_t)b * cin * T;
  const uint16_t *wr = w + (size_t)co * cin * CN_K;
  for (int ci = 0; ci < cin; ci++) {
    const float *xc = xb + (size_t)ci * T;
#pragma unroll
    for (int k = 0; k < CN_K; k++) {
      const int s = t - dil * (CN_K - 1 - k);
      if (s >= 0) acc += bf16_to_f32(wr[ci * CN_K + k]) * xc[s];
    }
  }
  z[id] = acc;
}

// a = tanh(z); h_out = h_in + a (residual) or a.
__global__ void __launch_bounds__(256) cn_act_fwd(const float *__restrict__ z, float *__restrict__ a,
                                                  const float *__restrict__ h_in,
                                                  float *__restrict__ h_out, size_t n,
                                                  int residual) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = tanhf(z[i]);
  a[i] = v;
  if (residual)
    h_out[i] = h_in[i] + v;
  else
    h_out[i] = v;
}

// Mean squared error over t >= t0 and its gradient; per-block partial sums.
__global__ void __launch_bounds__(256) cn_loss(const float *__restrict__ out,
                                               const float *__restrict__ y, float *__restrict__ g,
                                               size_t n, int T, int t0, float inv_count,
                                               float *__restrict__ partial) {
  __shared__ float red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (i < n) {
    const int t = (int)(i % T);
    const float d = t >= t0 ? out[i] - y[i] : 0.0f;
    g[i] = 2.0f * d * inv_count;
    s = d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// The spectral term: per sample and output channel, the error d over
// samples [t0, t0 + len) adds w d^T K d to the loss and 2 w K d to its
// gradient (K symmetric, len x len, row-major). A thread per (sample,
// channel, t); per-block sums of d_t (K d)_t.
__global__ void __launch_bounds__(256) cn_spec(const float *__restrict__ out,
                                               const float *__restrict__ y,
                                               const float *__restrict__ K, float *__restrict__ g,
                                               int BC, int T, int t0, int len, float w,
                                               float *__restrict__ partial) {
  __shared__ float red[256];
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (id < (size_t)BC * len) {
    const int t = (int)(id % len);
    const size_t row = (id / len) * T + t0;
    const float *kr = K + (size_t)t * len;
    float acc = 0.0f;
    for (int u = 0; u < len; u++) acc += kr[u] * (out[row + u] - y[row + u]);
    g[row + t] += 2.0f * w * acc;
    s = (out[row + t] - y[row + t]) * acc;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = w * red[0];
}

// dz = g (1 - a^2).
__global__ void __launch_bounds__(256) cn_act_bwd(const float *__restrict__ g,
                                                  const float *__restrict__ a,
                                                  float *__restrict__ dz, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = a[i];
  dz[i] = g[i] * (1.0f - v * v);
}

// dx[b][ci][t] (+)= sum_co sum_k w[co][ci][k] dz[b][co][t + dil (K-1-k)].
__global__ void __launch_bounds__(256) cn_conv_bwd_x(const float *__restrict__ dz, int cout,
                                                     const uint16_t *__restrict__ w, int cin,
                                                     float *__restrict__ dx, int B, int T, int dil,
                                                     int accumulate) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)B * cin * T) return;
  const int t = (int)(id % T);
  const int ci = (int)((id / T) % cin);
  const int b = (int)(id / ((size_t)T * cin));
  float acc = 0.0f;
  const float *db = dz + (size_t)b * cout * T;
  for (int co = 0; co < cout; co++) {
    const float *dc = db + (size_t)co * T;
    const uint16_t *wr = w + ((size_t)co * cin + ci) * CN_K;
#pragma unroll
    for (int k = 0; k < CN_K; k++) {
      const int s = t + dil * (CN_K - 1 - k);
      if (s < T) acc += bf16_to_f32(wr[k]) * dc[s];
    }
  }
  dx[id] = accumulate ? dx[id] + acc : acc;
}

// dw[co][ci][k] = sum_b sum_t dz[b][co][t] x[b][ci][t - dil (K-1-k)], and
// (ci == 0 blocks) db[co] = sum dz[b][co][t]. A block per (co, ci).
__global__ void __launch_bounds__(256) cn_conv_bwd_w(const float *__restrict__ dz, int cout,
                                                     const float *__restrict__ x, int cin,
                                                     float *__restrict__ dw,
                                                     float *__restrict__ dbias, int B, int T,
                                                     int dil) {
  const int co = blockIdx.x / cin, ci = blockIdx.x % cin;
  float acc[CN_K + 1];
#pragma unroll
  for (int k = 0; k <= CN_K; k++) acc[k] = 0.0f;
  const size_t n = (size_t)B * T;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
    const int t = (int)(i % T), b = (int)(i / T);
    const float d = dz[((size_t)b * cout + co) * T + t];
    const float *xc = x + ((size_t)b * cin + ci) * T;
#pragma unroll
    for (int k = 0; k < CN_K; k++) {
      const int s = t - dil * (CN_K - 1 - k);
      if (s >= 0) acc[k] += d * xc[s];
    }
    acc[CN_K] += d;
  }
  __shared__ float red[CN_K + 1][256];
#pragma unroll
  for (int k = 0; k <= CN_K; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k <= CN_K; k++) red[k][threadIdx.x] += red[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < CN_K) dw[((size_t)co * cin + ci) * CN_K + threadIdx.x] = red[threadIdx.x][0];
  if (threadIdx.x == CN_K && ci == 0) dbias[co] = red[CN_K][0];
}

// Per-block sums of g^2 (the gradient norm, for clipping).
__global__ void __launch_bounds__(256) cn_sqsum(const float *__restrict__ g, size_t n,
                                                float *__restrict__ partial) {
  __shared__ float red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const float x = i < n ? g[i] : 0.0f;
  red[threadIdx.x] = x * x;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on the f32 master parameters (gradient scaled by gscale, the
// clipping factor), then their bf16 copy.
__global__ void __launch_bounds__(256) cn_adam(float *__restrict__ p, const float *__restrict__ g,
                                               float *__restrict__ m, float *__restrict__ v,
                                               size_t n, float lr, float b1, float b2, float eps,
                                               float c1, float c2, float gscale,
                                               uint16_t *__restrict__ pbf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * gscale;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float q = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = q;
  if (pbf) pbf[i] = f32_to_bf16(q);
}

__global__ void __launch_bounds__(256) cn_to_bf16(const float *__restrict__ p,
                                                  uint16_t *__restrict__ q, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = f32_to_bf16(p[i]);
}

// Gathers samples idx[] of a dataset [n][c][T] into a batch.
__global__ void __launch_bounds__(256) cn_gather(const float *__restrict__ src,
                                                 const int32_t *__restrict__ idx,
                                                 float *__restrict__ dst, int c, int T, int B) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)c * T;
  if (i >= per * B) return;
  const int b = (int)(i / per);
  dst[i] = src[(size_t)idx[b] * per + (i % per)];
}

unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct ems_cn {
  int cin = 0, cout = 0, C = 0, L = 0;
  std::vector<int> dil;
  std::vector<int> li, lo;          // per layer in/out channels
  std::vector<size_t> woff, boff;   // offsets into the parameter vector
  size_t np = 0;                    // parameters
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *pbf = nullptr;
  int step = 0;
  hipStream_t stream = nullptr;
  // dataset
  float *dx = nullptr, *dy = nullptr;
  int dn = 0, dT = 0;
  // work buffers for (B, T)
  size_t cap = 0;
  float *hx = nullptr;              // input batch
  float *hy = nullptr;              // target batch
  std::vector<float *> h, a, z;     // per layer: input activation, tanh, pre-activation
  float *out = nullptr, *gout = nullptr, *gh = nullptr, *dz = nullptr, *part = nullptr;
  float *gpart = nullptr;  // per-block sums of g^2
  float *spart = nullptr;  // per-block sums of the spectral loss
  // spectral loss term (len x len kernel over samples [st0, st0 + slen))
  float *dK = nullptr;
  int st0 = 0, slen = 0;
  float sw = 0.0f;
  int32_t *didx = nullptr;
  std::string err;
};

namespace {

int cfail(ems_cn *n, int code, const std::string &m) {
  n->err = m;
  return code;
}

#define CCK(n, x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return cfail((n), EMS_E_DEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

int ensure_work(ems_cn *n, int B, int T) {
  const size_t need = (size_t)B * T;
  if (need <= n->cap) return 0;
  auto fr = [](float *&q) {
    if (q) (void)hipFree(q);
    q = nullptr;
  };
  fr(n->hx), fr(n->hy), fr(n->out), fr(n->gout), fr(n->gh), fr(n->dz), fr(n->part), fr(n->spart);
  // h[0] is the input batch (hx), freed above
  for (size_t l = 1; l < n->h.size(); l++) fr(n->h[l]);
  for (auto *&q : n->a) fr(q);
  for (auto *&q : n->z) fr(q);
  if (n->didx) (void)hipFree(n->didx);
  const int C = std::max(n->C, std::max(n->cin, n->cout));
  CCK(n, hipMalloc(&n->hx, need * n->cin * sizeof(float)));
  CCK(n, hipMalloc(&n->hy, need * n->cout * sizeof(float)));
  CCK(n, hipMalloc(&n->out, need * n->cout * sizeof(float)));
  CCK(n, hipMalloc(&n->gout, need * n->cout * sizeof(float)));
  CCK(n, hipMalloc(&n->gh, need * C * sizeof(float)));
  CCK(n, hipMalloc(&n->dz, need * C * sizeof(float)));
  CCK(n, hipMalloc(&n->part, (blocks(need * n->cout) + 1) * sizeof(float)));
  CCK(n, hipMalloc(&n->spart, (blocks(need * n->cout) + 1) * sizeof(float)));
  CCK(n, hipMalloc(&n->didx, (size_t)B * sizeof(int32_t)));
  n->h.assign(n->L, nullptr), n->a.assign(n->L, nullptr), n->z.assign(n->L, nullptr);
  for (int l = 0; l < n->L; l++) {
    if (l > 0) CCK(n, hipMalloc(&n->h[l], need * n->li[l] * sizeof(float)));
    if (l < n->L - 1) {
      CCK(n, hipMalloc(&n->a[l], need * n->lo[l] * sizeof(float)));
      CCK(n, hipMalloc(&n->z[l], need * n->lo[l] * sizeof(float)));
    }
  }
  n->cap = need;
  return 0;
}

// Forward pass of the batch in n->hx (h[0]); output in n->out.
int forward(ems_cn *n, int B, int T) {
  const size_t BT = (size_t)B * T;
  n->h[0] = n->hx;
  for (int l = 0; l < n->L; l++) {
    const bool last = l == n->L - 1;
    float *zl = last ? n->out : n->z[l];
    const size_t no = BT * n->lo[l];
    hipLaunchKernelGGL(cn_conv_fwd, dim3(blocks(no)), dim3(256), 0, n->stream, n->h[l], n->li[l],
                       n->pbf + n->woff[l], n->p + n->boff[l], zl, n->lo[l], B, T, n->dil[l]);
    if (!last)
      hipLaunchKernelGGL(cn_act_fwd, dim3(blocks(no)), dim3(256), 0, n->stream, zl, n->a[l],
                         l > 0 ? n->h[l] : nullptr, n->h[l + 1], no, l > 0 ? 1 : 0);
  }
  CCK(n, hipGetLastError());
  return 0;
}

// Loss over t >= t0 against n->hy and the parameter gradient in n->g;
// *gnorm (when not NULL) receives the gradient's 2-norm.
int backward(ems_cn *n, int B, int T, int t0, float *loss, double *gnorm = nullptr) {
  const size_t BT = (size_t)B * T;
  const size_t no = BT * n->cout;
  const int Tv = T - t0 > 0 ? T - t0 : 1;
  const float inv = 1.0f / (float)((size_t)B * n->cout * Tv);
  const unsigned nb = blocks(no);
  hipLaunchKernelGGL(cn_loss, dim3(nb), dim3(256), 0, n->stream, n->out, n->hy, n->gout, no, T, t0,
                     inv, n->part);
  unsigned nsb = 0;
  if (n->sw != 0.0f) {
    if (n->st0 + n->slen > T) return cfail(n, EMS_E_INVALID, "spectral term past the sequence end");
    const int BC = B * n->cout;
    nsb = blocks((size_t)BC * n->slen);
    hipLaunchKernelGGL(cn_spec, dim3(nsb), dim3(256), 0, n->stream, n->out, n->hy, n->dK, n->gout,
                       BC, T, n->st0, n->slen, n->sw / (float)BC, n->spart);
  }
  for (int l = n->L - 1; l >= 0; l--) {
    const bool last = l == n->L - 1;
    const size_t nl = BT * n->lo[l];
    const float *dzl;
    if (last) {
      dzl = n->gout;
    } else {
      hipLaunchKernelGGL(cn_act_bwd, dim3(blocks(nl)), dim3(256), 0, n->stream, n->gh, n->a[l],
                         n->dz, nl);
      dzl = n->dz;
    }
    hipLaunchKernelGGL(cn_conv_bwd_w, dim3(n->lo[l] * n->li[l]), dim3(256), 0, n->stream, dzl,
                       n->lo[l], n->h[l], n->li[l], n->g + n->woff[l], n->g + n->boff[l], B, T,
                       n->dil[l]);
    if (l > 0) {
      // gradient of h[l]: through the conv, plus the skip path of a
      // residual layer (h[l+1] = h[l] + a[l]).
      const size_t ni = BT * n->li[l];
      hipLaunchKernelGGL(cn_conv_bwd_x, dim3(blocks(ni)), dim3(256), 0, n->stream, dzl, n->lo[l],
                         n->pbf + n->woff[l], n->li[l], n->gh, B, T, n->dil[l],
                         (!last && l > 0) ? 1 : 0);
    }
  }
  const unsigned ng = blocks(n->np);
  if (gnorm)
    hipLaunchKernelGGL(cn_sqsum, dim3(ng), dim3(256), 0, n->stream, n->g, n->np, n->gpart);
  CCK(n, hipGetLastError());
  std::vector<float> part(nb), gp(gnorm ? ng : 0), sp(nsb);
  CCK(n, hipMemcpyAsync(part.data(), n->part, nb * sizeof(float), hipMemcpyDeviceToHost,
                        n->stream));
  if (nsb)
    CCK(n, hipMemcpyAsync(sp.data(), n->spart, nsb * sizeof(float), hipMemcpyDeviceToHost,
                          n->stream));
  if (gnorm)
    CCK(n, hipMemcpyAsync(gp.data(), n->gpart, ng * sizeof(float), hipMemcpyDeviceToHost,
                          n->stream));
  CCK(n, hipStreamSynchronize(n->stream));
  double s = 0.0;
  for (float v : part) s += v;
  double ss = 0.0;
  for (float v : sp) ss += v;
  if (loss) *loss = (float)(s * inv + ss);
  if (gnorm) {
    double g2 = 0.0;
    for (float v : gp) g2 += v;
    *gnorm = std::sqrt(g2);
  }
  return 0;
}

}  // namespace

extern "C" {

int ems_cn_create(int cin, int cout, int channels, int layers, const int32_t *dilation,
                  uint64_t seed, ems_cn **out) {
  if (!out || cin < 1 || cout < 1 || channels < 1 || layers < 2 || !dilation) return EMS_E_INVALID;
  ems_cn *n = new ems_cn();
  n->cin = cin, n->cout = cout, n->C = channels, n->L = layers;
  for (int l = 0; l < layers; l++) {
    if (dilation[l] < 1) {
      delete n;
      return EMS_E_INVALID;
    }
    n->dil.push_back(dilation[l]);
    n->li.push_back(l == 0 ? cin : channels);
    n->lo.push_back(l == layers - 1 ? cout : channels);
  }
  for (int l = 0; l < layers; l++) {
    n->woff.push_back(n->np);
    n->np += (size_t)n->lo[l] * n->li[l] * CN_K;
    n->boff.push_back(n->np);
    n->np += n->lo[l];
  }
  // Initial weights: normal(0, 1/sqrt(fan_in)), the output layer 0.1 of
  // that; biases 0.
  std::vector<float> p(n->np, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int l = 0; l < layers; l++) {
    const float sc = (l == layers - 1 ? 0.1f : 1.0f) / std::sqrt((float)(n->li[l] * CN_K));
    for (size_t q = 0; q < (size_t)n->lo[l] * n->li[l] * CN_K; q++) p[n->woff[l] + q] = nd(rng) * sc;
  }
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&n->gpart, blocks(n->np) * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->p, n->np * 4) != hipSuccess || hipMalloc(&n->g, n->np * 4) != hipSuccess ||
      hipMalloc(&n->m, n->np * 4) != hipSuccess || hipMalloc(&n->v, n->np * 4) != hipSuccess ||
      hipMalloc(&n->pbf, n->np * 2) != hipSuccess) {
    ems_cn_destroy(n);
    return EMS_E_DEVICE;
  }
  *out = n;
  return ems_cn_set_params(n, p.data(), n->np);
}

void ems_cn_destroy(ems_cn *n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void *q : {(void *)n->p, (void *)n->g, (void *)n->m, (void *)n->v, (void *)n->pbf,
                  (void *)n->dx, (void *)n->dy, (void *)n->hx, (void *)n->hy, (void *)n->out,
                  (void *)n->gout, (void *)n->gh, (void *)n->dz, (void *)n->part, (void *)n->gpart,
                  (void *)n->spart, (void *)n->dK,
                  (void *)n->didx})
    if (q) (void)hipFree(q);
  for (int l = 1; l < (int)n->h.size(); l++)
    if (n->h[l]) (void)hipFree(n->h[l]);
  for (auto *q : n->a)
    if (q) (void)hipFree(q);
  for (auto *q : n->z)
    if (q) (void)hipFree(q);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

const char *ems_cn_last_error(const ems_cn *n) { return n ? n->err.c_str() : "null network"; }

uint64_t ems_cn_param_count(const ems_cn *n) { return n ? n->np : 0; }

int ems_cn_set_params(ems_cn *n, const float *p, uint64_t count) {
  if (!n || !p || count != n->np) return EMS_E_INVALID;
  CCK(n, hipMemcpyAsync(n->p, p, n->np * 4, hipMemcpyHostToDevice, n->stream));
  CCK(n, hipMemsetAsync(n->m, 0, n->np * 4, n->stream));
  CCK(n, hipMemsetAsync(n->v, 0, n->np * 4, n->stream));
  hipLaunchKernelGGL(cn_to_bf16, dim3(blocks(n->np)), dim3(256), 0, n->stream, n->p, n->pbf, n->np);
  CCK(n, hipGetLastError());
  CCK(n, hipStreamSynchronize(n->stream));
  n->step = 0;
  return EMS_OK;
}

int ems_cn_get_params(ems_cn *n, float *p, uint64_t count) {
  if (!n || !p || count != n->np) return EMS_E_INVALID;
  CCK(n, hipStreamSynchronize(n->stream));
  CCK(n, hipMemcpyAsync(p, n->p, n->np * 4, hipMemcpyDeviceToHost, n->stream));
  CCK(n, hipStreamSynchronize(n->stream));
  return EMS_OK;
}

int ems_cn_set_spectral(ems_cn *n, const float *K, int t0, int len, float weight) {
  if (!n || t0 < 0 || len < 0 || (weight != 0.0f && (!K || len < 1))) return EMS_E_INVALID;
  if (n->dK) CCK(n, hipFree(n->dK));
  n->dK = nullptr, n->sw = 0.0f, n->st0 = 0, n->slen = 0;
  if (weight == 0.0f) return EMS_OK;
  const size_t bytes = (size_t)len * len * sizeof(float);
  CCK(n, hipMalloc(&n->dK, bytes));
  CCK(n, hipMemcpyAsync(n->dK, K, bytes, hipMemcpyHostToDevice, n->stream));
  CCK(n, hipStreamSynchronize(n->stream));
  n->sw = weight, n->st0 = t0, n->slen = len;
  return EMS_OK;
}

int ems_cn_set_data(ems_cn *n, const float *x, const float *y, int count, int T) {
  if (!n || !x || !y || count < 1 || T < 1) return EMS_E_INVALID;
  if (n->dx) CCK(n, hipFree(n->dx));
  if (n->dy) CCK(n, hipFree(n->dy));
  n->dx = n->dy = nullptr;
  const size_t nx = (size_t)count * n->cin * T, ny = (size_t)count * n->cout * T;
  CCK(n, hipMalloc(&n->dx, nx * 4));
  CCK(n, hipMalloc(&n->dy, ny * 4));
  CCK(n, hipMemcpyAsync(n->dx, x, nx * 4, hipMemcpyHostToDevice, n->stream));
  CCK(n, hipMemcpyAsync(n->dy, y, ny * 4, hipMemcpyHostToDevice, n->stream));
  CCK(n, hipStreamSynchronize(n->stream));
  n->dn = count, n->dT = T;
  return EMS_OK;
}

int ems_cn_train_batch(ems_cn *n, const int32_t *idx, int B, int t0, float lr, float clip,
                       float *loss, float *grad_norm) {
  if (!n || !idx || B < 1 || !n->dx) return EMS_E_INVALID;
  for (int b = 0; b < B; b++)
    if (idx[b] < 0 || idx[b] >= n->dn) return cfail(n, EMS_E_INVALID, "sample index out of range");
  const int T = n->dT;
  int rc = ensure_work(n, B, T);
  if (rc) return rc;
  CCK(n, hipMemcpyAsync(n->didx, idx, B * sizeof(int32_t), hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(cn_gather, dim3(blocks((size_t)B * n->cin * T)), dim3(256), 0, n->stream,
                     n->dx, n->didx, n->hx, n->cin, T, B);
  hipLaunchKernelGGL(cn_gather, dim3(blocks((size_t)B * n->cout * T)), dim3(256), 0, n->stream,
                     n->dy, n->didx, n->hy, n->cout, T, B);
  if ((rc = forward(n, B, T))) return rc;
  double gn = 0.0;
  if ((rc = backward(n, B, T, t0, loss, &gn))) return rc;
  if (grad_norm) *grad_norm = (float)gn;
  n->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
  const float gscale = (clip > 0.0f && gn > clip) ? (float)(clip / gn) : 1.0f;
  hipLaunchKernelGGL(cn_adam, dim3(blocks(n->np)), dim3(256), 0, n->stream, n->p, n->g, n->m, n->v,
                     n->np, lr, b1, b2, 1e-8f, c1, c2, gscale, n->pbf);
  CCK(n, hipGetLastError());
  return EMS_OK;
}

int ems_cn_forward(ems_cn *n, const float *x, int B, int T, float *out) {
  if (!n || !x || !out || B < 1 || T < 1) return EMS_E_INVALID;
  int rc = ensure_work(n, B, T);
  if (rc) return rc;
  CCK(n, hipMemcpyAsync(n->hx, x, (size_t)B * n->cin * T * 4, hipMemcpyHostToDevice, n->stream));
  if ((rc = forward(n, B, T))) return rc;
  CCK(n, hipMemcpyAsync(out, n->out, (size_t)B * n->cout * T * 4, hipMemcpyDeviceToHost,
                        n->stream));
  CCK(n, hipStreamSynchronize(n->stream));
  return EMS_OK;
}

int ems_cn_loss_grad(ems_cn *n, const float *x, const float *y, int B, int T, int t0, float *loss,
                     float *grad) {
  if (!n || !x || !y || B < 1 || T < 1) return EMS_E_INVALID;
  int rc = ensure_work(n, B, T);
  if (rc) return rc;
  CCK(n, hipMemcpyAsync(n->hx, x, (size_t)B * n->cin * T * 4, hipMemcpyHostToDevice, n->stream));
  CCK(n, hipMemcpyAsync(n->hy, y, (size_t)B * n->cout * T * 4, hipMemcpyHostToDevice, n->stream));
  if ((rc = forward(n, B, T))) return rc;
  if ((rc = backward(n, B, T, t0, loss))) return rc;
  if (grad) {
    CCK(n, hipMemcpyAsync(grad, n->g, n->np * 4, hipMemcpyDeviceToHost, n->stream));
    CCK(n, hipStreamSynchronize(n->stream));
  }
  return EMS_OK;
}

}  // extern "C"
