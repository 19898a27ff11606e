// Total-variation regularisation + Adam update of per-cell parameter fields
// on gfx950, and its C ABI (ernn_tv_adam_* in include/ernn.h).
//
// Reference (paths relative to the reference repository root):
//   DI = Damage_inversion/Inversion_adam.py
//   fields            C_p = clamp(exp(theta_p), lo_p, hi_p)   DI:304-322 (log parameters DI:245-251)
//   TV                sum |C[i+1,j] - C[i,j]| + sum |C[i,j+1] - C[i,j]|   DI:179-195
//   loss              data + lambda * sum_p TV(C_p)   DI:441-463, gradient by autograd DI:511
//   update            torch.optim.Adam over the six theta fields   DI:478-483, DI:514
//
// The reference builds TV and its gradient from ~20 tensor ops per field and
// steps Adam with ~10 more.  Here one call runs two passes over the fields:
//   tv_grad_kernel* g_theta += (lambda s(i,j)) m(i,j) e(i,j), the TV
//                   subgradient s (sign(0) = 0 as abs' backward) chained
//                   through the clamp mask m (inclusive bounds) and
//                   d exp = e; per-block FP64 TV partial sums
//   adam_kernel     m, v, theta in place with torch's arithmetic order
//                   (lerp, mul + addcmul, sqrt / bias correction + eps,
//                   addcdiv)
// Two passes because a cell's TV gradient reads its neighbours' theta,
// which an in-place update in the same pass would race with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "ernn.h"
#include "ernn_internal.h"

namespace ernn {

int set_last_error(int code, const std::string& msg);

constexpr int kTvThreads = 256;

struct TvArgs {
  const float* theta[ERNN_TV_MAX_FIELDS];
  float* grad[ERNN_TV_MAX_FIELDS];
  float lo[ERNN_TV_MAX_FIELDS], hi[ERNN_TV_MAX_FIELDS];
  double* partial;       // [n_fields][blocks_per_field] TV partial sums
  int nx, ny, log_params, blocks_per_field;
  int blocks_per_xcd;    // grid.x = 8 x blocks_per_xcd
  float lambda;
};

struct AdamArgs {
  float* theta[ERNN_TV_MAX_FIELDS];
  const float* grad[ERNN_TV_MAX_FIELDS];
  float* m[ERNN_TV_MAX_FIELDS];
  float* v[ERNN_TV_MAX_FIELDS];
  int64_t n;             // cells per field
  float beta1w;          // lerp weight 1 - beta1
  float beta2;
  float beta2w;          // addcmul value 1 - beta2
  float step_size;       // -lr / (1 - beta1^t)
  float bc2_sqrt;        // sqrt(1 - beta2^t)
  float eps;
  float weight_decay;
};

// C = clamp(exp(theta)) (or clamp(theta)); the clamp's backward passes the
// gradient where lo <= x <= hi, and exp's multiplies it by its result.
__device__ __forceinline__ float field_value(float th, int log_params, float lo, float hi, float* chain) {
  const float x = log_params ? expf(th) : th;
  const bool in = x >= lo && x <= hi;
  *chain = in ? (log_params ? x : 1.0f) : 0.0f;
  return fminf(fmaxf(x, lo), hi);
}

__device__ __forceinline__ float sgn(float d) { return d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f); }

__device__ __forceinline__ void load_cells(const float* p, float (&c)[4]) {
  const float4 t = *reinterpret_cast<const float4*>(p);
  c[0] = t.x; c[1] = t.y; c[2] = t.z; c[3] = t.w;
}

// Block sum of the threads' TV terms into partial[f][lb], fixed order:
// wavefront butterflies, then the four wave sums.
__device__ __forceinline__ void block_partial(double tv, double* partial, int f, int blocks_per_field, int lb) {
  for (int o = 32; o > 0; o >>= 1) tv += __shfl_xor(tv, o, kWave);
  __shared__ double ws[kTvThreads / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = tv;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kTvThreads / kWave; ++w) s += ws[w];
    partial[(int64_t)f * blocks_per_field + lb] = s;
  }
}

// One-cell variant (any alignment, any ny): thread = one cell.  Block b runs
// on XCD b % 8 (round-robin dispatch); it takes logical block
// (b % 8) x blocks_per_xcd + b / 8, so each XCD sweeps one contiguous band
// of rows and the rows above and below a block are in its own L2.
__global__ void __launch_bounds__(kTvThreads) tv_grad_kernel1(TvArgs a) {
  const int lb = (int)(blockIdx.x & 7) * a.blocks_per_xcd + (int)(blockIdx.x >> 3);
  if (lb >= a.blocks_per_field) return;   // whole block: the barrier below is not reached by anyone
  const int f = blockIdx.y;
  const float* __restrict__ th = a.theta[f];
  float* __restrict__ g = a.grad[f];
  const float lo = a.lo[f], hi = a.hi[f];
  const int lp = a.log_params;
  const int64_t k = (int64_t)lb * kTvThreads + threadIdx.x;
  double tv = 0.0;
  if (k < (int64_t)a.nx * a.ny) {
    const int i = (int)(k / a.ny), j = (int)(k % a.ny);
    float chain, dummy;
    const float c = field_value(th[k], lp, lo, hi, &chain);
    float s = 0.f;
    if (i > 0) s += sgn(c - field_value(th[k - a.ny], lp, lo, hi, &dummy));
    if (i + 1 < a.nx) {
      const float d = field_value(th[k + a.ny], lp, lo, hi, &dummy) - c;
      s -= sgn(d);
      tv += fabs((double)d);
    }
    if (j > 0) s += sgn(c - field_value(th[k - 1], lp, lo, hi, &dummy));
    if (j + 1 < a.ny) {
      const float d = field_value(th[k + 1], lp, lo, hi, &dummy) - c;
      s -= sgn(d);
      tv += fabs((double)d);
    }
    g[k] = g[k] + (a.lambda * s) * chain;
  }
  block_partial(tv, a.partial, f, a.blocks_per_field, lb);
}

// Strip variant (16-byte aligned fields, ny % 4 == 0): a thread sweeps a
// strip of kTvStrip rows x 4 columns top to bottom, so each row's fields are
// formed once (exp of the cell, of its two side neighbours, reused as the
// next row's upper neighbours): 1.75 exps per cell instead of 3.5 when every
// neighbour is formed on its own, which made the one-row version VALU-bound
// (90 % of the issue peak at 6 x 2048^2).  Same arithmetic per cell as above.
constexpr int kTvStrip = 8;

__global__ void __launch_bounds__(kTvThreads) tv_grad_kernel4(TvArgs a) {
  const int lb = (int)(blockIdx.x & 7) * a.blocks_per_xcd + (int)(blockIdx.x >> 3);
  if (lb >= a.blocks_per_field) return;
  const int f = blockIdx.y;
  const float* __restrict__ th = a.theta[f];
  float* __restrict__ g = a.grad[f];
  const float lo = a.lo[f], hi = a.hi[f];
  const int lp = a.log_params;
  const int groups = a.ny / 4;
  const int strips = (a.nx + kTvStrip - 1) / kTvStrip;
  const int64_t t = (int64_t)lb * kTvThreads + threadIdx.x;
  double tv = 0.0;
  if (t < (int64_t)strips * groups) {
    const int i0 = (int)(t / groups) * kTvStrip, j0 = (int)(t % groups) * 4;
    const int i1 = min(i0 + kTvStrip, a.nx);
    float up[4], cur[4], raw[4], dummy;
    if (i0 > 0) {
      load_cells(th + (int64_t)(i0 - 1) * a.ny + j0, raw);
#pragma unroll
      for (int e = 0; e < 4; ++e) up[e] = field_value(raw[e], lp, lo, hi, &dummy);
    }
    float chain[4];
    load_cells(th + (int64_t)i0 * a.ny + j0, raw);
#pragma unroll
    for (int e = 0; e < 4; ++e) cur[e] = field_value(raw[e], lp, lo, hi, &chain[e]);
#pragma unroll
    for (int r = 0; r < kTvStrip; ++r) {   // unrolled: the rows' loads are independent and issue early
      const int i = i0 + r;
      if (i >= i1) break;
      const int64_t k = (int64_t)i * a.ny + j0;
      float dn[4], dchain[4];
      if (i + 1 < a.nx) {
        load_cells(th + k + a.ny, raw);
#pragma unroll
        for (int e = 0; e < 4; ++e) dn[e] = field_value(raw[e], lp, lo, hi, &dchain[e]);
      }
      const float left = j0 > 0 ? field_value(th[k - 1], lp, lo, hi, &dummy) : 0.f;
      const float right = j0 + 4 < a.ny ? field_value(th[k + 4], lp, lo, hi, &dummy) : 0.f;
      float gv[4];
      load_cells(g + k, gv);
      float cv[6] = {left, cur[0], cur[1], cur[2], cur[3], right};   // cv[m]: column j0 - 1 + m
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float c = cv[e + 1];
        float s = 0.f;
        if (i > 0) s += sgn(c - up[e]);
        if (i + 1 < a.nx) {
          const float d = dn[e] - c;
          s -= sgn(d);
          tv += fabs((double)d);
        }
        if (j0 + e > 0) s += sgn(c - cv[e]);
        if (j0 + e + 1 < a.ny) {
          const float d = cv[e + 2] - c;
          s -= sgn(d);
          tv += fabs((double)d);
        }
        gv[e] = gv[e] + (a.lambda * s) * chain[e];
      }
      *reinterpret_cast<float4*>(g + k) = make_float4(gv[0], gv[1], gv[2], gv[3]);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        up[e] = cur[e];
        cur[e] = dn[e];
        chain[e] = dchain[e];
      }
    }
  }
  block_partial(tv, a.partial, f, a.blocks_per_field, lb);
}

// One block per field: TV = sum of the partials in block order.
__global__ void __launch_bounds__(kTvThreads) tv_finalize_kernel(const double* partial, int blocks_per_field,
                                                                  double* out) {
  const int f = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < blocks_per_field; b += kTvThreads) s += partial[(int64_t)f * blocks_per_field + b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
  __shared__ double ws[kTvThreads / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kTvThreads / kWave; ++w) t += ws[w];
    out[f] = t;
  }
}

// torch.optim.Adam (foreach form, torch/optim/adam.py _multi_tensor_adam) for one element:
//   g = g + wd theta                       (_foreach_add, alpha = weight_decay)
//   m = lerp(m, g, w), w = 1 - b1          (w < 0.5: m + w (g - m); else g - (g - m)(1 - w))
//   v = v b2 + (1 - b2) (g g)              (_foreach_mul_, _foreach_addcmul_)
//   theta = theta + step_size (m / (sqrt(v) / sqrt(1 - b2^t) + eps))   (_foreach_addcdiv_)
__device__ __forceinline__ void adam_cell(float& th, float g, float& m, float& v, const AdamArgs& a) {
  const float p = th;
  if (a.weight_decay != 0.f) g = g + a.weight_decay * p;
  const float mk = a.beta1w < 0.5f ? m + a.beta1w * (g - m) : g - (g - m) * (1.0f - a.beta1w);
  const float vk = v * a.beta2 + a.beta2w * (g * g);
  const float denom = sqrtf(vk) / a.bc2_sqrt + a.eps;
  m = mk;
  v = vk;
  th = p + a.step_size * (mk / denom);
}

// V = 4: 16-byte accesses over the first n - n % 4 cells (fields 16-byte
// aligned), block 0 takes the tail; V = 1: every cell.
template <int V>
__global__ void __launch_bounds__(kTvThreads) adam_kernel(AdamArgs a) {
  const int f = blockIdx.y;
  float* __restrict__ th = a.theta[f];
  const float* __restrict__ gr = a.grad[f];
  float* __restrict__ m = a.m[f];
  float* __restrict__ v = a.v[f];
  const int64_t groups = a.n / V;
  for (int64_t q = (int64_t)blockIdx.x * kTvThreads + threadIdx.x; q < groups; q += (int64_t)gridDim.x * kTvThreads) {
    if constexpr (V == 4) {
      float t4[4], g4[4], m4[4], v4[4];
      load_cells(th + 4 * q, t4);
      load_cells(gr + 4 * q, g4);
      load_cells(m + 4 * q, m4);
      load_cells(v + 4 * q, v4);
#pragma unroll
      for (int e = 0; e < 4; ++e) adam_cell(t4[e], g4[e], m4[e], v4[e], a);
      *reinterpret_cast<float4*>(th + 4 * q) = make_float4(t4[0], t4[1], t4[2], t4[3]);
      *reinterpret_cast<float4*>(m + 4 * q) = make_float4(m4[0], m4[1], m4[2], m4[3]);
      *reinterpret_cast<float4*>(v + 4 * q) = make_float4(v4[0], v4[1], v4[2], v4[3]);
    } else {
      adam_cell(th[q], gr[q], m[q], v[q], a);
    }
  }
  if (V > 1 && blockIdx.x == 0) {
    const int64_t k = groups * V + threadIdx.x;
    if (k < a.n) adam_cell(th[k], gr[k], m[k], v[k], a);
  }
}

}  // namespace ernn

using namespace ernn;

// Partial sums of the allocation-free entry point: one array per GPU, held
// until the process ends; its device address comes from hipGetSymbolAddress
// (the name itself is a host-side stand-in).
constexpr int64_t kTvScratchDoubles = int64_t(1) << 16;   // 512 KiB
static __device__ __align__(256) double scratch[kTvScratchDoubles];

namespace {

int64_t tv_blocks(int nx, int ny) { return ((int64_t)nx * ny + kTvThreads - 1) / kTvThreads; }

#define TV_HIP(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess)                                                                      \
      return set_last_error(ERNN_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

int check_desc(const ernn_tv_adam_desc_t* d) {
  if (!d) return set_last_error(ERNN_E_INVALID, "null descriptor");
  if (d->nx < 1 || d->ny < 1) return set_last_error(ERNN_E_INVALID, "nx and ny must be >= 1");
  if (d->n_fields < 1 || d->n_fields > ERNN_TV_MAX_FIELDS)
    return set_last_error(ERNN_E_INVALID, "n_fields must be in 1 .. " + std::to_string(ERNN_TV_MAX_FIELDS));
  if ((int64_t)d->nx * d->ny >= (int64_t(1) << 31)) return set_last_error(ERNN_E_CAPACITY, "field too large");
  if (d->step < 1) return set_last_error(ERNN_E_INVALID, "step (the Adam step count) must be >= 1");
  if (!(d->lr >= 0.0) || !(d->eps > 0.0) || !(d->beta1 >= 0.0 && d->beta1 < 1.0) ||
      !(d->beta2 >= 0.0 && d->beta2 < 1.0) || !std::isfinite(d->lambda_tv) || !std::isfinite(d->weight_decay))
    return set_last_error(ERNN_E_INVALID, "bad Adam hyper-parameters (lr >= 0, eps > 0, 0 <= beta < 1)");
  for (int f = 0; f < d->n_fields; ++f)
    if (!(d->lo[f] <= d->hi[f])) return set_last_error(ERNN_E_INVALID, "bounds of field " + std::to_string(f));
  return ERNN_OK;
}

int check_fields(const ernn_tv_adam_desc_t* d, const ernn_tv_adam_call_t* c) {
  if (!c) return set_last_error(ERNN_E_INVALID, "null call");
  for (int f = 0; f < d->n_fields; ++f)
    if (!c->theta[f] || !c->grad[f] || !c->m[f] || !c->v[f])
      return set_last_error(ERNN_E_INVALID, "theta, grad, m and v of field " + std::to_string(f) + " are required");
  return ERNN_OK;
}

// The TV pass, the optional TV finalize and the Adam pass on `st`; `partial`
// is device memory of at least tv_blocks(nx, ny) x n_fields doubles that every
// launch writes before it reads (no clearing needed).
int enqueue_tv_adam(const ernn_tv_adam_desc_t* d, const ernn_tv_adam_call_t* c, double* partial, hipStream_t st) {
  const int64_t n = (int64_t)d->nx * d->ny;
  // 16-byte accesses when every field pointer allows them
  bool aligned = true;
  for (int f = 0; f < d->n_fields; ++f)
    for (const void* q : {(const void*)c->theta[f], (const void*)c->grad[f], (const void*)c->m[f], (const void*)c->v[f]})
      aligned = aligned && (reinterpret_cast<uintptr_t>(q) & 15) == 0;
  const bool tv4 = aligned && d->ny % 4 == 0;
  const int64_t strips = (d->nx + kTvStrip - 1) / kTvStrip;
  const int64_t nb = tv4 ? (strips * (d->ny / 4) + kTvThreads - 1) / kTvThreads : tv_blocks(d->nx, d->ny);
  TvArgs t{};
  for (int f = 0; f < d->n_fields; ++f) {
    t.theta[f] = c->theta[f];
    t.grad[f] = c->grad[f];
    t.lo[f] = d->lo[f];
    t.hi[f] = d->hi[f];
  }
  t.partial = partial;
  t.nx = d->nx;
  t.ny = d->ny;
  t.log_params = d->log_params ? 1 : 0;
  t.blocks_per_field = (int)nb;
  t.blocks_per_xcd = (int)((nb + 7) / 8);
  t.lambda = (float)d->lambda_tv;   // the mul's backward scales by the FP32 scalar
  const unsigned grid = 8u * (unsigned)t.blocks_per_xcd;
  if (tv4)
    hipLaunchKernelGGL(tv_grad_kernel4, dim3(grid, (unsigned)d->n_fields), dim3(kTvThreads), 0, st, t);
  else
    hipLaunchKernelGGL(tv_grad_kernel1, dim3(grid, (unsigned)d->n_fields), dim3(kTvThreads), 0, st, t);
  TV_HIP(hipGetLastError());
  if (c->tv) {
    hipLaunchKernelGGL(tv_finalize_kernel, dim3((unsigned)d->n_fields), dim3(kTvThreads), 0, st,
                       static_cast<const double*>(partial), (int)nb, c->tv);
    TV_HIP(hipGetLastError());
  }
  // Adam's scalars as torch forms them: Python floats (double), cast to FP32 where a kernel uses them
  const double bc1 = 1.0 - std::pow(d->beta1, (double)d->step);
  const double bc2 = 1.0 - std::pow(d->beta2, (double)d->step);
  AdamArgs a{};
  for (int f = 0; f < d->n_fields; ++f) {
    a.theta[f] = c->theta[f];
    a.grad[f] = c->grad[f];
    a.m[f] = c->m[f];
    a.v[f] = c->v[f];
  }
  a.n = n;
  a.beta1w = (float)(1.0 - d->beta1);
  a.beta2 = (float)d->beta2;
  a.beta2w = (float)(1.0 - d->beta2);
  a.step_size = (float)((d->lr / bc1) * -1.0);
  a.bc2_sqrt = (float)std::pow(bc2, 0.5);
  a.eps = (float)d->eps;
  a.weight_decay = (float)d->weight_decay;
  const int V = aligned ? 4 : 1;
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n / V + kTvThreads - 1) / kTvThreads, 16384));
  if (V == 4)
    hipLaunchKernelGGL(adam_kernel<4>, dim3((unsigned)blocks, (unsigned)d->n_fields), dim3(kTvThreads), 0, st, a);
  else
    hipLaunchKernelGGL(adam_kernel<1>, dim3((unsigned)blocks, (unsigned)d->n_fields), dim3(kTvThreads), 0, st, a);
  TV_HIP(hipGetLastError());
  return ERNN_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(double** out) {
  int dev = -1;
  TV_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_last_error(ERNN_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    TV_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_last_error(ERNN_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return ERNN_OK;
}

}  // namespace

extern "C" {

int ernn_tv_adam_workspace_bytes(const ernn_tv_adam_desc_t* desc, int64_t* bytes) {
  if (!bytes) return set_last_error(ERNN_E_INVALID, "null argument");
  const int r = check_desc(desc);
  if (r != ERNN_OK) return r;
  *bytes = tv_blocks(desc->nx, desc->ny) * desc->n_fields * (int64_t)sizeof(double);
  return ERNN_OK;
}

int ernn_tv_adam_step(const ernn_tv_adam_desc_t* d, const ernn_tv_adam_call_t* c, void* stream) {
  int r = check_desc(d);
  if (r != ERNN_OK) return r;
  if ((r = check_fields(d, c)) != ERNN_OK) return r;
  if (!c->workspace) return set_last_error(ERNN_E_INVALID, "null workspace");
  if ((reinterpret_cast<uintptr_t>(c->workspace) & 7) != 0)
    return set_last_error(ERNN_E_INVALID, "workspace must be 8-byte aligned");
  return enqueue_tv_adam(d, c, static_cast<double*>(c->workspace), static_cast<hipStream_t>(stream));
}

int ernn_tv_adam_step_static(const ernn_tv_adam_desc_t* d, const ernn_tv_adam_call_t* c, void* stream) {
  int r = check_desc(d);
  if (r != ERNN_OK) return r;
  if ((r = check_fields(d, c)) != ERNN_OK) return r;
  if (tv_blocks(d->nx, d->ny) * d->n_fields > kTvScratchDoubles)
    return set_last_error(ERNN_E_CAPACITY, "ernn_tv_adam_step_static: n_fields x ceil(nx * ny / " + std::to_string(kTvThreads) +
                                               ") must be <= " + std::to_string(kTvScratchDoubles) +
                                               " (its static scratch); use ernn_tv_adam_step");
  double* partial = nullptr;
  if ((r = scratch_of_current_device(&partial)) != ERNN_OK) return r;
  return enqueue_tv_adam(d, c, partial, static_cast<hipStream_t>(stream));
}


}  // extern "C"
