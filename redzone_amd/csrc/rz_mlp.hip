// rz_mlp.hip -- learned dispatch: list features and a small FP32 MLP with
// hand-written inference and training kernels.
//
// The MLP maps RZ_MLP_FEATURES features of a (list, removal list) pair to the
// predicted log2 time (us) of each of RZ_MLP_OUTPUTS methods; the dispatcher
// runs the method with the smallest prediction.  Layers: F -> 32 -> 32 -> 32
// -> C, ReLU between them, parameters row-major W[out][in] followed by b[out]
// for each layer, in order.  Training minimises the masked mean squared error
// on measured log times with Adam.  Per-sample gradients are written out and
// summed in a fixed order, so a gradient is the same bits on every run.
#include <cstring>
#include <mutex>
#include <vector>

#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

constexpr int F = RZ_MLP_FEATURES, H = RZ_MLP_HIDDEN, C = RZ_MLP_OUTPUTS;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + C * H, NP = OB4 + C;
static_assert(NP == RZ_MLP_PARAMS, "parameter count");

// ---------------------------------------------------------------------------
// features: one wave per list, a 256-entry strided sample of the removal list
// ---------------------------------------------------------------------------
// The features of list d, by one 64-lane workgroup (s_x: 256 floats of LDS);
// lane 0 writes them to o[0..F).
__device__ __forceinline__ void list_features(const rz_list_desc& d, float* s_x, float* o) {
  const int lane = threadIdx.x;
  const double n = (double)(d.n > 0 ? d.n : 1);
  const int64_t k = d.k;
  const int64_t sz = d.n - d.k;
  const int m = k < 256 ? (int)k : 256;  // sample size
  float sum = 0.f, sum2 = 0.f, tail = 0.f;
  for (int j = lane; j < m; j += 64) {
    const int64_t i = (int64_t)j * k / m;
    const int64_t r = d.removals[i];
    const float x = (float)((double)r / n);
    s_x[j] = x;
    sum += x;
    sum2 += x * x;
    tail += (r >= sz) ? 1.f : 0.f;
  }
  // adjacent pairs (R[i], R[i+1]) at 128 strided positions: runs of the removal list
  float consec = 0.f, near = 0.f;
  const int mp = k < 2 ? 0 : (k - 1 < 128 ? (int)(k - 1) : 128);
  for (int j = lane; j < mp; j += 64) {
    const int64_t i = (int64_t)j * (k - 1) / mp;
    const int64_t a = d.removals[i], b = d.removals[i + 1];
    consec += (b == a + 1) ? 1.f : 0.f;
    near += (b - a < 64 && a - b < 64) ? 1.f : 0.f;
  }
  __syncthreads();
  float up = 0.f;
  for (int j = lane; j + 1 < m; j += 64) up += s_x[j + 1] > s_x[j] ? 1.f : 0.f;
  sum = wave_sum(sum);
  sum2 = wave_sum(sum2);
  tail = wave_sum(tail);
  consec = wave_sum(consec);
  near = wave_sum(near);
  up = wave_sum(up);
  if (lane == 0) {
    const float fm = m > 0 ? (float)m : 1.f;
    const float mean = sum / fm;
    const float var = fmaxf(sum2 / fm - mean * mean, 0.f);
    const double frac = (double)k / n;
    o[0] = (float)(log2(n) / 32.0);
    o[1] = k > 0 ? (float)(log2((double)k) / 32.0) : 0.f;
    o[2] = (float)frac;
    o[3] = k > 0 ? (float)fmax(log2(frac) / 20.0, -1.0) : -1.f;
    o[4] = (float)d.elem_bytes / 16.f;
    o[5] = m > 0 ? tail / fm : 0.f;
    o[6] = mean;
    o[7] = sqrtf(var);
    o[8] = mp > 0 ? consec / (float)mp : 0.f;
    o[9] = mp > 0 ? near / (float)mp : 0.f;
    o[10] = m > 1 ? up / (float)(m - 1) : 0.f;
    o[11] = 1.f;
  }
}

__global__ __launch_bounds__(64) void rz_features_kernel(const rz_list_desc* __restrict__ lists, int32_t nlists,
                                                          float* __restrict__ out) {
  const int li = blockIdx.x;
  if (li >= nlists) return;
  __shared__ float s_x[256];
  list_features(lists[li], s_x, out + (size_t)li * F);
}

// ---------------------------------------------------------------------------
// MLP
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dense_relu(const float* __restrict__ W, const float* __restrict__ b,
                                           const float* in, int nin, float* out_pre, float* out, int nout) {
  for (int o = 0; o < nout; ++o) {
    float acc = b[o];
    for (int i = 0; i < nin; ++i) acc = fmaf(W[o * nin + i], in[i], acc);
    out_pre[o] = acc;
    out[o] = acc > 0.f ? acc : 0.f;
  }
}

struct Acts {
  float x[F];
  float a1[H], h1[H], a2[H], h2[H], a3[H], h3[H];
  float y[C];
};

__device__ __forceinline__ void forward(const float* __restrict__ P, const float* __restrict__ xin, Acts& t) {
  for (int f = 0; f < F; ++f) t.x[f] = xin[f];
  dense_relu(P + OW1, P + OB1, t.x, F, t.a1, t.h1, H);
  dense_relu(P + OW2, P + OB2, t.h1, H, t.a2, t.h2, H);
  dense_relu(P + OW3, P + OB3, t.h2, H, t.a3, t.h3, H);
  for (int c = 0; c < C; ++c) {
    float acc = P[OB4 + c];
    for (int i = 0; i < H; ++i) acc = fmaf(P[OW4 + c * H + i], t.h3[i], acc);
    t.y[c] = acc;
  }
}

// One layer by one wave: lane o's pre-activation b[o] + sum_i W[o][i] in_i,
// in_i broadcast from lane i, summed in the order dense_relu uses (so the
// result is the same bits as forward()).
__device__ __forceinline__ float layer_wave(const float* __restrict__ W, const float* __restrict__ b, float in,
                                            int nin, int o) {
  float acc = b[o];
  for (int i = 0; i < nin; ++i) acc = fmaf(W[o * nin + i], __shfl(in, i), acc);
  return acc;
}

static_assert((H & (H - 1)) == 0 && (C & (C - 1)) == 0 && H <= 64 && F <= 64, "one unit per lane");

// The network on one row by one 64-lane wave: lane f < F holds feature f on
// entry; y_out (may be null) gets the C outputs.  Returns the argmin (every
// lane), ties to the lower index as in rz_mlp_forward_kernel's contract.
__device__ __forceinline__ int forward_wave(const float* __restrict__ P, float x, float* __restrict__ y_out) {
  const int lane = threadIdx.x & 63, o = lane & (H - 1), c = lane & (C - 1);
  float a = layer_wave(P + OW1, P + OB1, x, F, o);
  float h = a > 0.f ? a : 0.f;
  a = layer_wave(P + OW2, P + OB2, h, H, o);
  h = a > 0.f ? a : 0.f;
  a = layer_wave(P + OW3, P + OB3, h, H, o);
  h = a > 0.f ? a : 0.f;
  const float y = layer_wave(P + OW4, P + OB4, h, H, c);
  if (y_out && lane < C) y_out[lane] = y;
  int best = 0;
  float yb = __shfl(y, 0);
  for (int cc = 1; cc < C; ++cc) {
    const float v = __shfl(y, cc);
    if (v < yb) {
      yb = v;
      best = cc;
    }
  }
  return best;
}

// One 64-lane workgroup per row.
__global__ __launch_bounds__(64) void rz_mlp_forward_kernel(const float* __restrict__ params,
                                                             const float* __restrict__ X, int32_t B,
                                                             float* __restrict__ Y, int32_t* __restrict__ choice) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const int lane = threadIdx.x;
  const float x = lane < F ? X[(size_t)b * F + lane] : 0.f;
  const int best = forward_wave(params, x, Y ? Y + (size_t)b * C : nullptr);
  if (choice && lane == 0) choice[b] = best;
}

// rz_dispatch_choose for up to kChooseArgMax lists: the descriptors come as a
// kernel argument (no upload), one 64-lane workgroup per list computes its
// features and the forward pass, and lane 0 stores the choice straight into
// the caller-visible pinned staging with a system-scope store.
constexpr int kChooseArgMax = 48;
struct ChooseArgs {
  rz_list_desc d[kChooseArgMax];
};

__global__ __launch_bounds__(64) void rz_choose_kernel(ChooseArgs a, int32_t nlists, const float* __restrict__ params,
                                                        int32_t* __restrict__ h_choice) {
  const int li = blockIdx.x;
  if (li >= nlists) return;
  __shared__ float s_x[256];
  __shared__ float s_f[F];
  list_features(a.d[li], s_x, s_f);
  __syncthreads();
  const float x = threadIdx.x < F ? s_f[threadIdx.x] : 0.f;
  const int best = forward_wave(params, x, nullptr);
  if (threadIdx.x == 0) __hip_atomic_store(&h_choice[li], best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Per-sample gradient of the masked MSE loss sum_b sum_c m_bc (y_bc - t_bc)^2 / sum(m)
// into G[b][NP]; per-sample loss terms into Lb[b].
__global__ __launch_bounds__(64) void rz_mlp_grad_kernel(const float* __restrict__ params, const float* __restrict__ X,
                                                          const float* __restrict__ T, const float* __restrict__ M,
                                                          int32_t B, float inv_msum, float* __restrict__ G,
                                                          float* __restrict__ Lb) {
  __shared__ float P[NP];
  for (int i = threadIdx.x; i < NP; i += blockDim.x) P[i] = params[i];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  Acts t;
  forward(P, X + (size_t)b * F, t);
  float* g = G + (size_t)b * NP;
  float dy[C];
  float loss = 0.f;
  for (int c = 0; c < C; ++c) {
    const float m = M[(size_t)b * C + c];
    const float e = t.y[c] - T[(size_t)b * C + c];
    loss += m * e * e;
    dy[c] = 2.f * m * e * inv_msum;
  }
  Lb[b] = loss * inv_msum;
  float d3[H], d2[H], d1[H];
  // layer 4
  for (int c = 0; c < C; ++c) {
    g[OB4 + c] = dy[c];
    for (int i = 0; i < H; ++i) g[OW4 + c * H + i] = dy[c] * t.h3[i];
  }
  for (int i = 0; i < H; ++i) {
    float acc = 0.f;
    for (int c = 0; c < C; ++c) acc = fmaf(P[OW4 + c * H + i], dy[c], acc);
    d3[i] = t.a3[i] > 0.f ? acc : 0.f;
  }
  // layer 3
  for (int o = 0; o < H; ++o) {
    g[OB3 + o] = d3[o];
    for (int i = 0; i < H; ++i) g[OW3 + o * H + i] = d3[o] * t.h2[i];
  }
  for (int i = 0; i < H; ++i) {
    float acc = 0.f;
    for (int o = 0; o < H; ++o) acc = fmaf(P[OW3 + o * H + i], d3[o], acc);
    d2[i] = t.a2[i] > 0.f ? acc : 0.f;
  }
  // layer 2
  for (int o = 0; o < H; ++o) {
    g[OB2 + o] = d2[o];
    for (int i = 0; i < H; ++i) g[OW2 + o * H + i] = d2[o] * t.h1[i];
  }
  for (int i = 0; i < H; ++i) {
    float acc = 0.f;
    for (int o = 0; o < H; ++o) acc = fmaf(P[OW2 + o * H + i], d2[o], acc);
    d1[i] = t.a1[i] > 0.f ? acc : 0.f;
  }
  // layer 1
  for (int o = 0; o < H; ++o) {
    g[OB1 + o] = d1[o];
    for (int f = 0; f < F; ++f) g[OW1 + o * F + f] = d1[o] * t.x[f];
  }
}

// grad[p] = sum_b G[b][p] in ascending b (fixed order); loss = sum_b Lb[b].
__global__ __launch_bounds__(256) void rz_mlp_reduce_kernel(const float* __restrict__ G, const float* __restrict__ Lb,
                                                             int32_t B, float* __restrict__ grad,
                                                             float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NP) {
    float acc = 0.f;
    for (int b = 0; b < B; ++b) acc += G[(size_t)b * NP + p];
    grad[p] = acc;
  } else if (p == NP && loss) {
    float acc = 0.f;
    for (int b = 0; b < B; ++b) acc += Lb[b];
    *loss = acc;
  }
}

__global__ __launch_bounds__(256) void rz_mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ g,
                                                           float* __restrict__ m, float* __restrict__ v, float lr,
                                                           float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.f - b1) * gi;
  const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace rz

using namespace rz;

extern "C" {

rz_status rz_features(rz_ctx* ctx, const rz_list_desc* d_lists, int32_t nlists, float* d_features, void* stream) {
  if (!ctx || nlists < 0 || (nlists > 0 && (!d_lists || !d_features))) return RZ_EINVAL;
  if (nlists == 0) return RZ_OK;
  hipLaunchKernelGGL(rz_features_kernel, dim3((uint32_t)nlists), dim3(64), 0, (hipStream_t)stream, d_lists, nlists,
                     d_features);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status rz_mlp_forward(rz_ctx* ctx, const float* d_params, const float* d_x, int32_t batch, float* d_y,
                         int32_t* d_choice, void* stream) {
  if (!ctx || !d_params || !d_x || batch < 0) return RZ_EINVAL;
  if (batch == 0) return RZ_OK;
  hipLaunchKernelGGL(rz_mlp_forward_kernel, dim3((uint32_t)batch), dim3(64), 0, (hipStream_t)stream, d_params, d_x,
                     batch, d_y, d_choice);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status rz_mlp_grad(rz_ctx* ctx, const float* d_params, const float* d_x, const float* d_target,
                      const float* d_mask, int32_t batch, float mask_sum, float* d_grad, float* d_loss,
                      float* d_work, void* stream) {
  if (!ctx || !d_params || !d_x || !d_target || !d_mask || !d_grad || !d_work || batch < 1 || !(mask_sum > 0.f))
    return RZ_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  float* G = d_work;
  float* Lb = d_work + (size_t)batch * NP;
  hipLaunchKernelGGL(rz_mlp_grad_kernel, dim3((uint32_t)((batch + 63) / 64)), dim3(64), 0, s, d_params, d_x, d_target,
                     d_mask, batch, 1.f / mask_sum, G, Lb);
  hipLaunchKernelGGL(rz_mlp_reduce_kernel, dim3((NP + 1 + 255) / 256), dim3(256), 0, s, G, Lb, batch, d_grad, d_loss);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status rz_mlp_adam(rz_ctx* ctx, float* d_params, const float* d_grad, float* d_m, float* d_v, int32_t step,
                      float lr, void* stream) {
  if (!ctx || !d_params || !d_grad || !d_m || !d_v || step < 1) return RZ_EINVAL;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(rz_mlp_adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_params, d_grad,
                     d_m, d_v, lr, b1, b2, eps, c1, c2);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status rz_dispatch_method_config(int32_t method, uint32_t elem_bytes, rz_tile_config* out) {
  if (!out || esz_index(elem_bytes) < 0) return RZ_EINVAL;
  if (method == RZ_METHOD_ORDERED) {
    *out = rz_tile_config{256, 4, 4096, 3, 0, 1};
    return RZ_OK;
  }
  if (method == RZ_METHOD_DENSE) {
    // 16-byte items stage half as many entries per step (LDS)
    *out = rz_tile_config{1024, elem_bytes == 16 ? 4 : 8, 1024, 5, 11, 2};
    return RZ_OK;
  }
  return RZ_EINVAL;
}

// dispatch buffer: [params NP floats | pad to 256 B | descriptor | features | choice]
static constexpr size_t kDispParams = ((size_t)NP * 4 + 255) & ~(size_t)255;
static constexpr size_t kDispDesc = kDispParams;
static constexpr size_t kDispFeat = kDispDesc + 256;
static constexpr size_t kDispChoice = kDispFeat + 256;
static constexpr size_t kDispBytes = kDispChoice + 256;

rz_status rz_dispatch_set_params(rz_ctx* ctx, const float* h_params) {
  if (!ctx) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  rz_status st = RZ_OK;
  if (!h_params) {
    ctx->dispatch_ready = false;
  } else if ((st = ensure(ctx, ctx->dispatch, kDispBytes, nullptr)) == RZ_OK) {
    if (hipMemcpy(ctx->dispatch.ptr, h_params, (size_t)NP * 4, hipMemcpyHostToDevice) != hipSuccess) st = RZ_EHIP;
    ctx->dispatch_ready = st == RZ_OK;
  }
  if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
  return st;
}

// The network's choice for one list (stream-ordered, then read back).
static rz_status auto_choose(rz_ctx* ctx, void* d_data, int64_t n, uint32_t esz, int32_t* d_removals, int64_t k,
                             hipStream_t s, int32_t* method) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!ctx->dispatch_ready) {
    *method = 2 * k < n ? RZ_METHOD_REDZONE : RZ_METHOD_PARTITION;
    return RZ_OK;
  }
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  rz_status st = RZ_OK;
  if (!ctx->auto_host && hipHostMalloc((void**)&ctx->auto_host, 256, hipHostMallocDefault) != hipSuccess) {
    ctx->auto_host = nullptr;
    st = RZ_ENOMEM;
  }
  if (st == RZ_OK) {
    char* base = (char*)ctx->dispatch.ptr;
    rz_list_desc* hd = (rz_list_desc*)ctx->auto_host;
    *hd = rz_list_desc{d_data, d_removals, n, k, esz, 0u};
    int32_t* hc = (int32_t*)(ctx->auto_host + 128);
    *hc = -1;
    if (hipMemcpyAsync(base + kDispDesc, hd, sizeof(rz_list_desc), hipMemcpyHostToDevice, s) != hipSuccess) st = RZ_EHIP;
    if (st == RZ_OK)
      st = rz_features(ctx, (const rz_list_desc*)(base + kDispDesc), 1, (float*)(base + kDispFeat), s);
    if (st == RZ_OK)
      st = rz_mlp_forward(ctx, (const float*)base, (const float*)(base + kDispFeat), 1, nullptr,
                          (int32_t*)(base + kDispChoice), s);
    if (st == RZ_OK && hipMemcpyAsync(hc, base + kDispChoice, 4, hipMemcpyDeviceToHost, s) != hipSuccess) st = RZ_EHIP;
    if (st == RZ_OK && hipStreamSynchronize(s) != hipSuccess) st = RZ_EHIP;
    if (st == RZ_OK) *method = (*hc >= 0 && *hc < RZ_MLP_OUTPUTS) ? *hc : (2 * k < n ? 0 : 3);
  }
  if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
  return st;
}

static int32_t rule_method(const rz_list_desc& d) {
  return 2 * d.k < d.n ? RZ_METHOD_REDZONE : RZ_METHOD_PARTITION;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

rz_status rz_dispatch_choose(rz_ctx* ctx, const float* d_params, const rz_list_desc* h_lists, int32_t nlists,
                             void* stream) {
  if (!ctx || nlists < 0 || (nlists > 0 && !h_lists)) return RZ_EINVAL;
  for (int32_t i = 0; i < nlists; ++i) {
    const rz_list_desc& d = h_lists[i];
    if (d.n < 0 || d.k < 0 || d.k > d.n || esz_index(d.elem_bytes) < 0 || (d.k > 0 && !d.removals))
      return RZ_EINVAL;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  hipStream_t s = (hipStream_t)stream;
  // an uncollected choice still copying back must leave the staging area first
  if (ctx->choose_n >= 0 && ctx->choose_dev) RZ_HIP(hipEventSynchronize(ctx->choose_ev));
  ctx->choose_n = -1;
  const size_t dbytes = round256((size_t)nlists * sizeof(rz_list_desc));
  const size_t fbytes = round256((size_t)nlists * RZ_MLP_FEATURES * 4);
  const size_t cbytes = round256((size_t)nlists * 4 + 4);
  if (dbytes + cbytes > ctx->choose_host_bytes) {
    if (ctx->choose_host) RZ_HIP(hipHostFree(ctx->choose_host));
    ctx->choose_host = ctx->choose_host_dev = nullptr;
    ctx->choose_host_bytes = 0;
    size_t cap = 4096;
    while (cap < dbytes + cbytes) cap *= 2;
    // mapped and coherent: the choice kernel stores into it directly
    if (hipHostMalloc((void**)&ctx->choose_host, cap, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
      ctx->choose_host = nullptr;
      return RZ_ENOMEM;
    }
    if (hipHostGetDevicePointer((void**)&ctx->choose_host_dev, ctx->choose_host, 0) != hipSuccess) {
      (void)hipHostFree(ctx->choose_host);
      ctx->choose_host = ctx->choose_host_dev = nullptr;
      return RZ_EHIP;
    }
    ctx->choose_host_bytes = cap;
  }
  if (nlists) memcpy(ctx->choose_host, h_lists, (size_t)nlists * sizeof(rz_list_desc));
  const float* params = d_params ? d_params : (ctx->dispatch_ready ? (const float*)ctx->dispatch.ptr : nullptr);
  if (!params || nlists == 0) {  // no network: the rule, on the host, at collection
    ctx->choose_dev = false;
    ctx->choose_n = nlists;
    return RZ_OK;
  }
  if (!ctx->choose_ev) RZ_HIP(hipEventCreateWithFlags(&ctx->choose_ev, hipEventDisableTiming));
  if (nlists <= kChooseArgMax) {  // one launch, no copies
    ChooseArgs a;
    memcpy(a.d, h_lists, (size_t)nlists * sizeof(rz_list_desc));
    hipLaunchKernelGGL(rz_choose_kernel, dim3((uint32_t)nlists), dim3(64), 0, s, a, nlists, params,
                       (int32_t*)(ctx->choose_host_dev + dbytes));
    RZ_HIP(hipGetLastError());
    RZ_HIP(hipEventRecord(ctx->choose_ev, s));
    ctx->choose_dev = true;
    ctx->choose_n = nlists;
    return RZ_OK;
  }
  rz_status st;
  if ((st = ensure(ctx, ctx->choose, dbytes + fbytes + cbytes, s)) != RZ_OK) return st;
  char* base = (char*)ctx->choose.ptr;
  RZ_HIP(hipMemcpyAsync(base, ctx->choose_host, (size_t)nlists * sizeof(rz_list_desc), hipMemcpyHostToDevice, s));
  if ((st = rz_features(ctx, (const rz_list_desc*)base, nlists, (float*)(base + dbytes), s)) != RZ_OK) return st;
  if ((st = rz_mlp_forward(ctx, params, (const float*)(base + dbytes), nlists, nullptr,
                           (int32_t*)(base + dbytes + fbytes), s)) != RZ_OK)
    return st;
  RZ_HIP(hipMemcpyAsync(ctx->choose_host + dbytes, base + dbytes + fbytes, (size_t)nlists * 4, hipMemcpyDeviceToHost,
                        s));
  RZ_HIP(hipEventRecord(ctx->choose_ev, s));
  ctx->choose_dev = true;
  ctx->choose_n = nlists;
  return RZ_OK;
}

rz_status rz_dispatch_collect(rz_ctx* ctx, int32_t* h_choice, int32_t nlists) {
  if (!ctx || nlists < 0 || (nlists > 0 && !h_choice)) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ctx->choose_n < 0 || ctx->choose_n != nlists) return RZ_EINVAL;
  DeviceGuard g(ctx->device);
  if (ctx->choose_dev) RZ_HIP(hipEventSynchronize(ctx->choose_ev));
  const rz_list_desc* hd = (const rz_list_desc*)ctx->choose_host;
  const int32_t* hc = (const int32_t*)(ctx->choose_host + round256((size_t)nlists * sizeof(rz_list_desc)));
  for (int32_t i = 0; i < nlists; ++i) {
    const rz_list_desc& d = hd[i];
    int32_t m = ctx->choose_dev ? hc[i] : rule_method(d);
    if (m < 0 || m >= RZ_MLP_OUTPUTS) m = rule_method(d);
    if (d.k == 0 || d.k == d.n) m = RZ_METHOD_REDZONE;  // nothing to choose: no items move
    h_choice[i] = m;
  }
  ctx->choose_n = -1;
  return RZ_OK;
}

rz_status rz_remove_auto_batched(rz_ctx* ctx, const float* d_params, const rz_list_desc* lists, int32_t nlists,
                                 int64_t batch_max_k, uint32_t flags, int32_t* methods, void* stream) {
  if (!ctx || nlists < 0 || (nlists > 0 && !lists) || nlists >= (1 << 28) || batch_max_k < 0) return RZ_EINVAL;
  std::vector<rz_list_desc> small, large;
  std::vector<int32_t> large_idx;
  small.reserve((size_t)nlists);
  for (int32_t i = 0; i < nlists; ++i) {
    rz_list_desc d = lists[i];
    d.method = 0;
    if (d.k < batch_max_k) {
      small.push_back(d);
    } else {
      large.push_back(d);
      large_idx.push_back(i);
    }
    if (methods) methods[i] = RZ_METHOD_REDZONE;
  }
  rz_status st;
  const int32_t nl = (int32_t)large.size();
  // the choice first: the device computes it while the batch below is planned
  if (nl && (st = rz_dispatch_choose(ctx, d_params, large.data(), nl, stream)) != RZ_OK) return st;
  if (!small.empty() && (st = rz_remove_batched(ctx, small.data(), (int32_t)small.size(), flags, stream)) != RZ_OK)
    return st;
  if (!nl) return RZ_OK;
  std::vector<int32_t> m((size_t)nl);
  if ((st = rz_dispatch_collect(ctx, m.data(), nl)) != RZ_OK) return st;
  for (int32_t j = 0; j < nl; ++j) {  // queued behind the batch on the stream
    const rz_list_desc& d = large[(size_t)j];
    if (m[j] == RZ_METHOD_PARTITION) {
      st = rz_partition_on(ctx, d.data, d.n, d.elem_bytes, d.removals, d.k, flags, stream);
    } else if (m[j] == RZ_METHOD_REDZONE) {
      st = rz_remove(ctx, d.data, d.n, d.elem_bytes, d.removals, d.k, flags, stream);
    } else {
      rz_tile_config cfg;
      if ((st = rz_dispatch_method_config(m[j], d.elem_bytes, &cfg)) == RZ_OK)
        st = rz_remove_config(ctx, d.data, d.n, d.elem_bytes, d.removals, d.k, flags, &cfg, stream);
    }
    if (st != RZ_OK) return st;
    if (methods) methods[large_idx[(size_t)j]] = m[j];
  }
  return RZ_OK;
}

rz_status rz_remove_auto(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals, int64_t k,
                         uint32_t flags, int32_t* method, void* stream) {
  if (!ctx || n < 0 || k < 0 || k > n || esz_index(elem_bytes) < 0) return RZ_EINVAL;
  if (k > 0 && (!d_data || !d_removals)) return RZ_EINVAL;
  int32_t m = RZ_METHOD_REDZONE;
  rz_status st = RZ_OK;
  if (k > 0 && k < n) st = auto_choose(ctx, d_data, n, elem_bytes, d_removals, k, (hipStream_t)stream, &m);
  if (st != RZ_OK) return st;
  if (method) *method = m;
  if (m == RZ_METHOD_PARTITION) return rz_partition_on(ctx, d_data, n, elem_bytes, d_removals, k, flags, stream);
  if (m == RZ_METHOD_REDZONE) return rz_remove(ctx, d_data, n, elem_bytes, d_removals, k, flags, stream);
  rz_tile_config cfg;
  if ((st = rz_dispatch_method_config(m, elem_bytes, &cfg)) != RZ_OK) return st;
  return rz_remove_config(ctx, d_data, n, elem_bytes, d_removals, k, flags, &cfg, stream);
}

}  // extern "C"
