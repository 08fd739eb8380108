// The learned scheduler's device side: the batch features it reads and its network, a
// 12 -> 64 -> 64 -> 5 multilayer perceptron (ReLU hidden layers) with FP32 master weights,
// trained here (forward, backward and Adam in FP32) and evaluated in BF16 (weights and
// activations rounded to bfloat16, products accumulated in FP32).
//
// The network predicts log2 of the time of each of five schedules of an update on a batch
// (repair with the default schedule, repair with a 4x wider or 4x narrower near-far width,
// repair with fine-grained tiles, recompute from scratch); the scheduler runs the one predicted
// fastest. Every schedule yields the same tree. The kernels are small and exact in the sense that
// matters (the same arithmetic as the CPU oracle of the network, up to FP32 summation order).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace dyng_hip {
namespace {

constexpr int kIn = DYNG_HIP_MLP_INPUTS;
constexpr int kHidden = DYNG_HIP_MLP_HIDDEN;
constexpr int kOut = DYNG_HIP_MLP_OUTPUTS;
// parameter layout (floats): W1[kHidden][kIn], b1[kHidden], W2[kHidden][kHidden], b2[kHidden],
// W3[kOut][kHidden], b3[kOut], then the input normalization x_mean[kIn], x_scale[kIn] (not trained)
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHidden * kIn;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kOut * kHidden;
constexpr int kTrained = kB3 + kOut;
constexpr int kMean = kTrained;
constexpr int kScale = kMean + kIn;
constexpr int kAll = kScale + kIn;
static_assert(kTrained == DYNG_HIP_MLP_TRAINED, "trained parameter count");
static_assert(kAll == DYNG_HIP_MLP_PARAMS, "parameter count");

// ---- bfloat16: the upper 16 bits of an IEEE binary32, rounded to nearest, ties to even ----
DYNG_DEV uint16_t to_bf16(float f) {
  uint32_t bits = __float_as_uint(f);
  if ((bits & 0x7f800000u) == 0x7f800000u && (bits & 0x007fffffu) != 0) return static_cast<uint16_t>((bits >> 16) | 0x40u);  // NaN stays NaN
  bits += 0x7fffu + ((bits >> 16) & 1u);
  return static_cast<uint16_t>(bits >> 16);
}
DYNG_DEV float from_bf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
DYNG_DEV float round_bf16(float f) { return from_bf16(to_bf16(f)); }

DYNG_DEV float normalized(const float* P, const float* x, int i) { return (x[i] - P[kMean + i]) * P[kScale + i]; }

// ---- training (FP32) ----

/// Layer 1 of a minibatch: h1[s][j] = relu(b1[j] + W1[j] . xn[s]) (rows of X picked by idx).
__global__ void k_fwd1(const float* P, const float* X, const int32_t* idx, int B, float* XN, float* H1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kHidden) return;
  const int s = t / kHidden, j = t % kHidden;
  const float* x = X + static_cast<int64_t>(idx[s]) * kIn;
  float acc = P[kB1 + j];
  for (int i = 0; i < kIn; ++i) {
    const float xn = normalized(P, x, i);
    if (j == 0) XN[s * kIn + i] = xn;
    acc += P[kW1 + j * kIn + i] * xn;
  }
  H1[t] = acc > 0.0f ? acc : 0.0f;
}

/// Layer 2: h2[s][j] = relu(b2[j] + W2[j] . h1[s]).
__global__ void k_fwd2(const float* P, const float* H1, int B, float* H2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kHidden) return;
  const int s = t / kHidden, j = t % kHidden;
  float acc = P[kB2 + j];
  for (int k = 0; k < kHidden; ++k) acc += P[kW2 + j * kHidden + k] * H1[s * kHidden + k];
  H2[t] = acc > 0.0f ? acc : 0.0f;
}

/// Output layer and the loss gradient: y[s][o] = b3[o] + W3[o] . h2[s];
/// dy = 2 (y - t) / (B * kOut) (mean squared error over the minibatch).
__global__ void k_fwd3(const float* P, const float* H2, const float* T, const int32_t* idx, int B, float* Y,
                       float* DY) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kOut) return;
  const int s = t / kOut, o = t % kOut;
  float acc = P[kB3 + o];
  for (int k = 0; k < kHidden; ++k) acc += P[kW3 + o * kHidden + k] * H2[s * kHidden + k];
  Y[t] = acc;
  if (T != nullptr) {
    // a NaN target (a schedule not measured on that sample) adds nothing to the loss
    const float target = T[static_cast<int64_t>(idx[s]) * kOut + o];
    DY[t] = isnan(target) ? 0.0f : 2.0f * (acc - target) / static_cast<float>(B * kOut);
  }
}

/// dh2[s][k] = [h2 > 0] * sum_o dy[s][o] W3[o][k].
__global__ void k_bwd3(const float* P, const float* H2, const float* DY, int B, float* DH2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kHidden) return;
  const int s = t / kHidden, k = t % kHidden;
  float acc = 0.0f;
  for (int o = 0; o < kOut; ++o) acc += DY[s * kOut + o] * P[kW3 + o * kHidden + k];
  DH2[t] = H2[t] > 0.0f ? acc : 0.0f;
}

/// dh1[s][k] = [h1 > 0] * sum_j dh2[s][j] W2[j][k].
__global__ void k_bwd2(const float* P, const float* H1, const float* DH2, int B, float* DH1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * kHidden) return;
  const int s = t / kHidden, k = t % kHidden;
  float acc = 0.0f;
  for (int j = 0; j < kHidden; ++j) acc += DH2[s * kHidden + j] * P[kW2 + j * kHidden + k];
  DH1[t] = H1[t] > 0.0f ? acc : 0.0f;
}

/// One thread per trained parameter: its gradient summed over the minibatch.
__global__ void k_grad(const float* XN, const float* H1, const float* H2, const float* DH1, const float* DH2,
                       const float* DY, int B, float* G) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kTrained) return;
  float acc = 0.0f;
  if (q < kB1) {
    const int j = q / kIn, i = q % kIn;
    for (int s = 0; s < B; ++s) acc += DH1[s * kHidden + j] * XN[s * kIn + i];
  } else if (q < kW2) {
    const int j = q - kB1;
    for (int s = 0; s < B; ++s) acc += DH1[s * kHidden + j];
  } else if (q < kB2) {
    const int j = (q - kW2) / kHidden, k = (q - kW2) % kHidden;
    for (int s = 0; s < B; ++s) acc += DH2[s * kHidden + j] * H1[s * kHidden + k];
  } else if (q < kW3) {
    const int j = q - kB2;
    for (int s = 0; s < B; ++s) acc += DH2[s * kHidden + j];
  } else if (q < kB3) {
    const int o = (q - kW3) / kHidden, k = (q - kW3) % kHidden;
    for (int s = 0; s < B; ++s) acc += DY[s * kOut + o] * H2[s * kHidden + k];
  } else {
    const int o = q - kB3;
    for (int s = 0; s < B; ++s) acc += DY[s * kOut + o];
  }
  G[q] = acc;
}

/// Adam (Kingma and Ba, 2015) on the trained parameters, step t >= 1.
__global__ void k_adam(float* P, const float* G, float* M, float* V, float lr, float beta1, float beta2, float eps,
                       int t) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kTrained) return;
  const float g = G[q];
  const float m = beta1 * M[q] + (1.0f - beta1) * g;
  const float v = beta2 * V[q] + (1.0f - beta2) * g * g;
  M[q] = m;
  V[q] = v;
  const float mhat = m / (1.0f - powf(beta1, static_cast<float>(t)));
  const float vhat = v / (1.0f - powf(beta2, static_cast<float>(t)));
  P[q] -= lr * mhat / (sqrtf(vhat) + eps);
}

// ---- inference (BF16) ----

/// The BF16 copy of the trained parameters (the normalization stays FP32).
__global__ void k_to_bf16(const float* P, uint16_t* H) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < kTrained) H[q] = to_bf16(P[q]);
}

/// One workgroup of kHidden threads per sample: the normalized input, both hidden layers and the
/// outputs rounded to BF16 at each layer, every dot product accumulated in FP32 (in index order).
/// The BF16 parameters (10.6 KB) are first copied to LDS with coalesced 4-byte loads, so the dot
/// products read LDS instead of issuing one strided global load per term.
__global__ void __launch_bounds__(kHidden) k_predict(const uint16_t* H, const float* P, const float* X, int B, float* Y) {
  __shared__ uint32_t hw[(kTrained + 1) / 2];
  __shared__ float xs[kIn];
  __shared__ float h1[kHidden];
  __shared__ float h2[kHidden];
  const int s = blockIdx.x;
  if (s >= B) return;
  const int j = threadIdx.x;
  const uint32_t* H2 = reinterpret_cast<const uint32_t*>(H);  // the BF16 buffer is 4-byte aligned
  for (int q = j; q < kTrained / 2; q += kHidden) hw[q] = H2[q];
  if ((kTrained & 1) && j == 0) hw[kTrained / 2] = H[kTrained - 1];
  if (j < kIn) xs[j] = round_bf16(normalized(P, X + static_cast<int64_t>(s) * kIn, j));
  __syncthreads();
  const uint16_t* h = reinterpret_cast<const uint16_t*>(hw);
  float acc = from_bf16(h[kB1 + j]);
  for (int i = 0; i < kIn; ++i) acc += from_bf16(h[kW1 + j * kIn + i]) * xs[i];
  h1[j] = round_bf16(acc > 0.0f ? acc : 0.0f);
  __syncthreads();
  acc = from_bf16(h[kB2 + j]);
  for (int k = 0; k < kHidden; ++k) acc += from_bf16(h[kW2 + j * kHidden + k]) * h1[k];
  h2[j] = round_bf16(acc > 0.0f ? acc : 0.0f);
  __syncthreads();
  if (j < kOut) {
    acc = from_bf16(h[kB3 + j]);
    for (int k = 0; k < kHidden; ++k) acc += from_bf16(h[kW3 + j * kHidden + k]) * h2[k];
    Y[static_cast<int64_t>(s) * kOut + j] = acc;
  }
}

// ---- batch features ----

struct feature_acc {
  unsigned long long tree_deletions;
  unsigned long long touched;
  unsigned long long log_degree_sum;  ///< sum of log2(1 + out-degree) * 1024, rounded
  unsigned long long max_degree;
};

/// Over the batch (on G_t, before it is applied): deletions that hit the old tree (parent[v] ==
/// u) and the out-degrees of the operations' tail vertices.
__global__ void __launch_bounds__(256) k_batch_features(const int32_t* ins_src, int64_t num_ins, const int32_t* del_src,
                                 const int32_t* del_dst, int64_t num_del, const int32_t* parents, int64_t n_tree,
                                 const int2* meta, int64_t n, feature_acc* acc) {
  unsigned long long hits = 0, touched = 0, logs = 0, mx = 0;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < num_ins + num_del;
       i += int64_t{gridDim.x} * blockDim.x) {
    int32_t u;
    if (i < num_del) {
      u = del_src[i];
      const int32_t v = del_dst[i];
      if (parents != nullptr && v >= 0 && v < n_tree && parents[v] == u) ++hits;
    } else {
      u = ins_src[i - num_del];
    }
    if (u >= 0 && u < n) {
      const unsigned long long d = static_cast<unsigned long long>(meta[u].y);
      ++touched;
      logs += static_cast<unsigned long long>(log2f(1.0f + static_cast<float>(d)) * 1024.0f + 0.5f);
      mx = d > mx ? d : mx;
    }
  }
  // one atomic per workgroup and counter
  __shared__ unsigned long long part[4];
  hits = block_reduce(hits, part, op_sum{});
  touched = block_reduce(touched, part, op_sum{});
  logs = block_reduce(logs, part, op_sum{});
  mx = block_reduce(mx, part, op_max{});
  if (threadIdx.x == 0) {
    if (hits) atomicAdd(&acc->tree_deletions, hits);
    if (touched) atomicAdd(&acc->touched, touched);
    if (logs) atomicAdd(&acc->log_degree_sum, logs);
    if (mx) atomicMax(&acc->max_degree, mx);
  }
}

/// Input j of the network from the batch features (the formulas of the Python
/// `scheduler.features`).
DYNG_DEV float feature_input(const feature_acc& a, int j, int64_t num_ins, int64_t num_del, int64_t n, int64_t m,
                             int64_t last_invalidated, int64_t last_affected) {
  const double ops = static_cast<double>(num_ins + num_del);
  const double nn = static_cast<double>(n > 1 ? n : 1), mm = static_cast<double>(m > 1 ? m : 1);
  const double touched = static_cast<double>(a.touched);
  double x = 0.0;
  switch (j) {
    case 0: x = log2(1.0 + ops); break;
    case 1: x = static_cast<double>(num_del) / (ops > 1.0 ? ops : 1.0); break;
    case 2: x = log2(1.0 + static_cast<double>(num_del)); break;
    case 3: x = log2(1.0 + static_cast<double>(num_ins)); break;
    case 4: x = log2(nn); break;
    case 5: x = log2(mm / nn); break;
    case 6: x = static_cast<double>(a.tree_deletions) / static_cast<double>(num_del > 1 ? num_del : 1); break;
    case 7: x = log2(1.0 + static_cast<double>(a.tree_deletions)); break;
    case 8: x = touched > 0.0 ? static_cast<double>(a.log_degree_sum) / 1024.0 / touched : 0.0; break;
    case 9: x = log2(1.0 + static_cast<double>(a.max_degree)); break;
    case 10: x = log2(1.0 + static_cast<double>(last_invalidated)); break;
    default: x = log2(1.0 + static_cast<double>(last_affected)); break;
  }
  return static_cast<float>(x);
}

/// The network's 12 inputs from the batch features, on the device; the accumulators are cleared
/// for the next batch (they start at zero when allocated).
__global__ void k_assemble_input(feature_acc* acc, int64_t num_ins, int64_t num_del, int64_t n, int64_t m,
                                 int64_t last_invalidated, int64_t last_affected, float* x) {
  if (blockIdx.x != 0) return;
  const feature_acc a = *acc;
  if (threadIdx.x < kIn) x[threadIdx.x] = feature_input(a, threadIdx.x, num_ins, num_del, n, m, last_invalidated, last_affected);
  __syncthreads();
  if (threadIdx.x == 0) *acc = feature_acc{0, 0, 0, 0};
}

/// Features to prediction in one workgroup (the scheduled update's path): the 12 inputs from the
/// accumulators (then cleared), normalized, and the BF16 forward of k_predict.
__global__ void __launch_bounds__(kHidden) k_infer(const uint16_t* H, const float* P, feature_acc* acc,
                                                   int64_t num_ins, int64_t num_del, int64_t n, int64_t m,
                                                   int64_t last_invalidated, int64_t last_affected, float* Y) {
  __shared__ uint32_t hw[(kTrained + 1) / 2];
  __shared__ float xs[kIn];
  __shared__ float h1[kHidden];
  __shared__ float h2[kHidden];
  const int j = threadIdx.x;
  const uint32_t* H2 = reinterpret_cast<const uint32_t*>(H);
  for (int q = j; q < kTrained / 2; q += kHidden) hw[q] = H2[q];
  if ((kTrained & 1) && j == 0) hw[kTrained / 2] = H[kTrained - 1];
  if (j < kIn) {
    const float x = feature_input(*acc, j, num_ins, num_del, n, m, last_invalidated, last_affected);
    xs[j] = round_bf16((x - P[kMean + j]) * P[kScale + j]);
  }
  __syncthreads();
  if (j == 0) *acc = feature_acc{0, 0, 0, 0};
  const uint16_t* h = reinterpret_cast<const uint16_t*>(hw);
  float a = from_bf16(h[kB1 + j]);
  for (int i = 0; i < kIn; ++i) a += from_bf16(h[kW1 + j * kIn + i]) * xs[i];
  h1[j] = round_bf16(a > 0.0f ? a : 0.0f);
  __syncthreads();
  a = from_bf16(h[kB2 + j]);
  for (int k = 0; k < kHidden; ++k) a += from_bf16(h[kW2 + j * kHidden + k]) * h1[k];
  h2[j] = round_bf16(a > 0.0f ? a : 0.0f);
  __syncthreads();
  if (j < kOut) {
    a = from_bf16(h[kB3 + j]);
    for (int k = 0; k < kHidden; ++k) a += from_bf16(h[kW3 + j * kHidden + k]) * h2[k];
    Y[j] = a;
  }
}

unsigned int grid_for(int64_t work, int block, int64_t cap = 65535) {
  return static_cast<unsigned int>(std::max<int64_t>(1, std::min<int64_t>((work + block - 1) / block, cap)));
}

}  // namespace
}  // namespace dyng_hip

using namespace dyng_hip;

/// The network's state on one device: FP32 parameters, their BF16 copy, Adam's moments and the
/// scratch of a training minibatch.
/// The schedule to run from the predicted log2 engine times: the plain one (0) unless another is
/// predicted at least kSwitchMargin faster and at least kMinGainMs sooner (a misprediction costs
/// more than the small gains the other schedules offer on most batches, and a predicted saving
/// below one run's timing noise is no evidence: a repair of tens of microseconds stays plain).
constexpr float kSwitchMargin = 0.2016f;  ///< log2(1.15): a predicted gain of at least 15 %
constexpr float kMinGainMs = 0.02f;       ///< and of at least 20 us

int pick_schedule(const float* y) {
  int best = 0;
  for (int o = 1; o < kOut; ++o) best = y[o] < y[best] ? o : best;
  const bool gain = y[best] < y[0] - kSwitchMargin && std::exp2(y[0]) - std::exp2(y[best]) >= kMinGainMs;
  return (best != 0 && gain) ? best : 0;
}

struct dyng_hip_mlp {
  int device = 0;
  dbuf<float> params, grads, adam_m, adam_v;
  dbuf<uint16_t> bf16;
  int adam_t = 0;
  bool bf16_fresh = false;
  dbuf<float> X, T, XN, H1, H2, DH1, DH2, Y, DY;
  dbuf<int32_t> idx;
  dbuf<feature_acc> acc;
  struct prediction {
    float y[kOut];
  };
  pinned<prediction> host_y;  ///< the last prediction, read after the stream's next synchronization
  /// the scheduled update's features and inference run on this stream, beside the apply
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;  ///< recorded after the prediction's read-back
  dyng_hip_mlp() = default;
  dyng_hip_mlp(const dyng_hip_mlp&) = delete;
  dyng_hip_mlp& operator=(const dyng_hip_mlp&) = delete;
  ~dyng_hip_mlp() {
    if (stream != nullptr) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    if (done != nullptr) (void)hipEventDestroy(done);
  }
};

namespace {

void refresh_bf16(dyng_hip_mlp* m, hipStream_t s) {
  if (m->bf16_fresh) return;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(kTrained, 256)), dim3(256), 0, s, m->params.p, m->bf16.p);
  DYNG_HIP_TRY(hipGetLastError());
  m->bf16_fresh = true;
}

/// Forward and backward of minibatch rows idx[0, B) of the device sets X, T into m->grads.
void forward_backward(dyng_hip_mlp* m, int B) {
  m->XN.reserve(static_cast<int64_t>(B) * kIn);
  m->H1.reserve(static_cast<int64_t>(B) * kHidden);
  m->H2.reserve(static_cast<int64_t>(B) * kHidden);
  m->DH1.reserve(static_cast<int64_t>(B) * kHidden);
  m->DH2.reserve(static_cast<int64_t>(B) * kHidden);
  m->Y.reserve(static_cast<int64_t>(B) * kOut);
  m->DY.reserve(static_cast<int64_t>(B) * kOut);
  const unsigned gh = grid_for(static_cast<int64_t>(B) * kHidden, 256), go = grid_for(static_cast<int64_t>(B) * kOut, 256);
  hipLaunchKernelGGL(k_fwd1, dim3(gh), dim3(256), 0, 0, m->params.p, m->X.p, m->idx.p, B, m->XN.p, m->H1.p);
  hipLaunchKernelGGL(k_fwd2, dim3(gh), dim3(256), 0, 0, m->params.p, m->H1.p, B, m->H2.p);
  hipLaunchKernelGGL(k_fwd3, dim3(go), dim3(256), 0, 0, m->params.p, m->H2.p, m->T.p, m->idx.p, B, m->Y.p, m->DY.p);
  hipLaunchKernelGGL(k_bwd3, dim3(gh), dim3(256), 0, 0, m->params.p, m->H2.p, m->DY.p, B, m->DH2.p);
  hipLaunchKernelGGL(k_bwd2, dim3(gh), dim3(256), 0, 0, m->params.p, m->H1.p, m->DH2.p, B, m->DH1.p);
  hipLaunchKernelGGL(k_grad, dim3(grid_for(kTrained, 256)), dim3(256), 0, 0, m->XN.p, m->H1.p, m->H2.p, m->DH1.p,
                     m->DH2.p, m->DY.p, B, m->grads.p);
  DYNG_HIP_TRY(hipGetLastError());
}

void upload_set(dyng_hip_mlp* m, const float* x, const float* t, int64_t count) {
  m->X.reserve(count * kIn);
  m->T.reserve(count * kOut);
  DYNG_HIP_TRY(hipMemcpy(m->X.p, x, count * kIn * sizeof(float), hipMemcpyHostToDevice));
  DYNG_HIP_TRY(hipMemcpy(m->T.p, t, count * kOut * sizeof(float), hipMemcpyHostToDevice));
}

float host_loss(dyng_hip_mlp* m, const float* t, const std::vector<int32_t>& idx, int B) {
  std::vector<float> y(static_cast<size_t>(B) * kOut);
  DYNG_HIP_TRY(hipMemcpy(y.data(), m->Y.p, y.size() * sizeof(float), hipMemcpyDeviceToHost));
  double sum = 0.0;
  for (int s = 0; s < B; ++s) {
    for (int o = 0; o < kOut; ++o) {
      const float target = t[static_cast<int64_t>(idx[s]) * kOut + o];
      if (std::isnan(target)) continue;
      const double d = static_cast<double>(y[s * kOut + o]) - target;
      sum += d * d;
    }
  }
  return static_cast<float>(sum / (static_cast<double>(B) * kOut));
}

uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

namespace dyng_hip {

hipEvent_t mlp_predict_async(dyng_hip_mlp* m, const dyng_hip_graph* g, const batch_scan& sc, const int32_t* parents,
                             int64_t n_tree, int64_t last_invalidated, int64_t last_affected) {
  // on the network's own stream: the batch scan's synchronization left the batch on the device,
  // and the apply changes neither the parents nor the out-row lengths before its merges, which
  // wait for `done`
  const hipStream_t s = m->stream;
  refresh_bf16(m, s);
  m->X.reserve(kIn);
  m->Y.reserve(kOut);
  if (m->acc.p == nullptr) {
    m->acc.reserve(1);
    DYNG_HIP_TRY(hipMemsetAsync(m->acc.p, 0, sizeof(feature_acc), s));
  }
  const int64_t work = sc.num_ins + sc.num_del;
  if (work > 0) {
    hipLaunchKernelGGL(k_batch_features, dim3(grid_for(work, 256, 256)), dim3(256), 0, s, sc.ins_src, sc.num_ins,
                       sc.del_src, sc.del_dst, sc.num_del, parents, n_tree, g->out.meta.p, g->n, m->acc.p);
  }
  hipLaunchKernelGGL(k_infer, dim3(1), dim3(kHidden), 0, s, m->bf16.p, m->params.p, m->acc.p, sc.num_ins, sc.num_del,
                     g->n, g->m, last_invalidated, last_affected, m->Y.p);
  DYNG_HIP_TRY(hipGetLastError());
  DYNG_HIP_TRY(hipMemcpyAsync(m->host_y.p->y, m->Y.p, kOut * sizeof(float), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipEventRecord(m->done, s));
  return m->done;
}

int mlp_pick(const dyng_hip_mlp* m) { return pick_schedule(m->host_y.p->y); }

int mlp_choose(dyng_hip_mlp* m, const dyng_hip_graph* g, const batch_scan& sc, const int32_t* parents, int64_t n_tree,
               int64_t last_invalidated, int64_t last_affected, hipStream_t s) {
  refresh_bf16(m, s);
  m->X.reserve(kIn);
  m->Y.reserve(kOut);
  if (m->acc.p == nullptr) {
    m->acc.reserve(1);
    DYNG_HIP_TRY(hipMemsetAsync(m->acc.p, 0, sizeof(feature_acc), s));
  }
  const int64_t work = sc.num_ins + sc.num_del;
  if (work > 0) {
    hipLaunchKernelGGL(k_batch_features, dim3(grid_for(work, 256, 256)), dim3(256), 0, s, sc.ins_src, sc.num_ins,
                       sc.del_src, sc.del_dst, sc.num_del, parents, n_tree, g->out.meta.p, g->n, m->acc.p);
  }
  hipLaunchKernelGGL(k_assemble_input, dim3(1), dim3(64), 0, s, m->acc.p, sc.num_ins, sc.num_del, g->n, g->m,
                     last_invalidated, last_affected, m->X.p);
  hipLaunchKernelGGL(k_predict, dim3(1), dim3(kHidden), 0, s, m->bf16.p, m->params.p, m->X.p, 1, m->Y.p);
  DYNG_HIP_TRY(hipGetLastError());
  float y[kOut];
  DYNG_HIP_TRY(hipMemcpyAsync(y, m->Y.p, sizeof(y), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  return pick_schedule(y);
}

}  // namespace dyng_hip

extern "C" int dyng_hip_mlp_create(int device, uint64_t seed, dyng_hip_mlp** out) {
  return guarded([&] {
    if (out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: null output handle");
    scoped_device guard(device);
    auto m = std::make_unique<dyng_hip_mlp>();
    m->device = device;
    m->params.reserve(kAll);
    m->grads.reserve(kTrained);
    m->adam_m.reserve(kTrained);
    m->adam_v.reserve(kTrained);
    m->bf16.reserve(kTrained);
    DYNG_HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    DYNG_HIP_TRY(hipEventCreateWithFlags(&m->done, hipEventDisableTiming));
    // He initialization (uniform, variance 2 / fan_in), biases 0, identity normalization
    std::vector<float> h(kAll, 0.0f);
    uint64_t state = seed * 0x9E3779B97F4A7C15ULL + 1;
    auto uniform = [&]() {
      state += 0x9E3779B97F4A7C15ULL;
      return static_cast<float>(static_cast<double>(mix(state) >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
    };
    auto init = [&](int at, int rows, int cols) {
      const float a = std::sqrt(6.0f / static_cast<float>(cols));
      for (int q = 0; q < rows * cols; ++q) h[at + q] = a * uniform();
    };
    init(kW1, kHidden, kIn);
    init(kW2, kHidden, kHidden);
    init(kW3, kOut, kHidden);
    for (int i = 0; i < kIn; ++i) h[kScale + i] = 1.0f;
    DYNG_HIP_TRY(hipMemcpy(m->params.p, h.data(), kAll * sizeof(float), hipMemcpyHostToDevice));
    DYNG_HIP_TRY(hipMemset(m->adam_m.p, 0, kTrained * sizeof(float)));
    DYNG_HIP_TRY(hipMemset(m->adam_v.p, 0, kTrained * sizeof(float)));
    *out = m.release();
  });
}

extern "C" int dyng_hip_mlp_destroy(dyng_hip_mlp* m) {
  return guarded([&] { delete m; });
}

extern "C" int dyng_hip_mlp_get_params(const dyng_hip_mlp* m, float* out, int64_t count) {
  return guarded([&] {
    if (m == nullptr || out == nullptr || count != kAll) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: get_params needs DYNG_HIP_MLP_PARAMS floats");
    scoped_device guard(m->device);
    DYNG_HIP_TRY(hipStreamSynchronize(m->stream));  // a scheduled update may still read the buffers
    DYNG_HIP_TRY(hipMemcpy(out, m->params.p, kAll * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int dyng_hip_mlp_set_params(dyng_hip_mlp* m, const float* in, int64_t count) {
  return guarded([&] {
    if (m == nullptr || in == nullptr || count != kAll) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: set_params needs DYNG_HIP_MLP_PARAMS floats");
    for (int i = 0; i < kIn; ++i) {
      if (!(in[kScale + i] > 0.0f) || !std::isfinite(in[kScale + i])) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: input scales must be positive");
    }
    scoped_device guard(m->device);
    DYNG_HIP_TRY(hipStreamSynchronize(m->stream));  // a scheduled update may still read the buffers
    DYNG_HIP_TRY(hipMemcpy(m->params.p, in, kAll * sizeof(float), hipMemcpyHostToDevice));
    DYNG_HIP_TRY(hipMemset(m->adam_m.p, 0, kTrained * sizeof(float)));
    DYNG_HIP_TRY(hipMemset(m->adam_v.p, 0, kTrained * sizeof(float)));
    m->adam_t = 0;
    m->bf16_fresh = false;
  });
}

extern "C" int dyng_hip_mlp_gradients(dyng_hip_mlp* m, const float* x, const float* t, int64_t count, float* grad,
                                      float* loss) {
  return guarded([&] {
    if (m == nullptr || x == nullptr || t == nullptr || grad == nullptr || count < 1 || count > (1 << 20)) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: gradients need 1 to 2^20 samples and output arrays");
    }
    scoped_device guard(m->device);
    DYNG_HIP_TRY(hipStreamSynchronize(m->stream));  // a scheduled update may still read the buffers
    upload_set(m, x, t, count);
    std::vector<int32_t> idx(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) idx[i] = static_cast<int32_t>(i);
    m->idx.reserve(count);
    DYNG_HIP_TRY(hipMemcpy(m->idx.p, idx.data(), count * sizeof(int32_t), hipMemcpyHostToDevice));
    forward_backward(m, static_cast<int>(count));
    DYNG_HIP_TRY(hipDeviceSynchronize());
    DYNG_HIP_TRY(hipMemcpy(grad, m->grads.p, kTrained * sizeof(float), hipMemcpyDeviceToHost));
    if (loss != nullptr) *loss = host_loss(m, t, idx, static_cast<int>(count));
  });
}

extern "C" int dyng_hip_mlp_train(dyng_hip_mlp* m, const float* x, const float* t, int64_t count, int32_t steps,
                                  int32_t batch, float lr, uint64_t seed, float* final_loss) {
  return guarded([&] {
    if (m == nullptr || x == nullptr || t == nullptr || count < 1 || count > (1 << 24) || steps < 0 || batch < 1 ||
        batch > 65536 || !(lr > 0.0f)) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: train needs samples, steps >= 0, a minibatch in [1, 65536], lr > 0");
    }
    scoped_device guard(m->device);
    DYNG_HIP_TRY(hipStreamSynchronize(m->stream));  // a scheduled update may still read the buffers
    upload_set(m, x, t, count);
    const int B = static_cast<int>(std::min<int64_t>(batch, count));
    std::vector<int32_t> order(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) order[i] = static_cast<int32_t>(i);
    m->idx.reserve(B);
    uint64_t state = seed + 0x5DEECE66DULL;
    int64_t at = count;  // next unread position of the shuffled order
    std::vector<int32_t> pick(static_cast<size_t>(B));
    for (int step = 0; step < steps; ++step) {
      for (int s = 0; s < B; ++s) {
        if (at == count) {  // reshuffle (Fisher-Yates)
          for (int64_t i = count - 1; i > 0; --i) {
            state += 0x9E3779B97F4A7C15ULL;
            const int64_t j = static_cast<int64_t>(mix(state) % static_cast<uint64_t>(i + 1));
            std::swap(order[i], order[j]);
          }
          at = 0;
        }
        pick[s] = order[at++];
      }
      DYNG_HIP_TRY(hipMemcpy(m->idx.p, pick.data(), B * sizeof(int32_t), hipMemcpyHostToDevice));
      forward_backward(m, B);
      ++m->adam_t;
      hipLaunchKernelGGL(k_adam, dim3(grid_for(kTrained, 256)), dim3(256), 0, 0, m->params.p, m->grads.p,
                         m->adam_m.p, m->adam_v.p, lr, 0.9f, 0.999f, 1e-8f, m->adam_t);
      DYNG_HIP_TRY(hipGetLastError());
    }
    m->bf16_fresh = false;
    if (final_loss != nullptr) {
      // the loss over the whole set with the trained weights (FP32)
      std::vector<int32_t> all(static_cast<size_t>(count));
      for (int64_t i = 0; i < count; ++i) all[i] = static_cast<int32_t>(i);
      m->idx.reserve(count);
      DYNG_HIP_TRY(hipMemcpy(m->idx.p, all.data(), count * sizeof(int32_t), hipMemcpyHostToDevice));
      forward_backward(m, static_cast<int>(count));
      DYNG_HIP_TRY(hipDeviceSynchronize());
      *final_loss = host_loss(m, t, all, static_cast<int>(count));
    }
    DYNG_HIP_TRY(hipDeviceSynchronize());
  });
}

extern "C" int dyng_hip_mlp_predict(dyng_hip_mlp* m, const float* x, int64_t count, float* y) {
  return guarded([&] {
    if (m == nullptr || x == nullptr || y == nullptr || count < 1 || count > 65535) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: mlp: predict needs 1 to 65535 samples");
    }
    scoped_device guard(m->device);
    DYNG_HIP_TRY(hipStreamSynchronize(m->stream));  // a scheduled update may still read the buffers
    refresh_bf16(m, nullptr);
    m->X.reserve(count * kIn);
    m->Y.reserve(count * kOut);
    DYNG_HIP_TRY(hipMemcpy(m->X.p, x, count * kIn * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_predict, dim3(static_cast<unsigned>(count)), dim3(kHidden), 0, 0, m->bf16.p, m->params.p,
                       m->X.p, static_cast<int>(count), m->Y.p);
    DYNG_HIP_TRY(hipGetLastError());
    DYNG_HIP_TRY(hipMemcpy(y, m->Y.p, count * kOut * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int dyng_hip_batch_features(const dyng_hip_graph* g, const dyng_hip_batch* b, const int32_t* parents,
                                       int64_t n_tree, void* stream, dyng_hip_batch_stats* out) {
  return guarded([&] {
    if (g == nullptr || b == nullptr || out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: batch_features: null argument");
    if (b->num_ins < 0 || b->num_del < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: batch_features: negative sizes");
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int32_t *ins_src = b->ins_src, *del_src = b->del_src, *del_dst = b->del_dst;
    dbuf<int32_t> hi, hs, hd;
    if (!b->on_device) {
      hi.reserve(std::max<int64_t>(b->num_ins, 1));
      hs.reserve(std::max<int64_t>(b->num_del, 1));
      hd.reserve(std::max<int64_t>(b->num_del, 1));
      if (b->num_ins > 0) DYNG_HIP_TRY(hipMemcpyAsync(hi.p, ins_src, b->num_ins * 4, hipMemcpyHostToDevice, s));
      if (b->num_del > 0) {
        DYNG_HIP_TRY(hipMemcpyAsync(hs.p, del_src, b->num_del * 4, hipMemcpyHostToDevice, s));
        DYNG_HIP_TRY(hipMemcpyAsync(hd.p, del_dst, b->num_del * 4, hipMemcpyHostToDevice, s));
      }
      ins_src = hi.p;
      del_src = hs.p;
      del_dst = hd.p;
    }
    dbuf<feature_acc> acc;
    acc.reserve(1);
    DYNG_HIP_TRY(hipMemsetAsync(acc.p, 0, sizeof(feature_acc), s));
    const int64_t work = b->num_ins + b->num_del;
    if (work > 0) {
      hipLaunchKernelGGL(k_batch_features, dim3(grid_for(work, 256, 256)), dim3(256), 0, s, ins_src, b->num_ins, del_src,
                         del_dst, b->num_del, parents, n_tree, g->out.meta.p, g->n, acc.p);
      DYNG_HIP_TRY(hipGetLastError());
    }
    feature_acc h{};
    DYNG_HIP_TRY(hipMemcpyAsync(&h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    out->num_ins = b->num_ins;
    out->num_del = b->num_del;
    out->tree_deletions = static_cast<int64_t>(h.tree_deletions);
    out->touched = static_cast<int64_t>(h.touched);
    out->mean_log_degree = h.touched ? static_cast<double>(h.log_degree_sum) / 1024.0 / static_cast<double>(h.touched) : 0.0;
    out->max_degree = static_cast<int64_t>(h.max_degree);
    out->num_vertices = g->n;
    out->num_edges = g->m;
  });
}
