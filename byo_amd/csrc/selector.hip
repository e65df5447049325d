// selector.hip -- the learned edgeMap direction selector: an FP32 MLP
// 6->32->32->3 (ReLU, softmax cross-entropy) with forward, backward and Adam
// kernels of its own, plus the host-side forward used per BFS level.
//
// It replaces the fixed M/20 rule of edgeMapData (gbbs/edge_map_data.h:259,
// 291) as the per-level choice between sparse push (class 0), dense pull
// (class 1) and dense forward (class 2); the rule stays the fallback when no
// weights are loaded or the output is not finite.  Direction never changes
// BFS levels, so results stay bit-exact whatever the network says.
//
// Parameter layout (row-major, out x in): W1[32x6] b1[32] W2[32x32] b2[32]
// W3[3x32] b3[3] = 1379 floats.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace byo {

constexpr int kNP = BYO_MLP_NPARAM;
constexpr int oW1 = 0, ob1 = 192, oW2 = 224, ob2 = 1248, oW3 = 1280, ob3 = 1376;
static_assert(ob3 + 3 == kNP, "layout");

bool mlp_eval_host(const std::vector<float> &p, const float *xin, float *z, float range) {
  // p holds params followed by mean[6] and inv_std[6]
  float x[6];
  // normalised features are clamped to +-kSelClamp standard deviations, so
  // graphs unlike the training set cannot drive the network far outside
  // the region it was fitted on; the return value says none was clamped
  bool in_range = true;
  for (int i = 0; i < 6; i++) {
    const float v = (xin[i] - p[kNP + i]) * p[kNP + 6 + i];
    in_range = in_range && std::fabs(v) <= range;
    x[i] = std::min(std::max(v, -kSelClamp), kSelClamp);
  }
  float h1[32], h2[32];
  for (int o = 0; o < 32; o++) {
    float a = p[ob1 + o];
    for (int i = 0; i < 6; i++) a += p[oW1 + o * 6 + i] * x[i];
    h1[o] = a > 0 ? a : 0;
  }
  for (int o = 0; o < 32; o++) {
    float a = p[ob2 + o];
    for (int i = 0; i < 32; i++) a += p[oW2 + o * 32 + i] * h1[i];
    h2[o] = a > 0 ? a : 0;
  }
  for (int o = 0; o < 3; o++) {
    float a = p[ob3 + o];
    for (int i = 0; i < 32; i++) a += p[oW3 + o * 32 + i] * h2[i];
    z[o] = a;
  }
  return in_range;
}

namespace {

// One thread per sample: forward, loss, and this sample's gradient
// (already divided by the batch size) into g[b * kNP + k].
__global__ void __launch_bounds__(256) k_mlp_fwd_bwd(const float *P, const float *X, const int32_t *Y, int B,
                                                     float *g, float *loss, float *logits) {
  __shared__ float sp[kNP];
  for (int k = threadIdx.x; k < kNP; k += blockDim.x) sp[k] = P[k];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float x[6];
  for (int i = 0; i < 6; i++) x[i] = X[b * 6 + i];
  float h1[32], h2[32], z[3];
  for (int o = 0; o < 32; o++) {
    float a = sp[ob1 + o];
    for (int i = 0; i < 6; i++) a = fmaf(sp[oW1 + o * 6 + i], x[i], a);
    h1[o] = a > 0.f ? a : 0.f;
  }
  for (int o = 0; o < 32; o++) {
    float a = sp[ob2 + o];
    for (int i = 0; i < 32; i++) a = fmaf(sp[oW2 + o * 32 + i], h1[i], a);
    h2[o] = a > 0.f ? a : 0.f;
  }
  for (int o = 0; o < 3; o++) {
    float a = sp[ob3 + o];
    for (int i = 0; i < 32; i++) a = fmaf(sp[oW3 + o * 32 + i], h2[i], a);
    z[o] = a;
    if (logits) logits[b * 3 + o] = a;
  }
  if (!g) return;
  const float mx = fmaxf(z[0], fmaxf(z[1], z[2]));
  const float e0 = expf(z[0] - mx), e1 = expf(z[1] - mx), e2 = expf(z[2] - mx);
  const float se = e0 + e1 + e2;
  const int y = Y[b];
  loss[b] = -(z[y] - mx - logf(se));
  const float invB = 1.0f / (float)B;
  float dz[3] = {(e0 / se - (y == 0)) * invB, (e1 / se - (y == 1)) * invB, (e2 / se - (y == 2)) * invB};
  float *gb = g + (size_t)b * kNP;
  float dh2[32];
  for (int i = 0; i < 32; i++) dh2[i] = 0.f;
  for (int o = 0; o < 3; o++) {
    gb[ob3 + o] = dz[o];
    for (int i = 0; i < 32; i++) {
      gb[oW3 + o * 32 + i] = dz[o] * h2[i];
      dh2[i] = fmaf(sp[oW3 + o * 32 + i], dz[o], dh2[i]);
    }
  }
  float dh1[32];
  for (int i = 0; i < 32; i++) {
    dh2[i] = h2[i] > 0.f ? dh2[i] : 0.f;
    dh1[i] = 0.f;
  }
  for (int o = 0; o < 32; o++) {
    gb[ob2 + o] = dh2[o];
    for (int i = 0; i < 32; i++) {
      gb[oW2 + o * 32 + i] = dh2[o] * h1[i];
      dh1[i] = fmaf(sp[oW2 + o * 32 + i], dh2[o], dh1[i]);
    }
  }
  for (int o = 0; o < 32; o++) {
    const float d = h1[o] > 0.f ? dh1[o] : 0.f;
    gb[ob1 + o] = d;
    for (int i = 0; i < 6; i++) gb[oW1 + o * 6 + i] = d * x[i];
  }
}

// grads[k] = sum_b g[b][k] in sample order (deterministic); loss mean too.
__global__ void k_mlp_reduce(const float *g, const float *loss, int B, float *grads, float *mean_loss) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kNP) {
    float s = 0.f;
    for (int b = 0; b < B; b++) s += g[(size_t)b * kNP + k];
    grads[k] = s;
  }
  if (k == kNP) {
    float s = 0.f;
    for (int b = 0; b < B; b++) s += loss[b];
    *mean_loss = s / (float)B;
  }
}

__global__ void k_adam(float *P, const float *G, float *M, float *V, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNP) return;
  const float gk = G[k];
  const float m = b1 * M[k] + (1.f - b1) * gk;
  const float v = b2 * V[k] + (1.f - b2) * gk * gk;
  M[k] = m;
  V[k] = v;
  P[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace
}  // namespace byo

struct byo_trainer {
  int device = 0;
  hipStream_t stream = nullptr;
  byo::DevBuf P, G, M, V, X, Y, gs, loss, logits, scalar;
  int cap = 0;
  float lr = 1e-3f;
  int t = 0;
  void ensure(int B) {
    if (B <= cap) return;
    cap = B;
    X.alloc((size_t)B * 6 * 4);
    Y.alloc((size_t)B * 4);
    gs.alloc((size_t)B * byo::kNP * 4);
    loss.alloc((size_t)B * 4);
    logits.alloc((size_t)B * 3 * 4);
  }
};

using namespace byo;

namespace {
void trainer_fwd_bwd(byo_trainer *t, const float *x, const int32_t *y, int B, bool with_grad) {
  t->ensure(B);
  hipStream_t s = t->stream;
  BYO_HIP(hipMemcpyAsync(t->X.p, x, (size_t)B * 24, hipMemcpyHostToDevice, s));
  if (y) BYO_HIP(hipMemcpyAsync(t->Y.p, y, (size_t)B * 4, hipMemcpyHostToDevice, s));
  k_mlp_fwd_bwd<<<(B + 255) / 256, 256, 0, s>>>(t->P.as<float>(), t->X.as<float>(), t->Y.as<int32_t>(), B,
                                                with_grad ? t->gs.as<float>() : nullptr, t->loss.as<float>(),
                                                t->logits.as<float>());
  if (with_grad)
    k_mlp_reduce<<<(kNP + 1 + 255) / 256, 256, 0, s>>>(t->gs.as<float>(), t->loss.as<float>(), B,
                                                        t->G.as<float>(), t->scalar.as<float>());
  BYO_LAUNCH_CHECK();
}
}  // namespace

extern "C" {

int byo_selector_load_engine(byo_graph_t g, int engine, const float *params, const float *mean,
                             const float *inv_std) {
  return guard([&] {
    BYO_REQUIRE(g && params && mean && inv_std, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(engine == BYO_SEL_SINGLE || engine == BYO_SEL_BATCH, BYO_E_INVAL, "unknown traversal engine");
    if (engine == BYO_SEL_BATCH) {
      std::vector<float> h(params, params + kNP);
      h.insert(h.end(), mean, mean + 6);
      h.insert(h.end(), inv_std, inv_std + 6);
      bind(g);
      g->sel_params_b.ensure((kNP + 12) * sizeof(float));
      BYO_HIP(hipMemcpy(g->sel_params_b.p, h.data(), (kNP + 12) * sizeof(float), hipMemcpyHostToDevice));
      g->sel_loaded_b = true;
      return;
    }
    g->sel_host.assign(params, params + kNP);
    g->sel_host.insert(g->sel_host.end(), mean, mean + 6);
    g->sel_host.insert(g->sel_host.end(), inv_std, inv_std + 6);
    bind(g);
    g->sel_params.ensure((kNP + 12) * sizeof(float));
    BYO_HIP(hipMemcpy(g->sel_params.p, g->sel_host.data(), (kNP + 12) * sizeof(float), hipMemcpyHostToDevice));
    g->sel_loaded = true;
  });
}

int byo_selector_load(byo_graph_t g, const float *params, const float *mean, const float *inv_std) {
  return byo_selector_load_engine(g, BYO_SEL_SINGLE, params, mean, inv_std);
}

int byo_trainer_create(const float *init_params, float lr, byo_trainer_t *out) {
  return guard([&] {
    BYO_REQUIRE(init_params && out, BYO_E_INVAL, "null argument");
    require_device();
    auto *t = new byo_trainer();
    try {
      BYO_HIP(hipGetDevice(&t->device));
      BYO_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
      t->P.alloc(kNP * 4);
      t->G.alloc(kNP * 4);
      t->M.alloc(kNP * 4);
      t->V.alloc(kNP * 4);
      t->scalar.alloc(64);
      t->lr = lr;
      BYO_HIP(hipMemcpy(t->P.p, init_params, kNP * 4, hipMemcpyHostToDevice));
      BYO_HIP(hipMemset(t->M.p, 0, kNP * 4));
      BYO_HIP(hipMemset(t->V.p, 0, kNP * 4));
    } catch (...) {
      if (t->stream) (void)hipStreamDestroy(t->stream);
      delete t;
      throw;
    }
    *out = t;
  });
}

int byo_trainer_destroy(byo_trainer_t t) {
  if (!t) return BYO_OK;
  return guard([&] {
    (void)hipSetDevice(t->device);
    (void)hipStreamSynchronize(t->stream);
    (void)hipStreamDestroy(t->stream);
    delete t;
  });
}

int byo_trainer_grad(byo_trainer_t t, const float *x, const int32_t *labels, int batch, float *loss, float *grads,
                     float *logits) {
  return guard([&] {
    BYO_REQUIRE(t && x && labels && batch > 0, BYO_E_INVAL, "bad argument");
    for (int b = 0; b < batch; b++) BYO_REQUIRE(labels[b] >= 0 && labels[b] < 3, BYO_E_INVAL, "label not in 0..2");
    BYO_HIP(hipSetDevice(t->device));
    trainer_fwd_bwd(t, x, labels, batch, true);
    if (grads) BYO_HIP(hipMemcpyAsync(grads, t->G.p, kNP * 4, hipMemcpyDeviceToHost, t->stream));
    if (logits) BYO_HIP(hipMemcpyAsync(logits, t->logits.p, (size_t)batch * 12, hipMemcpyDeviceToHost, t->stream));
    if (loss) BYO_HIP(hipMemcpyAsync(loss, t->scalar.p, 4, hipMemcpyDeviceToHost, t->stream));
    BYO_HIP(hipStreamSynchronize(t->stream));
  });
}

int byo_trainer_step(byo_trainer_t t, const float *x, const int32_t *labels, int batch, float *loss) {
  return guard([&] {
    BYO_REQUIRE(t && x && labels && batch > 0, BYO_E_INVAL, "bad argument");
    for (int b = 0; b < batch; b++) BYO_REQUIRE(labels[b] >= 0 && labels[b] < 3, BYO_E_INVAL, "label not in 0..2");
    BYO_HIP(hipSetDevice(t->device));
    trainer_fwd_bwd(t, x, labels, batch, true);
    t->t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)t->t), c2 = 1.f - std::pow(b2, (float)t->t);
    k_adam<<<(kNP + 255) / 256, 256, 0, t->stream>>>(t->P.as<float>(), t->G.as<float>(), t->M.as<float>(),
                                                     t->V.as<float>(), t->lr, b1, b2, 1e-8f, c1, c2);
    BYO_LAUNCH_CHECK();
    if (loss) BYO_HIP(hipMemcpyAsync(loss, t->scalar.p, 4, hipMemcpyDeviceToHost, t->stream));
    BYO_HIP(hipStreamSynchronize(t->stream));
  });
}

int byo_trainer_params(byo_trainer_t t, float *params) {
  return guard([&] {
    BYO_REQUIRE(t && params, BYO_E_INVAL, "null argument");
    BYO_HIP(hipSetDevice(t->device));
    BYO_HIP(hipMemcpyAsync(params, t->P.p, kNP * 4, hipMemcpyDeviceToHost, t->stream));
    BYO_HIP(hipStreamSynchronize(t->stream));
  });
}

int byo_trainer_predict(byo_trainer_t t, const float *x, int batch, int32_t *cls) {
  return guard([&] {
    BYO_REQUIRE(t && x && cls && batch > 0, BYO_E_INVAL, "bad argument");
    BYO_HIP(hipSetDevice(t->device));
    trainer_fwd_bwd(t, x, nullptr, batch, false);
    std::vector<float> lg((size_t)batch * 3);
    BYO_HIP(hipMemcpyAsync(lg.data(), t->logits.p, lg.size() * 4, hipMemcpyDeviceToHost, t->stream));
    BYO_HIP(hipStreamSynchronize(t->stream));
    for (int b = 0; b < batch; b++) {
      int best = 0;
      for (int k = 1; k < 3; k++)
        if (lg[b * 3 + k] > lg[b * 3 + best]) best = k;
      cls[b] = best;
    }
  });
}

}  // extern "C"
