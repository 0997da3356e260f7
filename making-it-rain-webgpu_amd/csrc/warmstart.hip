// warmstart.hip -- training and evaluation of the learned SHAKE warm start
// (the network and its tile helpers are in warmstart.hpp).
//
// Training data are the engine's own: a recording run writes, for every H-bond
// cluster of up to 4 constraints at every step, the network's inputs and the
// multipliers plain SHAKE converged to.  Training is mean-squared error over
// the constraints present, minibatches drawn from a permutation of the
// samples, Adam on FP32 master weights.  Forward and backward passes run on
// the matrix cores in BF16 with FP32 accumulation:
//     dY  = 2 mask (Y - T) / (4 n)
//     dW3 = H2^T dY,  dA2 = (dY W3^T) (1 - H2^2)
//     dW2 = H1^T dA2, dA1 = (dA2 W2^T) (1 - H1^2)
//     dW1 = X^T dA1,  db = column sums of dY, dA2, dA1
// Each workgroup (one wave) accumulates its tiles' weight gradients in MFMA
// accumulators and writes them to its own row of a partial buffer; one pass
// then sums the rows in a fixed order, so a training run is deterministic.
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "warmstart.hpp"

// the "present" flag of constraint c is feature 4c + 3
__device__ __forceinline__ float md_ws_mask(const float *X, int row, int out) {
  return out < MD_WS_MAXC && X[row * MD_WS_F + 4 * out + 3] > 0.5f ? 1.0f : 0.0f;
}

// one wave per workgroup: LDS of one tile -- X, H1, H2, dY [16][16], dA2, dA1
#define MD_WS_TRAIN_LDS (16 * MD_WS_F + 4 * 16 * MD_WS_H + 16 * MD_WS_O)

// Gradients of the loss over samples idx[0..n) (rows of xs/ts), one
// workgroup per strided set of 16-sample tiles; this workgroup's sums go to
// partial[blockIdx.x][0..MD_WS_NPARAMS], its loss to partial[...][NPARAMS].
// With `ys`, also writes the network's outputs (for tests).
__global__ void __launch_bounds__(64) k_ws_gradients(md_ws_net N, const float *__restrict__ xs,
                                                      const float *__restrict__ ts, const int *__restrict__ idx, int n,
                                                      float inv_count, float *partial, float *ys) {
  __shared__ float lds[MD_WS_TRAIN_LDS];
  float *X = lds, *H1 = X + 16 * MD_WS_F, *H2 = H1 + 16 * MD_WS_H, *dA2 = H2 + 16 * MD_WS_H,
        *dA1 = dA2 + 16 * MD_WS_H, *dY = dA1 + 16 * MD_WS_H;
  const int l = threadIdx.x, row0 = 4 * (l >> 4), col = l & 15;
  md_f4 g1[2][4], g2[4][4], g3[4];
  for (int a = 0; a < 4; ++a) {
    g3[a] = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int b = 0; b < 4; ++b) g2[a][b] = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
    if (a < 2)
      for (int b = 0; b < 4; ++b) g1[a][b] = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
  }
  float db1 = 0.0f, db2 = 0.0f, db3 = 0.0f, loss = 0.0f;
  const int tiles = (n + 15) / 16;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    // load the tile's inputs (rows past n are zero: no mask, no gradient)
    for (int k = l; k < 16 * MD_WS_F; k += 64) {
      int r = k / MD_WS_F, q = k - r * MD_WS_F, s = tile * 16 + r;
      X[k] = s < n ? xs[(size_t)idx[s] * MD_WS_F + q] : 0.0f;
    }
    __syncthreads();
    md_f4 y = md_ws_forward(N, X, H1, H2, l);
    // dY, the loss and the outputs
    for (int j = 0; j < 4; ++j) {
      int r = row0 + j, s = tile * 16 + r;
      float t = 0.0f;
      if (s < n && col < MD_WS_MAXC) t = ts[(size_t)idx[s] * MD_WS_MAXC + col];
      float m = s < n ? md_ws_mask(X, r, col) : 0.0f;
      float diff = (y[j] - t) * m;
      loss += diff * diff;
      dY[r * MD_WS_O + col] = 2.0f * diff * inv_count;
      if (ys && s < n) ys[(size_t)s * MD_WS_O + col] = y[j];
    }
    __syncthreads();
    // dW3 += H2^T dY; db3
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) g3[mt] = md_mfma(md_b_rows(H2, MD_WS_H, 16 * mt, l), md_b_rows(dY, MD_WS_O, 0, l), g3[mt]);
    if (l < MD_WS_O)
      for (int r = 0; r < 16; ++r) db3 += dY[r * MD_WS_O + l];
    // dA2 = (dY W3^T) (1 - H2^2)
    {
      md_s4 a = md_a_rows(dY, MD_WS_O, 0, l);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        md_f4 d = md_mfma(a, N.W3ts[nt * 64 + l], md_f4{0.0f, 0.0f, 0.0f, 0.0f});
        for (int j = 0; j < 4; ++j) {
          int r = row0 + j, c = 16 * nt + col;
          float h = H2[r * MD_WS_H + c];
          dA2[r * MD_WS_H + c] = d[j] * (1.0f - h * h);
        }
      }
    }
    __syncthreads();
    // dW2 += H1^T dA2; db2
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      md_s4 a = md_b_rows(H1, MD_WS_H, 16 * mt, l);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) g2[mt][nt] = md_mfma(a, md_b_rows(dA2, MD_WS_H, 16 * nt, l), g2[mt][nt]);
    }
    for (int r = 0; r < 16; ++r) db2 += dA2[r * MD_WS_H + l];
    // dA1 = (dA2 W2^T) (1 - H1^2)
    {
      md_f4 acc[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc[nt] = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) {
        md_s4 a = md_a_rows(dA2, MD_WS_H, 16 * kt, l);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[nt] = md_mfma(a, N.W2ts[(kt * 4 + nt) * 64 + l], acc[nt]);
      }
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        for (int j = 0; j < 4; ++j) {
          int r = row0 + j, c = 16 * nt + col;
          float h = H1[r * MD_WS_H + c];
          dA1[r * MD_WS_H + c] = acc[nt][j] * (1.0f - h * h);
        }
    }
    __syncthreads();
    // dW1 += X^T dA1; db1
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      md_s4 a = md_b_rows(X, MD_WS_F, 16 * mt, l);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) g1[mt][nt] = md_mfma(a, md_b_rows(dA1, MD_WS_H, 16 * nt, l), g1[mt][nt]);
    }
    for (int r = 0; r < 16; ++r) db1 += dA1[r * MD_WS_H + l];
    __syncthreads();
  }
  // this workgroup's sums, in the parameter layout (row-major [K][N])
  float *out = partial + (size_t)blockIdx.x * (MD_WS_NPARAMS + 1);
  for (int mt = 0; mt < 2; ++mt)
    for (int nt = 0; nt < 4; ++nt)
      for (int j = 0; j < 4; ++j)
        out[MD_WS_OFF_W1 + (16 * mt + row0 + j) * MD_WS_H + 16 * nt + col] = g1[mt][nt][j];
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
      for (int j = 0; j < 4; ++j)
        out[MD_WS_OFF_W2 + (16 * mt + row0 + j) * MD_WS_H + 16 * nt + col] = g2[mt][nt][j];
  for (int mt = 0; mt < 4; ++mt)
    for (int j = 0; j < 4; ++j) out[MD_WS_OFF_W3 + (16 * mt + row0 + j) * MD_WS_O + col] = g3[mt][j];
  out[MD_WS_OFF_B1 + l] = db1;
  out[MD_WS_OFF_B2 + l] = db2;
  if (l < MD_WS_O) out[MD_WS_OFF_B3 + l] = db3;
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
  if (l == 0) out[MD_WS_NPARAMS] = loss * inv_count;
}

// Sum the partial rows in order into grad (and the loss into grad[NPARAMS]);
// with lr > 0, one Adam step on the master weights.
__global__ void k_ws_reduce_adam(const float *partial, int rows, float *grad, float *master, float *m1, float *m2,
                                 float lr, float b1t, float b2t) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > MD_WS_NPARAMS) return;
  float g = 0.0f;
  for (int r = 0; r < rows; ++r) g += partial[(size_t)r * (MD_WS_NPARAMS + 1) + p];
  grad[p] = g;
  if (p == MD_WS_NPARAMS || lr <= 0.0f) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  float a = beta1 * m1[p] + (1.0f - beta1) * g;
  float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  master[p] -= lr * (a / (1.0f - b1t)) / (sqrtf(b / (1.0f - b2t)) + eps);
}

// The matrix-core operand copies of the master weights.
__global__ void k_ws_pack(const float *master, md_s4 *W1s, md_s4 *W2s, md_s4 *W3s, md_s4 *W2ts, md_s4 *W3ts) {
  int tile = blockIdx.x, l = threadIdx.x;
  const float *W1 = master + MD_WS_OFF_W1, *W2 = master + MD_WS_OFF_W2, *W3 = master + MD_WS_OFF_W3;
  int kq = 4 * (l >> 4), n = l & 15;
  md_s4 v;
  if (tile < 8) {  // W1: kt 0..1, nt 0..3
    int kt = tile / 4, nt = tile % 4;
    for (int j = 0; j < 4; ++j) v[j] = md_bf16(W1[(16 * kt + kq + j) * MD_WS_H + 16 * nt + n]);
    W1s[tile * 64 + l] = v;
  } else if (tile < 24) {  // W2 and W2^T: kt 0..3, nt 0..3
    int t = tile - 8, kt = t / 4, nt = t % 4;
    for (int j = 0; j < 4; ++j) v[j] = md_bf16(W2[(16 * kt + kq + j) * MD_WS_H + 16 * nt + n]);
    W2s[t * 64 + l] = v;
    for (int j = 0; j < 4; ++j) v[j] = md_bf16(W2[(16 * nt + n) * MD_WS_H + 16 * kt + kq + j]);
    W2ts[t * 64 + l] = v;
  } else if (tile < 28) {  // W3: kt 0..3
    int kt = tile - 24;
    for (int j = 0; j < 4; ++j) v[j] = md_bf16(W3[(16 * kt + kq + j) * MD_WS_O + n]);
    W3s[kt * 64 + l] = v;
  } else {  // W3^T: nt 0..3 over the hidden units, k over the 16 outputs
    int nt = tile - 28;
    for (int j = 0; j < 4; ++j) v[j] = md_bf16(W3[(16 * nt + n) * MD_WS_O + kq + j]);
    W3ts[nt * 64 + l] = v;
  }
}

// ---- host ---------------------------------------------------------------------

struct md_warmstart {
  int device = 0;
  hipStream_t stream = nullptr;
  float *master = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *partial = nullptr;
  md_s4 *W1s = nullptr, *W2s = nullptr, *W3s = nullptr, *W2ts = nullptr, *W3ts = nullptr;
  md_ws_net *net_d = nullptr;  // device copy the integrator reads
  md_ws_net net{};
  float *xs = nullptr, *ts = nullptr;  // samples
  int64_t capacity = 0, count = 0;
  int *idx = nullptr;  // minibatch sample indices
  int64_t adam_steps = 0;
  uint64_t rng = 0;
  int rows = 0;  // workgroups of the gradient kernel
  std::vector<void *> allocations;

  template <typename T>
  T *alloc(size_t n) {
    void *p = nullptr;
    MD_HIP(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256)));
    MD_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n * sizeof(T), 256), stream));
    allocations.push_back(p);
    return static_cast<T *>(p);
  }
};

static const int WS_ROWS = 256;

static void ws_pack(md_warmstart *w) {
  k_ws_pack<<<32, 64, 0, w->stream>>>(w->master, w->W1s, w->W2s, w->W3s, w->W2ts, w->W3ts);
  MD_HIP(hipGetLastError());
}

static uint64_t ws_next(uint64_t &s) {  // splitmix64
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

md_warmstart *md_ws_create(int device, uint32_t seed, int64_t capacity) {
  md_warmstart *w = new md_warmstart();
  try {
    w->device = device;
    MD_HIP(hipSetDevice(device));
    MD_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    w->master = w->alloc<float>(MD_WS_NPARAMS);
    w->m1 = w->alloc<float>(MD_WS_NPARAMS);
    w->m2 = w->alloc<float>(MD_WS_NPARAMS);
    w->grad = w->alloc<float>(MD_WS_NPARAMS + 1);
    w->rows = WS_ROWS;
    w->partial = w->alloc<float>((size_t)WS_ROWS * (MD_WS_NPARAMS + 1));
    w->W1s = w->alloc<md_s4>(8 * 64);
    w->W2s = w->alloc<md_s4>(16 * 64);
    w->W3s = w->alloc<md_s4>(4 * 64);
    w->W2ts = w->alloc<md_s4>(16 * 64);
    w->W3ts = w->alloc<md_s4>(4 * 64);
    w->capacity = std::max<int64_t>(capacity, 16);
    w->xs = w->alloc<float>((size_t)w->capacity * MD_WS_F);
    w->ts = w->alloc<float>((size_t)w->capacity * MD_WS_MAXC);
    w->idx = w->alloc<int>((size_t)w->capacity);
    w->net = md_ws_net{w->W1s, w->W2s, w->W3s, w->W2ts, w->W3ts, w->master + MD_WS_OFF_B1,
                       w->master + MD_WS_OFF_B2, w->master + MD_WS_OFF_B3};
    w->net_d = w->alloc<md_ws_net>(1);
    MD_HIP(hipMemcpyAsync(w->net_d, &w->net, sizeof w->net, hipMemcpyHostToDevice, w->stream));
    // Glorot-uniform weights, zero biases, from the seed
    w->rng = 0x5DEECE66Dull ^ ((uint64_t)seed << 17);
    std::vector<float> p(MD_WS_NPARAMS, 0.0f);
    auto fill = [&](int off, int fan_in, int fan_out, int cols_used) {
      double lim = std::sqrt(6.0 / (fan_in + fan_out));
      for (int i = 0; i < fan_in; ++i)
        for (int j = 0; j < fan_out; ++j) {
          double u = (double)(ws_next(w->rng) >> 11) * (1.0 / 9007199254740992.0);
          p[off + i * fan_out + j] = j < cols_used ? (float)((2.0 * u - 1.0) * lim) : 0.0f;
        }
    };
    fill(MD_WS_OFF_W1, MD_WS_F, MD_WS_H, MD_WS_H);
    fill(MD_WS_OFF_W2, MD_WS_H, MD_WS_H, MD_WS_H);
    fill(MD_WS_OFF_W3, MD_WS_H, MD_WS_O, MD_WS_MAXC);
    MD_HIP(hipMemcpyAsync(w->master, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice, w->stream));
    ws_pack(w);
    MD_HIP(hipStreamSynchronize(w->stream));
  } catch (...) {
    md_ws_destroy(w);
    throw;
  }
  return w;
}

void md_ws_destroy(md_warmstart *w) {
  if (!w) return;
  hipSetDevice(w->device);
  if (w->stream) hipStreamSynchronize(w->stream);
  for (void *p : w->allocations) hipFree(p);
  if (w->stream) hipStreamDestroy(w->stream);
  delete w;
}

const md_ws_net *md_ws_device_net(md_warmstart *w) { return w->net_d; }
int md_ws_device(md_warmstart *w) { return w->device; }

void md_ws_reserve(md_warmstart *w, int64_t wanted, float **xs, float **ts, int64_t *room) {
  int64_t at = std::min(w->count, w->capacity);
  *xs = w->xs + at * MD_WS_F;
  *ts = w->ts + at * MD_WS_MAXC;
  *room = w->capacity - at;
  (void)wanted;
}

void md_ws_commit(md_warmstart *w, int64_t added) { w->count = std::min(w->capacity, w->count + added); }

int64_t md_ws_count(md_warmstart *w) { return w->count; }

void md_ws_get_params(md_warmstart *w, float *out) {
  MD_HIP(hipSetDevice(w->device));
  MD_HIP(hipMemcpy(out, w->master, sizeof(float) * MD_WS_NPARAMS, hipMemcpyDeviceToHost));
}

void md_ws_set_params(md_warmstart *w, const float *in) {
  MD_HIP(hipSetDevice(w->device));
  MD_HIP(hipMemcpyAsync(w->master, in, sizeof(float) * MD_WS_NPARAMS, hipMemcpyHostToDevice, w->stream));
  ws_pack(w);
  MD_HIP(hipStreamSynchronize(w->stream));
}

// gradients (and outputs) of given samples, no update
double md_ws_evaluate(md_warmstart *w, const float *x, const float *t, int n, float *y, float *grad) {
  MD_HIP(hipSetDevice(w->device));
  if (n <= 0) return 0.0;
  float *dx, *dt, *dy = nullptr;
  int *di;
  MD_HIP(hipMalloc(&dx, sizeof(float) * n * MD_WS_F));
  MD_HIP(hipMalloc(&dt, sizeof(float) * n * MD_WS_MAXC));
  MD_HIP(hipMalloc(&di, sizeof(int) * n));
  if (y) MD_HIP(hipMalloc(&dy, sizeof(float) * n * MD_WS_O));
  std::vector<int> id(n);
  for (int k = 0; k < n; ++k) id[k] = k;
  MD_HIP(hipMemcpyAsync(dx, x, sizeof(float) * n * MD_WS_F, hipMemcpyHostToDevice, w->stream));
  MD_HIP(hipMemcpyAsync(dt, t, sizeof(float) * n * MD_WS_MAXC, hipMemcpyHostToDevice, w->stream));
  MD_HIP(hipMemcpyAsync(di, id.data(), sizeof(int) * n, hipMemcpyHostToDevice, w->stream));
  int rows = std::min(w->rows, (n + 15) / 16);
  k_ws_gradients<<<rows, 64, 0, w->stream>>>(w->net, dx, dt, di, n, 1.0f / (4.0f * n), w->partial, dy);
  k_ws_reduce_adam<<<(MD_WS_NPARAMS + 256) / 256, 256, 0, w->stream>>>(w->partial, rows, w->grad, w->master, w->m1,
                                                                       w->m2, 0.0f, 0.0f, 0.0f);
  MD_HIP(hipGetLastError());
  std::vector<float> g(MD_WS_NPARAMS + 1);
  MD_HIP(hipMemcpyAsync(g.data(), w->grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost, w->stream));
  if (y) MD_HIP(hipMemcpyAsync(y, dy, sizeof(float) * n * MD_WS_O, hipMemcpyDeviceToHost, w->stream));
  MD_HIP(hipStreamSynchronize(w->stream));
  hipFree(dx);
  hipFree(dt);
  hipFree(di);
  if (dy) hipFree(dy);
  if (grad) std::memcpy(grad, g.data(), sizeof(float) * MD_WS_NPARAMS);
  return g[MD_WS_NPARAMS];
}

// `iterations` Adam steps on minibatches of `batch` recorded samples
double md_ws_train(md_warmstart *w, int iterations, int batch, double lr) {
  MD_HIP(hipSetDevice(w->device));
  const int64_t n = w->count;
  if (n < 16) throw md_error(MD_ERR_INVALID_ARGUMENT, "record at least 16 samples before training");
  batch = (int)std::min<int64_t>(std::max(16, batch), n);
  std::vector<int> perm(n);
  for (int64_t k = 0; k < n; ++k) perm[k] = (int)k;
  int64_t at = n;  // position in the current permutation
  double loss = 0.0;
  int rows = std::min(w->rows, (batch + 15) / 16);
  std::vector<float> g(MD_WS_NPARAMS + 1);
  for (int it = 0; it < iterations; ++it) {
    if (at + batch > n) {  // reshuffle (Fisher-Yates), upload once per epoch
      for (int64_t k = n - 1; k > 0; --k) std::swap(perm[k], perm[ws_next(w->rng) % (uint64_t)(k + 1)]);
      MD_HIP(hipMemcpyAsync(w->idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, w->stream));
      MD_HIP(hipStreamSynchronize(w->stream));
      at = 0;
    }
    ++w->adam_steps;
    float b1t = (float)std::pow(0.9, (double)w->adam_steps), b2t = (float)std::pow(0.999, (double)w->adam_steps);
    k_ws_gradients<<<rows, 64, 0, w->stream>>>(w->net, w->xs, w->ts, w->idx + at, batch, 1.0f / (4.0f * batch),
                                              w->partial, nullptr);
    k_ws_reduce_adam<<<(MD_WS_NPARAMS + 256) / 256, 256, 0, w->stream>>>(w->partial, rows, w->grad, w->master,
                                                                         w->m1, w->m2, (float)lr, b1t, b2t);
    ws_pack(w);
    at += batch;
    if (it == iterations - 1) {
      MD_HIP(hipMemcpyAsync(g.data(), w->grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost, w->stream));
      MD_HIP(hipStreamSynchronize(w->stream));
      loss = g[MD_WS_NPARAMS];
    }
  }
  MD_HIP(hipStreamSynchronize(w->stream));
  return loss;
}
