// minimise.hip -- the device side of md_minimise: the L-BFGS vectors, the
// OpenMM objective's gradient with constraint springs, and the reductions
// the search reads back as scalars.
//
// Replaces the upload / evaluate / read-back loop of upstream's
// minimiseOnDevice (src/gpu/minimise.ts:33-54), whose L-BFGS runs on the
// host (src/reference/minimise.ts:164-280, lbfgs.ts:102-275): here every
// 3N-vector -- positions, gradient, direction, trial point, the six
// correction pairs -- stays in HBM as f64 planes [3][np], and the host only
// reads back the few scalars the line search decides on (energies, dot
// products, largest components).  Reductions run in a fixed order (per-block
// partial sums, then one block over the partials), so a minimisation is
// reproducible run to run.
#include "engine.hpp"

#define MD_MIN_BLOCK 256
#define MD_MIN_PARTIALS 1024  // most blocks a reduction uses (then grid-strided)

static int blocks_for(int count) { return std::max(1, std::min(MD_MIN_PARTIALS, (count + MD_MIN_BLOCK - 1) / MD_MIN_BLOCK)); }

__device__ __forceinline__ double md_block_sum(double v) {
  __shared__ double part[MD_MIN_BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < MD_MIN_BLOCK / 64; ++w) s += part[w];
  __syncthreads();
  return s;
}

__device__ __forceinline__ double md_block_max(double v) {
  __shared__ double part[MD_MIN_BLOCK / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < MD_MIN_BLOCK / 64; ++w) s = fmax(s, part[w]);
  __syncthreads();
  return s;
}

// ---- vector kernels over the 3 planes of n real atoms (pitch np) -------------

__global__ void k_min_dot_partial(const double *a, const double *b, int n, int np, double *partial) {
  double s = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * n; k += gridDim.x * blockDim.x) {
    const int d = k / n, i = k - d * n;
    s += a[d * np + i] * b[d * np + i];
  }
  s = md_block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_min_absmax_partial(const double *a, int n, int np, double *partial) {
  double s = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * n; k += gridDim.x * blockDim.x) {
    const int d = k / n, i = k - d * n;
    s = fmax(s, fabs(a[d * np + i]));
  }
  s = md_block_max(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[0 .. count) -> *out in order (sum or max)
__global__ void k_min_finish(const double *partial, int count, int max, double *out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < count; ++b) s = max ? fmax(s, partial[b]) : s + partial[b];
  *out = s;
}

// out = a x + b y (x or y may alias out)
__global__ void k_min_lincomb(double *out, double a, const double *x, double b, const double *y, int n, int np) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * n; k += gridDim.x * blockDim.x) {
    const int d = k / n, i = k - d * n;
    const int at = d * np + i;
    out[at] = a * x[at] + (y ? b * y[at] : 0.0);
  }
}

// Q32.32 positions <-> f64 nm
__global__ void k_min_from_fixed(const int64_t *pos, double *x, int n, int np) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * n; k += gridDim.x * blockDim.x) {
    const int d = k / n, i = k - d * n;
    x[d * np + i] = md_fixed_to_double_dev(pos[d * np + i]);
  }
}

// positions from x (rounded to Q32.32 as md_set_positions rounds); the drift
// residual cleared; *bad set when a coordinate is outside +/-2^31 nm
__global__ void k_min_to_fixed(const double *x, int64_t *pos, float *resid, int n, int np, int *bad) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * n; k += gridDim.x * blockDim.x) {
    const int d = k / n, i = k - d * n;
    const double v = x[d * np + i];
    if (!(fabs(v) < 2147483648.0)) {
      *bad = 1;
      continue;
    }
    pos[d * np + i] = (int64_t)rint(v * 4294967296.0);
    resid[d * np + i] = 0.0f;
  }
}

// the potential's gradient: -F for moving atoms, 0 for frozen ones
__global__ void k_min_gradient(const int64_t *force, const float *inv_mass, double *g, int n, int np) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 3 * n; k += gridDim.x * blockDim.x) {
    const int d = k / n, i = k - d * n;
    g[d * np + i] = inv_mass[i] == 0.0f ? 0.0 : -md_fixed_to_double_dev(force[d * np + i]);
  }
}

// Constraint springs (k/2)(r - d)^2 of OpenMM's minimiser objective: each
// spring's pull stored per spring (k_min_spring_gather adds them to the
// atoms in a fixed order, so the gradient's bits never depend on thread
// timing, as f64 atomics would make them), their energy into partial[block].
__global__ void k_min_springs(const double *x, const int *p1, const int *p2, const double *dist, int nc,
                              double kspring, double *spring_f, int np, double *partial) {
  double e = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const int a = p1[c], b = p2[c];
    double v[3], r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      v[d] = x[d * np + a] - x[d * np + b];
      r2 += v[d] * v[d];
    }
    const double r = sqrt(r2), dr = r - dist[c];
    e += 0.5 * kspring * dr * dr;
    const double f = kspring * dr / (r == 0.0 ? 1.0 : r);
    for (int d = 0; d < 3; ++d) spring_f[3 * (size_t)c + d] = f * v[d];
  }
  e = md_block_sum(e);
  if (threadIdx.x == 0) partial[blockIdx.x] = e;
}

// Each moving atom's springs, in the order of its list, onto its gradient.
__global__ void k_min_spring_gather(const double *spring_f, const int *cstart, const int *clist,
                                    const float *inv_mass, double *g, int n, int np) {
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    if (inv_mass[a] == 0.0f || cstart[a] == cstart[a + 1]) continue;
    double sum[3] = {0.0, 0.0, 0.0};
    for (int k = cstart[a]; k < cstart[a + 1]; ++k) {
      const int code = clist[k];
      const int c = code >= 0 ? code : -code - 1;
      const double sign = code >= 0 ? 1.0 : -1.0;
      for (int d = 0; d < 3; ++d) sum[d] += sign * spring_f[3 * (size_t)c + d];
    }
    for (int d = 0; d < 3; ++d) g[d * np + a] += sum[d];
  }
}

// the largest relative constraint deviation |r - d| / d, per block
__global__ void k_min_constraint_error(const double *x, const int *p1, const int *p2, const double *dist, int nc,
                                       int np, double *partial) {
  double w = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const int a = p1[c], b = p2[c];
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double v = x[d * np + a] - x[d * np + b];
      r2 += v * v;
    }
    w = fmax(w, fabs(sqrt(r2) - dist[c]) / dist[c]);
  }
  w = md_block_max(w);
  if (threadIdx.x == 0) partial[blockIdx.x] = w;
}

// ---- host launchers (engine.hpp) ---------------------------------------------

md_min_work *md_min_work_create(md_system *s, const std::vector<int> &p1, const std::vector<int> &p2,
                                const std::vector<double> &dist) {
  auto *w = new md_min_work();
  const size_t len = 3 * (size_t)s->np;
  w->vec = md_alloc<double>(s, len * MD_MIN_VECTORS);
  w->partial = md_alloc<double>(s, MD_MIN_PARTIALS);
  w->scalar = md_alloc<double>(s, 1);
  w->bad = md_alloc<int>(s, 1);
  w->nc = (int)p1.size();
  w->p1 = md_alloc<int>(s, std::max<size_t>(1, p1.size()));
  w->p2 = md_alloc<int>(s, std::max<size_t>(1, p2.size()));
  w->dist = md_alloc<double>(s, std::max<size_t>(1, dist.size()));
  // by atom, the springs that pull it (first atoms +c, second atoms -c - 1)
  std::vector<int> cstart(s->np + 1, 0), clist(2 * p1.size());
  for (size_t c = 0; c < p1.size(); ++c) {
    ++cstart[p1[c] + 1];
    ++cstart[p2[c] + 1];
  }
  for (int a = 0; a < s->np; ++a) cstart[a + 1] += cstart[a];
  {
    std::vector<int> fill(cstart.begin(), cstart.end() - 1);
    for (size_t c = 0; c < p1.size(); ++c) {
      clist[fill[p1[c]]++] = (int)c;
      clist[fill[p2[c]]++] = -(int)c - 1;
    }
  }
  w->spring_f = md_alloc<double>(s, std::max<size_t>(1, 3 * p1.size()));
  w->cstart = md_alloc<int>(s, cstart.size());
  w->clist = md_alloc<int>(s, std::max<size_t>(1, clist.size()));
  MD_HIP(hipMemcpyAsync(w->cstart, cstart.data(), sizeof(int) * cstart.size(), hipMemcpyHostToDevice, s->stream));
  if (!clist.empty())
    MD_HIP(hipMemcpyAsync(w->clist, clist.data(), sizeof(int) * clist.size(), hipMemcpyHostToDevice, s->stream));
  if (w->nc) {
    MD_HIP(hipMemcpyAsync(w->p1, p1.data(), sizeof(int) * p1.size(), hipMemcpyHostToDevice, s->stream));
    MD_HIP(hipMemcpyAsync(w->p2, p2.data(), sizeof(int) * p2.size(), hipMemcpyHostToDevice, s->stream));
    MD_HIP(hipMemcpyAsync(w->dist, dist.data(), sizeof(double) * dist.size(), hipMemcpyHostToDevice, s->stream));
  }
  MD_HIP(hipStreamSynchronize(s->stream));
  return w;
}

double *md_min_vec(md_system *s, md_min_work *w, int k) { return w->vec + (size_t)k * 3 * s->np; }

static double read_scalar(md_system *s, md_min_work *w) {
  double v = 0.0;
  MD_HIP(hipMemcpyAsync(&v, w->scalar, sizeof v, hipMemcpyDeviceToHost, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  return v;
}

double md_min_dot(md_system *s, md_min_work *w, const double *a, const double *b) {
  const int nb = blocks_for(3 * s->n);
  k_min_dot_partial<<<nb, MD_MIN_BLOCK, 0, s->stream>>>(a, b, s->n, s->np, w->partial);
  k_min_finish<<<1, 64, 0, s->stream>>>(w->partial, nb, 0, w->scalar);
  MD_HIP(hipGetLastError());
  return read_scalar(s, w);
}

double md_min_absmax(md_system *s, md_min_work *w, const double *a) {
  const int nb = blocks_for(3 * s->n);
  k_min_absmax_partial<<<nb, MD_MIN_BLOCK, 0, s->stream>>>(a, s->n, s->np, w->partial);
  k_min_finish<<<1, 64, 0, s->stream>>>(w->partial, nb, 1, w->scalar);
  MD_HIP(hipGetLastError());
  return read_scalar(s, w);
}

void md_min_lincomb(md_system *s, double *out, double a, const double *x, double b, const double *y) {
  k_min_lincomb<<<blocks_for(3 * s->n), MD_MIN_BLOCK, 0, s->stream>>>(out, a, x, b, y, s->n, s->np);
  MD_HIP(hipGetLastError());
}

void md_min_positions_out(md_system *s, double *x) {
  k_min_from_fixed<<<blocks_for(3 * s->n), MD_MIN_BLOCK, 0, s->stream>>>(s->pos, x, s->n, s->np);
  MD_HIP(hipGetLastError());
}

bool md_min_positions_in(md_system *s, md_min_work *w, const double *x) {
  MD_HIP(hipMemsetAsync(w->bad, 0, sizeof(int), s->stream));
  k_min_to_fixed<<<blocks_for(3 * s->n), MD_MIN_BLOCK, 0, s->stream>>>(x, s->pos, s->resid, s->n, s->np, w->bad);
  MD_HIP(hipGetLastError());
  int bad = 0;
  MD_HIP(hipMemcpyAsync(&bad, w->bad, sizeof bad, hipMemcpyDeviceToHost, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  return bad == 0;
}

double md_min_gradient(md_system *s, md_min_work *w, const double *x, double kspring, double *g) {
  k_min_gradient<<<blocks_for(3 * s->n), MD_MIN_BLOCK, 0, s->stream>>>(s->force, s->inv_mass, g, s->n, s->np);
  double spring = 0.0;
  if (w->nc) {
    const int nb = blocks_for(w->nc);
    k_min_springs<<<nb, MD_MIN_BLOCK, 0, s->stream>>>(x, w->p1, w->p2, w->dist, w->nc, kspring, w->spring_f, s->np,
                                                     w->partial);
    k_min_spring_gather<<<blocks_for(s->n), MD_MIN_BLOCK, 0, s->stream>>>(w->spring_f, w->cstart, w->clist,
                                                                          s->inv_mass, g, s->n, s->np);
    k_min_finish<<<1, 64, 0, s->stream>>>(w->partial, nb, 0, w->scalar);
    MD_HIP(hipGetLastError());
    spring = read_scalar(s, w);
  }
  MD_HIP(hipGetLastError());
  return spring;
}

double md_min_constraint_error(md_system *s, md_min_work *w, const double *x) {
  if (!w->nc) return 0.0;
  const int nb = blocks_for(w->nc);
  k_min_constraint_error<<<nb, MD_MIN_BLOCK, 0, s->stream>>>(x, w->p1, w->p2, w->dist, w->nc, s->np, w->partial);
  k_min_finish<<<1, 64, 0, s->stream>>>(w->partial, nb, 1, w->scalar);
  MD_HIP(hipGetLastError());
  return read_scalar(s, w);
}
