// pme.hip -- smooth particle-mesh Ewald for gfx950.
//
// Replaces the upstream WGSL `spread_charge` (src/gpu/kernels/pme.wgsl:123-154),
// `load_grid` (pme-grid.wgsl:13-19), `fft_stage` (fft.wgsl:59-107), `convolve`
// (pme-convolve.wgsl:33-91) and `interpolate_force` (pme-interpolate.wgsl
// :104-155).  The mathematics is src/reference/forces/pme.ts:94-263: order-5
// cardinal B-splines at fractional coordinates (grid points floor(u)-4 ..
// floor(u)), E = (2 pi K / V) sum_{m != 0} exp(-k^2/4 alpha^2) / (k^2 |b(m)|^2)
// |Q(m)|^2 with k = 2 pi m B^-T, the grid multiplied by twice the factor, and
// forces -q B^-1 (n_x dtheta_x theta_y theta_z, ...) . phi.
//
// gfx950 design: the charge grid is int64 Q32.32 (125 atomic adds per charged
// atom, order independent); the transform is real-to-complex, so only
// nz/2+1 planes of k-space are stored and transformed; each 1D transform runs
// whole inside LDS (a batch of lines per workgroup, mixed-radix Stockham
// stages, twiddles from an f64-computed table); the x-axis pass does forward
// transform, convolution, energy and inverse transform in one kernel; the
// z pass that reads the integer grid also clears it for the next step.
#include <climits>
#include <cmath>
#include <vector>

#include "engine.hpp"

#define PME_MAX_STAGES 12
#define PME_BATCH 16

struct md_fft_plan {
  int n;
  int nstages;
  int radix[PME_MAX_STAGES];
  int tw;  // offset of this length's twiddle table
};

struct md_pme_args {
  int nx, ny, nz, nzc;
  md_fft_plan px, py, pz;
  const float2 *tw;  // exp(-2 pi i t / n), per length
  const float *mod_x, *mod_y, *mod_z;
  double inv[9];
  float invf[9];
  float prefactor;   // K 2 pi / V
  float exp_scale;   // -1/(4 alpha^2)
};

// one replica's table entry
struct md_pme_table {
  md_pme_args P;
  int n, np, nslots;
  const int64_t *pos;
  const float4 *param;
  const int *slot_atom;  // the nonbonded list's spatial order, 64-slot super-clusters
  int64_t *grid;
  float2 *cgrid;
  float *phi;
  int64_t *force;
  int64_t *energy;
};

__device__ __forceinline__ const md_pme_table &md_pme_entry(const md_pme_table *T, const md_bmap &map, int &blk) {
  return T[md_locate(map, blockIdx.x, blk)];
}

// ---- B-splines (src/reference/forces/bspline.ts:55-91, order 5) ------------

__device__ __forceinline__ void md_bspline5(float w, float *th, float *dth) {
  float l0 = 1.0f - w, l1 = w, l2, l3;
  // order 3
  l2 = 0.5f * w * l1;
  l1 = 0.5f * ((w + 1.0f) * l0 + (2.0f - w) * l1);
  l0 = 0.5f * (1.0f - w) * l0;
  // order 4
  const float third = 1.0f / 3.0f;
  l3 = third * w * l2;
  l2 = third * ((w + 1.0f) * l1 + (3.0f - w) * l2);
  l1 = third * ((w + 2.0f) * l0 + (2.0f - w) * l1);
  l0 = third * (1.0f - w) * l0;
  if (dth) {
    dth[0] = -l0;
    dth[1] = l0 - l1;
    dth[2] = l1 - l2;
    dth[3] = l2 - l3;
    dth[4] = l3;
  }
  // order 5
  th[4] = 0.25f * w * l3;
  th[3] = 0.25f * ((w + 1.0f) * l2 + (4.0f - w) * l3);
  th[2] = 0.25f * ((w + 2.0f) * l1 + (3.0f - w) * l2);
  th[1] = 0.25f * ((w + 3.0f) * l0 + (2.0f - w) * l1);
  th[0] = 0.25f * (1.0f - w) * l0;
}

// fractional position on the mesh, from the exact fixed-point coordinate
__device__ __forceinline__ void md_mesh_coord(const int64_t *pos, int np, int i, const double *M, int nx, int ny,
                                              int nz, int *base, float *w) {
  const double s = 2.3283064365386963e-10;
  double rx = (double)pos[i] * s, ry = (double)pos[np + i] * s, rz = (double)pos[2 * np + i] * s;
  double f[3] = {rx * M[0] + ry * M[3] + rz * M[6], ry * M[4] + rz * M[7], rz * M[8]};
  int n[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    double fr = f[d] - floor(f[d]);
    double u = fr * n[d];
    double cell = floor(u);
    w[d] = (float)(u - cell);
    base[d] = (int)cell - (MD_PME_ORDER - 1);
  }
}

__device__ __forceinline__ int md_wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Charge spreading, one workgroup per 64-slot super-cluster of the nonbonded
// list (spatially compact: its atoms' stencils overlap), one thread per
// (slot, x-plane of the stencil).  Each thread's 25 contributions -- the
// same Q32.32 values as one global atomic each would add -- go into an int64
// tile in LDS covering the super-cluster's stencil box; the tile is then
// added to the grid, one global atomic per non-zero point: about a quarter
// of the global atomics of spreading atom by atom.  Integer sums do not
// depend on order, so the grid is the same either way.  A super-cluster
// whose box exceeds the tile spreads straight to the grid.
#define MD_SPREAD_THREADS (64 * MD_PME_ORDER)
#define MD_SPREAD_TILE 6144  // int64 points (48 KB of LDS)

__device__ __forceinline__ int md_wrap_near(int d, int n) {
  d = md_wrap(d, n);
  return d >= n / 2 ? d - n : d;  // the representative in [-n/2, n/2)
}

__global__ void __launch_bounds__(MD_SPREAD_THREADS) k_pme_spread(const md_pme_table *__restrict__ T, md_bmap map) {
  // super-clusters laid out XCD by XCD: neighbouring ones add into the same
  // mesh region from one L2
  int S;
  const md_pme_table &E = T[md_locate(map, md_xcd_block(blockIdx.x, gridDim.x), S)];
  const md_pme_args &P = E.P;
  int64_t *grid = E.grid;
  __shared__ unsigned long long tile[MD_SPREAD_TILE];
  __shared__ int box[9];  // reference stencil base, then lowest and highest offsets from it
  const int t = threadIdx.x, lane = t & 63, a = t >> 6;
  const int i = E.slot_atom[S * 64 + lane];
  const float q = i >= 0 ? E.param[i].w : 0.0f;
  const bool live = q != 0.0f;
  int base[3] = {0, 0, 0};
  float w[3] = {0.0f, 0.0f, 0.0f};
  if (live) md_mesh_coord(E.pos, E.np, i, P.inv, P.nx, P.ny, P.nz, base, w);
  const int n[3] = {P.nx, P.ny, P.nz};
  // the first wave (plane 0 of every slot) finds the stencil box: offsets of
  // each base from the first live slot's, taken to the nearest periodic
  // representative (any representative lands on the same grid points; the
  // nearest keeps the box small)
  if (a == 0) {
    const uint64_t m = __ballot(live);
    const int first = m ? __ffsll((unsigned long long)m) - 1 : 0;
    for (int d = 0; d < 3; ++d) {
      const int ref = __shfl(base[d], first, 64);
      const int off = live ? md_wrap_near(base[d] - ref, n[d]) : 0;
      int lo = live ? off : INT_MAX, hi = live ? off : INT_MIN;
      for (int k = 32; k > 0; k >>= 1) {
        lo = min(lo, __shfl_xor(lo, k, 64));
        hi = max(hi, __shfl_xor(hi, k, 64));
      }
      if (lane == 0) {
        box[d] = ref;
        box[3 + d] = m ? lo : 0;
        box[6 + d] = m ? hi : -MD_PME_ORDER;  // no charge: an empty box
      }
    }
  }
  __syncthreads();
  int ext[3], off[3];
  for (int d = 0; d < 3; ++d) {
    ext[d] = box[6 + d] - box[3 + d] + MD_PME_ORDER;
    off[d] = live ? md_wrap_near(base[d] - box[d], n[d]) - box[3 + d] : 0;
  }
  const int volume = ext[0] * ext[1] * ext[2];
  if (volume == 0) return;  // no charged atom in this super-cluster (uniform)
  float tx[5], ty[5], tz[5];
  md_bspline5(w[0], tx, nullptr);
  md_bspline5(w[1], ty, nullptr);
  md_bspline5(w[2], tz, nullptr);
  const float wx = q * tx[a];
  if (volume > MD_SPREAD_TILE) {
    if (!live) return;
    const int gx = md_wrap(base[0] + a, P.nx);
    for (int b = 0; b < 5; ++b) {
      const int gy = md_wrap(base[1] + b, P.ny);
      const float wxy = wx * ty[b];
      int64_t *row = grid + ((size_t)gx * P.ny + gy) * P.nz;
      for (int c = 0; c < 5; ++c) md_atomic_add_i64(&row[md_wrap(base[2] + c, P.nz)], md_to_fixed(wxy * tz[c]));
    }
    return;
  }
  for (int k = t; k < volume; k += MD_SPREAD_THREADS) tile[k] = 0;
  __syncthreads();
  if (live) {
    const int plane = (off[0] + a) * ext[1];
    for (int b = 0; b < 5; ++b) {
      const float wxy = wx * ty[b];
      unsigned long long *row = tile + (size_t)(plane + off[1] + b) * ext[2] + off[2];
      for (int c = 0; c < 5; ++c) atomicAdd(&row[c], (unsigned long long)md_to_fixed(wxy * tz[c]));
    }
  }
  __syncthreads();
  const int yz = ext[1] * ext[2];
  for (int k = t; k < volume; k += MD_SPREAD_THREADS) {
    const long long v = (long long)tile[k];
    if (v == 0) continue;
    const int lx = k / yz, r = k - lx * yz, ly = r / ext[2], lz = r - ly * ext[2];
    const int gx = md_wrap(box[0] + box[3] + lx, P.nx), gy = md_wrap(box[1] + box[4] + ly, P.ny),
              gz = md_wrap(box[2] + box[5] + lz, P.nz);
    md_atomic_add_i64(grid + ((size_t)gx * P.ny + gy) * P.nz + gz, v);
  }
}

// ---- in-LDS mixed-radix Stockham transform ---------------------------------

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham stage of radix R (a compile-time constant, so a butterfly's R
// values stay in registers): output j of sub-transform k gathers the R
// inputs j + q m, multiplies input q by the twiddle w^(k q n / (ns R)) (an
// index below n, as k < ns and q < R), and takes their length-R DFT.
template <int R>
__device__ __forceinline__ void md_fft_stage(const float2 *in, float2 *out, int nlines, int n, int ns,
                                             const float2 *tw, float sign) {
  const int m = n / R;
  const int twstride = n / (ns * R);
  const int dftstride = n / R;
  for (int idx = threadIdx.x; idx < nlines * m; idx += blockDim.x) {
    const int line = idx / m, j = idx - line * m;
    const int k = j % ns;
    float2 v[R];
    const float2 *src = in + line * n;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 x = src[j + q * m];
      if (q > 0 && k > 0) {
        float2 w = tw[k * q * twstride];
        w.y *= -sign;  // table holds exp(-2 pi i t/n); inverse conjugates
        x = cmul(x, w);
      }
      v[q] = x;
    }
    float2 *dst = out + line * n + (j / ns) * ns * R + k;
    if constexpr (R == 2) {
      dst[0] = make_float2(v[0].x + v[1].x, v[0].y + v[1].y);
      dst[ns] = make_float2(v[0].x - v[1].x, v[0].y - v[1].y);
    } else if constexpr (R == 4) {
      float2 s0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
      float2 s1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
      float2 s2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
      float2 s3 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
      // multiply s3 by sign*i: forward (-i), inverse (+i)
      float2 t3 = make_float2(-sign * s3.y, sign * s3.x);
      dst[0] = make_float2(s0.x + s2.x, s0.y + s2.y);
      dst[ns] = make_float2(s1.x + t3.x, s1.y + t3.y);
      dst[2 * ns] = make_float2(s0.x - s2.x, s0.y - s2.y);
      dst[3 * ns] = make_float2(s1.x - t3.x, s1.y - t3.y);
    } else {
#pragma unroll
      for (int qo = 0; qo < R; ++qo) {
        float2 acc = v[0];
#pragma unroll
        for (int q = 1; q < R; ++q) {
          float2 w = tw[((qo * q) % R) * dftstride];
          w.y *= -sign;
          float2 p = cmul(v[q], w);
          acc.x += p.x;
          acc.y += p.y;
        }
        dst[qo * ns] = acc;
      }
    }
  }
}

// `nlines` contiguous lines of length P.n in `a` (pitch P.n), transformed with
// sign -1 (forward) or +1 (inverse); returns the buffer holding the result.
// Each stage dispatches on its radix to the specialised butterfly.
__device__ __forceinline__ float2 *md_lds_fft(float2 *a, float2 *b, int nlines, const md_fft_plan &P,
                                              const float2 *tw_all, float sign) {
  const int n = P.n;
  const float2 *tw = tw_all + P.tw;
  int ns = 1;
  float2 *in = a, *out = b;
  for (int s = 0; s < P.nstages; ++s) {
    const int r = P.radix[s];
    switch (r) {
      case 2: md_fft_stage<2>(in, out, nlines, n, ns, tw, sign); break;
      case 3: md_fft_stage<3>(in, out, nlines, n, ns, tw, sign); break;
      case 4: md_fft_stage<4>(in, out, nlines, n, ns, tw, sign); break;
      case 5: md_fft_stage<5>(in, out, nlines, n, ns, tw, sign); break;
      case 7: md_fft_stage<7>(in, out, nlines, n, ns, tw, sign); break;
      case 11: md_fft_stage<11>(in, out, nlines, n, ns, tw, sign); break;
      default: md_fft_stage<13>(in, out, nlines, n, ns, tw, sign); break;
    }
    __syncthreads();
    float2 *t = in;
    in = out;
    out = t;
    ns *= r;
  }
  return in;
}

// z pass, forward: int64 rows -> f32 -> complex FFT -> first nz/2+1 values.
// Clears the integer grid behind it.  Block = PME_BATCH rows.
__global__ void __launch_bounds__(256) k_pme_z_forward(const md_pme_table *__restrict__ T, md_bmap map) {
  extern __shared__ float2 lds[];
  int blk;
  const md_pme_table &E = md_pme_entry(T, map, blk);
  const md_pme_args &P = E.P;
  int64_t *grid = E.grid;
  float2 *cgrid = E.cgrid;
  int rows = P.nx * P.ny;
  int r0 = blk * PME_BATCH;
  int nl = min(PME_BATCH, rows - r0);
  float2 *a = lds, *b = lds + PME_BATCH * P.nz;
  for (int idx = threadIdx.x; idx < nl * P.nz; idx += blockDim.x) {
    int l = idx / P.nz, z = idx - l * P.nz;
    size_t at = (size_t)(r0 + l) * P.nz + z;
    int64_t v = grid[at];
    grid[at] = 0;
    a[l * P.nz + z] = make_float2((float)((double)v * 2.3283064365386963e-10), 0.0f);
  }
  __syncthreads();
  float2 *res = md_lds_fft(a, b, nl, P.pz, P.tw, -1.0f);
  for (int idx = threadIdx.x; idx < nl * P.nzc; idx += blockDim.x) {
    int l = idx / P.nzc, z = idx - l * P.nzc;
    cgrid[(size_t)(r0 + l) * P.nzc + z] = res[l * P.nz + z];
  }
}

// y pass: lines along y at fixed (x, kz); a block takes PME_BATCH
// consecutive kz at one x, so global reads are contiguous runs.
__global__ void __launch_bounds__(256) k_pme_y(const md_pme_table *__restrict__ T, md_bmap map, float sign) {
  extern __shared__ float2 lds[];
  int blk;
  const md_pme_table &E = md_pme_entry(T, map, blk);
  const md_pme_args &P = E.P;
  float2 *cgrid = E.cgrid;
  int kzb = (P.nzc + PME_BATCH - 1) / PME_BATCH;
  int x = blk / kzb;
  int z0 = (blk - x * kzb) * PME_BATCH;
  int nl = min(PME_BATCH, P.nzc - z0);
  float2 *a = lds, *b = lds + PME_BATCH * P.ny;
  for (int idx = threadIdx.x; idx < nl * P.ny; idx += blockDim.x) {
    int y = idx / nl, l = idx - y * nl;
    a[l * P.ny + y] = cgrid[((size_t)x * P.ny + y) * P.nzc + z0 + l];
  }
  __syncthreads();
  float2 *res = md_lds_fft(a, b, nl, P.py, P.tw, sign);
  for (int idx = threadIdx.x; idx < nl * P.ny; idx += blockDim.x) {
    int y = idx / nl, l = idx - y * nl;
    cgrid[((size_t)x * P.ny + y) * P.nzc + z0 + l] = res[l * P.ny + y];
  }
}

// x pass with the convolution: forward transform along x, multiply by the
// Ewald influence function (accumulating the energy), inverse transform.
template <bool ENERGY>
__global__ void __launch_bounds__(256) k_pme_x_convolve(const md_pme_table *__restrict__ T, md_bmap map) {
  extern __shared__ float2 lds[];
  int blk;
  const md_pme_table &E = md_pme_entry(T, map, blk);
  const md_pme_args &P = E.P;
  float2 *cgrid = E.cgrid;
  int64_t *energy = E.energy;
  int kzb = (P.nzc + PME_BATCH - 1) / PME_BATCH;
  int y = blk / kzb;
  int z0 = (blk - y * kzb) * PME_BATCH;
  int nl = min(PME_BATCH, P.nzc - z0);
  float2 *a = lds, *b = lds + PME_BATCH * P.nx;
  size_t xstride = (size_t)P.ny * P.nzc;
  for (int idx = threadIdx.x; idx < nl * P.nx; idx += blockDim.x) {
    int x = idx / nl, l = idx - x * nl;
    a[l * P.nx + x] = cgrid[x * xstride + (size_t)y * P.nzc + z0 + l];
  }
  __syncthreads();
  float2 *res = md_lds_fft(a, b, nl, P.px, P.tw, -1.0f);
  float e = 0.0f;
  int mb = y <= P.ny / 2 ? y : y - P.ny;
  const float two_pi = 6.283185307179586f;
  for (int idx = threadIdx.x; idx < nl * P.nx; idx += blockDim.x) {
    int l = idx / P.nx, x = idx - l * P.nx;
    int kz = z0 + l;
    int ma = x <= P.nx / 2 ? x : x - P.nx;
    int mc = kz;
    float2 v = res[l * P.nx + x];
    float2 outv = make_float2(0.0f, 0.0f);
    if (!(x == 0 && y == 0 && kz == 0)) {
      float kx = two_pi * (ma * P.invf[0]);
      float ky = two_pi * (ma * P.invf[3] + mb * P.invf[4]);
      float kzv = two_pi * (ma * P.invf[6] + mb * P.invf[7] + mc * P.invf[8]);
      float k2 = kx * kx + ky * ky + kzv * kzv;
      float mod = P.mod_x[x] * P.mod_y[y] * P.mod_z[kz];
      float factor = P.prefactor * __expf(k2 * P.exp_scale) / (k2 * mod);
      if (ENERGY) {
        // the half grid stands for both k and -k except on the kz = 0 and
        // kz = nz/2 planes, which are their own partners
        float weight = (kz == 0 || (2 * kz == P.nz)) ? 1.0f : 2.0f;
        e += weight * factor * (v.x * v.x + v.y * v.y);
      }
      outv = make_float2(2.0f * factor * v.x, 2.0f * factor * v.y);
    }
    res[l * P.nx + x] = outv;
  }
  __syncthreads();
  float2 *other = res == a ? b : a;
  float2 *res2 = md_lds_fft(res, other, nl, P.px, P.tw, 1.0f);
  for (int idx = threadIdx.x; idx < nl * P.nx; idx += blockDim.x) {
    int x = idx / nl, l = idx - x * nl;
    cgrid[x * xstride + (size_t)y * P.nzc + z0 + l] = res2[l * P.nx + x];
  }
  if (ENERGY) {
    int64_t f = md_wave_sum_i64(md_to_fixed(e));
    if ((threadIdx.x & 63) == 0) md_atomic_add_i64(energy + MD_E_RECIPROCAL, f);
  }
}

// z pass, inverse: Hermitian half line -> full line -> inverse FFT -> real.
__global__ void __launch_bounds__(256) k_pme_z_inverse(const md_pme_table *__restrict__ T, md_bmap map) {
  extern __shared__ float2 lds[];
  int blk;
  const md_pme_table &E = md_pme_entry(T, map, blk);
  const md_pme_args &P = E.P;
  const float2 *cgrid = E.cgrid;
  float *phi = E.phi;
  int rows = P.nx * P.ny;
  int r0 = blk * PME_BATCH;
  int nl = min(PME_BATCH, rows - r0);
  float2 *a = lds, *b = lds + PME_BATCH * P.nz;
  for (int idx = threadIdx.x; idx < nl * P.nz; idx += blockDim.x) {
    int l = idx / P.nz, z = idx - l * P.nz;
    float2 v;
    if (z < P.nzc) {
      v = cgrid[(size_t)(r0 + l) * P.nzc + z];
    } else {
      // x and y are back in real space, so each row is the z transform of a
      // real line: g(x, y, -kz) = conj(g(x, y, kz)), same row
      float2 w = cgrid[(size_t)(r0 + l) * P.nzc + (P.nz - z)];
      v = make_float2(w.x, -w.y);
    }
    a[l * P.nz + z] = v;
  }
  __syncthreads();
  float2 *res = md_lds_fft(a, b, nl, P.pz, P.tw, 1.0f);
  for (int idx = threadIdx.x; idx < nl * P.nz; idx += blockDim.x) {
    int l = idx / P.nz, z = idx - l * P.nz;
    phi[(size_t)(r0 + l) * P.nz + z] = res[l * P.nz + z].x;
  }
}

// The z and y passes of one x-plane in one workgroup, the plane kept in LDS
// between them (meshes whose plane fits: 2 ny nz complex values).  The same
// line transforms as k_pme_z_forward then k_pme_y, with the same inputs, so
// the same bits; one launch and one trip through memory fewer.
__global__ void __launch_bounds__(256) k_pme_yz_forward(const md_pme_table *__restrict__ T, md_bmap map) {
  extern __shared__ float2 lds[];
  int x;
  const md_pme_table &E = md_pme_entry(T, map, x);
  const md_pme_args &P = E.P;
  int64_t *grid = E.grid;
  float2 *cgrid = E.cgrid;
  const int ny = P.ny, nz = P.nz, nzc = P.nzc;
  float2 *a = lds, *b = lds + ny * nz;
  for (int idx = threadIdx.x; idx < ny * nz; idx += blockDim.x) {
    const int y = idx / nz, z = idx - y * nz;
    const size_t at = ((size_t)x * ny + y) * nz + z;
    const int64_t v = grid[at];
    grid[at] = 0;
    a[y * nz + z] = make_float2((float)((double)v * 2.3283064365386963e-10), 0.0f);
  }
  __syncthreads();
  float2 *res = md_lds_fft(a, b, ny, P.pz, P.tw, -1.0f);  // [y][kz]
  float2 *other = res == a ? b : a;
  for (int idx = threadIdx.x; idx < nzc * ny; idx += blockDim.x) {
    const int kz = idx / ny, y = idx - kz * ny;
    other[kz * ny + y] = res[y * nz + kz];
  }
  __syncthreads();
  float2 *res2 = md_lds_fft(other, res, nzc, P.py, P.tw, -1.0f);  // [kz][ky]
  for (int idx = threadIdx.x; idx < ny * nzc; idx += blockDim.x) {
    const int ky = idx / nzc, kz = idx - ky * nzc;
    cgrid[((size_t)x * ny + ky) * nzc + kz] = res2[kz * ny + ky];
  }
}

// The inverse y and z passes of one x-plane (k_pme_y with +1, then
// k_pme_z_inverse), the plane kept in LDS between them.
__global__ void __launch_bounds__(256) k_pme_yz_inverse(const md_pme_table *__restrict__ T, md_bmap map) {
  extern __shared__ float2 lds[];
  int x;
  const md_pme_table &E = md_pme_entry(T, map, x);
  const md_pme_args &P = E.P;
  const float2 *cgrid = E.cgrid;
  float *phi = E.phi;
  const int ny = P.ny, nz = P.nz, nzc = P.nzc;
  float2 *a = lds, *b = lds + ny * nz;
  for (int idx = threadIdx.x; idx < nzc * ny; idx += blockDim.x) {
    const int kz = idx / ny, y = idx - kz * ny;
    a[kz * ny + y] = cgrid[((size_t)x * ny + y) * nzc + kz];
  }
  __syncthreads();
  float2 *res = md_lds_fft(a, b, nzc, P.py, P.tw, 1.0f);  // [kz][y]
  float2 *other = res == a ? b : a;
  for (int idx = threadIdx.x; idx < ny * nz; idx += blockDim.x) {
    const int y = idx / nz, z = idx - y * nz;
    float2 v;
    if (z < nzc) {
      v = res[z * ny + y];
    } else {
      // x and y are back in real space: g(x, y, -kz) = conj(g(x, y, kz))
      const float2 w = res[(nz - z) * ny + y];
      v = make_float2(w.x, -w.y);
    }
    other[y * nz + z] = v;
  }
  __syncthreads();
  float2 *res2 = md_lds_fft(other, res, ny, P.pz, P.tw, 1.0f);  // [y][z]
  for (int idx = threadIdx.x; idx < ny * nz; idx += blockDim.x) {
    const int y = idx / nz, z = idx - y * nz;
    phi[((size_t)x * ny + y) * nz + z] = res2[y * nz + z].x;
  }
}

// interpolate dE/dQ back onto the atoms: one thread per slot of the
// nonbonded list's spatial order, blocks laid out XCD by XCD, so the atoms
// of one XCD's blocks read one region of the mesh through one L2
__global__ void k_pme_gather(const md_pme_table *__restrict__ T, md_bmap map) {
  int blk;
  const md_pme_table &E = T[md_locate(map, md_xcd_block(blockIdx.x, gridDim.x), blk)];
  const md_pme_args &P = E.P;
  const float *phi = E.phi;
  int64_t *force = E.force;
  const int np = E.np;
  const int s = blk * blockDim.x + threadIdx.x;
  if (s >= E.nslots) return;
  const int i = E.slot_atom[s];
  if (i < 0) return;
  float q = E.param[i].w;
  if (q == 0.0f) return;
  int base[3];
  float w[3];
  md_mesh_coord(E.pos, np, i, P.inv, P.nx, P.ny, P.nz, base, w);
  float tx[5], ty[5], tz[5], dx[5], dy[5], dz[5];
  md_bspline5(w[0], tx, dx);
  md_bspline5(w[1], ty, dy);
  md_bspline5(w[2], tz, dz);
  int gz[5];
  for (int c = 0; c < 5; ++c) gz[c] = md_wrap(base[2] + c, P.nz);
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  for (int a = 0; a < 5; ++a) {
    int gx = md_wrap(base[0] + a, P.nx);
    for (int b = 0; b < 5; ++b) {
      int gy = md_wrap(base[1] + b, P.ny);
      const float *row = phi + ((size_t)gx * P.ny + gy) * P.nz;
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
      for (int c = 0; c < 5; ++c) {
        float p = row[gz[c]];
        sz += p * dz[c];
        sx += p * tz[c];
      }
      fx += dx[a] * ty[b] * sx;
      fy += tx[a] * dy[b] * sx;
      fz += tx[a] * ty[b] * sz;
    }
  }
  // d/df chain rule back to Cartesian (src/reference/forces/pme.ts:246-260)
  float dfx = -q * fx * P.nx, dfy = -q * fy * P.ny, dfz = -q * fz * P.nz;
  const float *M = P.invf;
  float Fx = dfx * M[0] + dfy * M[1] + dfz * M[2];
  float Fy = dfx * M[3] + dfy * M[4] + dfz * M[5];
  float Fz = dfx * M[6] + dfy * M[7] + dfz * M[8];
  // one thread per atom here (a slot holds one atom), and kernels on one
  // stream run in order: a plain read-add-write, no atomics at memory
  force[i] += md_to_fixed(Fx);
  force[np + i] += md_to_fixed(Fy);
  force[2 * np + i] += md_to_fixed(Fz);
}

// ---- host -------------------------------------------------------------------

struct md_pme_host {
  md_fft_plan px, py, pz;
  float *mod = nullptr;  // x | y | z
  float2 *tw = nullptr;
};

static md_fft_plan make_plan(int n, int tw_offset) {
  md_fft_plan p{};
  p.n = n;
  p.tw = tw_offset;
  int r = n;
  // radix 4 first, then 2, 3, 5, 7, then any remaining prime (direct DFT)
  while (r % 4 == 0) {
    if (p.nstages >= PME_MAX_STAGES) break;
    p.radix[p.nstages++] = 4;
    r /= 4;
  }
  for (int f : {2, 3, 5, 7, 11, 13})
    while (r % f == 0) {
      if (p.nstages >= PME_MAX_STAGES) throw md_error(MD_ERR_NOT_SMOOTH, "mesh size has too many factors");
      p.radix[p.nstages++] = f;
      r /= f;
    }
  if (r != 1)
    throw md_error(MD_ERR_NOT_SMOOTH, std::to_string(n) + " has a prime factor above 13 (" + std::to_string(r) +
                                          " is left); PME mesh sizes must be 13-smooth");
  return p;
}

// |b(m)|^2 of the order-5 spline with the 1e-7 neighbour guard
// (src/reference/forces/bspline.ts:102-132), as its reciprocal's factor
static std::vector<double> moduli(int n) {
  const double knots[5] = {1.0 / 24, 11.0 / 24, 11.0 / 24, 1.0 / 24, 0.0};
  std::vector<double> m(n);
  for (int k = 0; k < n; ++k) {
    double re = 0, im = 0;
    for (int j = 0; j < 5 && j < n; ++j) {
      double a = 2.0 * M_PI * (double)k * j / n;
      re += knots[j] * std::cos(a);
      im += knots[j] * std::sin(a);
    }
    m[k] = re * re + im * im;
  }
  for (int k = 0; k < n; ++k)
    if (m[k] < 1e-7) m[k] = 0.5 * (m[(k - 1 + n) % n] + m[(k + 1) % n]);
  return m;
}

static md_pme_host &pme_host(md_system *s) { return *reinterpret_cast<md_pme_host *>(s->pme_ext); }

void md_pme_setup(md_system *s) {
  s->pme_ext = new md_pme_host();
  auto &H = pme_host(s);
  int nx = s->grid[0], ny = s->grid[1], nz = s->grid[2];
  s->pme_nx = nx;
  s->pme_ny = ny;
  s->pme_nz = nz;
  s->pme_nzc = nz / 2 + 1;
  int lens[3] = {nx, ny, nz};
  std::vector<float2> tw;
  int off[3];
  for (int d = 0; d < 3; ++d) {
    off[d] = (int)tw.size();
    for (int t = 0; t < lens[d]; ++t) {
      double a = -2.0 * M_PI * t / lens[d];
      tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
    }
  }
  H.px = make_plan(nx, off[0]);
  H.py = make_plan(ny, off[1]);
  H.pz = make_plan(nz, off[2]);
  std::vector<float> mod;
  for (int d = 0; d < 3; ++d)
    for (double v : moduli(lens[d])) mod.push_back((float)v);
  H.mod = md_alloc<float>(s, mod.size());
  H.tw = md_alloc<float2>(s, tw.size());
  MD_HIP(hipMemcpyAsync(H.mod, mod.data(), sizeof(float) * mod.size(), hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipMemcpyAsync(H.tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice, s->stream));
  size_t g = (size_t)nx * ny * nz;
  s->pme_grid = md_alloc<int64_t>(s, g);
  s->pme_cgrid = md_alloc<float2>(s, (size_t)nx * ny * s->pme_nzc);
  s->pme_phi = md_alloc<float>(s, g);
  MD_HIP(hipStreamSynchronize(s->stream));
}

void md_pme_free(md_system *s) {
  delete reinterpret_cast<md_pme_host *>(s->pme_ext);
  s->pme_ext = nullptr;
}

// The mesh's box-dependent arguments, from a reduced box b and its
// constants: the same on the host (table set-up) and on the device (a box
// changed by the Monte Carlo barostat).
__host__ __device__ inline void md_pme_box_fields(md_pme_args &P, const double *b, const md_box &d) {
#pragma clang fp contract(off)
  for (int k = 0; k < 9; ++k) {
    P.inv[k] = d.invd[k];
    P.invf[k] = d.inv[k];
  }
  P.prefactor = (float)(MD_ONE_4PI_EPS0 * 2.0 * M_PI / md_box_volume(b));
}

static md_pme_args pme_args(md_system *s) {
  auto &H = pme_host(s);
  md_pme_args P;
  P.nx = s->pme_nx;
  P.ny = s->pme_ny;
  P.nz = s->pme_nz;
  P.nzc = s->pme_nzc;
  P.px = H.px;
  P.py = H.py;
  P.pz = H.pz;
  P.tw = H.tw;
  P.mod_x = H.mod;
  P.mod_y = H.mod + P.nx;
  P.mod_z = H.mod + P.nx + P.ny;
  md_pme_box_fields(P, s->box, s->dbox);
  P.exp_scale = (float)(-1.0 / (4.0 * s->alpha * s->alpha));
  return P;
}

static md_pme_table pme_table(md_system *s) {
  md_pme_table E;
  E.P = pme_args(s);
  E.n = s->n;
  E.np = s->np;
  E.nslots = s->nb.nslots;
  E.pos = s->pos;
  E.param = s->atom_param;
  E.slot_atom = s->slot_atom;
  E.grid = s->pme_grid;
  E.cgrid = s->pme_cgrid;
  E.phi = s->pme_phi;
  E.force = s->force;
  E.energy = s->energy;
  return E;
}

struct md_pme_group {
  md_pme_table *table = nullptr;
  md_launch spread, zrows, ylines, xlines, gather, planes;
  size_t lz = 0, ly = 0, lx = 0;  // dynamic LDS of the three passes
  size_t lyz = 0;                 // dynamic LDS of a whole x-plane
  bool fused_yz = false;          // the y and z passes per x-plane in one launch
};

static md_pme_group &pg(md_group *g) { return *reinterpret_cast<md_pme_group *>(g->pme); }

void md_pme_group_setup(md_group *g) {
  auto *G = new md_pme_group();
  g->pme = G;
  std::vector<int> spread, zrows, ylines, xlines, gather, planes;
  int mx = 0, my = 0, mz = 0;
  for (md_system *s : g->reps) {
    bool on = s->has_nonbonded;
    int nx = s->pme_nx, ny = s->pme_ny, nz = s->pme_nz, kzb = (s->pme_nzc + PME_BATCH - 1) / PME_BATCH;
    spread.push_back(on ? s->nb.nslots / 64 : 0);
    zrows.push_back(on ? (nx * ny + PME_BATCH - 1) / PME_BATCH : 0);
    ylines.push_back(on ? nx * kzb : 0);
    xlines.push_back(on ? ny * kzb : 0);
    gather.push_back(on ? (s->nb.nslots + 255) / 256 : 0);
    planes.push_back(on ? nx : 0);
    if (on) {
      mx = std::max(mx, nx);
      my = std::max(my, ny);
      mz = std::max(mz, nz);
    }
  }
  G->spread = md_make_launch(g, spread);
  G->zrows = md_make_launch(g, zrows);
  G->ylines = md_make_launch(g, ylines);
  G->xlines = md_make_launch(g, xlines);
  G->gather = md_make_launch(g, gather);
  G->planes = md_make_launch(g, planes);
  G->lyz = sizeof(float2) * 2 * (size_t)my * mz;
  G->fused_yz = G->lyz <= 64 * 1024;  // up to 64^3 meshes
  G->lz = sizeof(float2) * 2 * PME_BATCH * mz;
  G->ly = sizeof(float2) * 2 * PME_BATCH * my;
  G->lx = sizeof(float2) * 2 * PME_BATCH * mx;
  size_t lds = std::max(G->lz, std::max(G->ly, G->lx));
  if (lds > 64 * 1024) {
    for (auto *k : {(const void *)k_pme_z_forward, (const void *)k_pme_y, (const void *)k_pme_x_convolve<true>,
                    (const void *)k_pme_x_convolve<false>, (const void *)k_pme_z_inverse})
      MD_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  G->table = md_group_alloc<md_pme_table>(g, g->reps.size());
  md_pme_group_refresh(g);
}

void md_pme_group_refresh(md_group *g) {
  std::vector<md_pme_table> host;
  for (md_system *s : g->reps) host.push_back(s->has_nonbonded ? pme_table(s) : md_pme_table{});
  MD_HIP(hipMemcpyAsync(pg(g).table, host.data(), sizeof(md_pme_table) * host.size(), hipMemcpyHostToDevice,
                        g->stream));
  MD_HIP(hipStreamSynchronize(g->stream));
}

// Replica 0's box-dependent mesh arguments from a box in device memory (a
// barostat move).
__global__ void k_pme_box(md_pme_table *T, const double *box) {
  md_box d{};
  bool bad = false;
  md_box_from_matrix(box, d, bad);
  md_pme_box_fields(T[0].P, box, d);
}

void md_pme_box_from_device(md_group *g, const double *box) {
  k_pme_box<<<1, 1, 0, g->stream>>>(pg(g).table, box);
  MD_HIP(hipGetLastError());
}

void md_pme_group_free(md_group *g) {
  delete &pg(g);
  g->pme = nullptr;
}

void md_pme(md_group *g, bool energy) {
  auto &G = pg(g);
  if (G.spread.total == 0) return;
  hipStream_t st = g->stream;
  const md_pme_table *T = G.table;
  k_pme_spread<<<G.spread.total, MD_SPREAD_THREADS, 0, st>>>(T, G.spread.map);
  md_mark(g, MD_PHASE_FFT);
  if (G.fused_yz) {
    k_pme_yz_forward<<<G.planes.total, 256, G.lyz, st>>>(T, G.planes.map);
  } else {
    k_pme_z_forward<<<G.zrows.total, 256, G.lz, st>>>(T, G.zrows.map);
    k_pme_y<<<G.ylines.total, 256, G.ly, st>>>(T, G.ylines.map, -1.0f);
  }
  if (energy)
    k_pme_x_convolve<true><<<G.xlines.total, 256, G.lx, st>>>(T, G.xlines.map);
  else
    k_pme_x_convolve<false><<<G.xlines.total, 256, G.lx, st>>>(T, G.xlines.map);
  if (G.fused_yz) {
    k_pme_yz_inverse<<<G.planes.total, 256, G.lyz, st>>>(T, G.planes.map);
  } else {
    k_pme_y<<<G.ylines.total, 256, G.ly, st>>>(T, G.ylines.map, 1.0f);
    k_pme_z_inverse<<<G.zrows.total, 256, G.lz, st>>>(T, G.zrows.map);
  }
  md_mark(g, MD_PHASE_GATHER);
  k_pme_gather<<<G.gather.total, 256, 0, st>>>(T, G.gather.map);
  MD_HIP(hipGetLastError());
}
