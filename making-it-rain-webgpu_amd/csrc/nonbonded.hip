// nonbonded.hip -- direct-space LJ + erfc-Ewald for gfx950.
//
// Replaces the upstream WGSL passes `wrap`, `clear/assign/scan/scatter/
// sort_cells` (src/gpu/kernels/cell-list.wgsl:139-331), the Verlet-list
// build (verlet.wgsl:46-101, nonbonded-pairs.wgsl:417-497) and the pair
// kernels `direct_pairs_list` / `direct_pairs` (nonbonded-pairs.wgsl:237-391).
// The physics is the reference pair term, src/reference/forces/nonbonded.ts
// :279-321: E = K q_i q_j erfc(alpha r)/r + 4 eps ((s/r)^12 - (s/r)^6), with
// sigma the arithmetic and epsilon the geometric mean, for non-excluded pairs
// inside the cutoff.
//
// Layout (designed for 64-lane wavefronts, not a translation of the WGSL):
//   * Atoms are bucketed by an exact integer cell sort into cells ~0.3 nm wide
//     whose order follows a 3D Hilbert curve; the sorted order is cut into
//     clusters of 8 atoms and super-clusters of 64 (= one wavefront).
//   * The list holds, per super-cluster S, the 8-atom j-clusters J (J >= the
//     i-cluster, each unordered pair once) whose bounding boxes come within
//     cutoff + skin, with an 8-bit mask of which of S's 8 i-clusters interact,
//     the periodic shift of J into S's frame, and 8x8 exclusion masks.
//   * The pair kernel gives lane l the pair (i = l/8, j = l%8) of each
//     (i-cluster, j-cluster) block.  i-forces stay in registers across a
//     segment of entries; j-forces are reduced across the 8 lanes sharing a j
//     with cross-lane shuffles once per entry.  No LDS, no per-pair shuffle.
//   * Positions reach the kernel as int32 coordinates in 2^-26 nm from the
//     super-cluster's fixed-point frame origin, so a pair difference is exact
//     integer arithmetic, rounded once when it becomes f32 (the reference's
//     difference-then-convert rule, fixed64.wgsl:166-216).
//   * Every f32 partial sum covers one list entry or one unit of 8 entries and
//     is converted to Q32.32 before it is added anywhere, so two identical
//     runs are bit-identical whatever order the hardware schedules waves in.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "engine.hpp"

// erfc(z) = exp(-z^2) P(t), t = 1/(1 + z/2): a degree-8 fit of erfcx(z) on
// [0, 4] with minimax relative error ~3.4e-7 in f32 (fit_erfc.py beside this
// file regenerates the coefficients).
#define MD_ERFC_C0 -0.00045665048f
#define MD_ERFC_C1 0.28968847f
#define MD_ERFC_C2 0.22711307f
#define MD_ERFC_C3 0.47317088f
#define MD_ERFC_C4 -0.4014762f
#define MD_ERFC_C5 1.010818f
#define MD_ERFC_C6 -0.8866913f
#define MD_ERFC_C7 0.33737373f
#define MD_ERFC_C8 -0.049540002f

__device__ __forceinline__ float md_erfcx_poly(float t) {
  float p = MD_ERFC_C8;
  p = fmaf(p, t, MD_ERFC_C7);
  p = fmaf(p, t, MD_ERFC_C6);
  p = fmaf(p, t, MD_ERFC_C5);
  p = fmaf(p, t, MD_ERFC_C4);
  p = fmaf(p, t, MD_ERFC_C3);
  p = fmaf(p, t, MD_ERFC_C2);
  p = fmaf(p, t, MD_ERFC_C1);
  return fmaf(p, t, MD_ERFC_C0);
}

// ---------------------------------------------------------------------------
// Build: fold, bucket, rank
// ---------------------------------------------------------------------------

struct md_build_ctx {
  int n, np, ncells, gdim0, gdim1, gdim2;
  int nclusters, nsuper, nslots;
  int segment_capacity;
  int seg_len;  // list entries per pair-kernel work unit (MD_SEGMENT or a multiple)
  int entries_per_super;  // the entry region of each super-cluster
  float list_cutoff2;
  float half_min_width;
  float width[3];  // perpendicular widths along a, b, c
  float fmax[3];   // bound on |fractional coordinate| of a reduced difference
  float list_cutoff;
  md_box box;
};

// image shells enumerated per axis at most (9 x 9 x 9 images); only boxes
// not much wider than the cutoff with very sparse clusters come near it
#define MD_MAX_IMAGE_SHELLS 4
// candidate j super-clusters a list wave keeps in LDS (beyond: scan all)
#define MD_LIST_CANDIDATES 1024
// wavefronts that share one super-cluster's entries in the mask kernel, and
// wavefronts per workgroup (a replica that did not rebuild dispatches
// MD_MASK_WAVES / MD_MASK_GROUP workgroups per super-cluster that return at once)
#define MD_MASK_WAVES 16
// entry meta word: bits 0-7 the i-clusters of S that meet J, bits 8-15 those
// that are J itself at the zero image, and MD_META_CLEAN when no active block
// has an excluded, padding or lower-triangle pair (its masks are all ones and
// the pair kernel does not read them)
#define MD_META_CLEAN (1 << 16)
#define MD_MASK_GROUP 4

struct md_list_ptrs {
  const int64_t *cluster_center;
  const float4 *cluster_half;
  const int64_t *super_origin;
  const float *super_radius;  // half diagonal of each super-cluster's bounding box (nm)
  int *super_count;   // [nsuper] entries per super-cluster (in its region)
  int *super_segs;    // [nsuper+1] segments per super-cluster, then starts
  int *entry_j;
  int *entry_imask;
  int4 *entry_shift;    // J's frame origin in S's frame, MD_FRAME_UNIT units
  int4 *segments;     // (S, begin, end, 0)
  uint64_t *entry_excl;  // [capacity][8]
  // the transposed list: for j-cluster J, the entries naming it are
  // j_entries[j_start[J] .. j_start[J+1]) (any order); j_count counts them
  int *j_count, *j_start, *j_entries;
};

struct md_pair_args {
  const int4 *segments;
  const int *entry_j;
  const int *entry_imask;
  const int4 *entry_shift;
  const uint64_t *entry_excl;
  const int4 *slot_xyzq;  // frame coordinates (MD_FRAME_UNIT units) and the charge's bits
  const float2 *slot_lj;
  const int *totals;
  // Q32.32 forces by slot, [3][nslots]: a cluster's 8 atoms are 8 adjacent
  // words, so the 8 or 64 lanes of one atomic instruction hit one or a few
  // cache lines instead of as many random atoms; k_nb_forces moves them to
  // the atoms
  int64_t *slot_force;
  // per list entry, the f32 force on each of J's 8 atoms, [entry][x|y|z][8]:
  // one plain store per entry instead of 24 atomics; k_nb_forces sums them
  float *j_force;
  // each cluster's current bounding box in its frame (frame units; a cluster
  // of padding only holds an inverted box), written by k_nb_pack
  const int4 *cluster_lo, *cluster_hi;
  int64_t *energy;
  int64_t *census;  // in-cutoff pairs (md_control)
  int64_t *blocks;  // 8x8 cluster-pair blocks evaluated (md_control)
  int nslots, nclusters;
  // in the pair kernel's frame units (see md_pair): cutoff^2 in 2^-52 nm^2;
  // alpha/2 and 2 alpha / sqrt(pi) times 2^-26; -alpha^2 / ln 2 times 2^-52
  float cutoff2, half_alpha, beta, exp_scale;
  // cutoff^2 (1 + 1e-5) in 2^-52 nm^2: a block whose boxes are this far apart
  // has no pair within the cutoff
  float prune2;
  // the Coulomb polynomials of md_ewald_fit in frame units (v = fit_scale r2
  // - 1 with r2 in 2^-52 nm^2; fit_force carries 2^-78 and fit_energy 2^-26,
  // the units of 1/r^3 and 1/r), used when ewald_fit is set
  int ewald_fit;
  float fit_scale;
  float fit_force[MD_EWALD_TERMS], fit_energy[MD_EWALD_TERMS];
};

// Everything the nonbonded kernels of one replica read: its table entry.
struct md_nb_args {
  md_build_ctx c;
  md_list_ptrs L;
  md_pair_args P;
  const int64_t *pos;
  const int *cell_rank;
  int *cell_count, *cell_start, *cell_cursor, *atom_cell, *cell_member;
  int *slot_atom, *atom_slot;
  int64_t *folded, *build_pos, *slot_offset, *cluster_center, *super_origin;
  float *super_radius;
  float2 *slot_lj;
  int4 *slot_xyzq;
  float4 *cluster_half;
  const float4 *atom_param;
  const int *excl_start, *excl_list;
  const uint64_t *excl_window;
  int64_t *force;  // the atoms' Q32.32 forces, [3][np]
  int4 *cluster_lo, *cluster_hi;
  int *totals;
  md_control *ctl;
};

// this block's replica entry and its block index inside the replica
__device__ __forceinline__ const md_nb_args &md_nb_entry(const md_nb_args *T, const md_bmap &map, int &blk) {
  return T[md_locate(map, blockIdx.x, blk)];
}

// Pair-kernel frame coordinates: int32 in units of 2^-26 nm (+/-32 nm), so a
// pair's difference is exact integer arithmetic and is rounded once, when it
// becomes f32 -- the reference's difference-then-convert rule (fixed64.wgsl
// :166-216) -- instead of carrying the rounding of two offsets up to a few nm
// long.  Overflow code 3 reports a coordinate outside the range (a diverged
// system).
#define MD_FRAME_SHIFT 6  // Q32.32 -> 2^-26 nm
#define MD_FRAME_UNIT 1.4901161193847656e-08f   // 2^-26 nm
#define MD_FRAME_UNIT2 2.220446049250313e-16f   // 2^-52 nm^2
// the pair kernel's per-atom scalings (see md_pair): charge (with sqrt K)
// 2^26, sigma/2 2^26, 2 sqrt(eps) 2^13
#define MD_PAIR_Q_SCALE 67108864.0f
#define MD_PAIR_S_SCALE 67108864.0f
#define MD_PAIR_E_SCALE 8192.0f

__device__ __forceinline__ int md_frame_coord(int64_t v, bool &bad) {
  int64_t u = (v + (int64_t(1) << (MD_FRAME_SHIFT - 1))) >> MD_FRAME_SHIFT;
  bad |= u > INT32_MAX || u < INT32_MIN;
  return (int)u;
}

// Fold each atom into the cell and bucket it (cell-list.wgsl:139-250).  The
// fractional coordinate is computed in f64 from the exact fixed-point
// position; the folding subtracts whole lattice vectors in Q32.32.
__global__ void k_nb_assign(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  if (!A.ctl->rebuild_active) return;
  const md_build_ctx &c = A.c;
  int i = blk * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  const int64_t *pos = A.pos;
  int *atom_cell = A.atom_cell;
  int64_t *folded = A.folded, *build_pos = A.build_pos;
  int64_t x = pos[i], y = pos[c.np + i], z = pos[2 * c.np + i];
  const double s = 2.3283064365386963e-10;
  double rx = (double)x * s, ry = (double)y * s, rz = (double)z * s;
  const double *M = c.box.invd;
  double fa = rx * M[0] + ry * M[3] + rz * M[6];
  double fb = ry * M[4] + rz * M[7];
  double fc = rz * M[8];
  double na = floor(fa), nb = floor(fb), nc = floor(fc);
  int64_t ia = (int64_t)na, ib = (int64_t)nb, ic = (int64_t)nc;
  int64_t fx = x - ia * c.box.fa[0] - ib * c.box.fb[0] - ic * c.box.fc[0];
  int64_t fy = y - ib * c.box.fb[1] - ic * c.box.fc[1];
  int64_t fz = z - ic * c.box.fc[2];
  double ua = fa - na, ub = fb - nb, uc = fc - nc;
  int ga = min((int)(ua * c.gdim0), c.gdim0 - 1);
  int gb = min((int)(ub * c.gdim1), c.gdim1 - 1);
  int gc = min((int)(uc * c.gdim2), c.gdim2 - 1);
  ga = max(ga, 0);
  gb = max(gb, 0);
  gc = max(gc, 0);
  int key = A.cell_rank[(ga * c.gdim1 + gb) * c.gdim2 + gc];
  atom_cell[i] = key;
  atomicAdd(&A.cell_count[key], 1);
  folded[i] = fx;
  folded[c.np + i] = fy;
  folded[2 * c.np + i] = fz;
  build_pos[i] = x;
  build_pos[c.np + i] = y;
  build_pos[2 * c.np + i] = z;
}

// Exclusive scan by one workgroup of 1024 threads over `count` values; the
// total lands at out[count].  `in` may alias `out`.  Zeroes `clear` if given.
__device__ void md_block_scan(const int *in, int *out, int count, int *clear) {
  __shared__ int partial[1024];
  int t = threadIdx.x;
  int per = (count + 1023) / 1024;
  int begin = min(t * per, count), end = min(begin + per, count);
  int sum = 0;
  for (int k = begin; k < end; ++k) sum += in[k];
  partial[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? partial[t - off] : 0;
    __syncthreads();
    partial[t] += v;
    __syncthreads();
  }
  int run = partial[t] - sum;
  for (int k = begin; k < end; ++k) {
    int v = in[k];
    out[k] = run;
    run += v;
    if (clear) clear[k] = 0;
  }
  if (t == 1023) out[count] = partial[1023];
}

// one workgroup per replica
__global__ void __launch_bounds__(1024) k_nb_scan_cells(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  md_control *ctl = A.ctl;
  if (!ctl->rebuild_active) return;
  int ncells = A.c.ncells;
  md_block_scan(A.cell_count, A.cell_start, ncells, nullptr);
  __syncthreads();
  for (int k = threadIdx.x; k < ncells; k += 1024) {
    A.cell_count[k] = 0;
    A.cell_cursor[k] = 0;
  }
  if (threadIdx.x == 0) {
    ctl->rebuilds += 1;
    ctl->max_disp2_bits[0] = 0;
    ctl->cluster_radius_bits = 0;
    ctl->super_radius_bits = 0;
  }
}

__global__ void k_nb_scatter(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  if (!A.ctl->rebuild_active) return;
  int i = blk * blockDim.x + threadIdx.x;
  if (i >= A.c.n) return;
  int key = A.atom_cell[i];
  int at = A.cell_start[key] + atomicAdd(&A.cell_cursor[key], 1);
  A.cell_member[at] = i;
}

// Deterministic order inside a cell: rank by atom index over the cell's
// members (the reference's `sort_cells`, cell-list.wgsl:318-331).
__global__ void k_nb_rank(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  if (!A.ctl->rebuild_active) return;
  int i = blk * blockDim.x + threadIdx.x;
  if (i >= A.c.n) return;
  int key = A.atom_cell[i];
  int b = A.cell_start[key], e = A.cell_start[key + 1];
  int rank = 0;
  for (int k = b; k < e; ++k) rank += A.cell_member[k] < i;
  int slot = b + rank;
  A.slot_atom[slot] = i;
  A.atom_slot[i] = slot;
}

// Per 8-atom cluster: bounding box of the folded build positions.  Per
// 64-atom super-cluster: a frame origin, and for each slot the fixed-point
// offset that takes a current position to f32 coordinates in that frame.
__global__ void k_nb_clusters(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  md_control *ctl = A.ctl;
  if (!ctl->rebuild_active) return;
  const md_build_ctx &c = A.c;
  const int *slot_atom = A.slot_atom;
  const int64_t *folded = A.folded, *build_pos = A.build_pos;
  const float4 *atom_param = A.atom_param;
  int64_t *cluster_center = A.cluster_center, *super_origin = A.super_origin, *slot_offset = A.slot_offset;
  float *super_radius = A.super_radius;
  float4 *cluster_half = A.cluster_half;
  float2 *slot_lj = A.slot_lj;
  // one thread per cluster; the 8 threads of a super-cluster cooperate
  int k = blk * blockDim.x + threadIdx.x;  // cluster index
  int lane = threadIdx.x & 63;
  bool live = k < c.nclusters;
  int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
  if (live) {
    for (int q = 0; q < MD_CLUSTER; ++q) {
      int a = slot_atom[k * MD_CLUSTER + q];
      if (a < 0) continue;
      for (int d = 0; d < 3; ++d) {
        int64_t v = folded[d * c.np + a];
        lo[d] = min(lo[d], v);
        hi[d] = max(hi[d], v);
      }
    }
  }
  int64_t cen[3];
  float half[3];
  for (int d = 0; d < 3; ++d) {
    cen[d] = live && lo[d] <= hi[d] ? lo[d] + ((hi[d] - lo[d]) >> 1) : 0;
    half[d] = live && lo[d] <= hi[d] ? md_from_fixed(hi[d] - lo[d]) * 0.5f : 0.0f;
  }
  float radius = sqrtf(half[0] * half[0] + half[1] * half[1] + half[2] * half[2]);
  if (live) {
    for (int d = 0; d < 3; ++d) cluster_center[d * c.nclusters + k] = cen[d];
    cluster_half[k] = make_float4(half[0], half[1], half[2], 0.0f);
    A.L.j_count[k] = 0;
  }
  // the largest cluster radius (non-negative floats order as their bits):
  // reduced over the wave first, one atomic per wave on the shared word
  // rather than one per cluster, which all queue on one address
  {
    float rmax = live ? radius : -1.0f;
    for (int off = 32; off > 0; off >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, off, 64));
    if (lane == 0 && rmax >= 0.0f) atomicMax(&ctl->cluster_radius_bits, __float_as_uint(rmax));
  }
  // super-cluster frame: the bounding box centre over its 8 clusters
  int64_t slo[3], shi[3];
  for (int d = 0; d < 3; ++d) {
    int64_t a = lo[d], b = hi[d];
    for (int off = 1; off < 8; off <<= 1) {
      a = min(a, (int64_t)__shfl_xor(a, off, 64));
      b = max(b, (int64_t)__shfl_xor(b, off, 64));
    }
    slo[d] = a;
    shi[d] = b;
  }
  int s = k / 8;
  int64_t origin[3];
  for (int d = 0; d < 3; ++d) origin[d] = slo[d] <= shi[d] ? slo[d] + ((shi[d] - slo[d]) >> 1) : 0;
  float srad = -1.0f;
  if (live && (lane & 7) == 0 && s < c.nsuper) {
    float r2 = 0.0f;
    for (int d = 0; d < 3; ++d) {
      super_origin[d * c.nsuper + s] = origin[d];
      float h = 0.5f * md_from_fixed(shi[d] - slo[d]);
      r2 += h * h;
    }
    // rounded up a little: the list test that uses it must stay conservative
    srad = sqrtf(r2) * 1.0001f + 1e-4f;
    super_radius[s] = srad;
  }
  // the largest super-cluster radius, one atomic per wave as above
  for (int off = 32; off > 0; off >>= 1) srad = fmaxf(srad, __shfl_xor(srad, off, 64));
  if (lane == 0 && srad >= 0.0f) atomicMax(&ctl->super_radius_bits, __float_as_uint(srad));
  if (!live) return;
  for (int q = 0; q < MD_CLUSTER; ++q) {
    int slot = k * MD_CLUSTER + q;
    int a = slot_atom[slot];
    if (a < 0) {
      for (int d = 0; d < 3; ++d) slot_offset[d * c.nslots + slot] = 0;
      slot_lj[slot] = make_float2(0.0f, 0.0f);
      continue;
    }
    for (int d = 0; d < 3; ++d)
      slot_offset[d * c.nslots + slot] = (build_pos[d * c.np + a] - folded[d * c.np + a]) + origin[d];
    float4 p = atom_param[a];
    slot_lj[slot] = make_float2(p.y, p.z);
  }
}

// AABB distance test between cluster boxes separated by d (nm, f32)
__device__ __forceinline__ bool md_boxes_within(float dx, float dy, float dz, float4 h1, float4 h2, float r2) {
  float ex = fmaxf(fabsf(dx) - (h1.x + h2.x), 0.0f);
  float ey = fmaxf(fabsf(dy) - (h1.y + h2.y), 0.0f);
  float ez = fmaxf(fabsf(dz) - (h1.z + h2.z), 0.0f);
  return ex * ex + ey * ey + ez * ez < r2;
}

// The candidate periodic images of cluster J relative to cluster c: the
// minimum image of the centre difference, and -- when the box is narrow
// compared with the list radius plus the cluster sizes -- the images within
// K[0] x K[1] x K[2] lattice steps of it.  Visits, in a fixed order, the image
// offsets (as lattice multiples added to the reduced difference) with a code
// that is 0 for the reduced image and negates with the offset.  For a cluster of the
// super-cluster itself (`own`) the images are centred on the unshifted
// cluster instead, so that code 0 is the zero shift and the codes of a
// shift and of its negative have opposite signs: a super-cluster that spans
// a periodic boundary would otherwise centre them elsewhere, and keeping
// the positive codes would then keep both of some +/- pairs.
//
// Images whose f32 estimate lies at or beyond `skip` (nm) are not visited:
// the estimate (the reduced difference plus the lattice steps, in f32) is
// within 1e-5 nm of the exact image for boxes within 32 nm, so with `skip`
// widened by 1e-3 nm past the caller's own exact test no image that test
// would keep is skipped, and the int64 image arithmetic runs only for the
// few images near enough to matter.
template <typename F>
__device__ void md_for_each_image(const md_build_ctx &c, int64_t dx, int64_t dy, int64_t dz, bool enumerate,
                                  bool own, const int *K, float skip, F &&visit) {
  if (!own) md_min_image_fixed(dx, dy, dz, c.box);
  if (!enumerate) {
    visit(dx, dy, dz, 0);
    return;
  }
  const int na = 2 * K[0] + 1, nb = 2 * K[1] + 1;
  const int centre = (K[2] * nb + K[1]) * na + K[0];
  const float fx0 = md_from_fixed(dx), fy0 = md_from_fixed(dy), fz0 = md_from_fixed(dz);
  const float skip2 = skip * skip;
  int code = 0;
  for (int kc = -K[2]; kc <= K[2]; ++kc)
    for (int kb = -K[1]; kb <= K[1]; ++kb)
      for (int ka = -K[0]; ka <= K[0]; ++ka, ++code) {
        const float gx = fmaf((float)kc, c.box.c[0], fmaf((float)kb, c.box.b[0], fmaf((float)ka, c.box.a[0], fx0)));
        const float gy = fmaf((float)kc, c.box.c[1], fmaf((float)kb, c.box.b[1], fy0));
        const float gz = fmaf((float)kc, c.box.c[2], fz0);
        if (fmaf(gx, gx, fmaf(gy, gy, gz * gz)) >= skip2) continue;
        int64_t ex = dx + ka * c.box.fa[0] + kb * c.box.fb[0] + kc * c.box.fc[0];
        int64_t ey = dy + kb * c.box.fb[1] + kc * c.box.fc[1];
        int64_t ez = dz + kc * c.box.fc[2];
        visit(ex, ey, ez, code - centre);  // 0 = the reduced image itself
      }
}

// Super-cluster S as a list wave sees it: its clusters' boxes (wave-uniform,
// indexed only by unrolled constants so they stay in registers), the frame
// origin and the bounding-sphere radius of the whole super-cluster.
struct md_super_view {
  int first, nlocal;
  int64_t cx[8], cy[8], cz[8];
  float4 ch[8];
  int64_t ox, oy, oz;
  float rs;
};

// The entries one lane contributes for candidate cluster J: for each passing
// periodic image, `emit(imask, self, shift)` in a fixed order.
//
// Before the eight box tests, an image is tested against S as a whole: if
// some point p of an i-cluster box and q of the J box lie within the list
// radius rl, then p is within rs of S's origin O and q within |hJ| of J's
// centre, so |centre_J - O| < rl + rs + |hJ|.  Failing that bound, no box
// test can pass, so skipping the image leaves the entries unchanged; the
// bound is widened by 1e-3 nm against f32 rounding in both tests.
// Which of two different super-clusters lists the pairs between them: the
// lower-numbered one when the numbers' sum is even, the higher one when it is
// odd.  Each unordered pair is listed exactly once, and every super-cluster
// lists about half of its neighbours whatever its number (with "the lower
// one" the first super-clusters listed all their neighbours and the last
// almost none, and the search, one wave per super-cluster, waited on the
// first ones).
__device__ __forceinline__ bool md_lists_pair(int S, int SJ) { return ((S + SJ) & 1) ? S > SJ : S < SJ; }

template <typename F>
__device__ __forceinline__ void md_candidate(const md_build_ctx &c, const md_list_ptrs &L, const md_super_view &V,
                                             int J, bool enumerate, const int *K, F &&emit) {
  int64_t jx = L.cluster_center[J], jy = L.cluster_center[c.nclusters + J], jz = L.cluster_center[2 * c.nclusters + J];
  float4 jh = L.cluster_half[J];
  const float reach = c.list_cutoff + V.rs + sqrtf(jh.x * jh.x + jh.y * jh.y + jh.z * jh.z) + 1e-3f;
  const float reach2 = reach * reach;
  // the image is chosen against S's origin and shared by the entry; S's own
  // clusters keep their unshifted frame (see md_for_each_image)
  int64_t dx = jx - V.ox, dy = jy - V.oy, dz = jz - V.oz;
  bool own = J >= V.first && J < V.first + 8;
  md_for_each_image(c, dx, dy, dz, enumerate, own, K, reach + 1e-3f, [&](int64_t ex, int64_t ey, int64_t ez, int code) {
    // the lattice vector that moves J into this image
    int64_t shx = ex - dx, shy = ey - dy, shz = ez - dz;
    float ox = md_from_fixed(jx + shx - V.ox), oy = md_from_fixed(jy + shy - V.oy), oz = md_from_fixed(jz + shz - V.oz);
    if (ox * ox + oy * oy + oz * oz >= reach2) return;
    bool zero = shx == 0 && shy == 0 && shz == 0;
    int imask = 0, self = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int I = V.first + q;
      // within S, J >= I only; a cluster with its own images keeps one of
      // each +/- pair (J of another super-cluster comes only from one that
      // S lists, md_lists_pair)
      if (q < V.nlocal && (!own || J >= I) && !(J == I && !zero && code < 0)) {
        float ddx = md_from_fixed(jx + shx - V.cx[q]);
        float ddy = md_from_fixed(jy + shy - V.cy[q]);
        float ddz = md_from_fixed(jz + shz - V.cz[q]);
        if (md_boxes_within(ddx, ddy, ddz, V.ch[q], jh, c.list_cutoff2)) {
          imask |= 1 << q;
          if (J == I && zero) self |= 1 << q;
        }
      }
    }
    if (imask) emit(imask, self, shx, shy, shz);
  });
}

// Write the list entries of one super-cluster into its own region of the
// entry arrays (entries_per_super entries from S * entries_per_super) and
// their count: one wavefront per S, lanes over candidate J clusters,
// compacted with a wavefront prefix count so the order is deterministic.  A
// count past the region reports a list overflow.
__global__ void __launch_bounds__(64) k_nb_list(const md_nb_args *__restrict__ T, md_bmap map) {
  int S;
  const md_nb_args &A = md_nb_entry(T, map, S);
  md_control *ctl = A.ctl;
  if (!ctl->rebuild_active) return;
  const md_build_ctx &c = A.c;
  const md_list_ptrs &L = A.L;
  int lane = threadIdx.x;
  if (S >= c.nsuper) return;
  md_super_view V;
  const int first = S * 8;
  V.first = first;
  V.nlocal = min(8, c.nclusters - first);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int k = min(first + q, c.nclusters - 1);
    V.cx[q] = L.cluster_center[k];
    V.cy[q] = L.cluster_center[c.nclusters + k];
    V.cz[q] = L.cluster_center[2 * c.nclusters + k];
    V.ch[q] = L.cluster_half[k];
  }
  const int64_t ox = L.super_origin[S], oy = L.super_origin[c.nsuper + S], oz = L.super_origin[2 * c.nsuper + S];
  V.ox = ox;
  V.oy = oy;
  V.oz = oz;
  V.rs = L.super_radius[S];
  // Every image of a j-cluster that can come within the list radius of one of
  // S's clusters lies within R = list radius + largest cluster radius + r_S of
  // S's origin (some point of an i-cluster box, within r_S of the origin, comes
  // within the list radius of some point of the J box).  Below half the
  // narrowest width that is the minimum image alone.  Otherwise an image K_i +
  // 1 or more lattice steps along axis i has fractional coordinate at least
  // K_i + 1 - f_i there (f_i bounds the difference it was stepped from: `fmax`
  // for a reduced one, r_S/w_i for S's own clusters, which are not reduced),
  // so it lies at least (K_i + 1 - f_i) w_i away: K_i = ceil(R / w_i + f_i -
  // 1) shells suffice.
  const float R = c.list_cutoff + __uint_as_float(ctl->cluster_radius_bits) + V.rs;
  const bool enumerate = R > c.half_min_width;
  int K[3] = {1, 1, 1};
  bool too_far = false;
  if (enumerate)
    for (int d = 0; d < 3; ++d) {
      float f = fmaxf(c.fmax[d], V.rs / c.width[d]);
      int k = max(1, (int)ceilf(R / c.width[d] + f - 1.0f));
      too_far |= k > MD_MAX_IMAGE_SHELLS;
      K[d] = min(k, MD_MAX_IMAGE_SHELLS);
    }
  if (lane == 0 && too_far) ctl->overflow = 2;
  // Candidate j super-clusters SJ >= S, ascending.  A j-cluster of SJ can
  // reach an i-cluster of S only if the bounding spheres of the two
  // super-clusters come within the list radius at some image, so SJ is
  // skipped when no image of its centre comes within list radius + both
  // radii.  Without enumeration that image is the minimum image, which the
  // sequential reduction of md_min_image_fixed gives in a rectangular box; in
  // a reduced triclinic one it is one of the 27 images around the reduced
  // difference.  With enumeration, the images within Kp shells of it, Kp from
  // the same bound as K with the largest super-cluster radius in place of the
  // largest cluster radius.  The j-clusters then go through the same tests in
  // the same order as a scan of every J >= S's first cluster, so the list is
  // the one that scan builds.
  __shared__ int cand[MD_LIST_CANDIDATES];
  int ncand = 0;
  const float rs = V.rs;
  int Kp[3] = {c.box.rectangular ? 0 : 1, c.box.rectangular ? 0 : 1, c.box.rectangular ? 0 : 1};
  bool prefilter = true;
  if (enumerate) {
    const float Rp = c.list_cutoff + rs + __uint_as_float(ctl->super_radius_bits);
    for (int d = 0; d < 3; ++d) {
      const int k = max(1, (int)ceilf(Rp / c.width[d] + c.fmax[d] - 1.0f));
      prefilter &= k <= MD_MAX_IMAGE_SHELLS;
      Kp[d] = min(k, MD_MAX_IMAGE_SHELLS);
    }
  }
  if (prefilter) {
    for (int sb = 0; sb < c.nsuper; sb += 64) {
      const int SJ = sb + lane;
      bool near = false;
      if (SJ < c.nsuper && (SJ == S || md_lists_pair(S, SJ))) {
        const float lim = c.list_cutoff + rs + L.super_radius[SJ];
        if (SJ == S) {
          near = true;
        } else {
          int64_t dx = L.super_origin[SJ] - ox, dy = L.super_origin[c.nsuper + SJ] - oy,
                  dz = L.super_origin[2 * c.nsuper + SJ] - oz;
          md_min_image_fixed(dx, dy, dz, c.box);
          // f32 images of the reduced difference, with a 1e-3 nm margin over
          // their rounding (a candidate kept in excess makes no entry: each
          // of its clusters still meets the exact tests below)
          const float fx0 = md_from_fixed(dx), fy0 = md_from_fixed(dy), fz0 = md_from_fixed(dz);
          const float limm = lim + 1e-3f, lim2 = limm * limm;
          for (int kc = -Kp[2]; kc <= Kp[2]; ++kc)
            for (int kb = -Kp[1]; kb <= Kp[1]; ++kb)
              for (int ka = -Kp[0]; ka <= Kp[0]; ++ka) {
                const float fx = fmaf((float)kc, c.box.c[0], fmaf((float)kb, c.box.b[0], fmaf((float)ka, c.box.a[0], fx0)));
                const float fy = fmaf((float)kc, c.box.c[1], fmaf((float)kb, c.box.b[1], fy0));
                const float fz = fmaf((float)kc, c.box.c[2], fz0);
                near |= fmaf(fx, fx, fmaf(fy, fy, fz * fz)) < lim2;
              }
        }
      }
      const uint64_t m = __ballot(near);
      const int below = __popcll(m & ((uint64_t(1) << lane) - 1));
      if (near && ncand + below < MD_LIST_CANDIDATES) cand[ncand + below] = SJ;
      ncand += __popcll(m);
    }
    __syncthreads();
  }
  // past the candidate store, or images too far out to prefilter: scan every J
  const bool scan_all = !prefilter || ncand > MD_LIST_CANDIDATES;
  const int chunks = scan_all ? (c.nclusters + 63) / 64 : (ncand + 7) / 8;
  int written = 0;
  const int base_out = S * c.entries_per_super, region_end = base_out + c.entries_per_super;
  for (int chunk = 0; chunk < chunks; ++chunk) {
    int J;
    if (scan_all) {
      J = 64 * chunk + lane;
      if (J < c.nclusters && J / 8 != S && !md_lists_pair(S, J / 8)) J = c.nclusters;
    } else {
      int k = 8 * chunk + (lane >> 3);
      J = k < ncand ? cand[k] * 8 + (lane & 7) : c.nclusters;
    }
    bool valid = J < c.nclusters;
    // the count, and the entry itself when there is just one (always, without
    // image enumeration): only several images need a second pass to write
    int count = 0, one_mask = 0;
    int64_t one_shift[3] = {0, 0, 0};
    if (valid)
      md_candidate(c, L, V, J, enumerate, K, [&](int imask, int self, int64_t shx, int64_t shy, int64_t shz) {
        ++count;
        one_mask = imask | (self << 8);
        one_shift[0] = shx;
        one_shift[1] = shy;
        one_shift[2] = shz;
      });
    int incl = count;
    for (int off = 1; off < 64; off <<= 1) {
      int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (valid && count) {
      int at = base_out + written + incl - count;
      int SJ = J / 8;
      int64_t jox = L.super_origin[SJ], joy = L.super_origin[c.nsuper + SJ], joz = L.super_origin[2 * c.nsuper + SJ];
      auto write = [&](int meta, int64_t shx, int64_t shy, int64_t shz) {
        if (at < region_end) {
          L.entry_j[at] = J;
          L.entry_imask[at] = meta;
          // J's frame origin, moved into this image, in S's frame
          bool bad = false;
          L.entry_shift[at] = make_int4(md_frame_coord(jox + shx - ox, bad), md_frame_coord(joy + shy - oy, bad),
                                        md_frame_coord(joz + shz - oz, bad), 0);
          if (bad) ctl->overflow = 3;
        }
        ++at;
      };
      if (count == 1)
        write(one_mask, one_shift[0], one_shift[1], one_shift[2]);
      else
        md_candidate(c, L, V, J, enumerate, K, [&](int imask, int self, int64_t shx, int64_t shy, int64_t shz) {
          write(imask | (self << 8), shx, shy, shz);
        });
    }
    written += __shfl(incl, 63, 64);
  }
  if (lane == 0) {
    L.super_count[S] = written;
    if (written > c.entries_per_super) ctl->overflow = 1;
  }
}

// Starts of every super-cluster's entries and segments, and the totals.
// one workgroup per replica
__global__ void __launch_bounds__(1024) k_nb_list_scan(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  md_control *ctl = A.ctl;
  if (!ctl->rebuild_active) return;
  const md_list_ptrs &L = A.L;
  const int nsuper = A.c.nsuper, region = A.c.entries_per_super, segment_capacity = A.c.segment_capacity;
  int *totals = A.totals;
  __shared__ int entries;
  if (threadIdx.x == 0) entries = 0;
  __syncthreads();
  int mine = 0;
  for (int s = threadIdx.x; s < nsuper; s += 1024) {
    const int n = min(L.super_count[s], region);
    L.super_segs[s] = (n + A.c.seg_len - 1) / A.c.seg_len;
    mine += n;
  }
  atomicAdd(&entries, mine);
  __syncthreads();
  md_block_scan(L.super_segs, L.super_segs, nsuper, nullptr);
  __syncthreads();
  if (threadIdx.x == 0) {
    int segs = L.super_segs[nsuper];
    if (segs > segment_capacity) ctl->overflow = 1;
    totals[0] = entries;
    totals[1] = min(segs, segment_capacity);
  }
}

__global__ void k_nb_segments(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  if (!A.ctl->rebuild_active) return;
  const md_list_ptrs &L = A.L;
  const int segment_capacity = A.c.segment_capacity, region = A.c.entries_per_super;
  int S = blk * blockDim.x + threadIdx.x;
  if (S >= A.c.nsuper) return;
  const int b = S * region, e = b + min(L.super_count[S], region);
  int sg = L.super_segs[S];
  const int len = A.c.seg_len;
  for (int at = b; at < e; at += len, ++sg) {
    if (sg >= segment_capacity) break;
    L.segments[sg] = make_int4(S, at, min(at + len, e), 0);
  }
}

// Exclusion masks: MD_MASK_WAVES wavefronts per super-cluster S, each taking
// every MD_MASK_WAVES-th of S's entries, lane l = (i = l/8, j = l%8); for each
// of S's 8 i-clusters a 64-bit mask of the pairs that interact.  Clears
// padding, the lower triangle of a cluster with its own zero image, and every
// excluded pair (the CSR holds both directions; nonbonded.ts:159-182).  An
// exclusion is looked up in the i atom's 64-bit index window, and in the CSR
// only for the rare atom flagged as having exclusions outside it.
__global__ void __launch_bounds__(64 * MD_MASK_GROUP) k_nb_masks(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  if (!A.ctl->rebuild_active) return;
  const md_build_ctx &c = A.c;
  const md_list_ptrs &L = A.L;
  const int wave = blk * MD_MASK_GROUP + (int)(threadIdx.x >> 6);
  const int S = wave / MD_MASK_WAVES, part = wave % MD_MASK_WAVES;
  if (S >= c.nsuper) return;
  const int *slot_atom = A.slot_atom, *excl_start = A.excl_start, *excl_list = A.excl_list;
  const int lane = threadIdx.x & 63, il = lane >> 3, jl = lane & 7;
  int ia[8];
  uint64_t iw[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int I = S * 8 + q;
    ia[q] = I < c.nclusters ? slot_atom[I * MD_CLUSTER + il] : -1;
    iw[q] = ia[q] >= 0 ? A.excl_window[ia[q]] : 0;
  }
  const int eb = S * c.entries_per_super, ee = eb + min(L.super_count[S], c.entries_per_super);
  for (int e = eb + part; e < ee; e += MD_MASK_WAVES) {
    const int J = L.entry_j[e];
    const int meta = L.entry_imask[e];
    if (lane == 0) atomicAdd(&L.j_count[J], 1);
    const int ja = slot_atom[J * MD_CLUSTER + jl];
    bool clean = true;
    uint64_t masks[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      bool ok = ((meta >> q) & 1) && ia[q] >= 0 && ja >= 0;
      if (ok && ((meta >> (8 + q)) & 1)) ok = jl > il;  // same cluster, zero image
      if (ok) {
        const int d = ja - ia[q];
        if (d >= -32 && d < 32 && d != 0) {
          ok = !((iw[q] >> (d + 32)) & 1);
        } else if ((iw[q] >> 32) & 1) {
          for (int k = excl_start[ia[q]]; k < excl_start[ia[q] + 1]; ++k)
            if (excl_list[k] == ja) {
              ok = false;
              break;
            }
        }
      }
      const uint64_t m = __ballot(ok);
      masks[q] = m;
      if ((meta >> q) & 1) clean = clean && m == ~0ull;
    }
    // a clean entry's masks are never read: only its meta word changes
    if (clean) {
      if (lane == 0) L.entry_imask[e] = meta | MD_META_CLEAN;
    } else if (lane < 8) {
      uint64_t m = masks[0];
#pragma unroll
      for (int q = 1; q < 8; ++q)
        if (lane == q) m = masks[q];
      L.entry_excl[(size_t)e * 8 + lane] = m;
    }
  }
}

// the pair kernels' segment counters, one per XCD (md_seg_queue)
#define MD_XCDS 8
#define MD_CURSOR_STRIDE 16  // ints between two counters: one 64-byte line

// Per step: each slot's position as f32 in its super-cluster's frame.
// The launch's first thread also lays out the pair kernel's work: the start of
// each replica's list segments in one numbering, and the counters the pair
// kernel's waves take segments from, reset to 0.
__global__ void k_nb_pack(const md_nb_args *__restrict__ T, md_bmap map, int *seg_start, int *cursor) {
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    // the first wave: each lane one replica's segment count (two dependent
    // loads, issued by all lanes at once rather than one replica after
    // another), then a prefix sum across the wave, 64 replicas per pass
    const int lane = threadIdx.x;
    int run = 0;
    for (int r0 = 0; r0 < map.R; r0 += 64) {
      const int r = r0 + lane;
      const int v = r < map.R ? T[r].totals[1] : 0;
      int incl = v;
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
      }
      if (r < map.R) seg_start[r] = run + incl - v;
      run += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
      seg_start[map.R] = run;
      for (int c = 0; c < MD_XCDS; ++c) cursor[c * MD_CURSOR_STRIDE] = 0;
    }
  }
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  const int nslots = A.c.nslots, np = A.c.np;
  const int *slot_atom = A.slot_atom;
  const int64_t *pos = A.pos, *slot_offset = A.slot_offset;
  int4 *slot_xyzq = A.slot_xyzq;
  int s = blk * blockDim.x + threadIdx.x;
  if (s >= nslots) return;  // whole wavefronts: nslots is a multiple of 64
  const int a = slot_atom[s];
  int lo[3], hi[3];
  if (a < 0) {
    // far away (20 nm out) and uncharged; every pair with it is masked anyway
    const int far = 20 << 26;
    slot_xyzq[s] = make_int4(far + ((s & 63) << 20), far, far, 0);
    for (int d = 0; d < 3; ++d) {
      lo[d] = INT_MAX;
      hi[d] = INT_MIN;
    }
  } else {
    bool bad = false;
    int x = md_frame_coord(pos[a] - slot_offset[s], bad);
    int y = md_frame_coord(pos[np + a] - slot_offset[nslots + s], bad);
    int z = md_frame_coord(pos[2 * np + a] - slot_offset[2 * nslots + s], bad);
    slot_xyzq[s] = make_int4(x, y, z, __float_as_int(A.atom_param[a].x));
    if (bad) A.ctl->overflow = 3;
    lo[0] = hi[0] = x;
    lo[1] = hi[1] = y;
    lo[2] = hi[2] = z;
  }
  // the cluster's current box over its atoms: 8 consecutive lanes
  for (int d = 0; d < 3; ++d)
    for (int off = 1; off < MD_CLUSTER; off <<= 1) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = max(hi[d], __shfl_xor(hi[d], off, 64));
    }
  if ((s & (MD_CLUSTER - 1)) == 0) {
    A.cluster_lo[s / MD_CLUSTER] = make_int4(lo[0], lo[1], lo[2], 0);
    A.cluster_hi[s / MD_CLUSTER] = make_int4(hi[0], hi[1], hi[2], 0);
  }
}

// After the pair kernel: each slot's force onto its atom -- its i-force sums
// plus, as j atom, the force of every list entry naming its cluster, each
// rounded to Q32.32 on its own (the value an atomic per entry would add) --
// and the slot cleared for the next step.  Kernels on one stream run in order
// and a slot is the only one holding its atom, so a plain read-add-write is
// enough.
// Blocks laid out XCD by XCD: neighbouring clusters read j-force records of
// neighbouring entries, so they share one L2.
__global__ void k_nb_forces(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = T[md_locate(map, md_xcd_block(blockIdx.x, gridDim.x), blk)];
  const int nslots = A.c.nslots, np = A.c.np;
  const int s = blk * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  int64_t *slot_force = A.P.slot_force;
  const int a = A.slot_atom[s];
  int64_t f[3];
  for (int d = 0; d < 3; ++d) {
    f[d] = slot_force[d * nslots + s];
    slot_force[d * nslots + s] = 0;
  }
  if (a < 0) return;
  const int J = s / MD_CLUSTER, q = s - J * MD_CLUSTER;
  const float *jf = A.P.j_force + q * 3;
  const int *je = A.L.j_entries;
  int k = A.L.j_start[J];
  const int end = A.L.j_start[J + 1];
  // four entries' records in flight at a time: each iteration's loads are
  // independent of the last, so the loop waits once per four entries on the
  // entry index and once on the records (integer sums: order-free)
  for (; k + 4 <= end; k += 4) {
    const int e0 = je[k], e1 = je[k + 1], e2 = je[k + 2], e3 = je[k + 3];
    const float *r0 = jf + (size_t)e0 * 24, *r1 = jf + (size_t)e1 * 24, *r2 = jf + (size_t)e2 * 24,
                *r3 = jf + (size_t)e3 * 24;
    const float x0 = r0[0], y0 = r0[1], z0 = r0[2], x1 = r1[0], y1 = r1[1], z1 = r1[2];
    const float x2 = r2[0], y2 = r2[1], z2 = r2[2], x3 = r3[0], y3 = r3[1], z3 = r3[2];
    f[0] += md_to_fixed(x0) + md_to_fixed(x1) + md_to_fixed(x2) + md_to_fixed(x3);
    f[1] += md_to_fixed(y0) + md_to_fixed(y1) + md_to_fixed(y2) + md_to_fixed(y3);
    f[2] += md_to_fixed(z0) + md_to_fixed(z1) + md_to_fixed(z2) + md_to_fixed(z3);
  }
  for (; k < end; ++k) {
    const float *e = jf + (size_t)je[k] * 24;
    f[0] += md_to_fixed(e[0]);
    f[1] += md_to_fixed(e[1]);
    f[2] += md_to_fixed(e[2]);
  }
  for (int d = 0; d < 3; ++d) A.force[d * np + a] += f[d];
}

// The transposed list: starts from the counts k_nb_masks took (one
// workgroup per replica), then each entry's number placed under its J.
__global__ void __launch_bounds__(1024) k_nb_j_scan(const md_nb_args *__restrict__ T, md_bmap map) {
  int blk;
  const md_nb_args &A = md_nb_entry(T, map, blk);
  if (!A.ctl->rebuild_active) return;
  md_block_scan(A.L.j_count, A.L.j_start, A.c.nclusters, A.L.j_count);
}

// One wave per super-cluster (its ~200 entries are a few passes of 64
// lanes): a launch over replicas that did not rebuild then dispatches 16
// times fewer waves that return at once.
__global__ void __launch_bounds__(64) k_nb_j_fill(const md_nb_args *__restrict__ T, md_bmap map) {
  int S;
  const md_nb_args &A = md_nb_entry(T, map, S);
  if (!A.ctl->rebuild_active) return;
  const md_list_ptrs &L = A.L;
  if (S >= A.c.nsuper) return;
  const int eb = S * A.c.entries_per_super, ee = eb + min(L.super_count[S], A.c.entries_per_super);
  for (int e = eb + (int)threadIdx.x; e < ee; e += 64) {
    const int J = L.entry_j[e];
    L.j_entries[L.j_start[J] + atomicAdd(&L.j_count[J], 1)] = e;
  }
}

// ---------------------------------------------------------------------------
// The pair kernel
// ---------------------------------------------------------------------------

// One pair from lane's i atom to the entry's j atom, counted when `inc` (not
// excluded, inside the cutoff).  Everything is in frame units: r2 in 2^-52
// nm^2 (the integer difference squared), the charges carry 2^26, sigma/2
// 2^26 and 2 sqrt(eps) 2^13, and the constants their powers of two, so the
// result is the F/r factor per frame unit (F_i = -fs * d with d in frame
// units gives kJ/mol/nm) and the energy carries 2^26.  Scaling by powers of
// two is exact in f32, so this is the nm evaluation bit for bit, without
// its two unit conversions per pair.
//
// FIT: the Coulomb factor is 1/r^3 plus a degree-12 polynomial in r^2
// (md_ewald_fit), twelve FMAs in place of exp, a reciprocal and the erfc
// polynomial; otherwise erfc(z) = exp(-z^2) P(1/(1 + z/2)).  `top` is the
// force polynomial's leading coefficient, which the caller keeps in a VGPR
// (an FMA reads one scalar operand, and the first Horner step has two).
//
// The result is the pair's F/r when `inc`; on the other lanes it is
// undefined (r2 may be 0 there) and the caller must not add it.
template <int N>
__device__ __forceinline__ float md_horner(float top, const float *c, float v) {
  float p = top;
#pragma unroll
  for (int k = N - 2; k >= 0; --k) p = fmaf(p, v, c[k]);
  return p;
}

template <bool ENERGY, bool FIT>
__device__ __forceinline__ float md_pair(float r2, float qi, float si, float ei, float qj, float sj, float ej,
                                         bool inc, float top, const md_pair_args &A, float &en, int &count) {
  float rinv = __builtin_amdgcn_rsqf(r2);
  float rinv2 = rinv * rinv;
  float coul, ecoul = 0.0f;  // F/r and E per charge product
  if (FIT) {
    const float v = fmaf(r2, A.fit_scale, -1.0f);
    coul = fmaf(rinv, rinv2, md_horner<MD_EWALD_TERMS>(top, A.fit_force, v));
    if (ENERGY) ecoul = rinv + md_horner<MD_EWALD_TERMS>(A.fit_energy[MD_EWALD_TERMS - 1], A.fit_energy, v);
  } else {
    float r = r2 * rinv;
    float ex = __builtin_amdgcn_exp2f(A.exp_scale * r2);  // exp(-alpha^2 r^2)
    float t = __builtin_amdgcn_rcpf(fmaf(A.half_alpha, r, 1.0f));
    float erfc_v = ex * md_erfcx_poly(t);
    coul = fmaf(erfc_v, rinv, A.beta * ex) * rinv2;
    if (ENERGY) ecoul = erfc_v * rinv;
  }
  float qq = qi * qj;
  float sig = si + sj;
  float eps = ei * ej;  // 4 eps_ij
  float s2 = sig * sig * rinv2;
  float s6 = s2 * s2 * s2;
  float e6 = eps * s6;
  float flj = e6 * fmaf(12.0f, s6, -6.0f) * rinv2;  // 4 eps (12 s^12 - 6 s^6) / r^2
  float fs = fmaf(qq, coul, flj);
  if (ENERGY) {
    float e = fmaf(qq, ecoul, e6 * (s6 - 1.0f));
    en += inc ? e : 0.0f;
    count += inc ? 1 : 0;
  }
  return fs;
}

// Loads of wave-uniform addresses from list data no kernel of this launch
// writes: through the constant address space they become scalar loads into
// SGPRs (the list is built by earlier launches, so the scalar cache sees it).
template <typename T>
__device__ __forceinline__ T md_uniform(const T *p) {
  return *(const __attribute__((address_space(4))) T *)p;
}
// an int4 of which only x, y, z are used, as two 64-bit scalar loads (there
// is no three-dword scalar load, and a narrowed one would become a vector load)
__device__ __forceinline__ int4 md_uniform_xyz(const int4 *p) {
  const uint64_t a = md_uniform((const uint64_t *)p), b = md_uniform((const uint64_t *)p + 1);
  return make_int4((int)(uint32_t)a, (int)(uint32_t)(a >> 32), (int)(uint32_t)b, 0);
}

// DPP lane moves (row = 16 lanes): quad_perm [1,0,3,2] and [2,3,0,1] give
// lane l the value of l^1 and l^2; row_ror:n (rotate right by n) the value of
// lane (l-n) mod 16 of its row
template <int CTRL>
__device__ __forceinline__ float md_dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
#define MD_DPP_XOR1 0xB1
#define MD_DPP_XOR2 0x4E
#define MD_DPP_ROR4 0x124
#define MD_DPP_ROR8 0x128
#define MD_DPP_ROR12 0x12C

// x(l) + x(l^16) and x(l) + x(l^32) on every lane, by the gfx950 permlane
// swaps (rows 0<->1, 2<->3; lane halves)
__device__ __forceinline__ float md_add_xor16(float x) {
  auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ __forceinline__ float md_add_xor32(float x) {
  auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}

// The sum over the 8 lanes sharing l%8 (l, l^8, l^16, ... l^56), the same
// on each: ((x + x^8) + (.)^16) + (.)^32, so the bits match a shuffle
// butterfly of the same order.
__device__ __forceinline__ float md_sum_same_j(float x) {
  x += md_dpp<MD_DPP_ROR8>(x);  // within a row, lane (l-8) mod 16 is l^8
  x = md_add_xor16(x);
  return md_add_xor32(x);
}

// The sum over the 8 lanes sharing l/8, valid on the lanes with l%8 in
// 0-3 (HIGH false) or 4-7 (HIGH true): the butterfly's ((x + x^1) + (.)^2)
// then + (.)^4, whose partner l^4 is lane l+4 (l%8 < 4: rotate right by 12)
// or l-4 (l%8 >= 4: rotate right by 4).
template <bool HIGH>
__device__ __forceinline__ float md_sum_same_i(float x) {
  x += md_dpp<MD_DPP_XOR1>(x);
  x += md_dpp<MD_DPP_XOR2>(x);
  return x + md_dpp<HIGH ? MD_DPP_ROR4 : MD_DPP_ROR12>(x);
}

// gap between boxes [ilo, ihi] and [jlo, jhi] on one axis, in frame units
// (exact integer differences for boxes within 32 nm of each other)
__device__ __forceinline__ float md_box_gap(int ilo, int ihi, int jlo, int jhi) {
  const int a = (int)((unsigned)jlo - (unsigned)ihi), b = (int)((unsigned)ilo - (unsigned)jhi);
  return (float)max(0, max(a, b));
}

// One segment: consecutive list entries of super-cluster S against their j
// clusters, in units of MD_SEGMENT entries counted from S's first entry.  The
// i forces and the energy are summed in f32 over one unit and rounded to
// Q32.32 at its end, then summed as integers over the segment and added once:
// the rounding never depends on the segment length, so every launch group
// gives the same bits whatever length it picked.
template <bool ENERGY, bool FIT>
__device__ __forceinline__ void md_pair_segment(const md_pair_args &A, int seg, int lane, int il, int jl,
                                                int64_t &efix, int &count, int &blocks) {
  const int4 sg = md_uniform_xyz(A.segments + seg);
  // wave-uniform by construction; readfirstlane tells the compiler, so the
  // entry loop runs on scalar registers and its loads are scalar loads
  const int S = __builtin_amdgcn_readfirstlane(sg.x), eb = __builtin_amdgcn_readfirstlane(sg.y),
            ee = __builtin_amdgcn_readfirstlane(sg.z);
  // i data: lane's atom in each of S's 8 clusters (8 clusters x 8 atoms = 64
  // slots; lane il*8+jl reads slot S*64 + q*8 + il)
  int xi[8], yi[8], zi[8];
  float qi[8], si[8], ei[8];
  float fx[8], fy[8], fz[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int slot = S * MD_SUPER + q * MD_CLUSTER + il;
    int4 p = A.slot_xyzq[slot];
    float2 lj = A.slot_lj[slot];
    xi[q] = p.x;
    yi[q] = p.y;
    zi[q] = p.z;
    qi[q] = __int_as_float(p.w) * MD_PAIR_Q_SCALE;
    si[q] = lj.x * MD_PAIR_S_SCALE;
    ei[q] = lj.y * MD_PAIR_E_SCALE;
    fx[q] = fy[q] = fz[q] = 0.0f;
  }
  float en = 0.0f;
  // lane (il, jl) keeps the i force of atom il of S's cluster q = jl
  int64_t ax = 0, ay = 0, az = 0;
  int unit_end = min(eb + MD_SEGMENT, ee);
  // lanes 0-7: the current box of S's cluster q = lane
  int4 ilo = make_int4(0, 0, 0, 0), ihi = ilo;
  if (lane < 8) {
    ilo = A.cluster_lo[S * 8 + lane];
    ihi = A.cluster_hi[S * 8 + lane];
  }
  // the force polynomial's leading coefficient in a VGPR (an empty asm
  // with a VGPR operand puts it there once per segment)
  float top = A.fit_force[MD_EWALD_TERMS - 1];
  if (FIT) asm volatile("" : "+v"(top));
  for (int e = eb; e < ee; ++e) {
    const int J = md_uniform(A.entry_j + e);
    const int meta = md_uniform(A.entry_imask + e);
    const int4 sh = md_uniform_xyz(A.entry_shift + e);
    // Blocks whose current boxes lie more than the cutoff apart (by a 1e-5
    // relative margin over the f32 distance test) have no pair inside it and
    // would add exactly zero: drop them before any pair arithmetic.  Lane q
    // tests S's cluster q; integer box differences wrap exactly for boxes
    // within 32 nm, and a box of padding only (inverted) may test either way,
    // its pairs being masked.
    int active = 0;
    {
      const int4 jlo = md_uniform_xyz(A.cluster_lo + J), jhi = md_uniform_xyz(A.cluster_hi + J);
      const float dx = md_box_gap(ilo.x, ihi.x, (int)((unsigned)jlo.x + (unsigned)sh.x),
                                  (int)((unsigned)jhi.x + (unsigned)sh.x));
      const float dy = md_box_gap(ilo.y, ihi.y, (int)((unsigned)jlo.y + (unsigned)sh.y),
                                  (int)((unsigned)jhi.y + (unsigned)sh.y));
      const float dz = md_box_gap(ilo.z, ihi.z, (int)((unsigned)jlo.z + (unsigned)sh.z),
                                  (int)((unsigned)jhi.z + (unsigned)sh.z));
      const bool near = lane < 8 && fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < A.prune2;  // frame units^2
      active = meta & (int)(__ballot(near) & 0xffu);
    }
    if (ENERGY) blocks += __popc(active);
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (active) {
      const int4 pj = A.slot_xyzq[J * MD_CLUSTER + jl];
      const float2 lj = A.slot_lj[J * MD_CLUSTER + jl];
      const float qj = __int_as_float(pj.w) * MD_PAIR_Q_SCALE, sj = lj.x * MD_PAIR_S_SCALE,
                  ej = lj.y * MD_PAIR_E_SCALE;
      // wrapping (unsigned) integer arithmetic: exact for every real pair; the
      // padding slots far out may wrap, and their pairs are masked
      unsigned xj = (unsigned)pj.x + (unsigned)sh.x, yj = (unsigned)pj.y + (unsigned)sh.y,
               zj = (unsigned)pj.z + (unsigned)sh.z;
      const uint64_t *ex = A.entry_excl + (size_t)e * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if ((active >> q) & 1) {
          // bit l of the mask allows lane l's pair: with the in-cutoff ballot
          // it is the block's lane mask of counted pairs
          const uint64_t allowed = (meta & MD_META_CLEAN) ? ~0ull : md_uniform(ex + q);
          float dx = (float)(int)(xj - (unsigned)xi[q]), dy = (float)(int)(yj - (unsigned)yi[q]),
                dz = (float)(int)(zj - (unsigned)zi[q]);
          float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const uint64_t inc = allowed & __ballot(r2 < A.cutoff2);
          // a block whose 64 pairs all lie beyond the cutoff (or are masked)
          // adds exactly zero: skip its arithmetic, wave-uniformly
          if (inc == 0) continue;
          const bool counted = __builtin_amdgcn_inverse_ballot_w64(inc);
          float fs = md_pair<ENERGY, FIT>(r2, qi[q], si[q], ei[q], qj, sj, ej, counted, top, A, en, count);
          // the other lanes' fs is not a force: a divergent branch (the
          // exec mask, scalar instructions) leaves them out without a select
          if (counted) {
            fx[q] = fmaf(-fs, dx, fx[q]);
            fy[q] = fmaf(-fs, dy, fy[q]);
            fz[q] = fmaf(-fs, dz, fz[q]);
            gx = fmaf(fs, dx, gx);
            gy = fmaf(fs, dy, gy);
            gz = fmaf(fs, dz, gz);
          }
        }
      }
      // j force: sum the 8 lanes that share j (lane%8), fixed order
      gx = md_sum_same_j(gx);
      gy = md_sum_same_j(gy);
      gz = md_sum_same_j(gz);
    }
    // lanes 0-7 hold (x, y, z) of j atoms 0-7: one 96-byte record, j-major,
    // so the slot thread that adds j's share reads 12 contiguous bytes
    if (lane < 8) {
      float *r = A.j_force + (size_t)e * 24 + lane * 3;
      r[0] = gx;
      r[1] = gy;
      r[2] = gz;
    }
    if (e + 1 == unit_end) {  // wave-uniform: the end of a unit
      // i forces: sum the 8 lanes that share i (same il), fixed order; lane
      // jl keeps q = jl's sum
      float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float x, y, z;
        if (q < 4) {
          x = md_sum_same_i<false>(fx[q]);
          y = md_sum_same_i<false>(fy[q]);
          z = md_sum_same_i<false>(fz[q]);
        } else {
          x = md_sum_same_i<true>(fx[q]);
          y = md_sum_same_i<true>(fy[q]);
          z = md_sum_same_i<true>(fz[q]);
        }
        if (jl == q) {
          sx = x;
          sy = y;
          sz = z;
        }
        fx[q] = fy[q] = fz[q] = 0.0f;
      }
      ax += md_to_fixed(sx);
      ay += md_to_fixed(sy);
      az += md_to_fixed(sz);
      if (ENERGY) {
        efix += md_to_fixed(en * MD_FRAME_UNIT);  // the energy carried 2^26
        en = 0.0f;
      }
      unit_end = min(unit_end + MD_SEGMENT, ee);
    }
  }
  // 64 consecutive slots: S's cluster jl, atom il
  const int is = (S * 8 + jl) * MD_CLUSTER + il;
  md_atomic_add_i64(&A.slot_force[is], ax);
  md_atomic_add_i64(&A.slot_force[A.nslots + is], ay);
  md_atomic_add_i64(&A.slot_force[2 * A.nslots + is], az);
}

// ---- the compacted schedule ------------------------------------------------
//
// In an 8x8 block only about one lane in six holds a pair inside the cutoff,
// yet the pair arithmetic costs the whole wave.  The compacted schedule
// splits each entry's work in three:
//   A. per active block, the distance, the in-cutoff ballot and, for each
//      counted pair, its code (entry of the batch, block, lane) appended to a
//      queue in LDS at a place given by the ballot's prefix count;
//   C. the queue evaluated densely, 64 pairs per pass, each lane fetching its
//      pair's i and j data from LDS tables, the force factor written back to
//      LDS at the pair's own (entry, block, lane) place;
//   D. per active block, each lane reading its pair's factor back (zero for a
//      pair not counted) and accumulating exactly as the direct schedule does.
// Two consecutive entries (always in the same unit: units are 8 entries
// counted from the super-cluster's first) share one queue, so the dense
// passes run fuller.  The factors are the same floating-point expression of
// the same inputs, and D's sums run in the direct schedule's order, so both
// schedules give the same bits.

#define MD_COMPACT_BATCH 2

struct md_pair_lds {
  int4 ixyz[MD_SUPER];  // S's slots: frame x, y, z and the scaled charge's bits
  float2 ise[MD_SUPER]; // scaled sigma/2 and 2 sqrt(eps)
  int4 jxyz[MD_COMPACT_BATCH][MD_CLUSTER];  // the batch's j atoms in S's frame, scaled charge bits
  float2 jse[MD_COMPACT_BATCH][MD_CLUSTER];
  unsigned short queue[MD_COMPACT_BATCH * 512];  // (entry << 9) | (block << 6) | lane
  float fs[MD_COMPACT_BATCH][512];               // the pair's force factor at (entry, block, lane)
};

__device__ __forceinline__ void md_pair_segment_compact(const md_pair_args &A, int seg, int lane, int il, int jl,
                                                        md_pair_lds &L) {
  const int4 sg = md_uniform_xyz(A.segments + seg);
  const int S = __builtin_amdgcn_readfirstlane(sg.x), eb = __builtin_amdgcn_readfirstlane(sg.y),
            ee = __builtin_amdgcn_readfirstlane(sg.z);
  int xi[8], yi[8], zi[8];
  float fx[8], fy[8], fz[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int4 p = A.slot_xyzq[S * MD_SUPER + q * MD_CLUSTER + il];
    xi[q] = p.x;
    yi[q] = p.y;
    zi[q] = p.z;
    fx[q] = fy[q] = fz[q] = 0.0f;
  }
  {
    // the i table: lane l holds slot S*64 + l (block l/8, atom l%8)
    const int4 p = A.slot_xyzq[S * MD_SUPER + lane];
    const float2 lj = A.slot_lj[S * MD_SUPER + lane];
    L.ixyz[lane] = make_int4(p.x, p.y, p.z, __float_as_int(__int_as_float(p.w) * MD_PAIR_Q_SCALE));
    L.ise[lane] = make_float2(lj.x * MD_PAIR_S_SCALE, lj.y * MD_PAIR_E_SCALE);
  }
  int64_t ax = 0, ay = 0, az = 0;
  int unit_end = min(eb + MD_SEGMENT, ee);
  int4 ilo = make_int4(0, 0, 0, 0), ihi = ilo;
  if (lane < 8) {
    ilo = A.cluster_lo[S * 8 + lane];
    ihi = A.cluster_hi[S * 8 + lane];
  }
  for (int e0 = eb; e0 < ee; e0 += MD_COMPACT_BATCH) {
    const int nb = min(MD_COMPACT_BATCH, ee - e0);
    int active[MD_COMPACT_BATCH];
    uint64_t m[MD_COMPACT_BATCH][8];
    unsigned xj[MD_COMPACT_BATCH], yj[MD_COMPACT_BATCH], zj[MD_COMPACT_BATCH];
    int base = 0;
    // A: distances, ballots and the queue
#pragma unroll
    for (int b = 0; b < MD_COMPACT_BATCH; ++b) {
      active[b] = 0;
      xj[b] = yj[b] = zj[b] = 0u;
#pragma unroll
      for (int q = 0; q < 8; ++q) m[b][q] = 0;
      if (b >= nb) continue;
      const int e = e0 + b;
      const int J = md_uniform(A.entry_j + e);
      const int meta = md_uniform(A.entry_imask + e);
      const int4 sh = md_uniform_xyz(A.entry_shift + e);
      {
        const int4 jlo = md_uniform_xyz(A.cluster_lo + J), jhi = md_uniform_xyz(A.cluster_hi + J);
        const float dx = md_box_gap(ilo.x, ihi.x, (int)((unsigned)jlo.x + (unsigned)sh.x),
                                    (int)((unsigned)jhi.x + (unsigned)sh.x));
        const float dy = md_box_gap(ilo.y, ihi.y, (int)((unsigned)jlo.y + (unsigned)sh.y),
                                    (int)((unsigned)jhi.y + (unsigned)sh.y));
        const float dz = md_box_gap(ilo.z, ihi.z, (int)((unsigned)jlo.z + (unsigned)sh.z),
                                    (int)((unsigned)jhi.z + (unsigned)sh.z));
        const bool near = lane < 8 && fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < A.prune2;
        active[b] = meta & (int)(__ballot(near) & 0xffu);
      }
      if (!active[b]) continue;
      const int4 pj = A.slot_xyzq[J * MD_CLUSTER + jl];
      xj[b] = (unsigned)pj.x + (unsigned)sh.x;
      yj[b] = (unsigned)pj.y + (unsigned)sh.y;
      zj[b] = (unsigned)pj.z + (unsigned)sh.z;
      if (lane < MD_CLUSTER) {  // lanes 0-7 hold j atoms 0-7
        const float2 lj = A.slot_lj[J * MD_CLUSTER + jl];
        L.jxyz[b][jl] = make_int4((int)xj[b], (int)yj[b], (int)zj[b],
                                  __float_as_int(__int_as_float(pj.w) * MD_PAIR_Q_SCALE));
        L.jse[b][jl] = make_float2(lj.x * MD_PAIR_S_SCALE, lj.y * MD_PAIR_E_SCALE);
      }
      const uint64_t *ex = A.entry_excl + (size_t)e * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if ((active[b] >> q) & 1) {
          const uint64_t allowed = (meta & MD_META_CLEAN) ? ~0ull : md_uniform(ex + q);
          const float dx = (float)(int)(xj[b] - (unsigned)xi[q]), dy = (float)(int)(yj[b] - (unsigned)yi[q]),
                      dz = (float)(int)(zj[b] - (unsigned)zi[q]);
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const uint64_t inc = allowed & __ballot(r2 < A.cutoff2);
          m[b][q] = inc;
          if (inc) {
            if ((inc >> lane) & 1ull) {
              const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(inc >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)inc, 0u));
              L.queue[base + below] = (unsigned short)((b << 9) | (q << 6) | lane);
            }
            base += __popcll(inc);
          }
        }
      }
    }
    __syncthreads();
    // C: the counted pairs, 64 per pass
    for (int c0 = 0; c0 < base; c0 += 64) {
      const int k = c0 + lane;
      if (k < base) {
        const int code = L.queue[k];
        const int b = code >> 9, q = (code >> 6) & 7, pl = code & 63;
        const int4 I = L.ixyz[q * MD_CLUSTER + (pl >> 3)];
        const float2 Ise = L.ise[q * MD_CLUSTER + (pl >> 3)];
        const int4 Jd = L.jxyz[b][pl & 7];
        const float2 Jse = L.jse[b][pl & 7];
        const float dx = (float)(int)((unsigned)Jd.x - (unsigned)I.x), dy = (float)(int)((unsigned)Jd.y - (unsigned)I.y),
                    dz = (float)(int)((unsigned)Jd.z - (unsigned)I.z);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        float en = 0.0f;
        int cnt = 0;
        const float qi = __int_as_float(I.w), qj = __int_as_float(Jd.w);
        L.fs[b][q * 64 + pl] = A.ewald_fit
                                   ? md_pair<false, true>(r2, qi, Ise.x, Ise.y, qj, Jse.x, Jse.y, true,
                                                          A.fit_force[MD_EWALD_TERMS - 1], A, en, cnt)
                                   : md_pair<false, false>(r2, qi, Ise.x, Ise.y, qj, Jse.x, Jse.y, true, 0.0f, A,
                                                           en, cnt);
      }
    }
    __syncthreads();
    // D: back in the block layout, summed in the direct schedule's order
#pragma unroll
    for (int b = 0; b < MD_COMPACT_BATCH; ++b) {
      if (b >= nb) continue;
      const int e = e0 + b;
      float gx = 0.0f, gy = 0.0f, gz = 0.0f;
      if (active[b]) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const uint64_t inc = m[b][q];
          if (inc == 0) continue;
          const float f = ((inc >> lane) & 1ull) ? L.fs[b][q * 64 + lane] : 0.0f;
          const float dx = (float)(int)(xj[b] - (unsigned)xi[q]), dy = (float)(int)(yj[b] - (unsigned)yi[q]),
                      dz = (float)(int)(zj[b] - (unsigned)zi[q]);
          fx[q] = fmaf(-f, dx, fx[q]);
          fy[q] = fmaf(-f, dy, fy[q]);
          fz[q] = fmaf(-f, dz, fz[q]);
          gx = fmaf(f, dx, gx);
          gy = fmaf(f, dy, gy);
          gz = fmaf(f, dz, gz);
        }
        gx = md_sum_same_j(gx);
        gy = md_sum_same_j(gy);
        gz = md_sum_same_j(gz);
      }
      if (lane < 8) {
        float *r = A.j_force + (size_t)e * 24 + lane * 3;
        r[0] = gx;
        r[1] = gy;
        r[2] = gz;
      }
      if (e + 1 == unit_end) {
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          float x, y, z;
          if (q < 4) {
            x = md_sum_same_i<false>(fx[q]);
            y = md_sum_same_i<false>(fy[q]);
            z = md_sum_same_i<false>(fz[q]);
          } else {
            x = md_sum_same_i<true>(fx[q]);
            y = md_sum_same_i<true>(fy[q]);
            z = md_sum_same_i<true>(fz[q]);
          }
          if (jl == q) {
            sx = x;
            sy = y;
            sz = z;
          }
          fx[q] = fy[q] = fz[q] = 0.0f;
        }
        ax += md_to_fixed(sx);
        ay += md_to_fixed(sy);
        az += md_to_fixed(sz);
        unit_end = min(unit_end + MD_SEGMENT, ee);
      }
    }
    __syncthreads();  // the next batch rewrites the tables this one read
  }
  const int is = (S * 8 + jl) * MD_CLUSTER + il;
  md_atomic_add_i64(&A.slot_force[is], ax);
  md_atomic_add_i64(&A.slot_force[A.nslots + is], ay);
  md_atomic_add_i64(&A.slot_force[2 * A.nslots + is], az);
}

// The pair kernels' work queue.  The segments, numbered replica by replica
// and in list (Hilbert) order within each, are cut into MD_XCDS contiguous
// ranges with a counter each (one 64-byte line apart).  A workgroup takes
// segments from the range of the XCD it runs on first -- the dispatcher deals
// workgroups to the XCDs in turn, workgroup b to XCD b mod 8 -- then from the
// following ranges once that one is empty.  Each XCD's waves thus walk one
// neighbourhood of the list at a time and its L2 keeps that neighbourhood's
// j clusters, instead of all eight L2s fetching every cluster.  Which wave
// takes which segment never changes a result (each segment's sums are
// rounded to Q32.32 on their own).
//
// Every lane joins the atomic (lane 0 adds 1, the others 0) and readfirstlane
// makes the index wave-uniform, so the exits are scalar branches.  A
// lane-0-only atomic followed by a broadcast is not safe here: the optimiser
// threads the per-segment `lane == 0` energy flush into the next `lane == 0`
// atomic, after which lanes 1-63 loop back to the broadcast without lane 0
// and never leave.  The counters only grow and a wave leaves a range at its
// first index past the end, so every wave leaves.
struct md_seg_queue {
  int *cursor;
  int nseg, home, range, lo, hi, lane;
  __device__ md_seg_queue(int *c, int n, int l) : cursor(c), nseg(n), home((int)(blockIdx.x & (MD_XCDS - 1))),
                                                  range(0), lane(l) {
    bounds();
  }
  __device__ void bounds() {
    const int c = (home + range) & (MD_XCDS - 1);
    lo = (int)(((int64_t)nseg * c) / MD_XCDS);
    hi = (int)(((int64_t)nseg * (c + 1)) / MD_XCDS);
  }
  __device__ bool next(int &g) {
    while (range < MD_XCDS) {
      const int c = (home + range) & (MD_XCDS - 1);
      g = lo + __builtin_amdgcn_readfirstlane(atomicAdd(cursor + c * MD_CURSOR_STRIDE, lane == 0 ? 1 : 0));
      if (g < hi) return true;
      if (++range < MD_XCDS) bounds();
    }
    return false;
  }
};

// The compacted schedule's kernel (forces only; energies use the direct one).
template <bool BATCH>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_nb_pairs_compact(const md_nb_args *__restrict__ T, md_bmap seg_map, int *cursor) {
  __shared__ md_pair_lds L;
  const int nseg = seg_map.start[seg_map.R];
  const int lane = threadIdx.x, il = lane >> 3, jl = lane & 7;
  md_seg_queue Q(cursor, nseg, lane);
  int g;
  while (Q.next(g)) {
    int seg;
    const md_pair_args &A = T[md_locate(seg_map, g, seg)].P;
    md_pair_segment_compact(A, seg, lane, il, jl, L);
    __syncthreads();  // the next segment rewrites the i table
  }
}

// BATCH only names the instance: a launch over several replicas runs
// k_nb_pairs<·, true>, one replica <·, false>, the same code, so a kernel
// trace reports the batched launches on their own.
template <bool ENERGY, bool BATCH = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_nb_pairs(const md_nb_args *__restrict__ T, md_bmap seg_map, int *cursor) {
  const int nseg = seg_map.start[seg_map.R];
  const int lane = threadIdx.x, il = lane >> 3, jl = lane & 7;
  // A resident grid of waves takes segments from the queue until none are
  // left, across every replica's list.  Each segment's energy is rounded to
  // Q32.32 and added on its own, so the sums do not depend on which wave took
  // which segment.
  md_seg_queue Q(cursor, nseg, lane);
  int g;
  while (Q.next(g)) {
    int seg;
    const md_pair_args &A = T[md_locate(seg_map, g, seg)].P;
    int64_t efix = 0;
    int count = 0, blocks = 0;
    if (A.ewald_fit)
      md_pair_segment<ENERGY, true>(A, seg, lane, il, jl, efix, count, blocks);
    else
      md_pair_segment<ENERGY, false>(A, seg, lane, il, jl, efix, count, blocks);
    if (ENERGY) {
      int64_t ef = md_wave_sum_i64(efix);
      int64_t cf = md_wave_sum_i64((int64_t)count);
      if (lane == 0) {
        md_atomic_add_i64(A.energy + MD_E_PAIR, ef);
        md_atomic_add_i64(A.census, cf);
        md_atomic_add_i64(A.blocks, (int64_t)blocks);  // blocks is wave-uniform
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Generalized Hilbert walk over a W x H x D block of cells (the "gilbert"
// construction, J. Cerveny 2018): the block is split along its long axes
// into sub-blocks walked with rotated axes, each starting where the last
// ended.  Every cell is visited once; with even W, H and D every step goes
// to a face neighbour, so runs of consecutive cells -- the clusters and
// super-clusters of the list -- stay compact and never wrap the box.
namespace {
struct md_v3 {
  int x, y, z;
};
inline md_v3 operator+(md_v3 a, md_v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline md_v3 operator-(md_v3 a, md_v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline md_v3 operator-(md_v3 a) { return {-a.x, -a.y, -a.z}; }
inline int md_sgn(int v) { return (v > 0) - (v < 0); }
inline int md_floor_half(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }
inline md_v3 md_half(md_v3 a) { return {md_floor_half(a.x), md_floor_half(a.y), md_floor_half(a.z)}; }
inline md_v3 md_dir(md_v3 a) { return {md_sgn(a.x), md_sgn(a.y), md_sgn(a.z)}; }
inline int md_len(md_v3 a) { return std::abs(a.x + a.y + a.z); }

// walk the block at p spanned by the axis-aligned vectors a (major), b, c
void md_walk(md_v3 p, md_v3 a, md_v3 b, md_v3 c, std::vector<md_v3> &out) {
  const int w = md_len(a), h = md_len(b), d = md_len(c);
  const md_v3 da = md_dir(a), db = md_dir(b), dc = md_dir(c);
  if (h == 1 && d == 1) {
    for (int i = 0; i < w; ++i, p = p + da) out.push_back(p);
    return;
  }
  if (w == 1 && d == 1) {
    for (int i = 0; i < h; ++i, p = p + db) out.push_back(p);
    return;
  }
  if (w == 1 && h == 1) {
    for (int i = 0; i < d; ++i, p = p + dc) out.push_back(p);
    return;
  }
  md_v3 a2 = md_half(a), b2 = md_half(b), c2 = md_half(c);
  if ((md_len(a2) & 1) && w > 2) a2 = a2 + da;  // even halves where possible
  if ((md_len(b2) & 1) && h > 2) b2 = b2 + db;
  if ((md_len(c2) & 1) && d > 2) c2 = c2 + dc;
  if (2 * w > 3 * h && 2 * w > 3 * d) {  // long in a: two halves along a
    md_walk(p, a2, b, c, out);
    md_walk(p + a2, a - a2, b, c, out);
  } else if (3 * h > 4 * d) {  // flat in c: three blocks, c kept whole
    md_walk(p, b2, c, a2, out);
    md_walk(p + b2, a, b - b2, c, out);
    md_walk(p + (a - da) + (b2 - db), -b2, c, -(a - a2), out);
  } else if (3 * d > 4 * h) {  // flat in b: three blocks, b kept whole
    md_walk(p, c2, a2, b, out);
    md_walk(p + c2, a, b, c - c2, out);
    md_walk(p + (a - da) + (c2 - dc), -c2, -(a - a2), b, out);
  } else {  // the regular case: five blocks
    md_walk(p, b2, c2, a2, out);
    md_walk(p + b2, c, a2, b - b2, out);
    md_walk(p + (b2 - db) + (c - dc), a, -b2, -(c - c2), out);
    md_walk(p + (a - da) + b2 + (c - dc), -c, -(a - a2), b - b2, out);
    md_walk(p + (a - da) + (b2 - db), -b2, c2, -(a - a2), out);
  }
}
}  // namespace

std::vector<int> md_cell_walk(int W, int H, int D) {
  std::vector<md_v3> out;
  out.reserve((size_t)W * H * D);
  const md_v3 o{0, 0, 0}, X{W, 0, 0}, Y{0, H, 0}, Z{0, 0, D};
  if (W >= H && W >= D)
    md_walk(o, X, Y, Z, out);
  else if (H >= W && H >= D)
    md_walk(o, Y, X, Z, out);
  else
    md_walk(o, Z, X, Y, out);
  std::vector<int> cells;
  cells.reserve(out.size());
  for (const md_v3 &v : out) cells.push_back((v.x * H + v.y) * D + v.z);
  return cells;
}

struct md_nb_buffers {
  int64_t *cluster_center = nullptr, *super_origin = nullptr;
  float *super_radius = nullptr;
  int *super_count = nullptr, *super_segs = nullptr;
  int *entry_j = nullptr, *entry_imask = nullptr;
  int4 *entry_shift = nullptr;
  int4 *segments = nullptr;
  uint64_t *entry_excl = nullptr;
  uint64_t *excl_window = nullptr;
  int64_t *slot_force = nullptr;
  int *j_count = nullptr, *j_start = nullptr, *j_entries = nullptr;
  float *j_force = nullptr;
  int4 *cluster_lo = nullptr, *cluster_hi = nullptr;
  int *totals = nullptr;
  int nsuper = 0;
  int entry_capacity = 0, segment_capacity = 0, entries_per_super = 0;
  float half_min_width = 0, list_cutoff = 0;
};

static md_nb_buffers &nbuf(md_system *s) {
  static_assert(sizeof(md_nb_buffers) > 0, "");
  return *reinterpret_cast<md_nb_buffers *>(s->nb_ext);
}

// ---- the pair kernel's Coulomb polynomials ------------------------------------

namespace {
// g(u) = ((2/sqrt(pi)) e^-u - erf(sqrt u)/sqrt u) / u: below u = 1/2 by its
// Taylor series, (2/sqrt(pi)) sum_{n>=1} (-1)^n u^(n-1)/n! 2n/(2n+1), whose
// terms past n = 25 are below 1e-40
double md_ewald_g(double u) {
  const double sp = 2.0 / std::sqrt(M_PI);
  if (u < 0.5) {
    double term = -1.0, acc = 0.0;  // (-1)^n u^(n-1) / n!
    for (int n = 1; n <= 25; ++n) {
      if (n > 1) term *= -u / n;
      acc += term * (2.0 * n) / (2.0 * n + 1.0);
    }
    return sp * acc;
  }
  const double r = std::sqrt(u);
  return (sp * std::exp(-u) - std::erf(r) / r) / u;
}

// h(u) = erf(sqrt u)/sqrt u = (2/sqrt(pi)) sum_{n>=0} (-u)^n / (n! (2n+1))
double md_ewald_h(double u) {
  const double sp = 2.0 / std::sqrt(M_PI);
  if (u < 0.5) {
    double term = 1.0, acc = 0.0;
    for (int n = 0; n <= 25; ++n) {
      if (n > 0) term *= -u / n;
      acc += term / (2.0 * n + 1.0);
    }
    return sp * acc;
  }
  const double r = std::sqrt(u);
  return std::erf(r) / r;
}

float md_horner_host(const float *c, float v) {
  float p = c[MD_EWALD_TERMS - 1];
  for (int k = MD_EWALD_TERMS - 2; k >= 0; --k) p = std::fmaf(p, v, c[k]);
  return p;
}
}  // namespace

// Truncated Chebyshev series (from the 64-node interpolant on [0, U],
// U = alpha^2 rc^2 (1 + 1e-4)) of degree MD_EWALD_TERMS - 1, rewritten in
// powers of v = 2 r^2 / (rc^2 (1 + 1e-4)) - 1, coefficients rounded to f32.
// The fit is then evaluated the way the pair kernel does (v by one fma from
// the f32 r^2, Horner by fma) at 4,001 points of [0, rc] and compared with g
// and h in double; it is used only if the force error stays within 1e-6 of
// the bare 1/r^3 and the energy error within 1e-6 of 1/r.  That holds for
// alpha rc up to ~3 (Ewald tolerances down to ~1e-4, e.g. 2.63 for 5e-4);
// past it the kernel evaluates erfc and exp instead.
md_ewald_poly md_ewald_fit(double alpha, double cutoff) {
  md_ewald_poly P{};
  const int N = 64, D = MD_EWALD_TERMS - 1;
  const double rc2 = cutoff * cutoff * (1.0 + 1e-4), U = alpha * alpha * rc2;
  double cheb[2][MD_EWALD_TERMS] = {};
  for (int j = 0; j < N; ++j) {
    const double th = M_PI * (j + 0.5) / N, u = 0.5 * U * (1.0 + std::cos(th));
    const double fg = md_ewald_g(u), fh = md_ewald_h(u);
    for (int k = 0; k <= D; ++k) {
      const double w = (k == 0 ? 1.0 : 2.0) / N * std::cos(k * th);
      cheb[0][k] += w * fg;
      cheb[1][k] += w * fh;
    }
  }
  // T_k(v) in powers of v by T_{k+1} = 2 v T_k - T_{k-1}
  double t0[MD_EWALD_TERMS] = {1.0}, t1[MD_EWALD_TERMS] = {0.0, 1.0}, pw[2][MD_EWALD_TERMS] = {};
  for (int k = 0; k <= D; ++k) {
    const double *tk = k == 0 ? t0 : t1;
    for (int m = 0; m <= D; ++m) {
      pw[0][m] += cheb[0][k] * tk[m];
      pw[1][m] += cheb[1][k] * tk[m];
    }
    if (k >= 1) {
      double t2[MD_EWALD_TERMS] = {};
      for (int m = 0; m <= D; ++m) t2[m] = (m > 0 ? 2.0 * t1[m - 1] : 0.0) - t0[m];
      std::copy(t1, t1 + MD_EWALD_TERMS, t0);
      std::copy(t2, t2 + MD_EWALD_TERMS, t1);
    }
  }
  P.scale = (float)(2.0 / rc2);
  for (int m = 0; m <= D; ++m) {
    P.force[m] = (float)(pw[0][m] * alpha * alpha * alpha);
    P.energy[m] = (float)(-pw[1][m] * alpha);
  }
  const int M = 4000;
  for (int i = 1; i <= M; ++i) {
    const float r2f = (float)(cutoff * cutoff * i / M);
    const double r2 = r2f, r = std::sqrt(r2), u = alpha * alpha * r2;
    const float v = std::fmaf(r2f, P.scale, -1.0f);
    const double ef = std::fabs(md_horner_host(P.force, v) - alpha * alpha * alpha * md_ewald_g(u)) * r2 * r;
    const double ee = std::fabs(md_horner_host(P.energy, v) + alpha * md_ewald_h(u)) * r;
    P.force_err = std::max(P.force_err, ef);
    P.energy_err = std::max(P.energy_err, ee);
  }
  P.ok = P.force_err <= 1e-6 && P.energy_err <= 1e-6;
  return P;
}

void md_nb_setup(md_system *s) {
  auto &nb = s->nb;
  s->nb_ext = new md_nb_buffers();
  auto &B = nbuf(s);
  nb.nclusters = (s->n + MD_CLUSTER - 1) / MD_CLUSTER;
  B.nsuper = (nb.nclusters + 7) / 8;
  nb.nslots = B.nsuper * MD_SUPER;  // whole super-clusters
  double skin = s->opt.verlet_skin < 0 ? 0.15 : s->opt.verlet_skin;
  double rlist = s->cutoff + skin;
  nb.cutoff2 = (float)(s->cutoff * s->cutoff);
  nb.list_cutoff2 = (float)(rlist * rlist);
  nb.ewald = md_ewald_fit(s->alpha, s->cutoff);
  nb.half_skin2 = (float)(0.25 * skin * skin);
  B.list_cutoff = (float)rlist;
  // sorting cells ~0.3 nm along each fractional axis
  double widths[3];
  {
    const double *b = s->box;
    double v = std::fabs(b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                         b[2] * (b[3] * b[7] - b[4] * b[6]));
    auto cross_norm = [](const double *u, const double *w) {
      double x = u[1] * w[2] - u[2] * w[1], y = u[2] * w[0] - u[0] * w[2], z = u[0] * w[1] - u[1] * w[0];
      return std::sqrt(x * x + y * y + z * z);
    };
    widths[0] = v / cross_norm(b + 3, b + 6);
    widths[1] = v / cross_norm(b + 6, b);
    widths[2] = v / cross_norm(b, b + 3);
  }
  B.half_min_width = (float)(0.5 * std::min(widths[0], std::min(widths[1], widths[2])));
  // cells of at least 0.3 nm, an even number along each axis so that the
  // sorting walk is continuous
  for (int d = 0; d < 3; ++d) nb.gdim[d] = std::max(2, 2 * (int)std::floor(widths[d] / 0.6));
  nb.ncells = nb.gdim[0] * nb.gdim[1] * nb.gdim[2];
  std::vector<int> walk = md_cell_walk(nb.gdim[0], nb.gdim[1], nb.gdim[2]);
  if ((int)walk.size() != nb.ncells) throw md_error(MD_ERR_INVALID_ARGUMENT, "cell walk does not cover the grid");
  std::vector<int> rank(nb.ncells, -1);
  for (int k = 0; k < nb.ncells; ++k) rank[walk[k]] = k;
  for (int k = 0; k < nb.ncells; ++k)
    if (rank[k] < 0) throw md_error(MD_ERR_INVALID_ARGUMENT, "cell walk misses a cell");

  int np = s->np;
  s->cell_rank = md_alloc<int>(s, nb.ncells);
  MD_HIP(hipMemcpyAsync(s->cell_rank, rank.data(), sizeof(int) * nb.ncells, hipMemcpyHostToDevice, s->stream));
  s->cell_count = md_alloc<int>(s, nb.ncells);
  s->cell_cursor = md_alloc<int>(s, nb.ncells);
  s->cell_start = md_alloc<int>(s, nb.ncells + 1);
  s->atom_cell = md_alloc<int>(s, np);
  s->cell_member = md_alloc<int>(s, np);
  s->slot_atom = md_alloc<int>(s, nb.nslots);
  std::vector<int> pad(nb.nslots, -1);
  MD_HIP(hipMemcpyAsync(s->slot_atom, pad.data(), sizeof(int) * nb.nslots, hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
  s->atom_slot = md_alloc<int>(s, np);
  s->atom_folded = md_alloc<int64_t>(s, 3 * (size_t)np);
  s->build_pos = md_alloc<int64_t>(s, 3 * (size_t)np);
  s->slot_offset = md_alloc<int64_t>(s, 3 * (size_t)nb.nslots);
  s->slot_lj = md_alloc<float2>(s, nb.nslots);
  s->slot_xyzq = md_alloc<int4>(s, nb.nslots);
  s->cluster_half = md_alloc<float4>(s, nb.nclusters);
  B.cluster_center = md_alloc<int64_t>(s, 3 * (size_t)nb.nclusters);
  B.super_origin = md_alloc<int64_t>(s, 3 * (size_t)B.nsuper);
  B.super_radius = md_alloc<float>(s, B.nsuper);
  B.super_count = md_alloc<int>(s, B.nsuper);
  B.super_segs = md_alloc<int>(s, B.nsuper + 1);
  B.totals = md_alloc<int>(s, 4);
  // list capacity: the cluster pairs within the list radius plus the cluster
  // extents at the system's density, doubled for headroom
  double volume = std::fabs(s->box[0] * s->box[4] * s->box[8]);
  double density = s->n / volume;
  double ext = std::cbrt(MD_CLUSTER / std::max(density, 1.0));
  // a super-cluster (a compact run of the cell walk, side ~ (64/density)^1/3)
  // meets the j-clusters within the list radius plus a cluster's extent of its
  // box: 1.5x that count, or every cluster at most (729 images of each)
  double side = std::cbrt(MD_SUPER / std::max(density, 1.0));
  double span = side + 2.0 * (rlist + ext);
  double per_super = span * span * span * density / MD_CLUSTER;
  // per super-cluster region, a multiple of the shortest segment
  long region = (long)(1.5 * std::min(per_super, 729.0 * nb.nclusters)) + 64;
  region = std::min((region + MD_SEGMENT - 1) / MD_SEGMENT * MD_SEGMENT, (long)INT32_MAX / 16 / B.nsuper);
  B.entries_per_super = (int)region;
  B.entry_capacity = (int)(region * B.nsuper);
  B.segment_capacity = B.nsuper * (int)(region / MD_SEGMENT) + 64;
  B.entry_j = md_alloc<int>(s, B.entry_capacity);
  B.entry_imask = md_alloc<int>(s, B.entry_capacity);
  B.entry_shift = md_alloc<int4>(s, B.entry_capacity);
  B.entry_excl = md_alloc<uint64_t>(s, (size_t)B.entry_capacity * 8);
  B.j_count = md_alloc<int>(s, nb.nclusters);
  B.j_start = md_alloc<int>(s, nb.nclusters + 1);
  B.j_entries = md_alloc<int>(s, B.entry_capacity);
  B.j_force = md_alloc<float>(s, (size_t)B.entry_capacity * 24);
  B.cluster_lo = md_alloc<int4>(s, nb.nslots / MD_CLUSTER);
  B.cluster_hi = md_alloc<int4>(s, nb.nslots / MD_CLUSTER);
  B.segments = md_alloc<int4>(s, B.segment_capacity);
  s->excl_start_d = md_alloc<int>(s, s->n + 1);
  s->excl_list_d = md_alloc<int>(s, std::max<size_t>(1, s->excl_list.size()));
  MD_HIP(hipMemcpyAsync(s->excl_start_d, s->excl_start.data(), sizeof(int) * (s->n + 1), hipMemcpyHostToDevice,
                        s->stream));
  if (!s->excl_list.empty())
    MD_HIP(hipMemcpyAsync(s->excl_list_d, s->excl_list.data(), sizeof(int) * s->excl_list.size(),
                          hipMemcpyHostToDevice, s->stream));
  // each atom's exclusions within 32 indices as a bit window (bit d + 32 for
  // atom i + d, d in [-32, 31], d != 0); bit 32 marks an atom that also has
  // exclusions further away, which the mask kernel then looks up in the CSR
  std::vector<uint64_t> window(s->n, 0);
  for (int i = 0; i < s->n; ++i)
    for (int k = s->excl_start[i]; k < s->excl_start[i + 1]; ++k) {
      const int d = s->excl_list[k] - i;
      if (d == 0) continue;
      window[i] |= d >= -32 && d < 32 ? uint64_t(1) << (d + 32) : uint64_t(1) << 32;
    }
  B.excl_window = md_alloc<uint64_t>(s, s->n);
  B.slot_force = md_alloc<int64_t>(s, 3 * (size_t)nb.nslots);
  MD_HIP(hipMemcpyAsync(B.excl_window, window.data(), sizeof(uint64_t) * s->n, hipMemcpyHostToDevice, s->stream));
  MD_HIP(hipStreamSynchronize(s->stream));
}

void md_nb_free(md_system *s) {
  delete reinterpret_cast<md_nb_buffers *>(s->nb_ext);
  s->nb_ext = nullptr;
}

// The list context's box-dependent fields, from a reduced box b and its
// constants: the same on the host (table set-up) and on the device (a box
// changed by the Monte Carlo barostat).
__host__ __device__ inline void md_nb_box_fields(md_build_ctx &c, const double *b, const md_box &d) {
#pragma clang fp contract(off)
  double w[3];
  md_box_widths(b, w);
  // half the narrowest perpendicular width of the current box
  c.half_min_width = (float)(0.5 * fmin(w[0], fmin(w[1], w[2])));
  // md_min_image_fixed rounds along c, then b, then a on Cartesian components,
  // which leaves the fractional coordinates of the reduced difference within
  // fc = 1/2, fb = 1/2 + fc |cy|/by and fa = 1/2 + (fb |bx| + fc |cx|)/ax (plus
  // a margin for the f32 rounding); a rectangular box gives 1/2 on every axis
  const double fc = 0.5, fb = 0.5 + fc * fabs(b[7]) / b[4], fa = 0.5 + (fb * fabs(b[3]) + fc * fabs(b[6])) / b[0];
  const double f[3] = {fa, fb, fc};
  for (int k = 0; k < 3; ++k) {
    c.width[k] = (float)(w[k] * (1.0 - 1e-5));
    c.fmax[k] = (float)(f[k] + 1e-5);
  }
  c.box = d;
}

static md_build_ctx build_ctx(md_system *s) {
  auto &B = nbuf(s);
  md_build_ctx c{};
  c.n = s->n;
  c.np = s->np;
  c.ncells = s->nb.ncells;
  c.gdim0 = s->nb.gdim[0];
  c.gdim1 = s->nb.gdim[1];
  c.gdim2 = s->nb.gdim[2];
  c.nclusters = s->nb.nclusters;
  c.nsuper = B.nsuper;
  c.nslots = s->nb.nslots;
  c.entries_per_super = B.entries_per_super;
  c.segment_capacity = B.segment_capacity;
  c.list_cutoff2 = s->nb.list_cutoff2;
  c.list_cutoff = B.list_cutoff;
  md_nb_box_fields(c, s->box, s->dbox);
  return c;
}

// one replica's table entry
static md_nb_args nb_args(md_system *s, int seg_len) {
  auto &B = nbuf(s);
  md_nb_args a{};
  a.c = build_ctx(s);
  a.c.seg_len = seg_len;
  md_list_ptrs &L = a.L;
  L.cluster_center = B.cluster_center;
  L.cluster_half = s->cluster_half;
  L.super_origin = B.super_origin;
  L.super_radius = B.super_radius;
  L.super_count = B.super_count;
  L.super_segs = B.super_segs;
  L.entry_j = B.entry_j;
  L.entry_imask = B.entry_imask;
  L.entry_shift = B.entry_shift;
  L.segments = B.segments;
  L.entry_excl = B.entry_excl;
  L.j_count = B.j_count;
  L.j_start = B.j_start;
  L.j_entries = B.j_entries;
  md_pair_args &P = a.P;
  P.segments = B.segments;
  P.entry_j = B.entry_j;
  P.entry_imask = B.entry_imask;
  P.entry_shift = B.entry_shift;
  P.entry_excl = B.entry_excl;
  P.j_force = B.j_force;
  P.cluster_lo = B.cluster_lo;
  P.cluster_hi = B.cluster_hi;
  P.slot_xyzq = s->slot_xyzq;
  P.slot_lj = s->slot_lj;
  P.totals = B.totals;
  P.slot_force = B.slot_force;
  P.energy = s->energy;
  P.census = &s->ctl->pair_census;
  P.blocks = &s->ctl->pair_blocks;
  P.nslots = s->nb.nslots;
  P.nclusters = s->nb.nclusters;
  P.cutoff2 = std::ldexp(s->nb.cutoff2, 52);
  P.prune2 = std::ldexp((float)((double)s->nb.cutoff2 * (1.0 + 1e-5)), 52);
  P.half_alpha = std::ldexp(0.5f * (float)s->alpha, -26);
  P.beta = std::ldexp((float)(2.0 * s->alpha / std::sqrt(M_PI)), -26);
  P.exp_scale = std::ldexp((float)(-s->alpha * s->alpha / std::log(2.0)), -52);
  const md_ewald_poly &W = s->nb.ewald;
  P.ewald_fit = W.ok;
  P.fit_scale = std::ldexp(W.scale, -52);
  for (int k = 0; k < MD_EWALD_TERMS; ++k) {
    P.fit_force[k] = std::ldexp(W.force[k], -78);
    P.fit_energy[k] = std::ldexp(W.energy[k], -26);
  }
  a.pos = s->pos;
  a.cell_rank = s->cell_rank;
  a.cell_count = s->cell_count;
  a.cell_start = s->cell_start;
  a.cell_cursor = s->cell_cursor;
  a.atom_cell = s->atom_cell;
  a.cell_member = s->cell_member;
  a.slot_atom = s->slot_atom;
  a.atom_slot = s->atom_slot;
  a.folded = s->atom_folded;
  a.build_pos = s->build_pos;
  a.slot_offset = s->slot_offset;
  a.cluster_center = B.cluster_center;
  a.super_origin = B.super_origin;
  a.super_radius = B.super_radius;
  a.slot_lj = s->slot_lj;
  a.slot_xyzq = s->slot_xyzq;
  a.cluster_half = s->cluster_half;
  a.atom_param = s->atom_param;
  a.excl_start = s->excl_start_d;
  a.excl_list = s->excl_list_d;
  a.excl_window = B.excl_window;
  a.force = s->force;
  a.cluster_lo = B.cluster_lo;
  a.cluster_hi = B.cluster_hi;
  a.totals = B.totals;
  a.ctl = s->ctl;
  return a;
}

// a group's launch tables for the nonbonded kernels
struct md_nb_group {
  md_nb_args *table = nullptr;  // [R], device
  md_launch atoms, scan, clusters, supers, segments, masks, slots;
  md_launch seg;                // the pair kernel's segment numbering
  int *cursor = nullptr;
  int pair_grid[2] = {1, 1};
  int seg_len = MD_SEGMENT;
  bool compact = false;  // forces through the compacted schedule
};

static md_nb_group &nbg(md_group *g) { return *reinterpret_cast<md_nb_group *>(g->nb); }

void md_nb_group_setup(md_group *g) {
  auto *G = new md_nb_group();
  g->nb = G;
  const int R = (int)g->reps.size();
  std::vector<int> atoms, one, clusters, supers, segments, masks, slots;
  for (md_system *s : g->reps) {
    auto &B = nbuf(s);
    atoms.push_back((s->n + 255) / 256);
    one.push_back(1);
    clusters.push_back((s->nb.nclusters + 63) / 64);
    supers.push_back(B.nsuper);
    segments.push_back((B.nsuper + 63) / 64);
    masks.push_back(B.nsuper * (MD_MASK_WAVES / MD_MASK_GROUP));
    slots.push_back((s->nb.nslots + 255) / 256);
  }
  G->atoms = md_make_launch(g, atoms);
  G->scan = md_make_launch(g, one);
  G->clusters = md_make_launch(g, clusters);
  G->supers = md_make_launch(g, supers);
  G->segments = md_make_launch(g, segments);
  G->masks = md_make_launch(g, masks);
  G->slots = md_make_launch(g, slots);
  G->seg.map.start = md_group_alloc<int>(g, R + 1);
  G->seg.map.R = R;
  G->cursor = md_group_alloc<int>(g, MD_XCDS * MD_CURSOR_STRIDE);
  G->table = md_group_alloc<md_nb_args>(g, R);
  // the pair kernel keeps one resident grid busy from the segment counter
  int cus = 1, per_cu[2] = {1, 1};
  MD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
  MD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[0], k_nb_pairs<false>, 64, 0));
  MD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[1], k_nb_pairs<true>, 64, 0));
  long capacity = 0;
  for (md_system *s : g->reps) capacity += nbuf(s).segment_capacity;
  for (int k = 0; k < 2; ++k)
    G->pair_grid[k] = (int)std::max(1L, std::min(capacity, (long)std::max(1, cus) * std::max(1, per_cu[k])));
  // Longer segments flush each i-force once per more entries (the flushes
  // are half of the kernel's 64-bit atomics at 8 entries), but leave fewer
  // units to share among the resident waves: double the length while at
  // least 8 units per wave remain, at about 180 entries per super-cluster.
  double entries = 0.0;
  for (md_system *s : g->reps) entries += 180.0 * nbuf(s).nsuper;
  const double waves = (double)G->pair_grid[0];
  while (G->seg_len < 4 * MD_SEGMENT && entries / (2.0 * G->seg_len) >= 8.0 * waves) G->seg_len *= 2;
  md_nb_group_refresh(g);
}

// The pair kernel's schedule: list entries per work unit (a multiple of
// MD_SEGMENT; the i-force rounding happens per MD_SEGMENT entries, so any
// length gives the same bits) and the resident grid of waves that draws units
// from the counter.  waves <= 0 keeps the current grid.
void md_nb_set_schedule(md_group *g, int seg_len, int waves) {
  auto &G = nbg(g);
  if (seg_len % MD_SEGMENT || seg_len < MD_SEGMENT || seg_len > 32 * MD_SEGMENT)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "segment length must be a multiple of 8 entries, at most 256");
  G.seg_len = seg_len;
  if (waves > 0) G.pair_grid[0] = waves;
  md_nb_group_refresh(g);
}

void md_nb_get_schedule(md_group *g, int *seg_len, int *waves) {
  *seg_len = nbg(g).seg_len;
  *waves = nbg(g).pair_grid[0];
}

void md_nb_set_compact(md_group *g, bool compact) { nbg(g).compact = compact; }
bool md_nb_get_compact(md_group *g) { return nbg(g).compact; }

void md_nb_group_refresh(md_group *g) {
  std::vector<md_nb_args> host;
  for (md_system *s : g->reps) host.push_back(nb_args(s, nbg(g).seg_len));
  MD_HIP(hipMemcpyAsync(nbg(g).table, host.data(), sizeof(md_nb_args) * host.size(), hipMemcpyHostToDevice,
                        g->stream));
  MD_HIP(hipStreamSynchronize(g->stream));
}

// The list context of replica 0 from a box in device memory (a barostat move):
// the values md_nb_group_refresh would upload for that box.
__global__ void k_nb_box(md_nb_args *T, const double *box) {
  md_box d{};
  bool bad = false;
  md_box_from_matrix(box, d, bad);
  md_nb_box_fields(T[0].c, box, d);
}

void md_nb_box_from_device(md_group *g, const double *box) {
  k_nb_box<<<1, 1, 0, g->stream>>>(nbg(g).table, box);
  MD_HIP(hipGetLastError());
}

void md_nb_group_free(md_group *g) {
  delete &nbg(g);
  g->nb = nullptr;
}

void md_nb_build(md_group *g) {
  auto &G = nbg(g);
  hipStream_t st = g->stream;
  const md_nb_args *T = G.table;
  k_nb_assign<<<G.atoms.total, 256, 0, st>>>(T, G.atoms.map);
  k_nb_scan_cells<<<G.scan.total, 1024, 0, st>>>(T, G.scan.map);
  k_nb_scatter<<<G.atoms.total, 256, 0, st>>>(T, G.atoms.map);
  k_nb_rank<<<G.atoms.total, 256, 0, st>>>(T, G.atoms.map);
  k_nb_clusters<<<G.clusters.total, 64, 0, st>>>(T, G.clusters.map);
  md_mark(g, MD_PHASE_SEARCH);
  k_nb_list<<<G.supers.total, 64, 0, st>>>(T, G.supers.map);
  k_nb_list_scan<<<G.scan.total, 1024, 0, st>>>(T, G.scan.map);
  k_nb_segments<<<G.segments.total, 64, 0, st>>>(T, G.segments.map);
  md_mark(g, MD_PHASE_MASKS);
  k_nb_masks<<<G.masks.total, 64 * MD_MASK_GROUP, 0, st>>>(T, G.masks.map);
  k_nb_j_scan<<<G.scan.total, 1024, 0, st>>>(T, G.scan.map);
  k_nb_j_fill<<<G.supers.total, 64, 0, st>>>(T, G.supers.map);
  MD_HIP(hipGetLastError());
}

void md_nb_pairs(md_group *g, bool energy) {
  auto &G = nbg(g);
  hipStream_t st = g->stream;
  k_nb_pack<<<G.slots.total, 256, 0, st>>>(G.table, G.slots.map, const_cast<int *>(G.seg.map.start), G.cursor);
  md_mark(g, MD_PHASE_PAIRS);
  const bool batch = g->reps.size() > 1;
  if (energy) {
    if (batch)
      k_nb_pairs<true, true><<<G.pair_grid[1], 64, 0, st>>>(G.table, G.seg.map, G.cursor);
    else
      k_nb_pairs<true, false><<<G.pair_grid[1], 64, 0, st>>>(G.table, G.seg.map, G.cursor);
  } else if (G.compact) {
    if (batch)
      k_nb_pairs_compact<true><<<G.pair_grid[0], 64, 0, st>>>(G.table, G.seg.map, G.cursor);
    else
      k_nb_pairs_compact<false><<<G.pair_grid[0], 64, 0, st>>>(G.table, G.seg.map, G.cursor);
  } else {
    if (batch)
      k_nb_pairs<false, true><<<G.pair_grid[0], 64, 0, st>>>(G.table, G.seg.map, G.cursor);
    else
      k_nb_pairs<false, false><<<G.pair_grid[0], 64, 0, st>>>(G.table, G.seg.map, G.cursor);
  }
  md_mark(g, MD_PHASE_FORCES);
  k_nb_forces<<<G.slots.total, 256, 0, st>>>(G.table, G.slots.map);
  MD_HIP(hipGetLastError());
}

int64_t md_nb_entries(md_system *s) {
  int t[2];
  MD_HIP(hipMemcpy(t, nbuf(s).totals, sizeof t, hipMemcpyDeviceToHost));
  return t[0];
}

int64_t md_nb_capacity(md_system *s) { return nbuf(s).entry_capacity; }
