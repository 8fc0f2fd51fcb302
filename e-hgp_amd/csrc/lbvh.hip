// Canonical cloud and Morton LBVH on gfx950.
//
// Canonical PointIds (reference morsehgp3d/src/cpu/spatial/point_cloud.cpp:40-95):
// lexicographic order of the three binary64 total-order keys, -0 -> +0,
// duplicates rejected.  Morton codes (reference lbvh.cpp:78-90, 237-298):
// 21 bits per axis, q = floor(2^21 (x - l) / (u - l)) exactly, certified here
// with error-free FP64 transforms.  Tree: the reference's top-down
// highest-differing-bit split with midpoint splits inside runs of equal codes
// (lbvh.cpp:831-906), built in one pass with Karras' radix-tree construction
// over the augmented key (code, midpoint path bits).  Boxes carry PointId
// witnesses with the reference's tie rule (lbvh.cpp:46-76).
#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "exact_device.hpp"

namespace mhgp {

namespace {

__global__ void k_order_keys(const double *__restrict__ src, int64_t n, uint64_t *k0,
                             uint64_t *k1, uint64_t *k2, uint32_t *perm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  k0[i] = order_key(dbits(src[3 * i]));
  k1[i] = order_key(dbits(src[3 * i + 1]));
  k2[i] = order_key(dbits(src[3 * i + 2]));
  perm[i] = (uint32_t)i;
}

__global__ void k_gather_key(const uint64_t *__restrict__ from, const uint32_t *__restrict__ perm,
                             int64_t n, uint64_t *to) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) to[i] = from[perm[i]];
}

// After the three stable passes perm lists source indices in canonical order.
__global__ void k_canonical(const double *__restrict__ src, const uint32_t *__restrict__ perm,
                            int64_t n, double *X, double *Y, double *Z, uint32_t *source_index,
                            uint32_t *status) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s = perm[i];
  double x = src[3 * s] + 0.0, y = src[3 * s + 1] + 0.0, z = src[3 * s + 2] + 0.0;  // -0 -> +0
  X[i] = x;
  Y[i] = y;
  Z[i] = z;
  source_index[i] = s;
  if (i > 0) {
    uint32_t t = perm[i - 1];
    if (dbits(src[3 * t] + 0.0) == dbits(x) && dbits(src[3 * t + 1] + 0.0) == dbits(y) &&
        dbits(src[3 * t + 2] + 0.0) == dbits(z))
      atomicOr(&status[0], 1u);
  }
  double v[3] = {x, y, z};
  for (int a = 0; a < 3; ++a)
    if (!isfinite(v[a])) atomicOr(&status[0], 2u);
}

// Global AABB with PointId witnesses: the first (smallest) PointId attaining
// each extremum, as the reference's sequential scan (lbvh.cpp:105-151).
__global__ void k_bounds(const double *__restrict__ X, const double *__restrict__ Y,
                         const double *__restrict__ Z, int64_t n, double *bounds,
                         uint32_t *ids) {
  __shared__ double sv[6][512];
  __shared__ uint32_t si[6][512];
  const int t = threadIdx.x;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  uint32_t w[6] = {kInvalid, kInvalid, kInvalid, kInvalid, kInvalid, kInvalid};
  for (int64_t i = t; i < n; i += blockDim.x) {
    double p[3] = {X[i], Y[i], Z[i]};
    for (int a = 0; a < 3; ++a) {
      if (p[a] < v[a]) { v[a] = p[a]; w[a] = (uint32_t)i; }
      if (p[a] > v[3 + a]) { v[3 + a] = p[a]; w[3 + a] = (uint32_t)i; }
    }
  }
  for (int k = 0; k < 6; ++k) { sv[k][t] = v[k]; si[k][t] = w[k]; }
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) {
      for (int k = 0; k < 6; ++k) {
        double a = sv[k][t], b = sv[k][t + off];
        uint32_t ia = si[k][t], ib = si[k][t + off];
        bool take = (k < 3) ? (b < a || (b == a && ib < ia)) : (b > a || (b == a && ib < ia));
        if (take) { sv[k][t] = b; si[k][t] = ib; }
      }
    }
    __syncthreads();
  }
  if (t < 6) { bounds[t] = sv[t][0]; ids[t] = si[t][0]; }
}

// floor(2^21 (x - l) / (u - l)), certified exactly.
__device__ uint32_t quantize(double x, double l, double u, uint32_t *fallbacks) {
  const uint32_t qmax = (1u << 21) - 1;
  if (u == l) return 0;
  if (x == u) return qmax;
  double eh, el, dh, dl;
  two_diff(u, l, eh, el);
  two_diff(x, l, dh, dl);
  double t = (dh / eh) * 2097152.0;
  int64_t q = (int64_t)floor(t);
  if (q < 0) q = 0;
  if (q > (int64_t)qmax) q = qmax;
  // accept q when q*E <= 2^21 D < (q+1)*E holds exactly
  for (int iter = 0; iter < 4; ++iter) {
    double terms[6];
    double qd = (double)q;
    double p, pe, r, re;
    two_prod(qd, eh, p, pe);
    two_prod(qd, el, r, re);
    terms[0] = dh * 2097152.0; terms[1] = dl * 2097152.0;
    terms[2] = -p; terms[3] = -pe; terms[4] = -r; terms[5] = -re;
    int lo_sign = exact_sum_sign(terms, 6);  // 2^21 D - q E
    double qd1 = (double)(q + 1);
    two_prod(qd1, eh, p, pe);
    two_prod(qd1, el, r, re);
    terms[2] = -p; terms[3] = -pe; terms[4] = -r; terms[5] = -re;
    int hi_sign = exact_sum_sign(terms, 6);  // 2^21 D - (q+1) E
    if (lo_sign >= 0 && hi_sign < 0) {
      if (iter) atomicAdd(fallbacks, 1u);
      return (uint32_t)q;
    }
    if (lo_sign < 0) q--; else q++;
    if (q < 0) q = 0;
    if (q > (int64_t)qmax) return qmax;
  }
  atomicAdd(fallbacks + 1, 1u);  // unresolved: reported to the host as a failure
  return (uint32_t)q;
}

__device__ inline uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void k_morton(const double *__restrict__ X, const double *__restrict__ Y,
                         const double *__restrict__ Z, int64_t n, const double *__restrict__ b,
                         uint64_t *codes, uint32_t *ids, uint32_t *status) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t qx = quantize(X[i], b[0], b[3], status + 1);
  uint32_t qy = quantize(Y[i], b[1], b[4], status + 1);
  uint32_t qz = quantize(Z[i], b[2], b[5], status + 1);
  codes[i] = (spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz);
  ids[i] = (uint32_t)i;
}

__global__ void k_leaves(const uint32_t *__restrict__ ids, const double *__restrict__ X,
                         const double *__restrict__ Y, const double *__restrict__ Z,
                         const uint64_t *__restrict__ codes, int64_t n, double *LX, double *LY,
                         double *LZ, uint32_t *leaf_pos, uint64_t *path) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t id = ids[p];
  LX[p] = X[id];
  LY[p] = Y[id];
  LZ[p] = Z[id];
  leaf_pos[id] = (uint32_t)p;
  // midpoint path bits inside a run of equal codes (reference find_split)
  uint64_t c = codes[p];
  uint64_t bits = 0;
  bool inrun = (p > 0 && codes[p - 1] == c) || (p + 1 < n && codes[p + 1] == c);
  if (inrun) {
    int64_t a = p, e = p + 1;
    while (a > 0 && codes[a - 1] == c) a--;
    while (e < n && codes[e] == c) e++;
    int depth = 0;
    while (e - a > 1) {
      int64_t mid = a + (e - a) / 2;
      if (p >= mid) { bits |= 1ull << (63 - depth); a = mid; } else { e = mid; }
      depth++;
    }
  }
  path[p] = bits;
}

__device__ inline int delta(const uint64_t *codes, const uint64_t *path, int64_t n, int64_t i,
                            int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint64_t a = codes[i], b = codes[j];
  if (a != b) return __clzll(a ^ b);
  return 64 + __clzll(path[i] ^ path[j]);
}

__global__ void k_karras(const uint64_t *__restrict__ codes, const uint64_t *__restrict__ path,
                         int64_t n, Node *nodes, uint32_t *parent_int, uint32_t *parent_leaf,
                         uint32_t *counter) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(codes, path, n, i, i + 1) - delta(codes, path, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(codes, path, n, i, i - d);
  int64_t lmax = 2;
  while (delta(codes, path, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(codes, path, n, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(codes, path, n, i, j);
  int64_t s = 0;
  int64_t t = l;
  do {
    t = (t + 1) / 2;
    if (delta(codes, path, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  Node nd;
  nd.begin = (uint32_t)lo;
  nd.end = (uint32_t)(hi + 1);
  if (lo == gamma) { nd.left = kLeafBit | (uint32_t)gamma; parent_leaf[gamma] = (uint32_t)i; }
  else { nd.left = (uint32_t)gamma; parent_int[gamma] = (uint32_t)i; }
  if (hi == gamma + 1) { nd.right = kLeafBit | (uint32_t)(gamma + 1); parent_leaf[gamma + 1] = (uint32_t)i; }
  else { nd.right = (uint32_t)(gamma + 1); parent_int[gamma + 1] = (uint32_t)i; }
  for (int a = 0; a < 3; ++a) { nd.lo[a] = 0; nd.hi[a] = 0; }
  nodes[i] = nd;
  counter[i] = 0;
  if (i == 0) parent_int[0] = kInvalid;
}

__device__ inline void child_box(uint32_t c, const Node *nodes, const uint32_t *wit,
                                 const double *LX, const double *LY, const double *LZ,
                                 const uint32_t *leaf_id, double lo[3], double hi[3],
                                 uint32_t wl[3], uint32_t wh[3]) {
  if (c & kLeafBit) {
    uint32_t p = c & ~kLeafBit;
    double v[3] = {LX[p], LY[p], LZ[p]};
    uint32_t id = leaf_id[p];
    for (int a = 0; a < 3; ++a) { lo[a] = hi[a] = v[a]; wl[a] = wh[a] = id; }
  } else {
    // volatile loads: the sibling thread wrote this box before its fence
    const volatile Node *vn = nodes + c;
    const volatile uint32_t *vw = wit + 6 * (size_t)c;
    for (int a = 0; a < 3; ++a) {
      lo[a] = vn->lo[a]; hi[a] = vn->hi[a];
      wl[a] = vw[a]; wh[a] = vw[3 + a];
    }
  }
}

__global__ void k_refit(int64_t n, Node *nodes, uint32_t *wit, const uint32_t *__restrict__ parent_int,
                        const uint32_t *__restrict__ parent_leaf, uint32_t *counter,
                        const double *__restrict__ LX, const double *__restrict__ LY,
                        const double *__restrict__ LZ, const uint32_t *__restrict__ leaf_id) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t node = parent_leaf[p];
  while (node != kInvalid) {
    __threadfence();
    if (atomicAdd(&counter[node], 1u) == 0) return;  // the sibling subtree is not done yet
    __threadfence();
    Node nd = nodes[node];
    double l0[3], h0[3], l1[3], h1[3];
    uint32_t wl0[3], wh0[3], wl1[3], wh1[3];
    child_box(nd.left, nodes, wit, LX, LY, LZ, leaf_id, l0, h0, wl0, wh0);
    child_box(nd.right, nodes, wit, LX, LY, LZ, leaf_id, l1, h1, wl1, wh1);
    volatile Node *vn = nodes + node;
    volatile uint32_t *vw = wit + 6 * (size_t)node;
    for (int a = 0; a < 3; ++a) {
      bool tl = l1[a] < l0[a] || (l1[a] == l0[a] && wl1[a] < wl0[a]);
      bool th = h1[a] > h0[a] || (h1[a] == h0[a] && wh1[a] < wh0[a]);
      vn->lo[a] = tl ? l1[a] : l0[a];
      vw[a] = tl ? wl1[a] : wl0[a];
      vn->hi[a] = th ? h1[a] : h0[a];
      vw[3 + a] = th ? wh1[a] : wh0[a];
    }
    node = parent_int[node];
  }
}

// Postorder index and depth of every node (internal i < n-1, leaf n-1+p).
__global__ void k_postorder(int64_t n, const Node *__restrict__ nodes,
                            const uint32_t *__restrict__ parent_int,
                            const uint32_t *__restrict__ parent_leaf, uint64_t *post,
                            uint32_t *depth) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n - 1) return;
  if (n == 1) { post[0] = 0; depth[0] = 0; return; }
  bool leaf = t >= n - 1;
  uint32_t cur = leaf ? (uint32_t)(t - (n - 1)) | kLeafBit : (uint32_t)t;
  uint32_t par = leaf ? parent_leaf[t - (n - 1)] : parent_int[t];
  uint64_t w = 0;
  uint32_t dep = 0;
  while (par != kInvalid) {
    const Node &P = nodes[par];
    if (P.right == cur) {
      w += 1;
    } else {
      uint32_t r = P.right;
      uint64_t leaves_r = (r & kLeafBit) ? 1 : (uint64_t)(nodes[r].end - nodes[r].begin);
      w += 2 * leaves_r;
    }
    dep++;
    cur = par;
    par = parent_int[par];
  }
  post[t] = (uint64_t)(2 * n - 2) - w;
  depth[t] = dep;
}

__global__ void k_snapshot(int64_t n, const Node *__restrict__ nodes, const uint32_t *__restrict__ wit,
                           const uint32_t *__restrict__ leaf_id, const uint64_t *__restrict__ post,
                           uint64_t *out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n - 1) return;
  if (post[t] >= (uint64_t)(2 * n - 1)) return;  // malformed tree: the host sees a zero record
  uint64_t *o = out + 10 * post[t];
  if (t >= n - 1) {
    uint64_t p = t - (n - 1);
    uint64_t id = leaf_id[p];
    for (int a = 0; a < 6; ++a) o[a] = id;
    o[6] = ~0ull; o[7] = ~0ull; o[8] = p; o[9] = p + 1;
    return;
  }
  const Node &nd = nodes[t];
  for (int a = 0; a < 6; ++a) o[a] = wit[6 * t + a];
  uint32_t l = nd.left, r = nd.right;
  o[6] = post[(l & kLeafBit) ? (n - 1) + (l & ~kLeafBit) : l];
  o[7] = post[(r & kLeafBit) ? (n - 1) + (r & ~kLeafBit) : r];
  o[8] = nd.begin;
  o[9] = nd.end;
}

// FP32 copy of the LBVH for ball queries: each internal node lists its
// grandchildren (a leaf child stands for itself) with boxes rounded outward
// (a leaf's box is its point).
__device__ inline void wide_entry(const Node *nodes, const double *LX, const double *LY, const double *LZ,
                                  uint32_t x, WideNode &w, int slot) {
  double lo[3], hi[3];
  if (x & kLeafBit) {
    const uint32_t lp = x & ~kLeafBit;
    lo[0] = hi[0] = LX[lp];
    lo[1] = hi[1] = LY[lp];
    lo[2] = hi[2] = LZ[lp];
  } else {
    for (int a = 0; a < 3; ++a) { lo[a] = nodes[x].lo[a]; hi[a] = nodes[x].hi[a]; }
  }
  for (int a = 0; a < 3; ++a) {
    w.lo[slot][a] = __double2float_rd(lo[a]);
    w.hi[slot][a] = __double2float_ru(hi[a]);
  }
  w.child[slot] = x;
}

__global__ void k_wide_nodes(int64_t m, const Node *__restrict__ nodes, const double *__restrict__ LX,
                             const double *__restrict__ LY, const double *__restrict__ LZ, WideNode *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const Node &N = nodes[i];
  WideNode w;
  int slot = 0;
  for (int c = 0; c < 2; ++c) {
    const uint32_t x = c ? N.right : N.left;
    if (x & kLeafBit) {
      wide_entry(nodes, LX, LY, LZ, x, w, slot++);
    } else {
      wide_entry(nodes, LX, LY, LZ, nodes[x].left, w, slot++);
      wide_entry(nodes, LX, LY, LZ, nodes[x].right, w, slot++);
    }
  }
  for (; slot < 4; ++slot) {  // empty slot: a box no query meets
    for (int a = 0; a < 3; ++a) { w.lo[slot][a] = INFINITY; w.hi[slot][a] = -INFINITY; }
    w.child[slot] = kInvalid;
  }
  out[i] = w;
}

// Power-of-two change of unit (context.hpp, scale_exp): possible only when
// every scaled coordinate stays a normal (or zero) finite binary64.
__global__ void k_scale_check(const double *__restrict__ X, const double *__restrict__ Y,
                              const double *__restrict__ Z, int64_t n, int e, uint32_t *bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v[3] = {X[i], Y[i], Z[i]};
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const double t = fabs(ldexp(v[a], e));
    ok = ok && (v[a] == 0.0 || (t >= 0x1p-1022 && t < INFINITY));
  }
  if (!ok) atomicOr(bad, 1u);
}

__global__ void k_scale_values(double *v, int64_t n, int e) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = ldexp(v[i], e);
}

__global__ void k_scale_nodes(Node *nodes, int64_t m, int e) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int a = 0; a < 3; ++a) {
    nodes[i].lo[a] = ldexp(nodes[i].lo[a], e);
    nodes[i].hi[a] = ldexp(nodes[i].hi[a], e);
  }
}

// Unit exponent for a cloud of extent E: none inside [2^-16, 2^16], else the
// one that brings E into [1, 2).
int scale_exponent_for(double extent) {
  if (!(extent > 0.0) || !(extent < INFINITY)) return 0;
  if (extent >= 0x1p-16 && extent <= 0x1p16) return 0;
  return -std::ilogb(extent);
}

void apply_unit_scale(Context &c, const double bounds[6]) {
  const int64_t n = c.n;
  hipStream_t s = c.stream;
  c.scale_exp = 0;
  double extent = 0.0;
  for (int a = 0; a < 3; ++a) extent = std::max(extent, bounds[3 + a] - bounds[a]);
  const int e = scale_exponent_for(extent);
  if (e == 0) return;
  MHGP_HIP(hipMemsetAsync(c.status.p + 4, 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_scale_check, dim3(blocks_for(n, 256)), dim3(256), 0, s, c.X.p, c.Y.p, c.Z.p, n, e,
                     c.status.p + 4);
  uint32_t bad = 0;
  copy_wait(&bad, c.status.p + 4, 4, hipMemcpyDeviceToHost, s);
  if (bad) return;  // keep the input unit: exactness first
  for (double *v : {c.X.p, c.Y.p, c.Z.p, c.LX.p, c.LY.p, c.LZ.p})
    hipLaunchKernelGGL(k_scale_values, dim3(blocks_for(n, 256)), dim3(256), 0, s, v, n, e);
  hipLaunchKernelGGL(k_scale_values, dim3(1), dim3(64), 0, s, c.bounds.p, (int64_t)6, e);
  if (n > 1)
    hipLaunchKernelGGL(k_scale_nodes, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, c.nodes.p, n - 1, e);
  MHGP_LAUNCH_CHECK();
  c.scale_exp = e;
}

}  // namespace

void canonicalize_and_build(Context &c) {
  const int64_t n = c.n;
  c.scale_exp = 0;
  hipStream_t s = c.stream;
  const int T = 256;
  c.X.ensure(n); c.Y.ensure(n); c.Z.ensure(n);
  c.source_index.ensure(n);
  c.key0.ensure(n); c.key1.ensure(n); c.key2.ensure(n);
  c.perm.ensure(n);
  c.status.ensure(16);
  c.bounds.ensure(6);
  c.bound_ids.ensure(6);
  MHGP_HIP(hipMemsetAsync(c.status.p, 0, 16 * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_order_keys, dim3(blocks_for(n, T)), dim3(T), 0, s, c.src.p, n, c.key0.p,
                     c.key1.p, c.key2.p, c.perm.p);
  // stable LSD over (k0, k1, k2): least significant key first
  c.sorter.sort(c.key2.p, c.perm.p, n, 0, 64, s);
  hipLaunchKernelGGL(k_gather_key, dim3(blocks_for(n, T)), dim3(T), 0, s, c.key1.p, c.perm.p, n,
                     c.key2.p);
  c.sorter.sort(c.key2.p, c.perm.p, n, 0, 64, s);
  hipLaunchKernelGGL(k_gather_key, dim3(blocks_for(n, T)), dim3(T), 0, s, c.key0.p, c.perm.p, n,
                     c.key2.p);
  c.sorter.sort(c.key2.p, c.perm.p, n, 0, 64, s);
  hipLaunchKernelGGL(k_canonical, dim3(blocks_for(n, T)), dim3(T), 0, s, c.src.p, c.perm.p, n,
                     c.X.p, c.Y.p, c.Z.p, c.source_index.p, c.status.p);
  hipLaunchKernelGGL(k_bounds, dim3(1), dim3(512), 0, s, c.X.p, c.Y.p, c.Z.p, n, c.bounds.p,
                     c.bound_ids.p);
  c.codes.ensure(n);
  c.leaf_id.ensure(n);
  c.leaf_pos.ensure(n);
  c.LX.ensure(n); c.LY.ensure(n); c.LZ.ensure(n);
  c.path.ensure(n);
  hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, T)), dim3(T), 0, s, c.X.p, c.Y.p, c.Z.p, n,
                     c.bounds.p, c.codes.p, c.leaf_id.p, c.status.p);
  c.sorter.sort(c.codes.p, c.leaf_id.p, n, 0, 64, s);
  hipLaunchKernelGGL(k_leaves, dim3(blocks_for(n, T)), dim3(T), 0, s, c.leaf_id.p, c.X.p, c.Y.p,
                     c.Z.p, c.codes.p, n, c.LX.p, c.LY.p, c.LZ.p, c.leaf_pos.p, c.path.p);
  const int64_t ni = n > 1 ? n - 1 : 1;
  c.nodes.ensure(ni);
  c.parent_int.ensure(ni);
  c.parent_leaf.ensure(n);
  c.wit.ensure(6 * ni);
  c.counter.ensure(ni);
  if (n > 1) {
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, T)), dim3(T), 0, s, c.codes.p, c.path.p, n,
                       c.nodes.p, c.parent_int.p, c.parent_leaf.p, c.counter.p);
    hipLaunchKernelGGL(k_refit, dim3(blocks_for(n, T)), dim3(T), 0, s, n, c.nodes.p, c.wit.p,
                       c.parent_int.p, c.parent_leaf.p, c.counter.p, c.LX.p, c.LY.p, c.LZ.p,
                       c.leaf_id.p);
  } else {
    MHGP_HIP(hipMemsetAsync(c.parent_leaf.p, 0xff, sizeof(uint32_t), s));
  }
  MHGP_LAUNCH_CHECK();
  uint32_t st[4];
  MHGP_HIP(hipMemcpyAsync(st, c.status.p, sizeof(st), hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipStreamSynchronize(s));
  if (st[0] & 2u) throw std::invalid_argument("a point coordinate is not finite");
  if (st[0] & 1u) throw std::invalid_argument("the cloud contains duplicate geometric points");
  if (st[2]) throw std::runtime_error("an exact Morton bin could not be certified");
  double bounds[6];
  copy_wait(bounds, c.bounds.p, sizeof(bounds), hipMemcpyDeviceToHost, s);
  apply_unit_scale(c, bounds);
  const int64_t internal = std::max<int64_t>(n - 1, 0);
  c.wide.ensure(internal);
  if (internal)
    hipLaunchKernelGGL(k_wide_nodes, dim3(blocks_for(internal, 256)), dim3(256), 0, s, internal, c.nodes.p, c.LX.p,
                       c.LY.p, c.LZ.p, c.wide.p);
  MHGP_LAUNCH_CHECK();
  c.lbvh_ready = true;
}

void export_snapshot(Context &c, uint64_t *leaves, uint64_t *nodes_out, uint64_t *counters) {
  const int64_t n = c.n;
  hipStream_t s = c.stream;
  const int T = 256;
  DevBuf<uint64_t> post, out;
  DevBuf<uint32_t> depth;
  post.ensure(2 * n - 1);
  depth.ensure(2 * n - 1);
  out.ensure(10 * (2 * n - 1));
  hipLaunchKernelGGL(k_postorder, dim3(blocks_for(2 * n - 1, T)), dim3(T), 0, s, n, c.nodes.p,
                     c.parent_int.p, c.parent_leaf.p, post.p, depth.p);
  hipLaunchKernelGGL(k_snapshot, dim3(blocks_for(2 * n - 1, T)), dim3(T), 0, s, n, c.nodes.p,
                     c.wit.p, c.leaf_id.p, post.p, out.p);
  MHGP_LAUNCH_CHECK();
  std::vector<uint64_t> codes(n);
  std::vector<uint32_t> ids(n), dep(2 * n - 1);
  MHGP_HIP(hipMemcpyAsync(codes.data(), c.codes.p, n * 8, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(ids.data(), c.leaf_id.p, n * 4, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(dep.data(), depth.p, (2 * n - 1) * 4, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipMemcpyAsync(nodes_out, out.p, 10 * (2 * n - 1) * 8, hipMemcpyDeviceToHost, s));
  MHGP_HIP(hipStreamSynchronize(s));
  uint64_t groups = 0, maxg = 0, maxd = 0;
  for (int64_t p = 0; p < n; ++p) { leaves[2 * p] = codes[p]; leaves[2 * p + 1] = ids[p]; }
  for (int64_t b = 0; b < n;) {
    int64_t e = b + 1;
    while (e < n && codes[e] == codes[b]) e++;
    if (e - b > 1) { groups++; maxg = std::max<uint64_t>(maxg, e - b); }
    b = e;
  }
  for (auto d : dep) maxd = std::max<uint64_t>(maxd, d);
  counters[0] = n;
  counters[1] = 2 * n - 1;
  counters[2] = maxd;
  counters[3] = groups;
  counters[4] = maxg;
}

}  // namespace mhgp
