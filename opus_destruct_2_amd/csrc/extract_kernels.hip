// A structure's extraction on the device: the walk of World::Impl::extract (world.cpp:475-620) and
// append_nodes (world.cpp:739-865) for the world grid, in the reference's order.
//
// The reference walks a queue: member after member, each member's voxels in its chunk's scan order
// (x slowest), each voxel's six faces in turn; a fragment first met becomes the next member (or is
// held fixed: past the radius or the node limit, or a fresh structure's). Its order is that of a
// breadth-first search whose ties go to the earliest (member, voxel, face) that met a fragment, so
// the device takes it level by level:
//   k_walk_meet    a wavefront a member of the level: every fragment it meets that has no slot yet
//                  keeps the least key (member, voxel's place in the box's x-slowest scan, face)
//                  that met it (atomicMin) - the member whose key it keeps met it first
//   k_walk_own     a wavefront a member: how many fragments it met first may be taken
//   k_walk_scan    those counts scanned over the level: each member's first slot
//   k_walk_assign  a wavefront a member: its first-met fragments in the order it met them (a lane a
//                  voxel in scan order, faces in turn, a wavefront scan) take the next slots while
//                  there is room, the rest are held fixed - the order the queue appends them in
// The host runs levels in batches (five launches a level, a look at the level's size a batch).
// Then the SectionLog (world_internal.hpp:533-620) as records in the walk's order:
//   k_walk_emit    a wavefront a member, its voxels in scan order, a wavefront scan placing each
//                  face's record: a fine bond's key (acc_key, world_internal.hpp:447-450), the face, a
//                  support's material
//   sort           records by key (stable: a fine bond's faces in the order met); a fine bond's
//                  first record gives its place in the order fine bonds were made
// append_nodes: clusters (min-label rounds over the fine pair bonds inside a cell), the nodes
// (k_walk_nodes), the fine bonds merged into node-level bonds in the order their first fine bond was
// made (a stable sort of the node-level keys over the fine bonds in order), a pair's ends in the
// order of their identities (k_walk_heads), and every face placed (k_walk_faces). All of it integer
// and ordering work over the box (HBM-bound: a few bytes a voxel and a record a face).
#include <algorithm>
#include <cstring>

#include "svxg_kernels.hpp"

namespace svxg {

namespace {

constexpr int kFaceNone = 0, kFaceFixed = 1, kFaceFrag = 2;
constexpr uint64_t kDropped = (uint64_t(1) << 54) - 1;  // a fine pair inside one node (sorts last)

__device__ __forceinline__ uint8_t mat_of(uint8_t v) { return uint8_t((v & 0x7F) - 1); }

// SplitMix64's output function (Steele, Lea and Flood 2014), as world_internal.hpp:17-23 uses it
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t key3(int64_t x, int64_t y, int64_t z) {
  const int64_t off = int64_t(1) << 20;
  return (uint64_t(x + off) << 42) | (uint64_t(y + off) << 21) | uint64_t(z + off);
}
__device__ __forceinline__ int fdiv32(int a) { return a >= 0 ? a / 32 : -((-a + 31) / 32); }

// face d (-x, +x, -y, +y, -z, +z) of voxel v = (x, y, z) (box-local) of fragment F, as the walk
// sees it (world.cpp:525-559): nothing; held fixed (an anchored voxel, or outside the box when that
// is the unknown world); or fragment *G. *mb: the material a support takes from that side.
__device__ __forceinline__ int walk_face(const WalkDev& W, int F, int64_t v, int x, int y, int z, int d, uint8_t brk,
                                         int* G, uint8_t* mb, int* unloaded) {
  const int a = d >> 1, sg = (d & 1) ? 1 : -1;
  int q[3] = {x, y, z};
  q[a] += sg;
  const bool in = q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < W.dim[0] && q[1] < W.dim[1] && q[2] < W.dim[2];
  const int64_t w = in ? q[0] + int64_t(W.dim[0]) * (q[1] + int64_t(W.dim[1]) * q[2]) : -1;
  if (in && W.frag[w] == F) return kFaceNone;
  const uint8_t vq = in ? W.vox[w] : 0;
  const bool unl = !in && W.outside == 1;
  if (!unl) {
    if (vq == 0) return kFaceNone;
    const bool cut = sg > 0 ? ((brk >> a) & 1) != 0 : (W.broken && ((W.broken[w] >> a) & 1) != 0);
    if (cut) return kFaceNone;
  }
  *unloaded = unl ? 1 : 0;
  if ((vq & 0x80) || unl) {
    *mb = unl ? mat_of(W.vox[v]) : mat_of(vq);
    return kFaceFixed;
  }
  const int g2 = W.frag[w];
  if (g2 < 0) return kFaceNone;
  *G = g2;
  *mb = mat_of(vq);
  return kFaceFrag;
}

// a member's voxels: its fragment's box (box-local, inclusive), walked x slowest
struct MemberBox {
  int lo[3], e[3];
  int64_t total;
};
__device__ __forceinline__ MemberBox member_box(const WalkDev& W, int F) {
  MemberBox b;
  const int* bb = W.bbox + 6 * F;
  for (int c = 0; c < 3; ++c) {
    b.lo[c] = bb[c];
    b.e[c] = bb[3 + c] - bb[c] + 1;
  }
  b.total = (b.e[0] > 0 && b.e[1] > 0 && b.e[2] > 0) ? int64_t(b.e[0]) * b.e[1] * b.e[2] : 0;
  return b;
}
__device__ __forceinline__ void member_voxel(const WalkDev& W, const MemberBox& b, int64_t j, int* x, int* y, int* z,
                                             int64_t* v) {
  const int64_t yz = int64_t(b.e[1]) * b.e[2];
  const int r = int(j % yz);
  *x = b.lo[0] + int(j / yz);
  *y = b.lo[1] + r / b.e[2];
  *z = b.lo[2] + r % b.e[2];
  *v = *x + int64_t(W.dim[0]) * (*y + int64_t(W.dim[1]) * *z);
}

constexpr int kWaves = kWG / 64;  // wavefronts a workgroup: a member a wavefront

// an exclusive scan over the wavefront (every lane calls it); *total: the wavefront's sum
__device__ __forceinline__ int wave_scan(int v, int* total) {
  const int lane = threadIdx.x & 63;
  int incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  *total = __shfl(incl, 63, 64);
  return incl - v;
}

__device__ __forceinline__ uint64_t walk_key(int i, int x, int y, int z, int d, const WalkDev& W) {
  const uint64_t t = (uint64_t(x) * uint64_t(W.dim[1]) + uint64_t(y)) * uint64_t(W.dim[2]) + uint64_t(z);
  return (uint64_t(i) << 34) | (t * 6 + uint64_t(d));
}

// met first: taken while there is room if not a fresh structure's and within the radius (or another
// structure's), else held fixed (world.cpp:484-506)
__device__ __forceinline__ bool eligible(const WalkDev& W, int G) {
  const svxg_fragment& g = W.frags[G];
  const bool owned = g.owner != 0;
  const bool fresh = owned && g.owner >= W.fresh_from;
  if (fresh) return false;
  if (owned) return true;
  const svxg_fragment& s0 = W.frags[W.seed];
  const double dx = g.com[0] - s0.com[0], dy = g.com[1] - s0.com[1], dz = g.com[2] - s0.com[2];
  return sqrt(dx * dx + dy * dy + dz * dz) <= W.max_radius;
}

__global__ void __launch_bounds__(kWG) k_walk_init(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  const int f = int(blockIdx.x) * kWG + threadIdx.x;
  if (f >= W.nfrag) return;
  W.slot[f] = -1;
  W.meet[f] = ~uint64_t(0);
  for (int c = 0; c < 3; ++c) {
    W.bbox[6 * f + c] = 0x7FFFFFFF;
    W.bbox[6 * f + 3 + c] = -1;
  }
  if (f == 0) {
    WalkState s;
    memset(&s, 0, sizeof s);
    s.count = 1;
    s.level_end = 1;
    *W.st = s;
  }
}

// the inputs' rules, a thread a voxel: 1 a fragment index out of range, 2 a fragment on air or an
// anchored voxel, 4 a fragment's voxel outside the chunk of its first voxel
__global__ void __launch_bounds__(kWG) k_walk_check(const WalkDev* __restrict__ Wp, int32_t* __restrict__ bad) {
  const WalkDev& W = *Wp;
  const int64_t v = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (v >= W.nvox) return;
  const int f = W.frag[v];
  if (f == -1) return;
  if (f < -1 || f >= W.nfrag) {
    atomicOr(bad, 1);
    return;
  }
  const uint8_t c = W.vox[v];
  if (c == 0 || (c & 0x80)) atomicOr(bad, 2);
  const int nx = W.dim[0], ny = W.dim[1];
  const int g[3] = {int(v % nx) + W.lo[0], int((v / nx) % ny) + W.lo[1], int(v / (int64_t(nx) * ny)) + W.lo[2]};
  const svxg_fragment& fr = W.frags[f];
  if (fdiv32(g[0]) != fdiv32(fr.first[0]) || fdiv32(g[1]) != fdiv32(fr.first[1]) || fdiv32(g[2]) != fdiv32(fr.first[2]))
    atomicOr(bad, 4);
}

__global__ void __launch_bounds__(kWG) k_walk_bbox(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  const int64_t v = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (v >= W.nvox) return;
  const int f = W.frag[v];
  if (f < 0) return;
  const int nx = W.dim[0], ny = W.dim[1];
  const int x = int(v % nx), y = int((v / nx) % ny), z = int(v / (int64_t(nx) * ny));
  int* bb = W.bbox + 6 * f;
  atomicMin(bb + 0, x);
  atomicMin(bb + 1, y);
  atomicMin(bb + 2, z);
  atomicMax(bb + 3, x);
  atomicMax(bb + 4, y);
  atomicMax(bb + 5, z);
}

__global__ void k_walk_seed(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  if (threadIdx.x == 0) {
    W.slot[W.seed] = 0;
    W.members[0] = W.seed;
  }
}

// a wavefront a member of the level (the grid strides over the level): every fragment it meets
// that has no slot yet keeps the least key that met it
__global__ void __launch_bounds__(kWG) k_walk_meet(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  const int lane = threadIdx.x & 63;
  const int w = int(blockIdx.x) * kWaves + int(threadIdx.x >> 6), nw = int(gridDim.x) * kWaves;
  const int end = W.st->level_end;
  for (int i = W.st->level_begin + w; i < end; i += nw) {
    const int F = W.members[i];
    const MemberBox b = member_box(W, F);
    for (int64_t j = lane; j < b.total; j += 64) {
      int x, y, z;
      int64_t v;
      member_voxel(W, b, j, &x, &y, &z, &v);
      if (W.frag[v] != F) continue;
      const uint8_t brk = W.broken ? W.broken[v] : 0;
      for (int d = 0; d < 6; ++d) {
        int G = -1, unl = 0;
        uint8_t mb = 0;
        if (walk_face(W, F, v, x, y, z, d, brk, &G, &mb, &unl) != kFaceFrag || W.slot[G] != -1) continue;
        atomicMin(reinterpret_cast<unsigned long long*>(W.meet + G), (unsigned long long)walk_key(i, x, y, z, d, W));
      }
    }
  }
}

// a wavefront a member: the fragments it met first (its key is their least) that may be taken
__global__ void __launch_bounds__(kWG) k_walk_own(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  const int lane = threadIdx.x & 63;
  const int w = int(blockIdx.x) * kWaves + int(threadIdx.x >> 6), nw = int(gridDim.x) * kWaves;
  const int end = W.st->level_end;
  for (int i = W.st->level_begin + w; i < end; i += nw) {
    const int F = W.members[i];
    const MemberBox b = member_box(W, F);
    int ne = 0;
    for (int64_t j = lane; j < b.total; j += 64) {
      int x, y, z;
      int64_t v;
      member_voxel(W, b, j, &x, &y, &z, &v);
      if (W.frag[v] != F) continue;
      const uint8_t brk = W.broken ? W.broken[v] : 0;
      for (int d = 0; d < 6; ++d) {
        int G = -1, unl = 0;
        uint8_t mb = 0;
        if (walk_face(W, F, v, x, y, z, d, brk, &G, &mb, &unl) != kFaceFrag) continue;
        if (W.meet[G] == walk_key(i, x, y, z, d, W) && eligible(W, G)) ++ne;
      }
    }
    int total;
    wave_scan(ne, &total);
    if (lane == 0) W.ecnt[i] = total;
  }
}

// one workgroup: the level's members' counts scanned (their first slots among the level's), and the
// level's total
__global__ void __launch_bounds__(1024) k_walk_scan(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  __shared__ int lds[1024];
  __shared__ int carry;
  const int t = threadIdx.x;
  const int begin = W.st->level_begin, end = W.st->level_end;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = begin; base < end; base += 1024) {
    const int i = base + t;
    const int v = i < end ? W.ecnt[i] : 0;
    lds[t] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int add = t >= o ? lds[t - o] : 0;
      __syncthreads();
      lds[t] += add;
      __syncthreads();
    }
    if (i < end) W.ebase[i] = carry + lds[t] - v;
    __syncthreads();
    if (t == 0) carry += lds[1023];
    __syncthreads();
  }
  if (t == 0) W.st->eligible = carry;
}

// a wavefront a member: its first-met fragments in the order it met them (voxels in scan order,
// a lane a voxel, faces in turn): slots while there is room, else held fixed
__global__ void __launch_bounds__(kWG) k_walk_assign(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  const int lane = threadIdx.x & 63;
  const int w = int(blockIdx.x) * kWaves + int(threadIdx.x >> 6), nw = int(gridDim.x) * kWaves;
  const int end = W.st->level_end;
  const int count = W.st->count;
  for (int i = W.st->level_begin + w; i < end; i += nw) {
    const int F = W.members[i];
    const MemberBox b = member_box(W, F);
    int next = count + W.ebase[i];
    for (int64_t j0 = 0; j0 < b.total; j0 += 64) {
      const int64_t j = j0 + lane;
      int G6[6];
      uint8_t e6[6];
      int n = 0, ne = 0;
      if (j < b.total) {
        int x, y, z;
        int64_t v;
        member_voxel(W, b, j, &x, &y, &z, &v);
        if (W.frag[v] == F) {
          const uint8_t brk = W.broken ? W.broken[v] : 0;
          for (int d = 0; d < 6; ++d) {
            int G = -1, unl = 0;
            uint8_t mb = 0;
            if (walk_face(W, F, v, x, y, z, d, brk, &G, &mb, &unl) != kFaceFrag) continue;
            if (W.meet[G] != walk_key(i, x, y, z, d, W)) continue;
            G6[n] = G;
            e6[n] = eligible(W, G) ? 1 : 0;
            ne += e6[n];
            ++n;
          }
        }
      }
      int total;
      int s = next + wave_scan(ne, &total);
      for (int k = 0; k < n; ++k) {
        const int G = G6[k];
        if (e6[k] && s < W.max_nodes) {
          W.slot[G] = s;
          W.members[s] = G;
        } else {
          W.slot[G] = -2;
          W.st->truncated = 1;
        }
        if (e6[k]) ++s;
      }
      next += total;
    }
  }
}

// the next level: the members just taken
__global__ void k_walk_next(const WalkDev* __restrict__ Wp) {
  const WalkDev& W = *Wp;
  if (threadIdx.x != 0) return;
  WalkState& s = *W.st;
  const int taken = min(s.eligible, max(0, W.max_nodes - s.count));
  s.level_begin = s.count;
  s.count += taken;
  s.level_end = s.count;
  s.eligible = 0;
}

// a member's records: how many (pass 0) or written in order (pass 1)
struct Rec {
  uint64_t key;
  int p[3];
  int8_t axis, sign;
  uint8_t mb;
};

__device__ __forceinline__ int voxel_records(const WalkDev& W, int i, int F, int x, int y, int z, int64_t v, Rec* out,
                                             int* flags) {
  const uint8_t brk = W.broken ? W.broken[v] : 0;
  int c = 0;
  for (int d = 0; d < 6; ++d) {
    int G = -1, unl = 0;
    uint8_t mb = 0;
    const int kind = walk_face(W, F, v, x, y, z, d, brk, &G, &mb, &unl);
    if (kind == kFaceNone) continue;
    const int a = d >> 1, sg = (d & 1) ? 1 : -1;
    int s = -2;
    if (kind == kFaceFrag) {
      s = W.slot[G];
      if (s == -1) {
        *flags |= 4;  // (met by a member, so taken or held: cannot happen)
        continue;
      }
      if (s >= 0 && sg < 0) continue;  // (a pair's face is logged from its lower voxel)
    }
    if (unl) *flags |= 2;
    Rec& r = out[c++];
    r.p[0] = x + W.lo[0];
    r.p[1] = y + W.lo[1];
    r.p[2] = z + W.lo[2];
    if (sg < 0) r.p[a] -= 1;
    r.axis = int8_t(a);
    r.mb = mb;
    if (s >= 0) {
      r.key = (uint64_t(min(i, s)) << 32) | (uint64_t(max(i, s)) << 1) | 1u;
      r.sign = int8_t(i < s ? 1 : -1);
    } else {
      *flags |= 1;
      r.key = (uint64_t(i) << 32) | (uint64_t(a * 2 + (sg > 0 ? 1 : 0)) << 1);
      r.sign = 1;
    }
  }
  return c;
}

__device__ __forceinline__ void raise_flags(const WalkDev& W, int flags) {
  if (flags & 1) W.st->any_support = 1;
  if (flags & 2) W.st->touched_unloaded = 1;
  if (flags & 4) W.st->bad = 1;
}

// a wavefront a member (the grid strides over the members): its record count
__global__ void __launch_bounds__(kWG) k_walk_count(const WalkDev* __restrict__ Wp, int nm, int32_t* __restrict__ count) {
  const WalkDev& W = *Wp;
  const int lane = threadIdx.x & 63;
  const int w = int(blockIdx.x) * kWaves + int(threadIdx.x >> 6), nw = int(gridDim.x) * kWaves;
  for (int i = w; i < nm; i += nw) {
    const int F = W.members[i];
    const MemberBox b = member_box(W, F);
    int c = 0, flags = 0;
    Rec r[6];
    for (int64_t j = lane; j < b.total; j += 64) {
      int x, y, z;
      int64_t v;
      member_voxel(W, b, j, &x, &y, &z, &v);
      if (W.frag[v] != F) continue;
      c += voxel_records(W, i, F, x, y, z, v, r, &flags);
    }
    raise_flags(W, flags);
    int total;
    wave_scan(c, &total);
    if (lane == 0) count[i] = total;
  }
}

// a wavefront a member: its records written in order from off[i] (a lane a voxel, the voxels in
// scan order)
__global__ void __launch_bounds__(kWG) k_walk_emit(const WalkDev* __restrict__ Wp, int nm,
                                                   const int32_t* __restrict__ off, uint64_t* __restrict__ rkey,
                                                   svxg_section_face* __restrict__ rface, uint8_t* __restrict__ rmb) {
  const WalkDev& W = *Wp;
  const int lane = threadIdx.x & 63;
  const int w = int(blockIdx.x) * kWaves + int(threadIdx.x >> 6), nw = int(gridDim.x) * kWaves;
  for (int i = w; i < nm; i += nw) {
    const int F = W.members[i];
    const MemberBox b = member_box(W, F);
    int64_t at = off[i];
    for (int64_t j0 = 0; j0 < b.total; j0 += 64) {
      const int64_t j = j0 + lane;
      Rec r[6];
      int c = 0, flags = 0;
      if (j < b.total) {
        int x, y, z;
        int64_t v;
        member_voxel(W, b, j, &x, &y, &z, &v);
        if (W.frag[v] == F) c = voxel_records(W, i, F, x, y, z, v, r, &flags);
      }
      int total;
      const int excl = wave_scan(c, &total);
      for (int k = 0; k < c; ++k) {
        const int64_t o = at + excl + k;
        rkey[o] = r[k].key;
        svxg_section_face f;
        memset(&f, 0, sizeof f);
        f.p[0] = r[k].p[0];
        f.p[1] = r[k].p[1];
        f.p[2] = r[k].p[2];
        f.axis = r[k].axis;
        f.sign = r[k].sign;
        rface[o] = f;
        rmb[o] = r[k].mb;
      }
      at += total;
    }
  }
}

// first[k]: a new key starts at k (a dropped key starts nothing)
__global__ void __launch_bounds__(kWG) k_seg_first(const uint64_t* __restrict__ key, int n, int32_t* __restrict__ first) {
  const int k = int(blockIdx.x) * kWG + threadIdx.x;
  if (k >= n) return;
  first[k] = key[k] != kDropped && (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// fine bonds (key order j): their first record in the sorted list, their key, their first record's
// place in the walk (the order they were made in); fstart[nfine] = n
__global__ void __launch_bounds__(kWG) k_fine(const uint64_t* __restrict__ skey, const int32_t* __restrict__ sidx,
                                              const int32_t* __restrict__ first, const int32_t* __restrict__ seg, int n,
                                              int32_t* __restrict__ fstart, int32_t* __restrict__ fpos,
                                              uint64_t* __restrict__ fkey, int32_t* __restrict__ flag) {
  const int k = int(blockIdx.x) * kWG + threadIdx.x;
  if (k >= n) return;
  if (first[k]) {
    const int j = seg[k];
    fstart[j] = k;
    fpos[j] = sidx[k];
    fkey[j] = skey[k];
    flag[sidx[k]] = 1;
  }
  if (k == n - 1) fstart[seg[k] + first[k]] = n;
}

__global__ void __launch_bounds__(kWG) k_fine_order(const int32_t* __restrict__ fpos, const int32_t* __restrict__ rank,
                                                    int nfine, int32_t* __restrict__ fo, int32_t* __restrict__ byfo) {
  const int j = int(blockIdx.x) * kWG + threadIdx.x;
  if (j >= nfine) return;
  const int o = rank[fpos[j]];
  fo[j] = o;
  byfo[o] = j;
}

// a member's cell (append_nodes, world.cpp:744-755): its chunk and the low six bits of the cell
// code the reference's key keeps; labels start as the members themselves
__global__ void __launch_bounds__(kWG) k_walk_cells(const WalkDev* __restrict__ Wp, int cell, int cubes,
                                                    uint64_t* __restrict__ cellkey, int32_t* __restrict__ label, int nm) {
  const WalkDev& W = *Wp;
  const int i = int(blockIdx.x) * kWG + threadIdx.x;
  if (i >= nm) return;
  const svxg_fragment& g = W.frags[W.members[i]];
  int cc[3], l[3];
  for (int c = 0; c < 3; ++c) {
    cc[c] = fdiv32(g.first[c]);
    l[c] = g.first[c] - 32 * cc[c];
  }
  const uint64_t code = cubes ? uint64_t(((l[0] / cell) << 4) | ((l[1] / cell) << 2) | (l[2] / cell))
                              : uint64_t((l[0] / cell) * 64 + (l[1] / cell) * 8 + l[2] / cell);
  const uint64_t off = uint64_t(1) << 15;
  cellkey[i] = ((uint64_t(cc[0]) + off) << 38) | ((uint64_t(cc[1]) + off) << 22) | ((uint64_t(cc[2]) + off) << 6) |
               (code & 63u);
  label[i] = i;
}

// a round over the fine pair bonds (links inside a cell take both ends, and their labels' labels, to
// the lesser label), then each label follows its label once
__global__ void __launch_bounds__(kWG) k_walk_link(const uint64_t* __restrict__ fkey, int nfine,
                                                   const uint64_t* __restrict__ cellkey, int32_t* __restrict__ label,
                                                   int32_t* __restrict__ changed) {
  const int j = int(blockIdx.x) * kWG + threadIdx.x;
  if (j >= nfine) return;
  const uint64_t key = fkey[j];
  if (!(key & 1u)) return;
  const int a = int(key >> 32), b = int((key >> 1) & 0x7FFFFFFFu);
  if (cellkey[a] != cellkey[b]) return;
  const int la = label[a], lb = label[b];
  if (la == lb) return;
  const int m = min(la, lb);
  atomicMin(label + a, m);
  atomicMin(label + b, m);
  atomicMin(label + la, m);
  atomicMin(label + lb, m);
  *changed = 1;
}
__global__ void __launch_bounds__(kWG) k_walk_jump(int32_t* __restrict__ label, int n) {
  const int i = int(blockIdx.x) * kWG + threadIdx.x;
  if (i >= n) return;
  const int l = label[i];
  const int ll = label[l];
  if (ll < l) atomicMin(label + i, ll);
}

__global__ void __launch_bounds__(kWG) k_walk_iota(int32_t* __restrict__ a, int32_t* __restrict__ b, int n) {
  const int i = int(blockIdx.x) * kWG + threadIdx.x;
  if (i > n) return;
  if (i < n) a[i] = i;
  b[i] = i;
}

// a thread a node, over its members in walk order (world.cpp:771-809): masses summed, centres
// weighted by them then times 1 / mass, the heaviest's material (the first of equal ones), the
// least strength, the first member's identity (frag_ident, world_internal.hpp:24)
__global__ void __launch_bounds__(kWG) k_walk_nodes(const WalkDev* __restrict__ Wp, const int32_t* __restrict__ mlist,
                                                    const int32_t* __restrict__ start, int nn,
                                                    svxg_node* __restrict__ nodes, uint8_t* __restrict__ nmat,
                                                    double* __restrict__ nstr, uint64_t* __restrict__ ident) {
  const WalkDev& W = *Wp;
  const int c = int(blockIdx.x) * kWG + threadIdx.x;
  if (c >= nn) return;
  double mass = 0.0, heavy = -1.0, strength = 1e30, cx = 0.0, cy = 0.0, cz = 0.0;
  uint8_t mat = 0;
  for (int k = start[c]; k < start[c + 1]; ++k) {
    const svxg_fragment& g = W.frags[W.members[mlist[k]]];
    mass += g.mass;
    cx += g.com[0] * g.mass;
    cy += g.com[1] * g.mass;
    cz += g.com[2] * g.mass;
    if (g.mass > heavy) {
      heavy = g.mass;
      mat = g.mat;
    }
    strength = g.strength < strength ? g.strength : strength;
  }
  svxg_node nd;
  memset(&nd, 0, sizeof nd);
  if (mass > 0) {
    const double inv = 1.0 / mass;
    nd.c[0] = cx * inv;
    nd.c[1] = cy * inv;
    nd.c[2] = cz * inv;
  } else {
    nd.c[0] = cx;
    nd.c[1] = cy;
    nd.c[2] = cz;
  }
  nd.mass = mass;
  nodes[c] = nd;
  nmat[c] = mat;
  nstr[c] = strength;
  const svxg_fragment& f0 = W.frags[W.members[mlist[start[c]]]];
  int cc[3], l[3];
  for (int q = 0; q < 3; ++q) {
    cc[q] = fdiv32(f0.first[q]);
    l[q] = f0.first[q] - 32 * cc[q];
  }
  const int first = l[0] * 1024 + l[1] * 32 + l[2];
  ident[c] = mix64(key3(cc[0], cc[1], cc[2]) * 0x9E3779B97F4A7C15ull ^ uint64_t(int64_t(first)));
}

// node-level keys of the fine bonds in the order they were made (SectionLog::merge)
__global__ void __launch_bounds__(kWG) k_merge_keys(const uint64_t* __restrict__ fkey, const int32_t* __restrict__ byfo,
                                                    const int32_t* __restrict__ fnode, int nfine,
                                                    uint64_t* __restrict__ mkey, int32_t* __restrict__ mval) {
  const int o = int(blockIdx.x) * kWG + threadIdx.x;
  if (o >= nfine) return;
  const uint64_t key = fkey[byfo[o]];
  const int na = fnode[int(key >> 32)];
  const uint64_t low = (key >> 1) & 0x7FFFFFFFu;
  uint64_t mk;
  if (key & 1u) {
    const int nb = fnode[int(low)];
    mk = na == nb ? kDropped : (uint64_t(min(na, nb)) << 32) | (uint64_t(max(na, nb)) << 1) | 1u;
  } else {
    mk = (uint64_t(na) << 32) | (low << 1);
  }
  mkey[o] = mk;
  mval[o] = o;
}

// per sorted fine bond: its face count (0 dropped); its place in the sorted list by its order
__global__ void __launch_bounds__(kWG) k_merge_counts(const uint64_t* __restrict__ mkey, const int32_t* __restrict__ mval,
                                                      const int32_t* __restrict__ byfo, const int32_t* __restrict__ fstart,
                                                      int n, int32_t* __restrict__ cnt, int32_t* __restrict__ kinv) {
  const int k = int(blockIdx.x) * kWG + threadIdx.x;
  if (k >= n) return;
  const int j = byfo[mval[k]];
  cnt[k] = mkey[k] == kDropped ? 0 : fstart[j + 1] - fstart[j];
  kinv[mval[k]] = k;
}

// merged bonds (key order m): their first entry in the sorted list and first fine bond (the earliest
// made, the list being stable over the order); mstart[nmerged] = the entries not dropped
__global__ void __launch_bounds__(kWG) k_merged(const uint64_t* __restrict__ mkey, const int32_t* __restrict__ mval,
                                                const int32_t* __restrict__ first, const int32_t* __restrict__ seg, int n,
                                                int32_t* __restrict__ mstart, int32_t* __restrict__ mpos,
                                                int32_t* __restrict__ flag) {
  const int k = int(blockIdx.x) * kWG + threadIdx.x;
  if (k >= n || mkey[k] == kDropped) return;
  if (first[k]) {
    const int m = seg[k];
    mstart[m] = k;
    mpos[m] = mval[k];
    flag[mval[k]] = 1;
  }
  if (k == n - 1 || mkey[k + 1] == kDropped) mstart[seg[k] + first[k]] = k + 1;
}

__global__ void __launch_bounds__(kWG) k_merged_order(const int32_t* __restrict__ mpos, const int32_t* __restrict__ rank,
                                                      const int32_t* __restrict__ mstart, const int32_t* __restrict__ S,
                                                      int nmerged, int32_t* __restrict__ final_of,
                                                      int32_t* __restrict__ fcount) {
  const int m = int(blockIdx.x) * kWG + threadIdx.x;
  if (m >= nmerged) return;
  const int f = rank[mpos[m]];
  final_of[m] = f;
  fcount[f] = S[mstart[m + 1]] - S[mstart[m]];
}

// the bonds' heads in their order (append_nodes, world.cpp:824-862): a pair's ends in the order of
// their identities, ma / mb and strengths from the nodes (a support's mb the material its first fine
// bond met last), the bond identity (world_internal.hpp:25-27)
__global__ void __launch_bounds__(kWG) k_walk_heads(const uint64_t* __restrict__ mkey, const int32_t* __restrict__ mstart,
                                                    const int32_t* __restrict__ mpos, const int32_t* __restrict__ final_of,
                                                    const int32_t* __restrict__ face0, const int32_t* __restrict__ byfo,
                                                    const uint64_t* __restrict__ fkey, const int32_t* __restrict__ fstart,
                                                    const int32_t* __restrict__ fpos, const int32_t* __restrict__ sidx,
                                                    const svxg_section_face* __restrict__ rface,
                                                    const uint8_t* __restrict__ rmb, const uint8_t* __restrict__ nmat,
                                                    const double* __restrict__ nstr, const uint64_t* __restrict__ ident,
                                                    int nmerged, svxg_section_head* __restrict__ heads,
                                                    uint64_t* __restrict__ bid, uint8_t* __restrict__ swapped) {
  const int m = int(blockIdx.x) * kWG + threadIdx.x;
  if (m >= nmerged) return;
  const uint64_t mk = mkey[mstart[m]];
  const int j0 = byfo[mpos[m]];
  const int f = final_of[m];
  svxg_section_head h;
  memset(&h, 0, sizeof h);
  int a = int(mk >> 32);
  const uint64_t low = (mk >> 1) & 0x7FFFFFFFu;
  uint8_t sw = 0;
  if (mk & 1u) {
    int b = int(low);
    if (ident[a] > ident[b]) {
      const int t = a;
      a = b;
      b = t;
      sw = 1;
    }
    h.a = a;
    h.b = b;
    h.axis = rface[fpos[j0]].axis;
    h.sign = 1;
    h.mb = nmat[b];
    h.strength_b = nstr[b];
  } else {
    h.a = a;
    h.b = -1;
    h.axis = int(low >> 1);
    h.sign = (low & 1u) ? 1 : -1;
    h.mb = rmb[sidx[fstart[j0 + 1] - 1]];
    h.strength_b = 1e9;
  }
  (void)fkey;
  h.ma = nmat[a];
  h.strength_a = nstr[a];
  h.face0 = face0[f];
  h.nfaces = face0[f + 1] - face0[f];
  heads[f] = h;
  swapped[m] = sw;
  const uint64_t ia = ident[h.a], ib = h.b >= 0 ? ident[h.b] : 0;
  const int ax = h.b >= 0 ? 0 : h.axis, sn = h.b >= 0 ? 1 : h.sign;
  bid[f] = mix64(ia * 0xD6E8FEB86659FD93ull ^ mix64(ib + uint64_t(ax * 2 + (sn > 0 ? 1 : 0))));
}

// every face to its place: its bond's first face, then its fine bond's place among the bond's, then
// its own among the fine bond's; a pair face's sign relative to the bond's first node
__global__ void __launch_bounds__(kWG) k_walk_faces(
    const int32_t* __restrict__ sidx, const int32_t* __restrict__ seg, const int32_t* __restrict__ ffirst,
    const int32_t* __restrict__ fstart, const int32_t* __restrict__ fo, const uint64_t* __restrict__ fkey,
    const int32_t* __restrict__ kinv, const uint64_t* __restrict__ mkey, const int32_t* __restrict__ mseg,
    const int32_t* __restrict__ mfirst, const int32_t* __restrict__ mstart,
    const int32_t* __restrict__ final_of, const int32_t* __restrict__ face0, const int32_t* __restrict__ S,
    const uint8_t* __restrict__ swapped, const int32_t* __restrict__ fnode, const svxg_section_face* __restrict__ rface,
    int n, svxg_section_face* __restrict__ out) {
  const int k = int(blockIdx.x) * kWG + threadIdx.x;
  if (k >= n) return;
  // (an entry's run: the starts up to and including it, less one - the scans are exclusive)
  const int j = seg[k] + ffirst[k] - 1;
  const int km = kinv[fo[j]];
  const uint64_t mk = mkey[km];
  if (mk == kDropped) return;
  const int m = mseg[km] + mfirst[km] - 1;
  const int64_t o = int64_t(face0[final_of[m]]) + (S[km] - S[mstart[m]]) + (k - fstart[j]);
  svxg_section_face f = rface[sidx[k]];
  const uint64_t key = fkey[j];
  if (key & 1u) {
    const bool flip = fnode[int(key >> 32)] != int(mk >> 32);
    int s = f.sign;
    if (flip) s = -s;
    if (swapped[m]) s = -s;
    f.sign = int8_t(s);
  }
  out[o] = f;
}

inline unsigned blocks(int64_t n) { return unsigned((n + kWG - 1) / kWG); }

}  // namespace

constexpr unsigned kLevelGrid = 512;  // workgroups of a level's launches (2048 wavefronts)

static unsigned member_grid(int nm) { return unsigned(std::min<int64_t>(4096, (int64_t(nm) + kWaves - 1) / kWaves)); }

void launch_walk_check(hipStream_t s, const WalkDev* W, int64_t nvox, int32_t* bad) {
  if (nvox > 0) hipLaunchKernelGGL(k_walk_check, dim3(blocks(nvox)), dim3(kWG), 0, s, W, bad);
}
void launch_walk_init(hipStream_t s, const WalkDev* W, int nfrag, int64_t nvox) {
  hipLaunchKernelGGL(k_walk_init, dim3(blocks(nfrag)), dim3(kWG), 0, s, W);
  if (nvox > 0) hipLaunchKernelGGL(k_walk_bbox, dim3(blocks(nvox)), dim3(kWG), 0, s, W);
  hipLaunchKernelGGL(k_walk_seed, dim3(1), dim3(64), 0, s, W);
}
void launch_walk_level(hipStream_t s, const WalkDev* W) {
  hipLaunchKernelGGL(k_walk_meet, dim3(kLevelGrid), dim3(kWG), 0, s, W);
  hipLaunchKernelGGL(k_walk_own, dim3(kLevelGrid), dim3(kWG), 0, s, W);
  hipLaunchKernelGGL(k_walk_scan, dim3(1), dim3(1024), 0, s, W);
  hipLaunchKernelGGL(k_walk_assign, dim3(kLevelGrid), dim3(kWG), 0, s, W);
  hipLaunchKernelGGL(k_walk_next, dim3(1), dim3(64), 0, s, W);
}
void launch_walk_count(hipStream_t s, const WalkDev* W, int nm, int32_t* count) {
  if (nm > 0) hipLaunchKernelGGL(k_walk_count, dim3(member_grid(nm)), dim3(kWG), 0, s, W, nm, count);
}
void launch_walk_emit(hipStream_t s, const WalkDev* W, int nm, const int32_t* off, uint64_t* rkey,
                      svxg_section_face* rface, uint8_t* rmb) {
  if (nm > 0)
    hipLaunchKernelGGL(k_walk_emit, dim3(member_grid(nm)), dim3(kWG), 0, s, W, nm, off, rkey, rface, rmb);
}
void launch_seg_first(hipStream_t s, const uint64_t* key, int n, int32_t* first) {
  if (n > 0) hipLaunchKernelGGL(k_seg_first, dim3(blocks(n)), dim3(kWG), 0, s, key, n, first);
}
void launch_fine(hipStream_t s, const uint64_t* skey, const int32_t* sidx, const int32_t* first, const int32_t* seg,
                 int n, int32_t* fstart, int32_t* fpos, uint64_t* fkey, int32_t* flag) {
  if (n > 0)
    hipLaunchKernelGGL(k_fine, dim3(blocks(n)), dim3(kWG), 0, s, skey, sidx, first, seg, n, fstart, fpos, fkey, flag);
}
void launch_fine_order(hipStream_t s, const int32_t* fpos, const int32_t* rank, int nfine, int32_t* fo,
                       int32_t* byfo) {
  if (nfine > 0) hipLaunchKernelGGL(k_fine_order, dim3(blocks(nfine)), dim3(kWG), 0, s, fpos, rank, nfine, fo, byfo);
}
void launch_walk_cells(hipStream_t s, const WalkDev* W, int cell, int cubes, uint64_t* cellkey, int32_t* label,
                       int nm) {
  hipLaunchKernelGGL(k_walk_cells, dim3(blocks(nm)), dim3(kWG), 0, s, W, cell, cubes, cellkey, label, nm);
}
void launch_walk_link_round(hipStream_t s, const uint64_t* fkey, int nfine, const uint64_t* cellkey, int32_t* label,
                            int32_t* changed, int nm) {
  if (nfine > 0)
    hipLaunchKernelGGL(k_walk_link, dim3(blocks(nfine)), dim3(kWG), 0, s, fkey, nfine, cellkey, label, changed);
  hipLaunchKernelGGL(k_walk_jump, dim3(blocks(nm)), dim3(kWG), 0, s, label, nm);
}
void launch_walk_iota(hipStream_t s, int32_t* a, int32_t* b, int n) {
  hipLaunchKernelGGL(k_walk_iota, dim3(blocks(int64_t(n) + 1)), dim3(kWG), 0, s, a, b, n);
}
void launch_walk_nodes(hipStream_t s, const WalkDev* W, const int32_t* mlist, const int32_t* start, int nn,
                       svxg_node* nodes, uint8_t* nmat, double* nstr, uint64_t* ident) {
  if (nn > 0)
    hipLaunchKernelGGL(k_walk_nodes, dim3(blocks(nn)), dim3(kWG), 0, s, W, mlist, start, nn, nodes, nmat, nstr, ident);
}
void launch_merge_keys(hipStream_t s, const uint64_t* fkey, const int32_t* byfo, const int32_t* fnode, int nfine,
                       uint64_t* mkey, int32_t* mval) {
  if (nfine > 0)
    hipLaunchKernelGGL(k_merge_keys, dim3(blocks(nfine)), dim3(kWG), 0, s, fkey, byfo, fnode, nfine, mkey, mval);
}
void launch_merge_counts(hipStream_t s, const uint64_t* mkey, const int32_t* mval, const int32_t* byfo,
                         const int32_t* fstart, int n, int32_t* cnt, int32_t* kinv) {
  if (n > 0)
    hipLaunchKernelGGL(k_merge_counts, dim3(blocks(n)), dim3(kWG), 0, s, mkey, mval, byfo, fstart, n, cnt, kinv);
}
void launch_merged(hipStream_t s, const uint64_t* mkey, const int32_t* mval, const int32_t* first, const int32_t* seg,
                   int n, int32_t* mstart, int32_t* mpos, int32_t* flag) {
  if (n > 0)
    hipLaunchKernelGGL(k_merged, dim3(blocks(n)), dim3(kWG), 0, s, mkey, mval, first, seg, n, mstart, mpos, flag);
}
void launch_merged_order(hipStream_t s, const int32_t* mpos, const int32_t* rank, const int32_t* mstart,
                         const int32_t* S, int nmerged, int32_t* final_of, int32_t* fcount) {
  if (nmerged > 0)
    hipLaunchKernelGGL(k_merged_order, dim3(blocks(nmerged)), dim3(kWG), 0, s, mpos, rank, mstart, S, nmerged, final_of,
                       fcount);
}
void launch_walk_heads(hipStream_t s, const WalkMerge& M, int nmerged, svxg_section_head* heads, uint64_t* bid) {
  if (nmerged > 0)
    hipLaunchKernelGGL(k_walk_heads, dim3(blocks(nmerged)), dim3(kWG), 0, s, M.mkey, M.mstart, M.mpos, M.final_of,
                       M.face0, M.byfo, M.fkey, M.fstart, M.fpos, M.sidx, M.rface, M.rmb, M.nmat, M.nstr, M.ident,
                       nmerged, heads, bid, M.swapped);
}
void launch_walk_faces(hipStream_t s, const WalkMerge& M, int nrec, svxg_section_face* out) {
  if (nrec > 0)
    hipLaunchKernelGGL(k_walk_faces, dim3(blocks(nrec)), dim3(kWG), 0, s, M.sidx, M.seg, M.ffirst, M.fstart, M.fo,
                       M.fkey, M.kinv, M.mkey, M.mseg, M.mfirst, M.mstart, M.final_of, M.face0, M.S, M.swapped, M.fnode,
                       M.rface, nrec, out);
}

}  // namespace svxg
