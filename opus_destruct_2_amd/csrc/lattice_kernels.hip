// A structure's graph from one voxel lattice on the device (the walk of World::Impl::extract and
// append_nodes, core/src/world/world.cpp:425-865, for one unrotated lattice, a node a fragment):
//   k_face_flags   a slot a (voxel, axis): is the face between the voxel and its +axis neighbour a
//                  bond's face (two fragments, or a fragment and an anchored voxel)?
//   scan           the faces' places, in (voxel, axis) order
//   k_face_write   each face's record and its bond key (node a, pair / support, node b or axis and side)
//   sort           a stable radix sort of the keys: the faces of a bond contiguous, in scan order
//   k_heads        a section head at each change of key
//   k_node_sums    a thread a fragment: its voxels' mass, centre of mass and commonest material, summed
//                  in scan order (the fragment's voxels sorted stably by fragment first)
// Integer and ordering work over the lattice (HBM-bound: a few bytes a voxel), then the sections
// (section_kernels.hip). The order of nodes, bonds and faces is opus_destruct_2_amd/voxels.py's
// extract(), bit for bit.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "svxg_kernels.hpp"

namespace svxg {

namespace {

constexpr uint8_t kAnchor = 0x80;

__device__ __forceinline__ uint64_t face_key(int a, int kind, int key2) {
  return (uint64_t(uint32_t(a)) << 22) | (uint64_t(kind) << 21) | uint64_t(uint32_t(key2));
}

struct Nbr {
  int a, b;  // the two voxels' fragments
  uint8_t va, vb;
  int x, y, z;
};

__device__ __forceinline__ bool neighbour(const LatticeDev& L, int64_t slot, Nbr* o) {
  const int64_t v = slot / 3;
  const int axis = int(slot % 3);
  const int nx = L.dim[0], ny = L.dim[1], nz = L.dim[2];
  const int x = int(v % nx), y = int((v / nx) % ny), z = int(v / (int64_t(nx) * ny));
  int q[3] = {x, y, z};
  q[axis] += 1;
  if (q[0] >= nx || q[1] >= ny || q[2] >= nz) return false;
  const int64_t w = q[0] + int64_t(nx) * (q[1] + int64_t(ny) * q[2]);
  o->a = L.frag[v];
  o->b = L.frag[w];
  if (L.fnode) {
    if (o->a >= 0) o->a = L.fnode[o->a];
    if (o->b >= 0) o->b = L.fnode[o->b];
  }
  o->va = L.vox[v];
  o->vb = L.vox[w];
  o->x = x;
  o->y = y;
  o->z = z;
  return true;
}

// 1: two fragments; 2: a's fragment on an anchored voxel above it; 3: b's fragment on one below it
__device__ __forceinline__ int face_kind(const Nbr& n) {
  if (n.a >= 0 && n.b >= 0 && n.a != n.b) return 1;
  if (n.a >= 0 && n.b < 0 && (n.vb & kAnchor)) return 2;
  if (n.b >= 0 && n.a < 0 && (n.va & kAnchor)) return 3;
  return 0;
}

__global__ void __launch_bounds__(kWG) k_face_flags(const LatticeDev* __restrict__ Lp, int32_t* __restrict__ flag) {
  const LatticeDev& L = *Lp;
  const int64_t slot = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (slot >= 3 * L.nvox) return;
  Nbr n;
  flag[slot] = (neighbour(L, slot, &n) && face_kind(n) != 0) ? 1 : 0;
}

__global__ void __launch_bounds__(kWG) k_face_write(const LatticeDev* __restrict__ Lp, const int32_t* __restrict__ flag,
                                                    const int32_t* __restrict__ at, uint64_t* __restrict__ key,
                                                    int32_t* __restrict__ idx, svxg_section_face* __restrict__ face,
                                                    uint8_t* __restrict__ mb) {
  const LatticeDev& L = *Lp;
  const int64_t slot = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (slot >= 3 * L.nvox || !flag[slot]) return;
  Nbr n;
  neighbour(L, slot, &n);
  const int kind = face_kind(n);
  const int axis = int(slot % 3);
  const int k = at[slot];
  svxg_section_face f;
  f.p[0] = n.x + L.lo[0];
  f.p[1] = n.y + L.lo[1];
  f.p[2] = n.z + L.lo[2];
  f.axis = int8_t(axis);
  f.pad_[0] = f.pad_[1] = 0;
  uint8_t m = 0;
  if (kind == 1) {
    key[k] = face_key(min(n.a, n.b), 0, max(n.a, n.b));
    f.sign = n.a < n.b ? 1 : -1;
  } else if (kind == 2) {
    key[k] = face_key(n.a, 1, 2 * axis + 1);
    f.sign = 1;
    m = uint8_t((n.vb & 0x7F) - 1);
  } else {
    key[k] = face_key(n.b, 1, 2 * axis);
    f.sign = -1;
    m = uint8_t((n.va & 0x7F) - 1);
  }
  face[k] = f;
  mb[k] = m;
  idx[k] = k;
}

__global__ void __launch_bounds__(kWG) k_gather_faces(const int32_t* __restrict__ idx, const svxg_section_face* __restrict__ in,
                                                      const uint8_t* __restrict__ mb_in, svxg_section_face* __restrict__ out,
                                                      uint8_t* __restrict__ mb_out, const uint64_t* __restrict__ key,
                                                      int32_t* __restrict__ first, int n) {
  const int i = int(blockIdx.x) * kWG + threadIdx.x;
  if (i >= n) return;
  out[i] = in[idx[i]];
  mb_out[i] = mb_in[idx[i]];
  first[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void __launch_bounds__(kWG) k_heads(const uint64_t* __restrict__ key, const int32_t* __restrict__ first,
                                               const int32_t* __restrict__ hid, const uint8_t* __restrict__ mb,
                                               const uint8_t* __restrict__ nmat, int nfaces,
                                               svxg_section_head* __restrict__ heads) {
  const int i = int(blockIdx.x) * kWG + threadIdx.x;
  if (i >= nfaces || !first[i]) return;
  const uint64_t k = key[i];
  const int a = int(k >> 22), kind = int((k >> 21) & 1), key2 = int(k & ((1u << 21) - 1));
  svxg_section_head h;
  memset(&h, 0, sizeof h);
  h.a = a;
  h.b = kind == 0 ? key2 : -1;
  h.axis = kind == 0 ? 0 : key2 / 2;
  h.sign = kind == 0 ? 1 : ((key2 & 1) ? 1 : -1);
  h.face0 = i;
  h.nfaces = 0;  // (set by k_head_counts)
  h.ma = nmat[a];
  h.mb = kind == 0 ? nmat[key2] : mb[i];
  h.strength_a = 1.0;
  h.strength_b = kind == 0 ? 1.0 : 1e9;
  heads[hid[i]] = h;
}

__global__ void __launch_bounds__(kWG) k_head_counts(svxg_section_head* __restrict__ heads, int nheads, int nfaces) {
  const int j = int(blockIdx.x) * kWG + threadIdx.x;
  if (j >= nheads) return;
  heads[j].nfaces = (j + 1 < nheads ? heads[j + 1].face0 : nfaces) - heads[j].face0;
}

// a solid unanchored voxel's fragment key (the others sort first, as key 0)
__global__ void __launch_bounds__(kWG) k_voxel_keys(const LatticeDev* __restrict__ Lp, uint32_t* __restrict__ key,
                                                    int32_t* __restrict__ idx, int32_t* __restrict__ count) {
  const LatticeDev& L = *Lp;
  const int64_t v = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (v >= L.nvox) return;
  const int f = L.frag[v];
  key[v] = uint32_t(f + 1);
  idx[v] = int32_t(v);
  if (f >= 0) atomicAdd(&count[f], 1);
}

// a thread a fragment, over its voxels in scan order (world.cpp:775-790 with one fragment a node;
// the mass from the materials' densities, voxel centres at h times their integer coordinates)
__global__ void __launch_bounds__(kWG) k_node_sums(const LatticeDev* __restrict__ Lp, const int32_t* __restrict__ vidx,
                                                   const int32_t* __restrict__ start, int64_t skip,
                                                   svxg_node* __restrict__ nodes, uint8_t* __restrict__ nmat) {
  const LatticeDev& L = *Lp;
  const int f = int(blockIdx.x) * kWG + threadIdx.x;
  if (f >= L.nfrag) return;
  const int64_t s0 = skip + start[f], s1 = skip + start[f + 1];
  const double hhh = L.h * L.h * L.h;
  double mass = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  const int nx = L.dim[0], ny = L.dim[1];
  for (int64_t k = s0; k < s1; ++k) {
    const int64_t v = vidx[k];
    const int id = (L.vox[v] & 0x7F) - 1;
    const double m = L.mats[(id >= 0 && id < L.nmats) ? id : 0].rho * hhh;
    const int x = int(v % nx), y = int((v / nx) % ny), z = int(v / (int64_t(nx) * ny));
    mass += m;
    sx += m * (double(x + L.lo[0]) * L.h);
    sy += m * (double(y + L.lo[1]) * L.h);
    sz += m * (double(z + L.lo[2]) * L.h);
  }
  svxg_node nd;
  memset(&nd, 0, sizeof nd);
  nd.c[0] = sx / mass;
  nd.c[1] = sy / mass;
  nd.c[2] = sz / mass;
  nd.mass = mass;
  nodes[f] = nd;
  // its commonest material (the lower id on a tie): counted over the first few distinct materials
  // met, or over every material id when a fragment holds more
  constexpr int kFew = 8;
  int ids[kFew], cnts[kFew], ndist = 0;
  bool many = false;
  for (int64_t k = s0; k < s1 && !many; ++k) {
    const int id = (L.vox[vidx[k]] & 0x7F) - 1;
    int j = 0;
    while (j < ndist && ids[j] != id) ++j;
    if (j < ndist) {
      ++cnts[j];
    } else if (ndist < kFew) {
      ids[ndist] = id;
      cnts[ndist++] = 1;
    } else {
      many = true;
    }
  }
  int best = 0, best_n = -1;
  if (!many) {
    for (int j = 0; j < ndist; ++j)
      if (cnts[j] > best_n || (cnts[j] == best_n && ids[j] < best)) {
        best_n = cnts[j];
        best = ids[j];
      }
  } else {
    for (int m = 0; m < 128; ++m) {
      int cnt = 0;
      for (int64_t k = s0; k < s1; ++k) cnt += ((L.vox[vidx[k]] & 0x7F) - 1) == m ? 1 : 0;
      if (cnt > 0 && cnt > best_n) {
        best_n = cnt;
        best = m;
      }
    }
  }
  nmat[f] = uint8_t(best);
}

// ---- clusters of fragments ----------------------------------------------------------------------

constexpr int64_t kCoordBias = int64_t(1) << 20;

// a fragment's first voxel: the least (x, y, z) of its voxels, x first, in global coordinates (the
// complement kept, so one atomicMax over a zeroed array finds the least)
__global__ void __launch_bounds__(kWG) k_frag_first(const LatticeDev* __restrict__ Lp, uint64_t* __restrict__ nfirst) {
  const LatticeDev& L = *Lp;
  const int64_t v = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (v >= L.nvox) return;
  const int f = L.frag[v];
  if (f < 0) return;
  const int nx = L.dim[0], ny = L.dim[1];
  const int64_t x = v % nx + L.lo[0], y = (v / nx) % ny + L.lo[1], z = v / (int64_t(nx) * ny) + L.lo[2];
  const uint64_t packed = (uint64_t(x + kCoordBias) << 42) | (uint64_t(y + kCoordBias) << 21) | uint64_t(z + kCoordBias);
  atomicMax(reinterpret_cast<unsigned long long*>(nfirst + f), (unsigned long long)(~packed));
}

__device__ __forceinline__ int64_t floor_div(int64_t a, int64_t b) {
  const int64_t q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// its cell (append_nodes' key): cubes of `cell` voxels, or the reference's default key - a chunk
// (32 voxels) long in x, `cell` in y and z; each label starts as the fragment itself
__global__ void __launch_bounds__(kWG) k_frag_cells(const uint64_t* __restrict__ nfirst, int cell, int cubes,
                                                    uint64_t* __restrict__ cellkey, int32_t* __restrict__ label,
                                                    int nfrag) {
  const int f = int(blockIdx.x) * kWG + threadIdx.x;
  if (f >= nfrag) return;
  const uint64_t packed = ~nfirst[f];
  const int64_t mask = (int64_t(1) << 21) - 1;
  const int64_t x = int64_t(packed >> 42) - kCoordBias, y = int64_t((packed >> 21) & mask) - kCoordBias,
                z = int64_t(packed & mask) - kCoordBias;
  const int64_t cx = floor_div(x, cubes ? cell : 32), cy = floor_div(y, cell), cz = floor_div(z, cell);
  cellkey[f] = (uint64_t(cx + kCoordBias) << 42) | (uint64_t(cy + kCoordBias) << 21) | uint64_t(cz + kCoordBias);
  label[f] = f;
}

// a round: every link (two touching fragments of one cell) takes both ends, and their labels' own
// labels, to the lesser label; then each label follows its label once. Labels only fall, and only
// to labels of the same cluster, so a round that changes nothing leaves each cluster its least
// fragment.
__global__ void __launch_bounds__(kWG) k_cluster_link(const LatticeDev* __restrict__ Lp,
                                                      const uint64_t* __restrict__ cellkey, int32_t* __restrict__ label,
                                                      int32_t* __restrict__ changed) {
  const LatticeDev& L = *Lp;
  const int64_t slot = int64_t(blockIdx.x) * kWG + threadIdx.x;
  if (slot >= 3 * L.nvox) return;
  Nbr n;
  if (!neighbour(L, slot, &n) || n.a < 0 || n.b < 0 || n.a == n.b || cellkey[n.a] != cellkey[n.b]) return;
  const int la = label[n.a], lb = label[n.b];
  if (la == lb) return;
  const int m = min(la, lb);
  atomicMin(label + n.a, m);
  atomicMin(label + n.b, m);
  atomicMin(label + la, m);
  atomicMin(label + lb, m);
  *changed = 1;
}
__global__ void __launch_bounds__(kWG) k_cluster_jump(int32_t* __restrict__ label, int nfrag) {
  const int f = int(blockIdx.x) * kWG + threadIdx.x;
  if (f >= nfrag) return;
  const int l = label[f];
  const int ll = label[l];
  if (ll < l) atomicMin(label + f, ll);
}

__global__ void __launch_bounds__(kWG) k_cluster_roots(const int32_t* __restrict__ label, int32_t* __restrict__ root,
                                                       int nfrag) {
  const int f = int(blockIdx.x) * kWG + threadIdx.x;
  if (f < nfrag) root[f] = label[f] == f ? 1 : 0;
}

// a fragment's node: its cluster's rank (clusters in order of their least fragment); the sort keys
// of the fragments by node, and the nodes' member counts
__global__ void __launch_bounds__(kWG) k_cluster_map(const int32_t* __restrict__ label, const int32_t* __restrict__ rank,
                                                     int32_t* __restrict__ fnode, uint32_t* __restrict__ key,
                                                     int32_t* __restrict__ idx, int32_t* __restrict__ count, int nfrag) {
  const int f = int(blockIdx.x) * kWG + threadIdx.x;
  if (f >= nfrag) return;
  const int nd = rank[label[f]];
  fnode[f] = nd;
  key[f] = uint32_t(nd);
  idx[f] = f;
  atomicAdd(count + nd, 1);
}

// a thread a node, over its fragments in index order (world.cpp:771-797): the masses summed, the
// centres weighted by them, then times 1 / mass; the heaviest fragment's material (the first of
// equal ones)
__global__ void __launch_bounds__(kWG) k_cluster_nodes(const int32_t* __restrict__ member, const int32_t* __restrict__ start,
                                                       const svxg_node* __restrict__ fnodes,
                                                       const uint8_t* __restrict__ fmat, svxg_node* __restrict__ nodes,
                                                       uint8_t* __restrict__ nmat, int nnodes) {
  const int c = int(blockIdx.x) * kWG + threadIdx.x;
  if (c >= nnodes) return;
  double mass = 0.0, heavy = -1.0, cx = 0.0, cy = 0.0, cz = 0.0;
  uint8_t mat = 0;
  for (int k = start[c]; k < start[c + 1]; ++k) {
    const svxg_node& fn = fnodes[member[k]];
    mass += fn.mass;
    cx += fn.c[0] * fn.mass;
    cy += fn.c[1] * fn.mass;
    cz += fn.c[2] * fn.mass;
    if (fn.mass > heavy) {
      heavy = fn.mass;
      mat = fmat[member[k]];
    }
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
}

}  // namespace

void launch_frag_first(hipStream_t s, const LatticeDev* L, uint64_t* nfirst, int64_t nvox) {
  if (nvox > 0) hipLaunchKernelGGL(k_frag_first, dim3(unsigned((nvox + kWG - 1) / kWG)), dim3(kWG), 0, s, L, nfirst);
}
void launch_frag_cells(hipStream_t s, const LatticeDev*, const uint64_t* nfirst, int cell, int cubes, uint64_t* cellkey,
                       int32_t* label, int nfrag) {
  if (nfrag > 0)
    hipLaunchKernelGGL(k_frag_cells, dim3((nfrag + kWG - 1) / kWG), dim3(kWG), 0, s, nfirst, cell, cubes, cellkey,
                       label, nfrag);
}
void launch_cluster_round(hipStream_t s, const LatticeDev* L, const uint64_t* cellkey, int32_t* label, int32_t* changed,
                          int64_t nslots, int nfrag) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_cluster_link, dim3(unsigned((nslots + kWG - 1) / kWG)), dim3(kWG), 0, s, L, cellkey, label,
                       changed);
  if (nfrag > 0) hipLaunchKernelGGL(k_cluster_jump, dim3((nfrag + kWG - 1) / kWG), dim3(kWG), 0, s, label, nfrag);
}
void launch_cluster_ids(hipStream_t s, const int32_t* label, int32_t* root, int nfrag) {
  if (nfrag > 0) hipLaunchKernelGGL(k_cluster_roots, dim3((nfrag + kWG - 1) / kWG), dim3(kWG), 0, s, label, root, nfrag);
}
void launch_cluster_map(hipStream_t s, const int32_t* label, const int32_t* rank, int32_t* fnode, uint32_t* key,
                        int32_t* idx, int32_t* count, int nfrag) {
  if (nfrag > 0)
    hipLaunchKernelGGL(k_cluster_map, dim3((nfrag + kWG - 1) / kWG), dim3(kWG), 0, s, label, rank, fnode, key, idx,
                       count, nfrag);
}
void launch_cluster_nodes(hipStream_t s, const int32_t* member, const int32_t* start, const svxg_node* fnodes,
                          const uint8_t* fmat, svxg_node* nodes, uint8_t* nmat, int nnodes) {
  if (nnodes > 0)
    hipLaunchKernelGGL(k_cluster_nodes, dim3((nnodes + kWG - 1) / kWG), dim3(kWG), 0, s, member, start, fnodes, fmat,
                       nodes, nmat, nnodes);
}

void launch_face_flags(hipStream_t s, const LatticeDev* L, int32_t* flag, int64_t nslots) {
  if (nslots > 0) hipLaunchKernelGGL(k_face_flags, dim3(unsigned((nslots + kWG - 1) / kWG)), dim3(kWG), 0, s, L, flag);
}
void launch_face_write(hipStream_t s, const LatticeDev* L, const int32_t* flag, const int32_t* at, uint64_t* key,
                       int32_t* idx, svxg_section_face* face, uint8_t* mb, int64_t nslots) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_face_write, dim3(unsigned((nslots + kWG - 1) / kWG)), dim3(kWG), 0, s, L, flag, at, key, idx,
                       face, mb);
}
void launch_gather_faces(hipStream_t s, const int32_t* idx, const svxg_section_face* in, const uint8_t* mb_in,
                         svxg_section_face* out, uint8_t* mb_out, const uint64_t* key, int32_t* first, int n) {
  if (n > 0)
    hipLaunchKernelGGL(k_gather_faces, dim3((n + kWG - 1) / kWG), dim3(kWG), 0, s, idx, in, mb_in, out, mb_out, key,
                       first, n);
}
void launch_heads(hipStream_t s, const uint64_t* key, const int32_t* first, const int32_t* hid, const uint8_t* mb,
                  const uint8_t* nmat, int nfaces, svxg_section_head* heads, int nheads) {
  if (nfaces > 0) {
    hipLaunchKernelGGL(k_heads, dim3((nfaces + kWG - 1) / kWG), dim3(kWG), 0, s, key, first, hid, mb, nmat, nfaces, heads);
    hipLaunchKernelGGL(k_head_counts, dim3((nheads + kWG - 1) / kWG), dim3(kWG), 0, s, heads, nheads, nfaces);
  }
}
void launch_voxel_keys(hipStream_t s, const LatticeDev* L, uint32_t* key, int32_t* idx, int32_t* count, int64_t nvox) {
  if (nvox > 0)
    hipLaunchKernelGGL(k_voxel_keys, dim3(unsigned((nvox + kWG - 1) / kWG)), dim3(kWG), 0, s, L, key, idx, count);
}
void launch_node_sums(hipStream_t s, const LatticeDev* L, const int32_t* vidx, const int32_t* start, int64_t skip,
                      svxg_node* nodes, uint8_t* nmat, int nfrag) {
  if (nfrag > 0)
    hipLaunchKernelGGL(k_node_sums, dim3((nfrag + kWG - 1) / kWG), dim3(kWG), 0, s, L, vidx, start, skip, nodes, nmat);
}

// rocPRIM's device scan and stable radix sort; with tmp null, `bytes` receives the temporary storage
// they need (rocPRIM's own protocol)
hipError_t scan_exclusive_i32(void* tmp, size_t& bytes, const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
  return rocprim::exclusive_scan(tmp, bytes, in, out, 0, size_t(n), rocprim::plus<int32_t>(), s);
}
hipError_t sort_pairs_u64(void* tmp, size_t& bytes, const uint64_t* kin, uint64_t* kout, const int32_t* vin,
                          int32_t* vout, int64_t n, int bits, hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, size_t(n), 0, bits, s);
}
hipError_t sort_pairs_u32(void* tmp, size_t& bytes, const uint32_t* kin, uint32_t* kout, const int32_t* vin,
                          int32_t* vout, int64_t n, int bits, hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, size_t(n), 0, bits, s);
}

}  // namespace svxg
