// The judge of a solved structure or piece (World::Impl::judge, world.cpp:1162-1277; the piece
// check, world_pieces.cpp:647-671) on the device, for a batch of problems:
//   k_judge         every judged bond's loads from the resident solution, the DIF lerp against
//                   the baseline (lerp_load, world_internal.hpp:452-461), bond_utilization, the
//                   baseline update, and an unordered list of the bonds over strength (phi >= 1)
//   k_judge_select  a workgroup a problem: the round's picks - the list ordered by phi
//                   (descending, ties by index, world.cpp:1220), the band max(1, band x max phi),
//                   the quota, the cap - and the plastic-hinge cuts of the picks (hinge_of,
//                   world.cpp:1286-1334). Lists longer than the LDS sort take a radix selection
//                   of the cap-th largest key first.
// Integer and ordering work except the per-bond arithmetic, which follows the reference's
// operation order (bond_math.hpp): phi and the picks are the CPU's bits for the same u.
#include "bond_math.hpp"

namespace svxg {

namespace {

constexpr int kSelWG = 1024;

__device__ __forceinline__ bool ductile(const JudgeDev& c, int m) {
  return m >= 0 && m < 128 && ((c.ductile[m >> 6] >> (m & 63)) & 1ull);
}

__global__ void __launch_bounds__(kWG) k_judge(const JudgeDev* __restrict__ J, const Tile* __restrict__ tiles, int ntiles) {
  if (int(blockIdx.x) >= ntiles) return;
  const Tile tl = tiles[blockIdx.x];
  const JudgeDev& c = J[tl.g];
  const int t = threadIdx.x, lane = t & 63;
  const int bi = tl.row0 + t;
  const bool in = t < tl.nrows;
  const bool judged = in && !c.dead[bi];
  double phi = 0.0;
  uint8_t md = 0;
  if (judged) {
    const svxg_bond& b = c.bonds[bi];
    const svxg_material& MA = bond_mat(c.mats, c.nmats, b.ma);
    const svxg_material& MB = bond_mat(c.mats, c.nmats, b.mb);
    const svxg_bond_load L = bond_load_dev(b, c.nodes, MA, MB, c.u);
    svxg_bond_load Le = L;
    // (a shock: the change since the last judgement, amplified - bonds judged before only)
    if (c.shock && c.dif != 1.0 && c.base && c.has_base[bi]) Le = lerp_load_dev(c.base[bi], L, c.dif);
    phi = bond_phi_dev(b, bond_strength_dev(b, MA, MB, c.fragility), Le, &md);
    if (!c.transient && c.base) {
      c.base[bi] = L;
      c.has_base[bi] = 1;
    }
  }
  if (in && c.phi) c.phi[bi] = phi;
  // the wavefront's largest phi (phi > 0: its bits order as the values; NaN is not counted) and
  // its bonds over strength, appended with one atomic each
  unsigned long long bits = (judged && phi > 0.0) ? static_cast<unsigned long long>(__double_as_longlong(phi)) : 0ull;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long v = __shfl_xor(bits, o);
    bits = v > bits ? v : bits;
  }
  const bool over = judged && phi >= 1.0;
  const unsigned long long mask = __ballot(over);
  int base = 0;
  if (lane == 0) {
    if (bits) atomicMax(&c.out->max_bits, bits);
    if (mask) base = atomicAdd(&c.out->over, int(__popcll(mask)));
  }
  base = __shfl(base, 0);
  if (over) {
    const int k = base + int(__popcll(mask & ((1ull << lane) - 1ull)));
    c.over[k] = OverEntry{static_cast<uint64_t>(__double_as_longlong(phi)), bi, md};
  }
}

// a precedes b: larger phi, then smaller index
__device__ __forceinline__ bool precedes(uint64_t ka, int32_t ba, uint64_t kb, int32_t bb) {
  return ka > kb || (ka == kb && ba < bb);
}

// masks of the key bits fixed after `depth` 8-bit digits of the 96-bit key (phi bits : ~index)
__device__ __forceinline__ void key_masks(int depth, uint64_t* mh, uint32_t* ml) {
  if (depth <= 8) {
    *mh = depth == 0 ? 0ull : ~0ull << (64 - 8 * depth);
    *ml = 0u;
  } else {
    *mh = ~0ull;
    *ml = depth == 12 ? ~0u : ~0u << (32 - 8 * (depth - 8));
  }
}

__device__ __forceinline__ int key_digit(uint64_t hi, uint32_t lo, int d) {
  return d < 8 ? int((hi >> (56 - 8 * d)) & 255ull) : int((lo >> (24 - 8 * (d - 8))) & 255u);
}

__global__ void __launch_bounds__(kSelWG) k_judge_select(const JudgeDev* __restrict__ J, int count) {
  __shared__ uint64_t skey[kMaxPicks];
  __shared__ int32_t sbond[kMaxPicks];
  __shared__ int32_t sslot[kMaxPicks];
  __shared__ int32_t shist[256];
  __shared__ int32_t s_count, s_thr, s_krem, s_depth, s_done, s_hinges;
  __shared__ uint64_t s_ph;
  __shared__ uint32_t s_pl;
  if (int(blockIdx.x) >= count) return;
  const JudgeDev& c = J[blockIdx.x];
  const int t = threadIdx.x;
  const int n = c.out->over;
  if (c.skip || n == 0) {
    if (t == 0) {
      c.out->picked = 0;
      c.out->hinges = 0;
    }
    return;
  }
  const double phimax = __longlong_as_double(static_cast<long long>(c.out->max_bits));
  const double thr = fmax(1.0, c.band * phimax);
  const uint64_t thr_bits = static_cast<uint64_t>(__double_as_longlong(thr));
  if (t == 0) {
    s_thr = 0;
    s_count = 0;
    s_hinges = 0;
  }
  __syncthreads();
  int mine = 0;
  for (int i = t; i < n; i += kSelWG) mine += c.over[i].key >= thr_bits ? 1 : 0;
  if (mine) atomicAdd(&s_thr, mine);
  __syncthreads();
  // the round takes the bonds at or above the band and at least the quota, at most the cap
  const int quota = c.rule == 0 ? max(1, (n + 3) / 4) : int(ceil(c.quota_fraction * double(n)));
  const int take = min(min(n, c.max_breaks), max(s_thr, quota));
  int cnt;
  if (n <= kMaxPicks) {
    for (int i = t; i < n; i += kSelWG) {
      skey[i] = c.over[i].key;
      sbond[i] = c.over[i].bond;
      sslot[i] = i;
    }
    cnt = n;
  } else {
    // the take-th largest key, digit by digit from the top; then every key at or above it
    if (t == 0) {
      s_ph = 0ull;
      s_pl = 0u;
      s_krem = take;
      s_depth = 0;
      s_done = 0;
    }
    __syncthreads();
    for (int d = 0; d < 12; ++d) {
      if (s_done) break;
      for (int k = t; k < 256; k += kSelWG) shist[k] = 0;
      __syncthreads();
      uint64_t mh;
      uint32_t ml;
      key_masks(d, &mh, &ml);
      const uint64_t ph = s_ph;
      const uint32_t pl = s_pl;
      for (int i = t; i < n; i += kSelWG) {
        const uint64_t hi = c.over[i].key;
        const uint32_t lo = ~static_cast<uint32_t>(c.over[i].bond);
        if ((hi & mh) == ph && (lo & ml) == pl) atomicAdd(&shist[key_digit(hi, lo, d)], 1);
      }
      __syncthreads();
      if (t == 0) {
        int krem = s_krem, v = 255;
        for (; v > 0; --v) {
          if (shist[v] >= krem) break;
          krem -= shist[v];
        }
        if (d < 8)
          s_ph |= static_cast<uint64_t>(v) << (56 - 8 * d);
        else
          s_pl |= static_cast<uint32_t>(v) << (24 - 8 * (d - 8));
        s_krem = krem;
        s_depth = d + 1;
        if (shist[v] == krem) s_done = 1;
      }
      __syncthreads();
    }
    uint64_t mh;
    uint32_t ml;
    key_masks(s_depth, &mh, &ml);
    const uint64_t ph = s_ph;
    const uint32_t pl = s_pl;
    for (int i = t; i < n; i += kSelWG) {
      const uint64_t hi = c.over[i].key & mh;
      const uint32_t lo = ~static_cast<uint32_t>(c.over[i].bond) & ml;
      if (hi > ph || (hi == ph && lo >= pl)) {
        const int k = atomicAdd(&s_count, 1);
        if (k < kMaxPicks) {
          skey[k] = c.over[i].key;
          sbond[k] = c.over[i].bond;
          sslot[k] = i;
        }
      }
    }
    __syncthreads();
    cnt = min(s_count, kMaxPicks);
  }
  // bitonic sort of the candidates (padding sorts last), in LDS
  int N = 1;
  while (N < cnt) N <<= 1;
  for (int i = cnt + t; i < N; i += kSelWG) {
    skey[i] = 0ull;
    sbond[i] = 0x7fffffff;
    sslot[i] = -1;
  }
  __syncthreads();
  for (int k = 2; k <= N; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < N; i += kSelWG) {
        const int x = i ^ j;
        if (x > i) {
          const bool asc = (i & k) == 0;
          const bool sw = asc ? precedes(skey[x], sbond[x], skey[i], sbond[i]) : precedes(skey[i], sbond[i], skey[x], sbond[x]);
          if (sw) {
            const uint64_t tk = skey[i];
            skey[i] = skey[x];
            skey[x] = tk;
            const int32_t tb = sbond[i];
            sbond[i] = sbond[x];
            sbond[x] = tb;
            const int32_t ts = sslot[i];
            sslot[i] = sslot[x];
            sslot[x] = ts;
          }
        }
      }
      __syncthreads();
    }
  const int picked = min(take, cnt);
  for (int k = t; k < picked; k += kSelWG) {
    const OverEntry e = c.over[sslot[k]];
    svxg_break r = {};
    r.bond = e.bond;
    r.round = c.round;
    r.phi = __longlong_as_double(static_cast<long long>(e.key));
    r.mode = static_cast<uint8_t>(e.mode);
    const svxg_bond& b = c.bonds[e.bond];
    // (world.cpp:1251-1252; pieces only between two nodes, world_pieces.cpp:683)
    if (c.hinges && (c.rule == 0 || b.b >= 0) && ductile(c, b.ma) && ductile(c, b.mb)) {
      const svxg_material& MA = bond_mat(c.mats, c.nmats, b.ma);
      const svxg_material& MB = bond_mat(c.mats, c.nmats, b.mb);
      const svxg_bond_load L = bond_load_dev(b, c.nodes, MA, MB, c.u);
      const double zero[3] = {0.0, 0.0, 0.0};
      const double* ra = c.u + 6 * size_t(b.a) + 3;
      const double* rb = b.b >= 0 ? c.u + 6 * size_t(b.b) + 3 : zero;
      if (hinge_dev(b, L, bond_strength_dev(b, MA, MB, c.fragility), ra, rb, c.hinge_shape, &r.cut)) {
        r.hinge = 1;
        atomicAdd(&s_hinges, 1);
      }
    }
    c.picks[k] = r;
  }
  __syncthreads();
  if (t == 0) {
    c.out->picked = picked;
    c.out->hinges = s_hinges;
  }
}

__global__ void __launch_bounds__(kWG) k_zero_rows(const ZeroRowsJob* __restrict__ jobs, int count) {
  if (int(blockIdx.x) >= count) return;
  const ZeroRowsJob j = jobs[blockIdx.x];
  for (int e = threadIdx.x; e < 6 * j.n; e += kWG) {
    const size_t at = 6 * size_t(j.rows[e / 6]) + size_t(e % 6);
    j.x[at] = 0.0;
    j.b[at] = 0.0;
  }
}

}  // namespace

void launch_zero_rows(hipStream_t s, const ZeroRowsJob* jobs, int count) {
  if (count > 0) hipLaunchKernelGGL(k_zero_rows, dim3(count), dim3(kWG), 0, s, jobs, count);
}

void launch_judge(hipStream_t s, const JudgeDev* jobs, const Tile* tiles, int ntiles) {
  if (ntiles > 0) hipLaunchKernelGGL(k_judge, dim3(ntiles), dim3(kWG), 0, s, jobs, tiles, ntiles);
}

void launch_judge_select(hipStream_t s, const JudgeDev* jobs, int count) {
  if (count > 0) hipLaunchKernelGGL(k_judge_select, dim3(count), dim3(kSelWG), 0, s, jobs, count);
}

}  // namespace svxg
