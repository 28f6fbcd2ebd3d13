// Bond sections from voxel faces (SecAcc::finish_lattice, world_internal.hpp:129-249, and
// section_strengths, :374-400): a thread a bond walks its faces (the interface normal, the
// weighted centroid, the principal axes of the section, its moments and extreme fibres, then the
// face strengths), in the reference's operation order - this file is compiled with
// -ffp-contract=off, so a bond is the CPU's bits for the same faces. The faces of a bond are
// contiguous in the face list; a bond's work is a few passes over its own faces (latency-bound
// integer and FP64 work: a thread per bond, no matrix cores).
//
// The general case (finish_general, :250-362, section_strengths_general, :402-441: faces of a
// placed lattice and junction samples, up to S x S a face) runs in two phases: a thread an element
// turns each face or sample into its square in the bond's frame (two 3x3 products for a placed
// lattice) and its strength share, so a bond of thousands of samples does not leave one thread
// with the transforms; then a thread a bond sums its elements in their order, as the CPU does.
#include "svxg_kernels.hpp"

namespace svxg {

namespace {

struct V3d {
  double x, y, z;
};
__device__ __forceinline__ V3d v_add(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3d v_sub(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3d v_mul(V3d a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double v_dot(V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3d v_cross(V3d a, V3d b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3d v_unit(V3d a) {
  const double n = sqrt(v_dot(a, a));
  return n > 0 ? v_mul(a, 1.0 / n) : V3d{0.0, 0.0, 0.0};
}
__device__ __forceinline__ double& v_at(V3d& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ double v_get(const V3d& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// frame_of (world_internal.hpp:64-69): t1 square to n and to the axis n is least along
__device__ __forceinline__ void frame(V3d n, V3d* t1, V3d* t2) {
  const double ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
  const V3d e = (ax <= ay && ax <= az) ? V3d{1, 0, 0} : (ay <= az ? V3d{0, 1, 0} : V3d{0, 0, 1});
  *t1 = v_unit(v_cross(n, e));
  *t2 = v_cross(n, *t1);
}

// the face's centre: its lower voxel's centre moved half a voxel along the axis
__device__ __forceinline__ V3d face_centre(const svxg_section_face& f, double h) {
  V3d x{h * f.p[0], h * f.p[1], h * f.p[2]};
  v_at(x, f.axis) += 0.5 * h;
  return x;
}

// a voxel of a box (lo, dim; x fastest) and its damage; outside the box: air, intact
__device__ __forceinline__ void box_voxel(const uint8_t* vox, const uint8_t* damage, const int32_t* lo,
                                          const int32_t* dim, int x, int y, int z, uint8_t* v, uint8_t* d) {
  const int i = x - lo[0], j = y - lo[1], k = z - lo[2];
  if (i < 0 || j < 0 || k < 0 || i >= dim[0] || j >= dim[1] || k >= dim[2]) {
    *v = 0;
    *d = 0;
    return;
  }
  const size_t at = size_t(i) + size_t(dim[0]) * (size_t(j) + size_t(dim[1]) * size_t(k));
  *v = vox[at];
  *d = damage ? damage[at] : 0;
}

__device__ __forceinline__ const svxg_material& mat_of_vox(const svxg_material* mats, int nmats, uint8_t v) {
  const int id = (v & 0x7F) - 1;  // (air: -1, taken as material 0)
  return mats[(id >= 0 && id < nmats) ? id : 0];
}

// One face or junction sample's share of the section strengths (section_strengths(_general),
// world_internal.hpp:374-441): weight w times the condition of the more damaged voxel times the
// weaker material's strengths (an air side takes the other side's material); friction without
// the condition.
struct Share {
  double ft, fb, fc, coh, mu;
};
__device__ __forceinline__ Share strength_share(const svxg_material* mats, int nmats, uint8_t va, uint8_t da,
                                                uint8_t vb, uint8_t db, double w) {
  const svxg_material& MA = mat_of_vox(mats, nmats, va != 0 ? va : vb);
  const svxg_material& MM = mat_of_vox(mats, nmats, vb != 0 ? vb : va);
  const double cond = fmax(1e-6, 1.0 - double(da > db ? da : db) / 255.0);
  return {w * cond * fmin(MA.ft, MM.ft), w * cond * fmin(MA.fb, MM.fb), w * cond * fmin(MA.fc, MM.fc),
          w * cond * fmin(MA.cohesion, MM.cohesion), w * fmin(MA.friction, MM.friction)};
}

// t1, t2 turned to the principal axes of the section (the eigenvector of [[s11, s12], [s12, s22]]
// of the larger value, in the better conditioned of its two forms; sqrt only)
__device__ __forceinline__ void principal(double s11, double s22, double s12, V3d* t1, V3d* t2) {
  if (!(fabs(s12) > 1e-9 * (s11 + s22))) return;
  const double half = 0.5 * (s11 - s22);
  const double l1 = 0.5 * (s11 + s22) + sqrt(half * half + s12 * s12);
  double e1 = s12, e2 = l1 - s11;
  if (fabs(l1 - s22) > fabs(e2)) {
    e1 = l1 - s22;
    e2 = s12;
  }
  const double en = sqrt(e1 * e1 + e2 * e2);
  if (en > 0.0) {
    e1 /= en;
    e2 /= en;
    const V3d u1 = v_add(v_mul(*t1, e1), v_mul(*t2, e2));
    const V3d u2 = v_sub(v_mul(*t2, e1), v_mul(*t1, e2));
    *t1 = u1;
    *t2 = u2;
  }
}

// the spring lengths from the centres: a pair splits L = max(|d.n|, |d| / 2, h / 2) at the
// centroid (clamped to 10 .. 90 %); a support reaches h / 2 beyond it
__device__ __forceinline__ void springs(bool pair, V3d ca, V3d cb, V3d n, V3d pc, double h, svxg_bond* B) {
  if (pair) {
    const V3d d = v_sub(cb, ca);
    const double L = fmax(fmax(fabs(v_dot(d, n)), 0.5 * sqrt(v_dot(d, d))), 0.5 * h);
    const double t = v_dot(v_sub(pc, ca), n);
    B->la = t < 0.1 * L ? 0.1 * L : (0.9 * L < t ? 0.9 * L : t);
    B->lb = L - B->la;
  } else {
    B->la = fmax(0.25 * h, v_dot(v_sub(pc, ca), n));
    B->lb = 0.5 * h;
  }
}

__device__ __forceinline__ void put_frame(svxg_bond* B, V3d n, V3d t1, V3d t2, V3d pc) {
  for (int i = 0; i < 3; ++i) {
    B->n[i] = v_get(n, i);
    B->t1[i] = v_get(t1, i);
    B->t2[i] = v_get(t2, i);
    B->p[i] = v_get(pc, i);
  }
}

// The faces' accumulator of one unrotated lattice (finish_lattice + section_strengths): the faces'
// mean axis normal, their centres weighted by the projected area, fibres half a voxel beyond the
// outermost centres, the faces' mean strengths.
struct LatticeHead {
  int32_t a, b, axis, sign, nf, ma, mb;
  double strength_a, strength_b;
};
__device__ svxg_bond lattice_bond(const LatticeHead& hd, const svxg_section_face* __restrict__ F, double h,
                                  const svxg_node* __restrict__ nodes, const uint8_t* vox, const uint8_t* damage,
                                  const int32_t* lo, const int32_t* dim, const svxg_material* mats, int nmats,
                                  int tag) {
  const int nf = hd.nf;
  const double A1 = h * h;
  const V3d ca{nodes[hd.a].c[0], nodes[hd.a].c[1], nodes[hd.a].c[2]};
  const bool pair = hd.b >= 0;
  V3d cb{0, 0, 0};
  if (pair) cb = V3d{nodes[hd.b].c[0], nodes[hd.b].c[1], nodes[hd.b].c[2]};
  V3d n{0, 0, 0};
  if (pair) {
    for (int f = 0; f < nf; ++f) v_at(n, F[f].axis) += double(F[f].sign);
    if (v_dot(n, n) < 1e-12) n = v_sub(cb, ca);
    if (v_dot(n, n) < 1e-12) n = V3d{0, 0, 1};
    n = v_unit(n);
  } else {
    v_at(n, hd.axis) = double(hd.sign);
  }
  V3d t1, t2;
  frame(n, &t1, &t2);
  double W = 0.0;
  V3d pc{0, 0, 0};
  for (int f = 0; f < nf; ++f) {
    const double w = A1 * fmax(0.05, fabs(v_get(n, F[f].axis)));
    W += w;
    pc = v_add(pc, v_mul(face_centre(F[f], h), w));
  }
  pc = v_mul(pc, 1.0 / W);
  {
    double s11 = 0.0, s22 = 0.0, s12 = 0.0;
    for (int f = 0; f < nf; ++f) {
      const V3d d = v_sub(face_centre(F[f], h), pc);
      const double w = A1 * fmax(0.05, fabs(v_get(n, F[f].axis)));
      const double y1 = v_dot(d, t1), y2 = v_dot(d, t2);
      s11 += w * y1 * y1;
      s22 += w * y2 * y2;
      s12 += w * y1 * y2;
    }
    principal(s11, s22, s12, &t1, &t2);
  }
  svxg_bond B;
  memset(&B, 0, sizeof B);
  B.a = hd.a;
  B.b = hd.b;
  B.ma = uint8_t(hd.ma);
  B.mb = uint8_t(hd.mb);
  B.faces = nf;
  B.tag = tag;
  put_frame(&B, n, t1, t2, pc);
  B.area = fmax(W, 0.25 * nf * A1);
  double s1 = 0.0, s2 = 0.0, m1 = 0.0, m2 = 0.0;
  for (int f = 0; f < nf; ++f) {
    const V3d d = v_sub(face_centre(F[f], h), pc);
    const double w = A1 * fmax(0.05, fabs(v_get(n, F[f].axis)));
    const double y1 = v_dot(d, t1), y2 = v_dot(d, t2);
    s1 += w * y1 * y1;
    s2 += w * y2 * y2;
    m1 = fmax(m1, fabs(y1));
    m2 = fmax(m2, fabs(y2));
  }
  B.s1 = s1 + B.area * A1 / 12.0;
  B.s2 = s2 + B.area * A1 / 12.0;
  B.c1 = m1 + 0.5 * h;
  B.c2 = m2 + 0.5 * h;
  B.rmax = sqrt(B.c1 * B.c1 + B.c2 * B.c2);
  springs(pair, ca, cb, n, pc, h, &B);
  B.strength = fmin(hd.strength_a, hd.strength_b);
  double ft = 0, fb = 0, fc = 0, coh = 0, mu = 0;
  for (int f = 0; f < nf; ++f) {
    uint8_t va, da, vb, db;
    box_voxel(vox, damage, lo, dim, F[f].p[0], F[f].p[1], F[f].p[2], &va, &da);
    int q[3] = {F[f].p[0], F[f].p[1], F[f].p[2]};
    q[F[f].axis] += 1;
    box_voxel(vox, damage, lo, dim, q[0], q[1], q[2], &vb, &db);
    const Share sh = strength_share(mats, nmats, va, da, vb, db, 1.0);
    ft += sh.ft;
    fb += sh.fb;
    fc += sh.fc;
    coh += sh.coh;
    mu += sh.mu;
  }
  if (nf > 0) {
    const double w = 1.0 / double(nf);
    B.sectioned = 1;
    B.ft = float(ft * w);
    B.fb = float(fb * w);
    B.fc = float(fc * w);
    B.coh = float(coh * w);
    B.mu = float(mu * w);
  }
  return B;
}

__global__ void __launch_bounds__(kWG) k_bond_sections(const SectionDev* __restrict__ cp) {
  const SectionDev& c = *cp;
  const int k = int(blockIdx.x) * kWG + int(threadIdx.x);
  if (k >= c.m) return;
  const svxg_section_head& h = c.heads[k];
  const LatticeHead hd{h.a, h.b, h.axis, h.sign, h.nfaces, h.ma, h.mb, h.strength_a, h.strength_b};
  c.out[k] = lattice_bond(hd, c.faces + h.face0, c.h, c.nodes, c.vox, c.damage, c.lo, c.dim, c.mats, c.nmats, k);
}

// ---- the general case: placed lattices and junction samples ----------------------------------------

// LatticeXf::to / dir_to (rigid.hpp:91-93): R s by rows, then the offset; an unplaced lattice as is
__device__ __forceinline__ V3d lat_dir(const svxg_lattice& L, V3d s) {
  if (L.identity) return s;
  const double* R = L.R;
  return {R[0] * s.x + R[1] * s.y + R[2] * s.z, R[3] * s.x + R[4] * s.y + R[5] * s.z,
          R[6] * s.x + R[7] * s.y + R[8] * s.z};
}
__device__ __forceinline__ V3d lat_point(const svxg_lattice& L, V3d s) {
  if (L.identity) return s;
  const V3d r = lat_dir(L, s);
  return {L.off[0] + r.x, L.off[1] + r.y, L.off[2] + r.z};
}
__device__ __forceinline__ void lat_voxel(const svxg_lattice& L, const int32_t* p, uint8_t* v, uint8_t* d) {
  box_voxel(L.vox, L.damage, L.lo, L.dim, p[0], p[1], p[2], v, d);
}

// Phase 1, a thread an element (a face or a junction sample of a general bond): its square in the
// bond's frame - centre, normal from a to b, area, side - and its strength share.
__global__ void __launch_bounds__(kWG) k_section_elements(const SectionGenDev* __restrict__ cp) {
  const SectionGenDev& c = *cp;
  const int64_t e = int64_t(blockIdx.x) * kWG + int64_t(threadIdx.x);
  if (e >= c.E) return;
  // its bond: eoff[k] <= e < eoff[k + 1] (eoff[0] = 0, eoff[m] = E)
  int lo = 0, hi = c.m;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (c.eoff[mid] <= e) lo = mid;
    else hi = mid;
  }
  const svxg_section_head_general& hd = c.heads[lo];
  const int i = int(e - c.eoff[lo]);
  SecElem r;
  uint8_t va, da, vb, db;
  double w;
  if (i < hd.nfaces) {
    const svxg_section_face& f = c.faces[hd.face0 + i];
    const svxg_lattice& G = c.lat[hd.grid];
    const double h = G.h;
    const V3d x = lat_point(G, face_centre(f, h));
    V3d nn{0, 0, 0};
    v_at(nn, f.axis) = double(f.sign);
    const V3d nk = lat_dir(G, nn);
    r.x[0] = x.x, r.x[1] = x.y, r.x[2] = x.z;
    r.nk[0] = nk.x, r.nk[1] = nk.y, r.nk[2] = nk.z;
    r.ar = h * h;
    r.sd = h;
    lat_voxel(G, f.p, &va, &da);
    int32_t q[3] = {f.p[0], f.p[1], f.p[2]};
    q[f.axis] += 1;
    lat_voxel(G, q, &vb, &db);
    w = 1.0;
  } else {
    const svxg_junction_sample& js = c.samples[hd.sample0 + (i - hd.nfaces)];
    const svxg_lattice& X = c.lat[js.vg];
    const int S = c.S;
    const double hs = X.h, sj = hs / double(S > 1 ? S : 1);
    // junction_point (world_internal.hpp:94-103): the centre of cell (i, j) of an S x S split of
    // the face, on the face (no push)
    const int ax = js.face >> 1, a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    const double sg = (js.face & 1) ? 1.0 : -1.0;
    const int si = js.sub / S, sjx = js.sub % S;
    V3d pt{hs * js.v[0], hs * js.v[1], hs * js.v[2]};
    v_at(pt, ax) += sg * (0.5 * hs + 0.0);
    v_at(pt, a1) += ((si + 0.5) / double(S) - 0.5) * hs;
    v_at(pt, a2) += ((sjx + 0.5) / double(S) - 0.5) * hs;
    const V3d x = lat_point(X, pt);
    V3d fn{0, 0, 0};
    v_at(fn, ax) = sg;
    const V3d nk = v_mul(lat_dir(X, fn), double(js.sign));
    r.x[0] = x.x, r.x[1] = x.y, r.x[2] = x.z;
    r.nk[0] = nk.x, r.nk[1] = nk.y, r.nk[2] = nk.z;
    r.ar = sj * sj * double(js.weight);
    r.sd = sj;
    lat_voxel(X, js.v, &va, &da);
    if (js.kind == 1) {
      vb = va;
      db = da;
    } else {
      lat_voxel(c.lat[js.og], js.o, &vb, &db);
    }
    w = (1.0 / double(S * S > 1 ? S * S : 1)) * double(1.0f);
  }
  const Share sh = strength_share(c.mats, c.nmats, va, da, vb, db, w);
  r.st[0] = sh.ft;
  r.st[1] = sh.fb;
  r.st[2] = sh.fc;
  r.st[3] = sh.coh;
  r.st[4] = sh.mu;
  r.w = w;
  c.elem[e] = r;
}

// Phase 2, a thread a bond: the general bond's sums over its elements in their order (the CPU's
// bits), or the one-lattice bond of an unplaced lattice without samples.
__global__ void __launch_bounds__(kWG) k_section_general(const SectionGenDev* __restrict__ cp) {
  const SectionGenDev& c = *cp;
  const int k = int(blockIdx.x) * kWG + int(threadIdx.x);
  if (k >= c.m) return;
  const svxg_section_head_general& hd = c.heads[k];
  const svxg_lattice& G = c.lat[hd.grid];
  if (hd.nsamples == 0 && G.identity) {
    const LatticeHead lh{hd.a, hd.b, hd.axis, hd.sign, hd.nfaces, hd.ma, hd.mb, hd.strength_a, hd.strength_b};
    c.out[k] = lattice_bond(lh, c.faces + hd.face0, G.h, c.nodes, G.vox, G.damage, G.lo, G.dim, c.mats, c.nmats, k);
    return;
  }
  const SecElem* E = c.elem + c.eoff[k];
  const int ne = int(c.eoff[k + 1] - c.eoff[k]);
  auto X = [&](int q) { return V3d{E[q].x[0], E[q].x[1], E[q].x[2]}; };
  auto N = [&](int q) { return V3d{E[q].nk[0], E[q].nk[1], E[q].nk[2]}; };
  const double h = G.h;
  double atot = 0.0;
  for (int q = 0; q < ne; ++q) atot += E[q].ar;
  const bool pair = hd.b >= 0;
  const V3d ca{c.nodes[hd.a].c[0], c.nodes[hd.a].c[1], c.nodes[hd.a].c[2]};
  V3d cb{0, 0, 0};
  if (pair) cb = V3d{c.nodes[hd.b].c[0], c.nodes[hd.b].c[1], c.nodes[hd.b].c[2]};
  // the interface normal: the area-weighted mean normal (a support of the lattice: its face axis)
  V3d n{0, 0, 0};
  if (pair || hd.nsamples) {
    for (int q = 0; q < ne; ++q) n = v_add(n, v_mul(N(q), E[q].ar));
    const double tiny = 1e-12 * atot * atot;
    if (v_dot(n, n) <= tiny && pair) n = v_sub(cb, ca);
    if (v_dot(n, n) <= tiny) n = V3d{0, 0, 1};
    n = v_unit(n);
  } else {
    V3d ax{0, 0, 0};
    v_at(ax, hd.axis) = double(hd.sign);
    n = lat_dir(G, ax);
  }
  V3d t1, t2;
  frame(n, &t1, &t2);
  double W = 0.0;
  V3d pc{0, 0, 0};
  for (int q = 0; q < ne; ++q) {
    const double wq = E[q].ar * fmax(0.05, fabs(v_dot(n, N(q))));
    W += wq;
    pc = v_add(pc, v_mul(X(q), wq));
  }
  pc = v_mul(pc, 1.0 / W);
  {
    double s11 = 0.0, s22 = 0.0, s12 = 0.0;
    for (int q = 0; q < ne; ++q) {
      const double wq = E[q].ar * fmax(0.05, fabs(v_dot(n, N(q))));
      const V3d d = v_sub(X(q), pc);
      const double y1 = v_dot(d, t1), y2 = v_dot(d, t2);
      s11 += wq * y1 * y1;
      s22 += wq * y2 * y2;
      s12 += wq * y1 * y2;
    }
    principal(s11, s22, s12, &t1, &t2);
  }
  svxg_bond B;
  memset(&B, 0, sizeof B);
  B.a = hd.a;
  B.b = hd.b;
  B.ma = hd.ma;
  B.mb = hd.mb;
  B.faces = ne;
  B.tag = k;
  put_frame(&B, n, t1, t2, pc);
  B.area = fmax(W, 0.25 * atot);
  double s1 = 0.0, s2 = 0.0, m1 = 0.0, m2 = 0.0, own = 0.0;
  for (int q = 0; q < ne; ++q) {
    const double wq = E[q].ar * fmax(0.05, fabs(v_dot(n, N(q))));
    const V3d d = v_sub(X(q), pc);
    const double y1 = v_dot(d, t1), y2 = v_dot(d, t2);
    const double sd = E[q].sd;
    s1 += wq * y1 * y1;
    s2 += wq * y2 * y2;
    m1 = fmax(m1, fabs(y1) + 0.5 * sd);
    m2 = fmax(m2, fabs(y2) + 0.5 * sd);
    own += wq * sd * sd;
  }
  own /= W;  // (the mean square side of the section's squares)
  B.s1 = s1 + B.area * own / 12.0;
  B.s2 = s2 + B.area * own / 12.0;
  B.c1 = m1;
  B.c2 = m2;
  B.rmax = sqrt(B.c1 * B.c1 + B.c2 * B.c2);
  springs(pair, ca, cb, n, pc, h, &B);
  B.strength = fmin(hd.strength_a, hd.strength_b);
  double ft = 0, fb = 0, fc = 0, coh = 0, mu = 0, wsum = 0;
  for (int q = 0; q < ne; ++q) {
    ft += E[q].st[0];
    fb += E[q].st[1];
    fc += E[q].st[2];
    coh += E[q].st[3];
    mu += E[q].st[4];
    wsum += E[q].w;
  }
  if (ne > 0 && wsum > 0.0) {
    const double w = 1.0 / wsum;
    B.sectioned = 1;
    B.ft = float(ft * w);
    B.fb = float(fb * w);
    B.fc = float(fc * w);
    B.coh = float(coh * w);
    B.mu = float(mu * w);
  }
  c.out[k] = B;
}

}  // namespace

void launch_bond_sections(hipStream_t s, const SectionDev* dev, int m) {
  if (m > 0) hipLaunchKernelGGL(k_bond_sections, dim3(unsigned((m + kWG - 1) / kWG)), dim3(kWG), 0, s, dev);
}

void launch_bond_sections_general(hipStream_t s, const SectionGenDev* dev, int m, int64_t elements) {
  if (elements > 0)
    hipLaunchKernelGGL(k_section_elements, dim3(unsigned((elements + kWG - 1) / kWG)), dim3(kWG), 0, s, dev);
  if (m > 0) hipLaunchKernelGGL(k_section_general, dim3(unsigned((m + kWG - 1) / kWG)), dim3(kWG), 0, s, dev);
}

}  // namespace svxg
