// Bond sections from voxel faces (SecAcc::finish_lattice, world_internal.hpp:129-249, and
// section_strengths, :374-400): a thread a bond walks its faces (the interface normal, the
// weighted centroid, the principal axes of the section, its moments and extreme fibres, then the
// face strengths), in the reference's operation order - this file is compiled with
// -ffp-contract=off, so a bond is the CPU's bits for the same faces. The faces of a bond are
// contiguous in the face list; a bond's work is a few passes over its own faces (latency-bound
// integer and FP64 work: a thread per bond, no matrix cores).
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

__device__ __forceinline__ void voxel(const SectionDev& c, int x, int y, int z, uint8_t* v, uint8_t* d) {
  const int i = x - c.lo[0], j = y - c.lo[1], k = z - c.lo[2];
  if (i < 0 || j < 0 || k < 0 || i >= c.dim[0] || j >= c.dim[1] || k >= c.dim[2]) {
    *v = 0;
    *d = 0;
    return;
  }
  const size_t at = size_t(i) + size_t(c.dim[0]) * (size_t(j) + size_t(c.dim[1]) * size_t(k));
  *v = c.vox[at];
  *d = c.damage ? c.damage[at] : 0;
}

__device__ __forceinline__ const svxg_material& mat_of_vox(const SectionDev& c, uint8_t v) {
  const int id = (v & 0x7F) - 1;  // (air: -1, taken as material 0)
  return c.mats[(id >= 0 && id < c.nmats) ? id : 0];
}

__global__ void __launch_bounds__(kWG) k_bond_sections(const SectionDev* __restrict__ cp) {
  const SectionDev& c = *cp;
  const int k = int(blockIdx.x) * kWG + int(threadIdx.x);
  if (k >= c.m) return;
  const svxg_section_head hd = c.heads[k];
  const svxg_section_face* F = c.faces + hd.face0;
  const int nf = hd.nfaces;
  const double h = c.h, A1 = h * h;
  const V3d ca{c.nodes[hd.a].c[0], c.nodes[hd.a].c[1], c.nodes[hd.a].c[2]};
  const bool pair = hd.b >= 0;
  V3d cb{0, 0, 0};
  if (pair) cb = V3d{c.nodes[hd.b].c[0], c.nodes[hd.b].c[1], c.nodes[hd.b].c[2]};
  // the interface normal
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
  // projected weights and centroid
  double W = 0.0;
  V3d pc{0, 0, 0};
  for (int f = 0; f < nf; ++f) {
    const V3d x = face_centre(F[f], h);
    const double w = A1 * fmax(0.05, fabs(v_get(n, F[f].axis)));
    W += w;
    pc = v_add(pc, v_mul(x, w));
  }
  pc = v_mul(pc, 1.0 / W);
  // the section's principal axes: the eigenvector of [[s11, s12], [s12, s22]] of the larger value
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
    if (fabs(s12) > 1e-9 * (s11 + s22)) {
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
        const V3d u1 = v_add(v_mul(t1, e1), v_mul(t2, e2));
        const V3d u2 = v_sub(v_mul(t2, e1), v_mul(t1, e2));
        t1 = u1;
        t2 = u2;
      }
    }
  }
  svxg_bond B;
  memset(&B, 0, sizeof B);
  B.a = hd.a;
  B.b = hd.b;
  B.ma = hd.ma;
  B.mb = hd.mb;
  B.faces = nf;
  B.tag = k;
  const V3d nn = n;
  for (int i = 0; i < 3; ++i) {
    B.n[i] = v_get(nn, i);
    B.t1[i] = v_get(t1, i);
    B.t2[i] = v_get(t2, i);
    B.p[i] = v_get(pc, i);
  }
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
  if (pair) {
    const V3d d = v_sub(cb, ca);
    const double L = fmax(fmax(fabs(v_dot(d, n)), 0.5 * sqrt(v_dot(d, d))), 0.5 * h);
    const double t = v_dot(v_sub(pc, ca), n);
    B.la = t < 0.1 * L ? 0.1 * L : (0.9 * L < t ? 0.9 * L : t);
    B.lb = L - B.la;
  } else {
    B.la = fmax(0.25 * h, v_dot(v_sub(pc, ca), n));
    B.lb = 0.5 * h;
  }
  B.strength = fmin(hd.strength_a, hd.strength_b);
  // the section strengths: each face's weaker material times the condition of its more damaged side
  double ft = 0, fb = 0, fc = 0, coh = 0, mu = 0;
  for (int f = 0; f < nf; ++f) {
    uint8_t va, da, vb, db;
    voxel(c, F[f].p[0], F[f].p[1], F[f].p[2], &va, &da);
    int q[3] = {F[f].p[0], F[f].p[1], F[f].p[2]};
    q[F[f].axis] += 1;
    voxel(c, q[0], q[1], q[2], &vb, &db);
    const svxg_material& MA = mat_of_vox(c, va != 0 ? va : vb);
    const svxg_material& MM = mat_of_vox(c, vb != 0 ? vb : va);
    const double cond = fmax(1e-6, 1.0 - double(da > db ? da : db) / 255.0);
    ft += cond * fmin(MA.ft, MM.ft);
    fb += cond * fmin(MA.fb, MM.fb);
    fc += cond * fmin(MA.fc, MM.fc);
    coh += cond * fmin(MA.cohesion, MM.cohesion);
    mu += fmin(MA.friction, MM.friction);
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
  c.out[k] = B;
}

}  // namespace

void launch_bond_sections(hipStream_t s, const SectionDev* dev, int m) {
  if (m > 0) hipLaunchKernelGGL(k_bond_sections, dim3(unsigned((m + kWG - 1) / kWG)), dim3(kWG), 0, s, dev);
}

}  // namespace svxg
