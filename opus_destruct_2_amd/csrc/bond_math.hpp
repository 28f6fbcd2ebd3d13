// Per-bond arithmetic shared by the bond checks and the judge: bond_load, bond_strength and
// bond_utilization (stress.cpp:14-57, 477-502), the DIF lerp (lerp_load, world_internal.hpp:452-461)
// and the plastic-hinge cut (hinge_of, world.cpp:1286-1334), each in the reference's operation
// order (the kernels are compiled with -ffp-contract=off, so the bits match the CPU's).
#pragma once

#include "svxg_kernels.hpp"

namespace svxg {

__device__ __forceinline__ const svxg_material& bond_mat(const svxg_material* m, int nm, int id) {
  return (id >= 0 && id < nm) ? m[id] : m[0];
}

// element (r, c) of rigid_block(r): [[I, -[r]x], [0, I]] (amg.cpp:143-153)
__device__ __forceinline__ double rigid_elem(const double r[3], int R, int C) {
  if (R == C) return 1.0;
  switch (R * 6 + C) {
    case 4: return r[2];
    case 5: return -r[1];
    case 9: return -r[2];
    case 11: return r[0];
    case 15: return r[1];
    case 16: return -r[0];
    default: return 0.0;
  }
}

// y = rigid_block(r) u, each row summed left to right as blk6::mv6 does
__device__ __forceinline__ void rigid_mv6(const double r[3], const double* u, double* y) {
  for (int R = 0; R < 6; ++R) {
    double s = rigid_elem(r, R, 0) * u[0];
    for (int c = 1; c < 6; ++c) s = s + rigid_elem(r, R, c) * u[c];
    y[R] = s;
  }
}

// StressProblem::bond_load (stress.cpp:477-502): the section's forces from the two sides' motion
__device__ __forceinline__ svxg_bond_load bond_load_dev(const svxg_bond& b, const svxg_node* nodes,
                                                       const svxg_material& MA, const svxg_material& MB,
                                                       const double* u) {
  double xa[6] = {0, 0, 0, 0, 0, 0}, xb[6] = {0, 0, 0, 0, 0, 0};
  const svxg_node& na = nodes[b.a];
  if (!na.fixed) {
    const double r[3] = {b.p[0] - na.c[0], b.p[1] - na.c[1], b.p[2] - na.c[2]};
    rigid_mv6(r, u + 6 * size_t(b.a), xa);
  }
  if (b.b >= 0) {
    const svxg_node& nb = nodes[b.b];
    if (!nb.fixed) {
      const double r[3] = {b.p[0] - nb.c[0], b.p[1] - nb.c[1], b.p[2] - nb.c[2]};
      rigid_mv6(r, u + 6 * size_t(b.b), xb);
    }
  }
  const double dt[3] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]};
  const double dr[3] = {xb[3] - xa[3], xb[4] - xa[4], xb[5] - xa[5]};
  auto dot3 = [](const double* p, const double* q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
  const double la = fmax(b.la, 1e-4), lb = fmax(b.lb, 1e-4);
  const double invE = la / MA.E + lb / MB.E;
  const double invG = la / MA.G + lb / MB.G;
  svxg_bond_load L;
  L.N = b.area / invE * dot3(dt, b.n);
  L.V1 = (5.0 / 6.0) * b.area / invG * dot3(dt, b.t1);
  L.V2 = (5.0 / 6.0) * b.area / invG * dot3(dt, b.t2);
  L.T = 0.8 * (b.s1 + b.s2) / invG * dot3(dr, b.n);
  L.M1 = b.s2 / invE * dot3(dr, b.t1);
  L.M2 = b.s1 / invE * dot3(dr, b.t2);
  return L;
}

// lerp_load (world_internal.hpp:452-461)
__device__ __forceinline__ svxg_bond_load lerp_load_dev(const svxg_bond_load& p, const svxg_bond_load& q, double k) {
  svxg_bond_load r;
  r.N = p.N + k * (q.N - p.N);
  r.V1 = p.V1 + k * (q.V1 - p.V1);
  r.V2 = p.V2 + k * (q.V2 - p.V2);
  r.T = p.T + k * (q.T - p.T);
  r.M1 = p.M1 + k * (q.M1 - p.M1);
  r.M2 = p.M2 + k * (q.M2 - p.M2);
  return r;
}

struct BondStrengthDev {
  double ft, fb, fc, coh, mu;
};

// bond_strength (stress.cpp:14-21)
__device__ __forceinline__ BondStrengthDev bond_strength_dev(const svxg_bond& b, const svxg_material& MA,
                                                            const svxg_material& MB, double fragility) {
  const double sc = b.strength / fmax(1e-6, fragility);
  BondStrengthDev S;
  if (b.sectioned) {
    S.ft = sc * double(b.ft);
    S.fb = sc * double(b.fb);
    S.fc = sc * double(b.fc);
    S.coh = sc * double(b.coh);
    S.mu = double(b.mu);
  } else {
    S.ft = sc * fmin(MA.ft, MB.ft);
    S.fb = sc * fmin(MA.fb, MB.fb);
    S.fc = sc * fmin(MA.fc, MB.fc);
    S.coh = sc * fmin(MA.cohesion, MB.cohesion);
    S.mu = fmin(MA.friction, MB.friction);
  }
  return S;
}

// bond_utilization (stress.cpp:23-57): phi, and the FailMode (0 None, 1 Tension, 2 Crush, 3 Shear)
__device__ __forceinline__ double bond_phi_dev(const svxg_bond& b, const BondStrengthDev& S, const svxg_bond_load& L,
                                               uint8_t* mode) {
  const double kWeak = 1e-9;
  const double ft = fmax(S.ft, kWeak), fb = fmax(S.fb, kWeak), fc = fmax(S.fc, kWeak);
  const double A = fmax(b.area, 1e-12);
  const double sN = L.N / A;
  const double sb = (b.s2 > 0 ? fabs(L.M1) * b.c2 / b.s2 : 0.0) + (b.s1 > 0 ? fabs(L.M2) * b.c1 / b.s1 : 0.0);
  const double phi_t = sN >= 0.0 ? sN / ft + sb / fb : fmax(0.0, sb + sN) / fb;
  const double phi_c = fmax(0.0, sb - sN) / fc;
  const double Jt = fmax(b.s1 + b.s2, 1e-18);
  const double tau = 1.5 * sqrt(L.V1 * L.V1 + L.V2 * L.V2) / A + fabs(L.T) * b.rmax / Jt;
  const double cap = fmax(S.coh + S.mu * fmax(0.0, -sN), kWeak);
  const double phi_s = tau / cap;
  double phi = phi_t;
  uint8_t md = 1;
  if (phi_c > phi) {
    phi = phi_c;
    md = 2;
  }
  if (phi_s > phi) {
    phi = phi_s;
    md = 3;
  }
  *mode = phi > 0 ? md : 0;
  return phi;
}

// hinge_of (world.cpp:1286-1334): a ductile section failing in bending turns about its compression
// edge. rot_a, rot_b: the two sides' rotations (b < 0: a support, which does not turn).
__device__ __forceinline__ bool hinge_dev(const svxg_bond& b, const svxg_bond_load& L, const BondStrengthDev& S0,
                                          const double rot_a[3], const double rot_b[3], double hinge_shape,
                                          svxg_hinge_cut* out) {
  auto dot3 = [](const double* p, const double* q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
  const double A = fmax(b.area, 1e-12);
  const double sN = L.N / A;
  const double w1 = b.s2 > 0 ? b.c2 / b.s2 : 0.0, w2 = b.s1 > 0 ? b.c1 / b.s1 : 0.0;
  const double sb = fabs(L.M1) * w1 + fabs(L.M2) * w2;
  const double fb = fmax(S0.fb, 1e-9), ft = fmax(S0.ft, 1e-9), fc = fmax(S0.fc, 1e-9);
  const double phi_b = sb / fb;
  const double phi_n = sN >= 0.0 ? sN / ft : -sN / fc;
  const double J = fmax(b.s1 + b.s2, 1e-18);
  const double tau = 1.5 * sqrt(L.V1 * L.V1 + L.V2 * L.V2) / A + fabs(L.T) * b.rmax / J;
  const double phi_s = tau / fmax(S0.coh + S0.mu * fmax(0.0, -sN), 1e-9);
  if (!(phi_b > 0.0) || phi_b < 1.5 * phi_n || phi_b < 1.5 * phi_s) return false;
  // the part beyond turns on from where the solve has it (b's rotation less a's), square to the bond
  double w[3];
  for (int k = 0; k < 3; ++k) w[k] = (b.b < 0 ? 0.0 : rot_b[k]) - rot_a[k];
  const double wn = dot3(w, b.n);
  for (int k = 0; k < 3; ++k) w[k] -= b.n[k] * wn;
  double m[3];
  for (int k = 0; k < 3; ++k) m[k] = b.t1[k] * L.M1 + b.t2[k] * L.M2;
  svxg_hinge_cut h;
  h.a = b.a;
  h.b = b.b;
  for (int k = 0; k < 3; ++k) {
    h.n[k] = b.n[k];
    h.p[k] = b.p[k];
  }
  if (dot3(w, w) > 1e-30) {
    const double nw = sqrt(dot3(w, w));
    for (int k = 0; k < 3; ++k) h.axis[k] = w[k] * (1.0 / nw);
    // the pivot: the section's compression edge, so the parts fold about it
    double y[3] = {h.axis[1] * b.n[2] - h.axis[2] * b.n[1], h.axis[2] * b.n[0] - h.axis[0] * b.n[2],
                   h.axis[0] * b.n[1] - h.axis[1] * b.n[0]};
    const double yl = sqrt(dot3(y, y));
    if (yl > 1e-9) {
      for (int k = 0; k < 3; ++k) y[k] *= 1.0 / yl;
      const double c = fabs(dot3(y, b.t1)) * b.c1 + fabs(dot3(y, b.t2)) * b.c2;
      for (int k = 0; k < 3; ++k) h.p[k] = b.p[k] - y[k] * (0.9 * c);
    }
  } else if (dot3(m, m) > 1e-30) {
    const double nm = sqrt(dot3(m, m));
    for (int k = 0; k < 3; ++k) h.axis[k] = m[k] * (1.0 / nm);
  } else {
    return false;
  }
  const double a1 = fabs(dot3(h.axis, b.t1)), a2 = fabs(dot3(h.axis, b.t2));
  h.mp = fmax(0.0, hinge_shape) * fb / fmax(a1 * w1 + a2 * w2, 1e-12);
  h.pull = ft * A;
  *out = h;
  return true;
}

}  // namespace svxg
