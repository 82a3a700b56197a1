// Device bodies of the bonded terms, exception pairs and positional
// restraints: the per-term functions and `bonded`, one workgroup's share of
// the terms. Included by the bonded kernels and by the launch that runs the
// direct space and the bonded terms together.
#pragma once

#include "kernels.h"

// Multiply-adds fuse only within one expression, so each body rounds the same
// way wherever it is inlined (the single-system and the batched kernels).
#pragma clang fp contract(on)

namespace rmd {

__device__ __forceinline__ void addForce(const Dev &d, int atom, double fx, double fy, double fz) {
  atomicAddFixed(&d.force[atom], __double2ll_rn(fx * FORCE_SCALE));
  atomicAddFixed(&d.force[d.nPad + atom], __double2ll_rn(fy * FORCE_SCALE));
  atomicAddFixed(&d.force[2 * d.nPad + atom], __double2ll_rn(fz * FORCE_SCALE));
}

__device__ __forceinline__ void loadPos(const Dev &d, int a, double &x, double &y, double &z) {
  const float4 p = d.pos[a];
  x = p.x;
  y = p.y;
  z = p.z;
}

// Harmonic bond, (1/2) k (r - r0)^2 (cpu/forces/bonded.js:22-46).
// Each term also adds sum_a F_a . (r_a - r_ref) over its atoms to `vir` (the
// negative of its scaling derivative; translation invariant since its forces
// sum to zero).
__device__ double bondTerm(const Dev &d, int t, double &vir) {
  const int2 ab = d.bondAtoms[t];
  const double2 kp = d.bondParams[t];
  double ax, ay, az, bx, by, bz;
  loadPos(d, ab.x, ax, ay, az);
  loadPos(d, ab.y, bx, by, bz);
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double stretch = r - (double)kp.y;
  const double scale = r > 0 ? (double)kp.x * stretch / r : 0.0;
  addForce(d, ab.x, -scale * dx, -scale * dy, -scale * dz);
  addForce(d, ab.y, scale * dx, scale * dy, scale * dz);
  vir -= scale * r * r;
  return 0.5 * (double)kp.x * stretch * stretch;
}

// Harmonic angle via atan2 with the clamped cross product
// (cpu/forces/bonded.js:58-114).
__device__ double angleTerm(const Dev &d, int t, double &vir) {
  const int4 abc = d.angleAtoms[t];
  const double2 kp = d.angleParams[t];
  double ax, ay, az, bx, by, bz, cx, cy, cz;
  loadPos(d, abc.x, ax, ay, az);
  loadPos(d, abc.y, bx, by, bz);
  loadPos(d, abc.z, cx, cy, cz);
  const double v1x = ax - bx, v1y = ay - by, v1z = az - bz;
  const double v2x = cx - bx, v2y = cy - by, v2z = cz - bz;
  const double n1 = v1x * v1x + v1y * v1y + v1z * v1z, n2 = v2x * v2x + v2y * v2y + v2z * v2z;
  const double nx = v1y * v2z - v1z * v2y, ny = v1z * v2x - v1x * v2z, nz = v1x * v2y - v1y * v2x;
  const double cross = sqrt(nx * nx + ny * ny + nz * nz);
  const double dot = v1x * v2x + v1y * v2y + v1z * v2z;
  const double theta = atan2(cross, dot);
  const double dev = theta - (double)kp.y;
  const double dE = (double)kp.x * dev;
  const double safe = cross > 1e-12 ? cross : 1e-12;
  const double s1 = -dE / (n1 * safe), s2 = dE / (n2 * safe);
  const double fax = s1 * (v1y * nz - v1z * ny), fay = s1 * (v1z * nx - v1x * nz), faz = s1 * (v1x * ny - v1y * nx);
  const double fcx = s2 * (v2y * nz - v2z * ny), fcy = s2 * (v2z * nx - v2x * nz), fcz = s2 * (v2x * ny - v2y * nx);
  addForce(d, abc.x, fax, fay, faz);
  addForce(d, abc.z, fcx, fcy, fcz);
  addForce(d, abc.y, -(fax + fcx), -(fay + fcy), -(faz + fcz));
  vir += fax * v1x + fay * v1y + faz * v1z + fcx * v2x + fcy * v2y + fcz * v2z;
  return 0.5 * (double)kp.x * dev * dev;
}

// Periodic torsion k (1 + cos(n phi - phase)) with the reference's middle-atom
// coefficients (cpu/forces/bonded.js:126-207).
__device__ double torsionTerm(const Dev &d, int t, double &vir) {
  const int4 q = d.torsionAtoms[t];
  const double4 kp = d.torsionParams[t];
  double ax, ay, az, bx, by, bz, cx, cy, cz, dx_, dy_, dz_;
  loadPos(d, q.x, ax, ay, az);
  loadPos(d, q.y, bx, by, bz);
  loadPos(d, q.z, cx, cy, cz);
  loadPos(d, q.w, dx_, dy_, dz_);
  const double b1x = bx - ax, b1y = by - ay, b1z = bz - az;
  const double b2x = cx - bx, b2y = cy - by, b2z = cz - bz;
  const double b3x = dx_ - cx, b3y = dy_ - cy, b3z = dz_ - cz;
  const double n1x = b1y * b2z - b1z * b2y, n1y = b1z * b2x - b1x * b2z, n1z = b1x * b2y - b1y * b2x;
  const double n2x = b2y * b3z - b2z * b3y, n2y = b2z * b3x - b2x * b3z, n2z = b2x * b3y - b2y * b3x;
  const double n1s = n1x * n1x + n1y * n1y + n1z * n1z, n2s = n2x * n2x + n2y * n2y + n2z * n2z;
  const double b2l = sqrt(b2x * b2x + b2y * b2y + b2z * b2z);
  if (n1s < 1e-24 || n2s < 1e-24 || b2l < 1e-12) return 0.0;
  const double dn = n1x * n2x + n1y * n2y + n1z * n2z;
  const double cr = (n1y * n2z - n1z * n2y) * b2x + (n1z * n2x - n1x * n2z) * b2y + (n1x * n2y - n1y * n2x) * b2z;
  const double phi = atan2(cr / b2l, dn);
  const double n = (double)kp.y;
  const double arg = n * phi - (double)kp.z;
  const double k = (double)kp.x;
  const double dE = -k * n * sin(arg);
  const double fA = dE * b2l / n1s, fD = -dE * b2l / n2s;
  const double fax = fA * n1x, fay = fA * n1y, faz = fA * n1z;
  const double fdx = fD * n2x, fdy = fD * n2y, fdz = fD * n2z;
  const double inv = 1.0 / (b2l * b2l);
  const double tt = (b1x * b2x + b1y * b2y + b1z * b2z) * inv;
  const double u = (b3x * b2x + b3y * b2y + b3z * b2z) * inv;
  const double fbx = -(1 + tt) * fax + u * fdx, fby = -(1 + tt) * fay + u * fdy, fbz = -(1 + tt) * faz + u * fdz;
  const double fcx = tt * fax - (1 + u) * fdx, fcy = tt * fay - (1 + u) * fdy, fcz = tt * faz - (1 + u) * fdz;
  addForce(d, q.x, fax, fay, faz);
  addForce(d, q.w, fdx, fdy, fdz);
  addForce(d, q.y, fbx, fby, fbz);
  addForce(d, q.z, fcx, fcy, fcz);
  // relative to atom a: b - a = b1, c - a = b1 + b2, d - a = b1 + b2 + b3
  vir += fbx * b1x + fby * b1y + fbz * b1z + fcx * (b1x + b2x) + fcy * (b1y + b2y) + fcz * (b1z + b2z) +
         fdx * (b1x + b2x + b3x) + fdy * (b1y + b2y + b3y) + fdz * (b1z + b2z + b3z);
  return k * (1.0 + cos(arg));
}

// Exception pair: scaled 1-4 LJ and Coulomb plus the Ewald erf correction
// with the atoms' full charges, no cutoff (webgpu/kernels/nonbonded.js:128-200).
// FP64: a solvated system has one such pair per atom, and a rigid water's
// corrections are the same large number (~150 kJ/mol) for every water, so an
// FP32 rounding bias would add up coherently across the box.
// One exception pair at the given positions: its energies and virial added
// in, and its force on atom i (the negative on j) returned.
__device__ __forceinline__ void exceptionPair(const BoxF &box, double alpha, double4 prm, double xi, double yi,
                                              double zi, double xj, double yj, double zj, double &elj,
                                              double &ecoul, double &vir, double &fx, double &fy, double &fz) {
  double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  if (box.periodic) {
    const double *bx = box.box64;
    const double s2 = floor(dz / bx[8] + 0.5);
    dx -= s2 * bx[6];
    dy -= s2 * bx[7];
    dz -= s2 * bx[8];
    const double s1 = floor(dy / bx[4] + 0.5);
    dx -= s1 * bx[3];
    dy -= s1 * bx[4];
    const double s0 = floor(dx / bx[0] + 0.5);
    dx -= s0 * bx[0];
  }
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double r = sqrt(r2), invR = 1.0 / r, invR2 = invR * invR, invR6 = invR2 * invR2 * invR2;
  double scalar = 0.0;
  if (prm.y != 0.0 || prm.z != 0.0) {
    elj += (prm.y * invR6 - prm.z) * invR6;
    scalar += (12.0 * prm.y * invR6 - 6.0 * prm.z) * invR6 * invR2;
  }
  if (prm.x != 0.0) {
    ecoul += prm.x * invR;
    scalar += prm.x * invR * invR2;
  }
  if (alpha > 0.0 && prm.w != 0.0) {
    const double ar = alpha * r;
    const double unscreened = erf(ar);
    const double gaussian = 1.1283791670955126 * alpha * exp(-ar * ar);
    ecoul -= prm.w * unscreened * invR;
    scalar -= prm.w * (unscreened * invR - gaussian) * invR2;
  }
  fx = scalar * dx;
  fy = scalar * dy;
  fz = scalar * dz;
  vir += scalar * r2;
}

__device__ void exceptionTerm(const Dev &d, const BoxF &box, double alpha, int t, double &elj, double &ecoul,
                              double &vir) {
  const int2 ij = d.excAtoms[t];
  const double4 prm = d.excParams[t];  // qq scaled, A, B, qq full (both with 1/4 pi eps0)
  double xi, yi, zi, xj, yj, zj;
  loadPos(d, ij.x, xi, yi, zi);
  loadPos(d, ij.y, xj, yj, zj);
  double fx, fy, fz;
  exceptionPair(box, alpha, prm, xi, yi, zi, xj, yj, zj, elj, ecoul, vir, fx, fy, fz);
  addForce(d, ij.x, fx, fy, fz);
  addForce(d, ij.y, -fx, -fy, -fz);
}

// The three exception pairs of a three-atom molecule (a rigid water), at
// 3k .. 3k + 2: the atoms' positions loaded once, and each pair's force
// rounded to fixed point as exceptionTerm rounds it, the three atoms' sums
// added once - nine atomics for the molecule instead of eighteen, the same
// forces bit for bit.
__device__ void exceptionTriplet(const Dev &d, const BoxF &box, double alpha, int k, double &elj, double &ecoul,
                                 double &vir) {
  int2 e[3];
  double4 prm[3];
  for (int m = 0; m < 3; ++m) {
    e[m] = d.excAtoms[3 * k + m];
    prm[m] = d.excParams[3 * k + m];
  }
  const int atom[3] = {e[0].x, e[0].y, (e[1].x != e[0].x && e[1].x != e[0].y) ? e[1].x : e[1].y};
  double px[3], py[3], pz[3];
  for (int q = 0; q < 3; ++q) loadPos(d, atom[q], px[q], py[q], pz[q]);
  long long f[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int m = 0; m < 3; ++m) {
    const int qi = e[m].x == atom[0] ? 0 : e[m].x == atom[1] ? 1 : 2;
    const int qj = e[m].y == atom[0] ? 0 : e[m].y == atom[1] ? 1 : 2;
    // selects, not indexing: the arrays stay in registers
    const double xi = qi == 0 ? px[0] : qi == 1 ? px[1] : px[2];
    const double yi = qi == 0 ? py[0] : qi == 1 ? py[1] : py[2];
    const double zi = qi == 0 ? pz[0] : qi == 1 ? pz[1] : pz[2];
    const double xj = qj == 0 ? px[0] : qj == 1 ? px[1] : px[2];
    const double yj = qj == 0 ? py[0] : qj == 1 ? py[1] : py[2];
    const double zj = qj == 0 ? pz[0] : qj == 1 ? pz[1] : pz[2];
    double fx, fy, fz;
    exceptionPair(box, alpha, prm[m], xi, yi, zi, xj, yj, zj, elj, ecoul, vir, fx, fy, fz);
    const long long g[3] = {__double2ll_rn(fx * FORCE_SCALE), __double2ll_rn(fy * FORCE_SCALE),
                            __double2ll_rn(fz * FORCE_SCALE)};
    for (int q = 0; q < 3; ++q)
      for (int c = 0; c < 3; ++c) f[q][c] += (qi == q ? g[c] : 0) - (qj == q ? g[c] : 0);
  }
  for (int q = 0; q < 3; ++q) {
    atomicAddFixed(&d.force[atom[q]], f[q][0]);
    atomicAddFixed(&d.force[d.nPad + atom[q]], f[q][1]);
    atomicAddFixed(&d.force[2 * d.nPad + atom[q]], f[q][2]);
  }
}

// Positional restraint k |minImage(r - r0)|^2, no half
// (shared/topology/restraints.js:98-126), in FP64 against an FP64 reference.
__device__ double restraintTerm(const Dev &d, const BoxF &box, int t) {
  const int a = d.restraintAtoms[t];
  const double4 ref = d.restraintRef[t];
  double px, py, pz;
  loadPos(d, a, px, py, pz);
  double dx = px - ref.x, dy = py - ref.y, dz = pz - ref.z;
  if (box.periodic) {
    const double *bx = box.box64;
    // JavaScript Math.round semantics (halves towards +infinity), c -> b -> a.
    double s2 = floor(dz / bx[8] + 0.5);
    dx -= s2 * bx[6];
    dy -= s2 * bx[7];
    dz -= s2 * bx[8];
    double s1 = floor(dy / bx[4] + 0.5);
    dx -= s1 * bx[3];
    dy -= s1 * bx[4];
    double s0 = floor(dx / bx[0] + 0.5);
    dx -= s0 * bx[0];
  }
  const double k = ref.w;
  addForce(d, a, -2 * k * dx, -2 * k * dy, -2 * k * dz);
  return k * (dx * dx + dy * dy + dz * dz);
}

enum { BT_BOND = 0, BT_ANGLE, BT_TORSION, BT_LJX, BT_COULX, BT_RESTRAINT, BT_DSCALE, BT_COUNT };

// One thread per exception triplet, then one per remaining exception.
__host__ __device__ inline int exceptionThreads(const Dev &d) { return d.nException - 2 * d.nExcTriplets; }

// Threads of the bonded launch: one per term.
__host__ __device__ inline int bondedThreads(const Dev &d) {
  return d.nBond + d.nAngle + d.nTorsion + exceptionThreads(d) + d.nRestraint;
}

__device__ __forceinline__ void bonded(const Dev &d, double alphaD, int terms, int energies, int block) {
  const BoxF &box = *d.box;
  const int t = block * blockDim.x + threadIdx.x;
  double e[BT_COUNT] = {0, 0, 0, 0, 0, 0, 0};
  double vir = 0.0;
  int base = 0;
  const bool doBonded = terms & RMD_TERM_BONDED;
  if (t < base + d.nBond) {
    if (doBonded) e[BT_BOND] = bondTerm(d, t - base, vir);
  } else if ((base += d.nBond, t < base + d.nAngle)) {
    if (doBonded) e[BT_ANGLE] = angleTerm(d, t - base, vir);
  } else if ((base += d.nAngle, t < base + d.nTorsion)) {
    if (doBonded) e[BT_TORSION] = torsionTerm(d, t - base, vir);
  } else if ((base += d.nTorsion, t < base + exceptionThreads(d))) {
    if (terms & RMD_TERM_EXCEPTIONS) {
      double elj = 0.0, ecoul = 0.0;
      const int u = t - base;
      if (u < d.nExcTriplets)
        exceptionTriplet(d, box, alphaD, u, elj, ecoul, vir);
      else
        exceptionTerm(d, box, alphaD, 2 * d.nExcTriplets + u, elj, ecoul, vir);
      e[BT_LJX] = elj;
      e[BT_COULX] = ecoul;
    }
  } else if ((base += exceptionThreads(d), t < base + d.nRestraint)) {
    if (terms & RMD_TERM_RESTRAINT) e[BT_RESTRAINT] = restraintTerm(d, box, t - base);
  }
  if (!energies) return;
  e[BT_DSCALE] = -vir;
  __shared__ double partial[BT_COUNT][8];  // up to 512 threads
  const int wave = threadIdx.x >> 6;
  for (int k = 0; k < BT_COUNT; ++k) {
    const double v = waveSum(e[k]);
    if ((threadIdx.x & 63) == 0) partial[k][wave] = v;
  }
  __syncthreads();
  // One slot per 256 threads, numbered as a 256-thread launch numbers its
  // workgroups, whatever this launch's workgroup size: the slot sums, and so
  // the energies, are bit for bit those of the 256-thread kernels.
  const int group = threadIdx.x >> 8, k = threadIdx.x & 255;
  if (k < BT_COUNT) {
    const int slot = block * (int)(blockDim.x >> 8) + group;
    const int waves = min(4, (int)(blockDim.x >> 6) - 4 * group);
    double s = 0;
    for (int w = 4 * group; w < 4 * group + waves; ++w) s += partial[k][w];
    const int term = k == BT_BOND ? E_BOND
                     : k == BT_ANGLE ? E_ANGLE
                     : k == BT_TORSION ? E_TORSION
                     : k == BT_LJX ? E_LJ_EXCEPTION
                     : k == BT_COULX ? E_COUL_EXCEPTION
                     : k == BT_RESTRAINT ? E_RESTRAINT
                                         : E_DSCALE_BONDED;
    if (slot < (bondedThreads(d) + 255) / 256) d.slots[term * MAX_SLOTS + slot] = s;
  }
}

}  // namespace rmd
