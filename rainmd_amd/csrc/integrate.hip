// The BAOAB (LangevinMiddle) integrator with constraints, fused per
// constraint unit, plus the kinetic-energy and energy reductions, the
// barostat's centroid scaling and the minimiser's helpers.
//
// Replaces encodeLangevinStep and its kernels (webgpu/kernels/integrate.js:
// kick 126-147, drift 158-173, shake 188-334, rattle 337-398, thermostat
// 406-437, kinetic 440-453, step counter 456-464) and the barostat kernels
// (webgpu/kernels/barostat.js:25-107). The step order is the reference's
// (integrate.js:523-557; cpu/integrate/langevin.js:30-114):
//   B R A S O R A S [forces] B R
// Every constraint cluster is independent, so one thread owns one unit - a
// free atom, a rigid water (SETTLE, closed form, FP64) or a general cluster
// (iterative SHAKE/RATTLE in FP64, its working state in the block's LDS) -
// and runs the whole B R A S O R A S sequence for it with no
// synchronisation. The host lays the units out so that every wave holds one
// kind. The closing B R of step n is folded into the start of step n+1's
// launch, so a step is this kernel plus the force kernels.
#include "kernels.h"
#include "search.h"

// Multiply-adds fuse only within one expression, so each body rounds the same
// way wherever it is inlined (the single-system and the batched kernels).
#pragma clang fp contract(on)

namespace rmd {

struct Atom {
  double x[3], v[3], f[3];
  double im;  // inverse mass (0 for massless)
  // What storeAtom needs besides the new state, loaded with the state so
  // that no load waits at the end of the unit's long dependency chain: the
  // charge, the atom's slot in the cell order and its rebuild position.
  float q;
  int slot;
  float rx, ry, rz;
};

__device__ __forceinline__ void loadStoreTail(const Dev &d, int a, Atom &s) {
  const float4 r = d.posRebuild[a];
  s.slot = d.slotOf[a];
  s.rx = r.x;
  s.ry = r.y;
  s.rz = r.z;
}

__device__ __forceinline__ void loadAtom(const Dev &d, int a, Atom &s) {
  const float4 p = d.pos[a], v = d.vel[a];
  s.x[0] = p.x;
  s.x[1] = p.y;
  s.x[2] = p.z;
  s.q = p.w;
  s.v[0] = v.x;
  s.v[1] = v.y;
  s.v[2] = v.z;
  s.im = v.w;
  s.f[0] = (double)d.force[a] * FORCE_UNSCALE;
  s.f[1] = (double)d.force[d.nPad + a] * FORCE_UNSCALE;
  s.f[2] = (double)d.force[2 * d.nPad + a] * FORCE_UNSCALE;
  loadStoreTail(d, a, s);
}

// Stores position and velocity, clears the force accumulators for the coming
// evaluation and returns |x - x_rebuild|^2 in f32.
__device__ __forceinline__ float storeAtom(const Dev &d, int a, const Atom &s, bool clearForce) {
  const float4 p = make_float4((float)s.x[0], (float)s.x[1], (float)s.x[2], s.q);
  d.pos[a] = p;
  d.slotPos[s.slot] = p;
  d.vel[a] = make_float4((float)s.v[0], (float)s.v[1], (float)s.v[2], (float)s.im);
  if (clearForce) {
    d.force[a] = 0;
    d.force[d.nPad + a] = 0;
    d.force[2 * d.nPad + a] = 0;
  }
  const float dx = p.x - s.rx, dy = p.y - s.ry, dz = p.z - s.rz;
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ void kick(Atom &s, double h) {
  if (s.im <= 0) return;
  for (int c = 0; c < 3; ++c) s.v[c] += h * s.im * s.f[c];
}

__device__ __forceinline__ double kinetic(const Atom &s) {
  if (s.im <= 0) return 0.0;
  return 0.5 / s.im * (s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
}

// shared/random.js:24-87, in integers (bit-identical to the host) and FP64.
__device__ __forceinline__ unsigned int mix(unsigned int v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}
// ((h >> 9) + 0.5) * 2^-23: 24 significant bits at most, so exact in FP32.
__device__ __forceinline__ float uniformAt(unsigned int seed, unsigned int a, unsigned int b, unsigned int c) {
  unsigned int st = mix(seed ^ (a * 0x9e3779b1u));
  st = mix(st ^ (b * 0x85ebca6bu));
  st = mix(st ^ (c * 0xc2b2ae35u));
  return ((float)(st >> 9) + 0.5f) * (1.0f / 8388608.0f);
}

// sin(2 pi u) and cos(2 pi u) for a uniform u in (0, 1): the quadrant
// q = rint(4u), r = u - q/4 in [-1/8, 1/8] (exact in FP32), then
// sin(2 pi r) = r S(r^2) and cos(2 pi r) = C(r^2), minimax polynomials of
// degree 4 in r^2 made by sincos_fit.py (8.5e-8 largest absolute error in
// FP32 over the counter hash's uniforms), rotated by q: 18 instructions where
// a general sincospi takes 43.
__device__ __forceinline__ void sinCosTwoPi(float u, float &s, float &c) {
  const float q = rintf(4.0f * u);
  const float r = fmaf(q, -0.25f, u);
  const float r2 = r * r;
  float ps = 4.141468069e+01f;
  ps = fmaf(ps, r2, -7.669582170e+01f);
  ps = fmaf(ps, r2, 8.160518075e+01f);
  ps = fmaf(ps, r2, -4.134170205e+01f);
  ps = fmaf(ps, r2, 6.283185307e+00f);
  float pc = 5.924596419e+01f;
  pc = fmaf(pc, r2, -8.544374173e+01f);
  pc = fmaf(pc, r2, 6.493932647e+01f);
  pc = fmaf(pc, r2, -1.973920869e+01f);
  pc = fmaf(pc, r2, 1.000000000e+00f);
  const float sr = r * ps;
  const int k = (int)q & 3;
  const float sa = (k & 1) ? pc : sr, ca = (k & 1) ? sr : pc;
  s = (k & 2) ? -sa : sa;
  c = ((k + 1) & 2) ? -ca : ca;
}

// gaussianTriple (shared/random.js:66-87): three normals from the four
// counter-hash uniforms of (seed, atom, step) by Box-Muller, in FP32. The
// square roots are the hardware's (one ulp) and the angle's sine and cosine
// are sinCosTwoPi's; the normals stay within 5e-7 of FP64 Box-Muller on the
// same uniforms.
__device__ __forceinline__ void gaussianTriple(unsigned int seed, unsigned int atom, unsigned int step, float u[4],
                                               float g[3]) {
  for (int k = 0; k < 4; ++k) u[k] = uniformAt(seed, atom, step, k);
  const float r1 = __builtin_amdgcn_sqrtf(-2.0f * logf(u[0])), r2 = __builtin_amdgcn_sqrtf(-2.0f * logf(u[2]));
  float s1, c1, s2, c2;
  sinCosTwoPi(u[1], s1, c1);
  sinCosTwoPi(u[3], s2, c2);
  g[0] = r1 * c1;
  g[1] = r1 * s1;
  g[2] = r2 * c2;
}

// O: v <- a v + sqrt(1 - a^2) sqrt(kT/m) xi (integrate.js:406-437). The
// normals are drawn in FP32 - the noise is a random draw, and FP64 is kept
// for the velocity update itself.
// The draw depends only on (seed, atom, step), so the units take it right
// after issuing their state loads: its arithmetic runs while those loads are
// in flight instead of in the middle of the unit's dependency chain.
__device__ __forceinline__ void drawNoise(const Dev &d, const IntegratorParams &ip, int a, unsigned int step,
                                          float g[3]) {
  if (ip.suppliedNoise) {
    for (int c = 0; c < 3; ++c) g[c] = d.noise[3 * a + c];
  } else {
    float u[4];
    gaussianTriple(ip.seed, a, step, u, g);
  }
}

__device__ __forceinline__ void thermostat(const IntegratorParams &ip, Atom &s, const float g[3]) {
  if (s.im <= 0) return;
  const double amp = ip.noiseScale * sqrt(ip.kT * s.im);
  for (int c = 0; c < 3; ++c) s.v[c] = ip.decay * s.v[c] + amp * (double)g[c];
}

// ---- rigid water ----------------------------------------------------------

// SETTLE (Miyamoto & Kollman, J. Comput. Chem. 13, 952 (1992)): the rigid
// water's new positions in closed form, with the constraint displacements
// along the old bond vectors - the limit SHAKE converges to
// (cpu/integrate/constraints.js:93-171).
__device__ void settle(const double o[9], double n[9], double mO, double mH, double dOH, double dHH) {
  const double wohh = mO + 2.0 * mH;
  const double rc = 0.5 * dHH;
  const double hgt = sqrt(dOH * dOH - rc * rc);
  const double ra = 2.0 * mH * hgt / wohh, rb = hgt - ra;
  double b0[3], c0[3], com[3], a1[3], b1[3], c1[3];
  for (int k = 0; k < 3; ++k) {
    b0[k] = o[3 + k] - o[k];
    c0[k] = o[6 + k] - o[k];
    com[k] = (mO * n[k] + mH * (n[3 + k] + n[6 + k])) / wohh;
    a1[k] = n[k] - com[k];
    b1[k] = n[3 + k] - com[k];
    c1[k] = n[6 + k] - com[k];
  }
  double ez[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2], b0[0] * c0[1] - b0[1] * c0[0]};
  double ex[3] = {a1[1] * ez[2] - a1[2] * ez[1], a1[2] * ez[0] - a1[0] * ez[2], a1[0] * ez[1] - a1[1] * ez[0]};
  double ey[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
  const double lx = rsqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
  const double ly = rsqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
  const double lz = rsqrt(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
  for (int k = 0; k < 3; ++k) {
    ex[k] *= lx;
    ey[k] *= ly;
    ez[k] *= lz;
  }
  auto dot = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  const double xb0 = dot(b0, ex), yb0 = dot(b0, ey), xc0 = dot(c0, ex), yc0 = dot(c0, ey);
  const double za1 = dot(a1, ez);
  const double xb1 = dot(b1, ex), yb1 = dot(b1, ey), zb1 = dot(b1, ez);
  const double xc1 = dot(c1, ex), yc1 = dot(c1, ey), zc1 = dot(c1, ez);
  const double sinphi = za1 / ra;
  const double cosphi = sqrt(fmax(0.0, 1.0 - sinphi * sinphi));
  const double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi);
  const double cospsi = sqrt(fmax(0.0, 1.0 - sinpsi * sinpsi));
  const double ya2 = ra * cosphi, xb2 = -rc * cospsi;
  const double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
  const double yb2 = t1 - t2, yc2 = t1 + t2;
  const double alpa = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
  const double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
  const double gama = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
  const double al2be2 = alpa * alpa + beta * beta;
  const double sinthe = (alpa * gama - beta * sqrt(fmax(0.0, al2be2 - gama * gama))) / al2be2;
  const double costhe = sqrt(fmax(0.0, 1.0 - sinthe * sinthe));
  const double a3[3] = {-ya2 * sinthe, ya2 * costhe, za1};
  const double b3[3] = {xb2 * costhe - yb2 * sinthe, xb2 * sinthe + yb2 * costhe, zb1};
  const double c3[3] = {-xb2 * costhe - yc2 * sinthe, -xb2 * sinthe + yc2 * costhe, zc1};
  for (int k = 0; k < 3; ++k) {
    n[k] = com[k] + a3[0] * ex[k] + a3[1] * ey[k] + a3[2] * ez[k];
    n[3 + k] = com[k] + b3[0] * ex[k] + b3[1] * ey[k] + b3[2] * ez[k];
    n[6 + k] = com[k] + c3[0] * ex[k] + c3[1] * ey[k] + c3[2] * ez[k];
  }
}

// RATTLE for a rigid water, exactly: the three bond multipliers that remove
// every relative velocity along a bond solve a 3x3 system.
__device__ void rattleWater(Atom *w) {
  double r01[3], r02[3], r12[3], v01[3], v02[3], v12[3];
  for (int c = 0; c < 3; ++c) {
    r01[c] = w[0].x[c] - w[1].x[c];
    r02[c] = w[0].x[c] - w[2].x[c];
    r12[c] = w[1].x[c] - w[2].x[c];
    v01[c] = w[0].v[c] - w[1].v[c];
    v02[c] = w[0].v[c] - w[2].v[c];
    v12[c] = w[1].v[c] - w[2].v[c];
  }
  auto dot = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  const double iO = w[0].im, i1 = w[1].im, i2 = w[2].im;
  // dv0 = (l0 r01 + l1 r02) iO; dv1 = (-l0 r01 + l2 r12) i1; dv2 = (-l1 r02 - l2 r12) i2
  const double A[3][3] = {
      {(iO + i1) * dot(r01, r01), iO * dot(r01, r02), -i1 * dot(r01, r12)},
      {iO * dot(r02, r01), (iO + i2) * dot(r02, r02), i2 * dot(r02, r12)},
      {-i1 * dot(r12, r01), i2 * dot(r12, r02), (i1 + i2) * dot(r12, r12)}};
  const double bvec[3] = {-dot(r01, v01), -dot(r02, v02), -dot(r12, v12)};
  // Cramer's rule with each determinant written out (no array indexed in a
  // loop, so everything stays in registers).
  auto det3 = [](double a00, double a01, double a02, double a10, double a11, double a12, double a20, double a21,
                 double a22) {
    return a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
  };
  const double det = det3(A[0][0], A[0][1], A[0][2], A[1][0], A[1][1], A[1][2], A[2][0], A[2][1], A[2][2]);
  if (!(fabs(det) > 0)) return;
  const double inv = 1.0 / det;
  const double l[3] = {
      det3(bvec[0], A[0][1], A[0][2], bvec[1], A[1][1], A[1][2], bvec[2], A[2][1], A[2][2]) * inv,
      det3(A[0][0], bvec[0], A[0][2], A[1][0], bvec[1], A[1][2], A[2][0], bvec[2], A[2][2]) * inv,
      det3(A[0][0], A[0][1], bvec[0], A[1][0], A[1][1], bvec[1], A[2][0], A[2][1], bvec[2]) * inv};
  for (int c = 0; c < 3; ++c) {
    w[0].v[c] += (l[0] * r01[c] + l[1] * r02[c]) * iO;
    w[1].v[c] += (-l[0] * r01[c] + l[2] * r12[c]) * i1;
    w[2].v[c] += (-l[1] * r02[c] - l[2] * r12[c]) * i2;
  }
}

// A: half drift, SETTLE, then the velocity takes the constraint displacement
// (v += correction / h, integrate.js:237-253).
__device__ void driftWater(Atom *w, double h, const double4 geom) {
  double o[9], n[9], u[9];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      o[3 * a + c] = w[a].x[c];
      n[3 * a + c] = w[a].x[c] + h * w[a].v[c];
      u[3 * a + c] = n[3 * a + c];
    }
  settle(o, n, geom.z, geom.w, geom.x, geom.y);
  const double ih = 1.0 / h;
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      w[a].x[c] = n[3 * a + c];
      w[a].v[c] += (n[3 * a + c] - u[3 * a + c]) * ih;
    }
}

// ---- star clusters (registers, FP64) ----------------------------------------
// The same SHAKE/RATTLE sweeps as a general cluster, with the whole cluster
// in registers: slot 0 is the centre, constraint k joins slot k + 1 to it.

// Every star function takes SM, the widest star the kernel handles (3 or 4
// constraints; the host picks the smallest that covers the system's stars).
// Constraint slots past a star's own are the identity with a zero right-hand
// side, so a star of m <= 3 gets the same multipliers, bit for bit, in
// either width; the narrower one only does less arithmetic.

// 1/x from the hardware FP64 reciprocal refined by one Newton step (about one
// ulp): the sweeps below divide on their serial dependency chain, where a
// correctly rounded division is several times longer.
__device__ __forceinline__ double reciprocal(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return fma(fma(-x, y, 1.0), y, y);
}

// Every slot is loaded (slots past the cluster repeat its last atom) and the
// unused ones are then given zero inverse mass, velocity and force: no branch
// around the loads, so the compiler keeps the whole cluster in registers.
// t2 and invT2: each constraint's squared length and its reciprocal.
template <int SM>
__device__ void starLoad(const Dev &d, int cluster, int m, Atom *w, int *atom, double *t2, double *invT2) {
  const int ab = d.clusterAtomStart[cluster], cb = d.clusterConStart[cluster];
#pragma unroll
  for (int k = 0; k < SM + 1; ++k) {
    atom[k] = d.clusterAtoms[ab + (k <= m ? k : m)];
    loadAtom(d, atom[k], w[k]);
    const double use = k <= m ? 1.0 : 0.0;
    w[k].im *= use;
    for (int c = 0; c < 3; ++c) {
      w[k].v[c] *= use;
      w[k].f[c] *= use;
    }
  }
#pragma unroll
  for (int k = 0; k < SM; ++k) {
    const double target = d.clusterConD[cb + (k < m ? k : m - 1)];
    t2[k] = target * target;
    invT2[k] = reciprocal(t2[k]);
  }
}

// x = A^-1 b for a star's SM x SM system by Gaussian elimination
// without pivoting: A is the constraints' coupling matrix (or, for SHAKE, its
// Newton Jacobian, close to it), im_0 times a Gram matrix of a few bonds at
// tetrahedral angles plus a positive diagonal, so its pivots stay large.
// Rows and columns past the cluster's constraints are the identity.
template <int SM>
__device__ __forceinline__ void solveStar(double A[SM][SM], double b[SM], double x[SM]) {
#pragma unroll
  for (int p = 0; p < SM; ++p) {
    const double inv = reciprocal(A[p][p]);
#pragma unroll
    for (int i = p + 1; i < SM; ++i) {
      const double f = A[i][p] * inv;
#pragma unroll
      for (int j = p + 1; j < SM; ++j) A[i][j] -= f * A[p][j];
      b[i] -= f * b[p];
    }
  }
#pragma unroll
  for (int i = SM - 1; i >= 0; --i) {
    double t = b[i];
#pragma unroll
    for (int j = i + 1; j < SM; ++j) t -= A[i][j] * x[j];
    x[i] = t * reciprocal(A[i][i]);
  }
}

// R: the velocity constraints r_k . (v_k - v_0) = 0 of a star are linear in
// the multipliers, so they are solved at once (the limit of RATTLE's sweeps,
// cpu/integrate/constraints.js:173-222): v_k += lam_k im_k r_k and
// v_0 -= im_0 sum_j lam_j r_j with
// (im_0 r_k . r_j + [k = j] im_k |r_k|^2) lam_j = -r_k . (v_k - v_0).
template <int SM>
__device__ void starRattle(Atom *w, int m) {
  double r[SM][3], A[SM][SM], b[SM], lam[SM];
#pragma unroll
  for (int k = 0; k < SM; ++k)
    for (int c = 0; c < 3; ++c) r[k][c] = w[k + 1].x[c] - w[0].x[c];
  const double i0 = w[0].im;
#pragma unroll
  for (int k = 0; k < SM; ++k) {
    const bool live = k < m;
    b[k] = 0.0;
    if (live)
      for (int c = 0; c < 3; ++c) b[k] -= r[k][c] * (w[k + 1].v[c] - w[0].v[c]);
#pragma unroll
    for (int j = 0; j < SM; ++j) {
      double a = 0.0;
      if (live && j < m) {
        const double dot = r[k][0] * r[j][0] + r[k][1] * r[j][1] + r[k][2] * r[j][2];
        a = i0 * dot + (k == j ? w[k + 1].im * dot : 0.0);
      }
      A[k][j] = live ? a : (k == j ? 1.0 : 0.0);
    }
  }
  solveStar<SM>(A, b, lam);
#pragma unroll
  for (int k = 0; k < SM; ++k) {
    if (k >= m) continue;
    for (int c = 0; c < 3; ++c) {
      w[k + 1].v[c] += lam[k] * w[k + 1].im * r[k][c];
      w[0].v[c] -= lam[k] * i0 * r[k][c];
    }
  }
}

// A with SHAKE: half drift, then the position constraints with the
// displacements along the bond vectors before the drift
// (cpu/integrate/constraints.js:93-171): x_k += lam_k im_k r_k,
// x_0 -= im_0 sum_j lam_j r_j, |x_k - x_0|^2 = d_k^2. The sweeps' fixed point
// is found by Newton's method on the multipliers (each iteration a 3 x 3
// solve; quadratic convergence where the sweeps converge linearly), to the
// same relative tolerance on |x_k - x_0|^2. The velocity takes the
// constraint displacement.
template <int SM>
__device__ void starDrift(const Dev &d, const IntegratorParams &ip, Atom *w, int m, const double *t2,
                          const double *invT2, double h) {
  double r[SM][3], s[SM][3];
  // the bond vectors before the drift, and after it
#pragma unroll
  for (int k = 0; k < SM; ++k)
    for (int c = 0; c < 3; ++c) r[k][c] = w[k + 1].x[c] - w[0].x[c];
#pragma unroll
  for (int k = 0; k < SM + 1; ++k)
    for (int c = 0; c < 3; ++c) w[k].x[c] += h * w[k].v[c];
#pragma unroll
  for (int k = 0; k < SM; ++k)
    for (int c = 0; c < 3; ++c) s[k][c] = w[k + 1].x[c] - w[0].x[c];
  const double i0 = w[0].im;
  double lam[SM];
#pragma unroll
  for (int k = 0; k < SM; ++k) lam[k] = 0.0;
  int iterations = 0;
  bool converged = false, stalled = false;
  double previous = 1e300;
  for (int it = 0; it < ip.maxIterations; ++it) {
    double u[SM][3], g[SM], A[SM][SM];
    double worst = 0;
#pragma unroll
    for (int k = 0; k < SM; ++k)
      for (int c = 0; c < 3; ++c) {
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < SM; ++j) sum += (j < m ? lam[j] : 0.0) * r[j][c];
        u[k][c] = s[k][c] + lam[k] * w[k + 1].im * r[k][c] + i0 * sum;
      }
#pragma unroll
    for (int k = 0; k < SM; ++k) {
      const bool live = k < m;
      const double err = live ? u[k][0] * u[k][0] + u[k][1] * u[k][1] + u[k][2] * u[k][2] - t2[k] : 0.0;
      worst = fmax(worst, fabs(err) * (live ? invT2[k] : 0.0));
      g[k] = -err;
#pragma unroll
      for (int j = 0; j < SM; ++j) {
        double a = 0.0;
        if (live && j < m) {
          const double dot = u[k][0] * r[j][0] + u[k][1] * r[j][1] + u[k][2] * r[j][2];
          a = 2.0 * (i0 * dot + (k == j ? w[k + 1].im * dot : 0.0));
        }
        A[k][j] = live ? a : (k == j ? 1.0 : 0.0);
      }
    }
    iterations = it + 1;
    if (worst <= ip.tolerance) {
      converged = true;
      break;
    }
    if (it >= 2 && worst >= previous * 0.99) {
      stalled = true;
      break;
    }
    previous = worst;
    double step[SM];
    solveStar<SM>(A, g, step);
    for (int k = 0; k < SM; ++k) lam[k] += step[k];
  }
  const double ih = 1.0 / h;
#pragma unroll
  for (int k = 0; k < SM; ++k) {
    if (k >= m) continue;
    for (int c = 0; c < 3; ++c) {
      const double dk = lam[k] * w[k + 1].im * r[k][c], d0 = lam[k] * i0 * r[k][c];
      w[k + 1].x[c] += dk;
      w[0].x[c] -= d0;
      w[k + 1].v[c] += dk * ih;
      w[0].v[c] -= d0 * ih;
    }
  }
  atomicMax(&d.sc->shakeMaxIterations, (unsigned int)iterations);
  if (!converged && !stalled) atomicMax(&d.sc->shakeFailed, 1u);
}

// ---- general clusters (iterative, FP64) -------------------------------------
// Working state per cluster atom slot (CLUSTER_SLOT doubles): x[3], ref[3],
// v[3], corr[3], inverse mass. S points at the cluster's first slot, in LDS
// or (clusters too large for LDS) in global scratch.

__device__ __forceinline__ double *clusterState(const Dev &d, const int4 desc, double *lds) {
  const int ab = d.clusterAtomStart[desc.y];
  return desc.z ? d.clusterScratch + (size_t)CLUSTER_SLOT * ab : lds + CLUSTER_SLOT * desc.w;
}

__device__ void clusterLoad(const Dev &d, int cluster, double *S) {
  const int ab = d.clusterAtomStart[cluster], ae = d.clusterAtomStart[cluster + 1];
  for (int s = ab; s < ae; ++s) {
    const int a = d.clusterAtoms[s];
    const float4 p = d.pos[a], v = d.vel[a];
    double *x = S + CLUSTER_SLOT * (s - ab);
    x[0] = p.x;
    x[1] = p.y;
    x[2] = p.z;
    x[6] = v.x;
    x[7] = v.y;
    x[8] = v.z;
    x[12] = v.w;
  }
}

__device__ void clusterShake(const Dev &d, const IntegratorParams &ip, int cluster, double *S, double h) {
  const int cb = d.clusterConStart[cluster], ce = d.clusterConStart[cluster + 1];
  const int ab = d.clusterAtomStart[cluster], ae = d.clusterAtomStart[cluster + 1];
  for (int s = 0; s < ae - ab; ++s) {
    double *x = S + CLUSTER_SLOT * s;
    for (int c = 0; c < 3; ++c) {
      x[3 + c] = x[c];
      x[c] += h * x[6 + c];
      x[9 + c] = 0.0;
    }
  }
  int iterations = 0;
  bool converged = false, stalled = false;
  double previous = 1e300;
  for (int it = 0; it < ip.maxIterations; ++it) {
    double worst = 0;
    for (int k = cb; k < ce; ++k) {
      const int4 con = d.clusterCons[k];
      const double target = d.clusterConD[k];
      double *pa = S + CLUSTER_SLOT * (con.x - ab), *pb = S + CLUSTER_SLOT * (con.y - ab);
      const double ia = pa[12], ib = pb[12];
      const double dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
      const double err = dx * dx + dy * dy + dz * dz - target * target;
      const double rel = fabs(err) / (target * target);
      worst = fmax(worst, rel);
      if (rel <= ip.tolerance) continue;
      const double rx = pa[3] - pb[3], ry = pa[4] - pb[4], rz = pa[5] - pb[5];
      const double den = 2.0 * (ia + ib) * (dx * rx + dy * ry + dz * rz);
      if (fabs(den) < 1e-20) continue;
      const double lam = -err / den;
      const double r[3] = {rx, ry, rz};
      for (int c = 0; c < 3; ++c) {
        pa[c] += lam * ia * r[c];
        pb[c] -= lam * ib * r[c];
        pa[9 + c] += lam * ia * r[c];
        pb[9 + c] -= lam * ib * r[c];
      }
    }
    iterations = it + 1;
    if (worst <= ip.tolerance) {
      converged = true;
      break;
    }
    if (it >= 2 && worst >= previous * 0.99) {
      stalled = true;
      break;
    }
    previous = worst;
  }
  const double ih = 1.0 / h;
  for (int s = 0; s < ae - ab; ++s) {
    double *x = S + CLUSTER_SLOT * s;
    for (int c = 0; c < 3; ++c) x[6 + c] += x[9 + c] * ih;
  }
  atomicMax(&d.sc->shakeMaxIterations, (unsigned int)iterations);
  if (!converged && !stalled) atomicMax(&d.sc->shakeFailed, 1u);
}

__device__ void clusterRattle(const Dev &d, const IntegratorParams &ip, int cluster, double *S) {
  const int cb = d.clusterConStart[cluster], ce = d.clusterConStart[cluster + 1];
  const int ab = d.clusterAtomStart[cluster];
  bool converged = false, stalled = false;
  double previous = 1e300;
  for (int it = 0; it < ip.maxIterations; ++it) {
    double worst = 0;
    for (int k = cb; k < ce; ++k) {
      const int4 con = d.clusterCons[k];
      const double target = d.clusterConD[k];
      double *pa = S + CLUSTER_SLOT * (con.x - ab), *pb = S + CLUSTER_SLOT * (con.y - ab);
      const double ia = pa[12], ib = pb[12];
      const double r[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
      const double vr[3] = {pa[6] - pb[6], pa[7] - pb[7], pa[8] - pb[8]};
      const double proj = r[0] * vr[0] + r[1] * vr[1] + r[2] * vr[2];
      const double rel = fabs(proj) / (target * target);
      worst = fmax(worst, rel);
      if (rel <= ip.tolerance) continue;
      const double den = (ia + ib) * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      if (den < 1e-20) continue;
      const double lam = -proj / den;
      for (int c = 0; c < 3; ++c) {
        pa[6 + c] += lam * ia * r[c];
        pb[6 + c] -= lam * ib * r[c];
      }
    }
    if (worst <= ip.tolerance) {
      converged = true;
      break;
    }
    if (it >= 2 && worst >= previous * 0.99) {
      stalled = true;
      break;
    }
    previous = worst;
  }
  if (!converged && !stalled) atomicMax(&d.sc->rattleFailed, 1u);
}

__device__ void clusterKick(const Dev &d, int cluster, double *S, double h) {
  const int ab = d.clusterAtomStart[cluster], ae = d.clusterAtomStart[cluster + 1];
  for (int s = ab; s < ae; ++s) {
    double *x = S + CLUSTER_SLOT * (s - ab);
    const double im = x[12];
    if (im <= 0) continue;
    const int a = d.clusterAtoms[s];
    const double f[3] = {(double)d.force[a] * FORCE_UNSCALE, (double)d.force[d.nPad + a] * FORCE_UNSCALE,
                         (double)d.force[2 * d.nPad + a] * FORCE_UNSCALE};
    for (int c = 0; c < 3; ++c) x[6 + c] += h * im * f[c];
  }
}

// ---- the kernels ------------------------------------------------------------

// Block reductions: FP64 sum into a slot, float max into an atomic word.
__device__ void blockSumToSlot(double v, double *slot) {
  __shared__ double part[4];
  v = waveSum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *slot = part[0] + part[1] + part[2] + part[3];
  __syncthreads();
}

__device__ void blockMaxToWord(float v, unsigned int *word) {
  __shared__ float part[4];
  // NaN must win: compare as unsigned bits (non-negative floats order like
  // their bit patterns and every NaN sorts above every finite value).
  unsigned int bits = __float_as_uint(v);
  for (int offset = 32; offset > 0; offset >>= 1) {
    const unsigned int other = __shfl_xor(bits, offset, 64);
    bits = other > bits ? other : bits;
  }
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __uint_as_float(bits);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int m = 0;
    for (int w = 0; w < 4; ++w) m = __float_as_uint(part[w]) > m ? __float_as_uint(part[w]) : m;
    atomicMax(word, m);
  }
  __syncthreads();
}

// One step's opening half for a water unit.
__device__ __forceinline__ void openWater(const Dev &d, const IntegratorParams &ip, int u, int4 desc, int closing,
                                          unsigned int step, double h, double &ke, float &disp) {
  Atom w[3];
  const int atoms[3] = {desc.y, desc.z, desc.w};
  for (int a = 0; a < 3; ++a) loadAtom(d, atoms[a], w[a]);
  const double4 geom = d.waterGeom[u];
  float g[3][3];
  for (int a = 0; a < 3; ++a) drawNoise(d, ip, atoms[a], step, g[a]);
  if (closing) {
    for (int a = 0; a < 3; ++a) kick(w[a], h);
    rattleWater(w);
  }
  for (int a = 0; a < 3; ++a) kick(w[a], h);
  rattleWater(w);
  driftWater(w, h, geom);
  for (int a = 0; a < 3; ++a) thermostat(ip, w[a], g[a]);
  rattleWater(w);
  driftWater(w, h, geom);
  for (int a = 0; a < 3; ++a) {
    ke += kinetic(w[a]);
    disp = fmaxf(disp, storeAtom(d, atoms[a], w[a], true));
  }
}

// The water blocks (the unit table's leading blocks hold water units and
// padding only) run the water path alone, in a kernel of their own, so that
// its register budget is the water path's and not that of the star and
// cluster paths. Their displacement maximum and kinetic slots join those of
// the general blocks' kernel, which runs after them on the same stream and
// decides the rebuild.
__device__ __forceinline__ void integrateWater(const Dev &d, const IntegratorParams &ip, int closing, int early,
                                               int block) {
  const int u = block * blockDim.x + threadIdx.x;
  const double h = ip.halfStep;
  const unsigned int step = (unsigned int)d.sc->step;
  double ke = 0.0;
  float disp = 0.f;
  if (u < d.nUnits) {
    const int4 desc = d.units[u];
    if (desc.x == UNIT_WATER) openWater(d, ip, u, desc, closing, step, h, ke, disp);
  }
  blockMaxToWord(disp, &d.sc->maxDisp2Bits);
  if (early) blockSumToSlot(ke, &d.slots[E_KINETIC_EARLY * MAX_SLOTS + block]);
}

// (Two waves per SIMD: the water path fits 256 registers.)
__global__ void __launch_bounds__(INTEGRATE_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
    integrateWaterKernel(Dev d, IntegratorParams ip, int closing, int early) {
  integrateWater(d, ip, closing, early, blockIdx.x);
}

__global__ void __launch_bounds__(INTEGRATE_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
    integrateWaterBatchKernel(const BatchView *__restrict__ v, const int2 *__restrict__ task, int closing, int early) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  integrateWater(v[t.x].d, v[t.x].ip, closing, early, t.y);
}

// One step's opening half for the units of workgroup `block`, one of the
// `blocks` general workgroups (whose last one decides the rebuild).
template <int SM>
__device__ __forceinline__ void integrateOpen(const Dev &d, const IntegratorParams &ip, int closing, int early,
                                              int block, int blocks, double *clusterLds) {
  const int u = block * blockDim.x + threadIdx.x;
  const double h = ip.halfStep;
  const unsigned int step = (unsigned int)d.sc->step;
  double ke = 0.0;
  float disp = 0.f;
  if (u < d.nUnits) {
    const int4 desc = d.units[u];
    if (desc.x == UNIT_FREE) {
      Atom s;
      loadAtom(d, desc.y, s);
      float g[3];
      drawNoise(d, ip, desc.y, step, g);
      if (closing) kick(s, h);
      kick(s, h);
      for (int c = 0; c < 3; ++c) s.x[c] += h * s.v[c];
      thermostat(ip, s, g);
      for (int c = 0; c < 3; ++c) s.x[c] += h * s.v[c];
      ke = kinetic(s);
      disp = storeAtom(d, desc.y, s, true);
    } else if (desc.x == UNIT_WATER) {
      openWater(d, ip, u, desc, closing, step, h, ke, disp);
    } else if (desc.x == UNIT_STAR) {
      const int m = desc.z;
      Atom w[SM + 1];
      int atom[SM + 1];
      double t2[SM], invT2[SM];
      starLoad<SM>(d, desc.y, m, w, atom, t2, invT2);
      float g[SM + 1][3];
#pragma unroll
      for (int k = 0; k < SM + 1; ++k) drawNoise(d, ip, atom[k], step, g[k]);
      if (closing) {
#pragma unroll
        for (int k = 0; k < SM + 1; ++k) kick(w[k], h);
        starRattle<SM>(w, m);
      }
#pragma unroll
      for (int k = 0; k < SM + 1; ++k) kick(w[k], h);
      starRattle<SM>(w, m);
      starDrift<SM>(d, ip, w, m, t2, invT2, h);
#pragma unroll
      for (int k = 0; k < SM + 1; ++k)
        if (k <= m) thermostat(ip, w[k], g[k]);
      starRattle<SM>(w, m);
      starDrift<SM>(d, ip, w, m, t2, invT2, h);
#pragma unroll
      for (int k = 0; k < SM + 1; ++k)
        if (k <= m) {
          ke += kinetic(w[k]);
          disp = fmaxf(disp, storeAtom(d, atom[k], w[k], true));
        }
    } else if (desc.x == UNIT_CLUSTER) {
      const int cl = desc.y;
      const int ab = d.clusterAtomStart[cl], ae = d.clusterAtomStart[cl + 1];
      double *S = clusterState(d, desc, clusterLds);
      clusterLoad(d, cl, S);
      if (closing) {
        clusterKick(d, cl, S, h);
        clusterRattle(d, ip, cl, S);
      }
      clusterKick(d, cl, S, h);
      clusterRattle(d, ip, cl, S);
      clusterShake(d, ip, cl, S, h);
      for (int s = ab; s < ae; ++s) {
        double *x = S + CLUSTER_SLOT * (s - ab);
        Atom t;
        t.im = x[12];
        for (int c = 0; c < 3; ++c) t.v[c] = x[6 + c];
        float g[3];
        drawNoise(d, ip, d.clusterAtoms[s], step, g);
        thermostat(ip, t, g);
        for (int c = 0; c < 3; ++c) x[6 + c] = t.v[c];
      }
      clusterRattle(d, ip, cl, S);
      clusterShake(d, ip, cl, S, h);
      for (int s = ab; s < ae; ++s) {
        const double *x = S + CLUSTER_SLOT * (s - ab);
        Atom t;
        t.im = x[12];
        for (int c = 0; c < 3; ++c) {
          t.x[c] = x[c];
          t.v[c] = x[6 + c];
        }
        const int a = d.clusterAtoms[s];
        t.q = d.pos[a].w;
        loadStoreTail(d, a, t);
        ke += kinetic(t);
        disp = fmaxf(disp, storeAtom(d, a, t, true));
      }
    }
  }
  blockMaxToWord(disp, &d.sc->maxDisp2Bits);
  if (early) blockSumToSlot(ke, &d.slots[E_KINETIC_EARLY * MAX_SLOTS + block]);
  // Last block: the displacement maximum is final; decide the rebuild and
  // advance the noise counter (integrate.js:456-464).
  __shared__ bool last, build;
  __threadfence();
  if (threadIdx.x == 0) last = atomicAdd(&d.sc->ticket, 1u) == (unsigned int)blocks - 1;
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const unsigned int bits = atomicAdd(&d.sc->maxDisp2Bits, 0u);
    const float maxDisp2 = __uint_as_float(bits);
    const bool rebuild = d.sc->forceRebuild || displacementExceeded(d.sc, maxDisp2, ip.dispHalf);
    d.sc->rebuild = rebuild ? 1u : 0u;
    if (rebuild) {
      d.sc->maxDisp2Bits = 0u;
      d.sc->slack = 0.f;
      retireRequirement(d.sc);
      d.sc->rebuilds += 1u;
      d.sc->forceRebuild = 0u;
    }
    d.sc->ticket = 0u;
    d.sc->step += 1;
    build = rebuild && ip.cellsInIntegrator;
  }
  // The cell list of a rebuild, by the last workgroup (every workgroup's
  // positions were released before its ticket; each thread acquires them).
  if (ip.cellsInIntegrator && last) {
    __syncthreads();
    if (build) {
      __threadfence();
      cellsBuildFew<INTEGRATE_BLOCK>(d);
    }
  }
}

// The units of blocks firstBlock.. (0: every unit; the water blocks' count:
// the general units after the water kernel).
template <int SM>
__global__ void __launch_bounds__(INTEGRATE_BLOCK)
    integrateOpenKernel(Dev d, IntegratorParams ip, int closing, int early, int firstBlock) {
  extern __shared__ double clusterLds[];
  integrateOpen<SM>(d, ip, closing, early, firstBlock + blockIdx.x, gridDim.x, clusterLds);
}

template <int SM>
__global__ void __launch_bounds__(INTEGRATE_BLOCK)
    integrateOpenBatchKernel(const BatchView *__restrict__ v, const int2 *__restrict__ task, int closing, int early) {
  extern __shared__ double clusterLds[];
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  IntegratorParams ip = v[t.x].ip;
  ip.cellsInIntegrator = 0;  // a batch builds its replicas' cell lists in a launch of its own
  integrateOpen<SM>(v[t.x].d, ip, closing, early, v[t.x].d.waterBlocks + t.y, v[t.x].openBlocks, clusterLds);
}

// The closing kick + RATTLE of a block's last step.
template <int SM>
__device__ __forceinline__ void integrateClose(const Dev &d, const IntegratorParams &ip, int withKinetic, int block,
                                               double *clusterLds) {
  const int u = block * blockDim.x + threadIdx.x;
  const double h = ip.halfStep;
  double ke = 0.0;
  if (u < d.nUnits) {
    const int4 desc = d.units[u];
    if (desc.x == UNIT_FREE) {
      Atom s;
      loadAtom(d, desc.y, s);
      kick(s, h);
      ke = kinetic(s);
      d.vel[desc.y] = make_float4((float)s.v[0], (float)s.v[1], (float)s.v[2], (float)s.im);
    } else if (desc.x == UNIT_WATER) {
      Atom w[3];
      const int atoms[3] = {desc.y, desc.z, desc.w};
      for (int a = 0; a < 3; ++a) loadAtom(d, atoms[a], w[a]);
      for (int a = 0; a < 3; ++a) kick(w[a], h);
      rattleWater(w);
      for (int a = 0; a < 3; ++a) {
        ke += kinetic(w[a]);
        d.vel[atoms[a]] = make_float4((float)w[a].v[0], (float)w[a].v[1], (float)w[a].v[2], (float)w[a].im);
      }
    } else if (desc.x == UNIT_STAR) {
      const int m = desc.z;
      Atom w[SM + 1];
      int atom[SM + 1];
      double t2[SM], invT2[SM];
      starLoad<SM>(d, desc.y, m, w, atom, t2, invT2);
#pragma unroll
      for (int k = 0; k < SM + 1; ++k) kick(w[k], h);
      starRattle<SM>(w, m);
#pragma unroll
      for (int k = 0; k < SM + 1; ++k)
        if (k <= m) {
          ke += kinetic(w[k]);
          d.vel[atom[k]] = make_float4((float)w[k].v[0], (float)w[k].v[1], (float)w[k].v[2], (float)w[k].im);
        }
    } else if (desc.x == UNIT_CLUSTER) {
      const int cl = desc.y;
      const int ab = d.clusterAtomStart[cl], ae = d.clusterAtomStart[cl + 1];
      double *S = clusterState(d, desc, clusterLds);
      clusterLoad(d, cl, S);
      clusterKick(d, cl, S, h);
      clusterRattle(d, ip, cl, S);
      for (int s = ab; s < ae; ++s) {
        const double *x = S + CLUSTER_SLOT * (s - ab);
        Atom t;
        t.im = x[12];
        for (int c = 0; c < 3; ++c) t.v[c] = x[6 + c];
        ke += kinetic(t);
        d.vel[d.clusterAtoms[s]] = make_float4((float)t.v[0], (float)t.v[1], (float)t.v[2], (float)t.im);
      }
    }
  }
  if (withKinetic) blockSumToSlot(ke, &d.slots[E_KINETIC * MAX_SLOTS + block]);
}

template <int SM>
__global__ void __launch_bounds__(INTEGRATE_BLOCK) integrateCloseKernel(Dev d, IntegratorParams ip, int withKinetic) {
  extern __shared__ double clusterLds[];
  integrateClose<SM>(d, ip, withKinetic, blockIdx.x, clusterLds);
}

template <int SM>
__global__ void __launch_bounds__(INTEGRATE_BLOCK)
    integrateCloseBatchKernel(const BatchView *__restrict__ v, const int2 *__restrict__ task) {
  extern __shared__ double clusterLds[];
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  integrateClose<SM>(v[t.x].d, v[t.x].ip, 0, t.y, clusterLds);
}

__global__ void __launch_bounds__(256) kineticKernel(Dev d, int term) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0.0;
  if (a < d.n) {
    const float4 v = d.vel[a];
    if (v.w > 0.f) ke = 0.5 / (double)v.w * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
  }
  blockSumToSlot(ke, &d.slots[term * MAX_SLOTS + blockIdx.x]);
}

// Sum each term's slot row in a fixed order: one block per term.
__device__ __forceinline__ void reduceEnergies(const double *slots, int term, int count, double *out) {
  double s = 0.0;
  for (int k = threadIdx.x; k < count; k += blockDim.x) s += slots[term * MAX_SLOTS + k];
  __shared__ double part[256];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) part[threadIdx.x] += part[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[term] = part[0];
}

__global__ void __launch_bounds__(256) reduceEnergiesKernel(const double *slots, SlotCounts counts, double *out) {
  reduceEnergies(slots, blockIdx.x, counts.c[blockIdx.x], out);
}

// Batched: block (term, replica). The kinetic rows are summed only at a
// block's last step (withEarly), as the single-system block does.
__global__ void __launch_bounds__(256) reduceEnergiesBatchKernel(const BatchView *__restrict__ v, int withEarly) {
  const int term = blockIdx.x % E_TERM_COUNT, r = blockIdx.x / E_TERM_COUNT;
  const int count = term == E_KINETIC ? 0 : term == E_KINETIC_EARLY ? (withEarly ? v[r].openSlots : 0)
                                                                    : v[r].counts.c[term];
  reduceEnergies(v[r].d.slots, term, count, v[r].terms);
}

__device__ __forceinline__ void zeroForces(const Dev &d, int a) {
  if (a >= d.nPad) return;
  d.force[a] = 0;
  d.force[d.nPad + a] = 0;
  d.force[2 * d.nPad + a] = 0;
}

__global__ void __launch_bounds__(256) zeroForcesBatchKernel(const BatchView *__restrict__ v,
                                                             const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  zeroForces(v[t.x].d, t.y * blockDim.x + threadIdx.x);
}

// The rebuild decision of an evaluation outside the step loop, one thread per
// replica (the single-system form lives beside the context).
__global__ void decideRebuildBatchKernel(const BatchView *__restrict__ v, int replicas, int force) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= replicas) return;
  Scalars *sc = v[r].d.sc;
  const float maxDisp2 = __uint_as_float(sc->maxDisp2Bits);
  const bool rebuild = force || v[r].rebuild || sc->forceRebuild || displacementExceeded(sc, maxDisp2, v[r].ip.dispHalf);
  sc->rebuild = rebuild ? 1u : 0u;
  if (rebuild) {
    sc->maxDisp2Bits = 0u;
    sc->slack = 0.f;
    retireRequirement(sc);
    sc->rebuilds += 1u;
    sc->forceRebuild = 0u;
  }
}

// Barostat (webgpu/kernels/barostat.js:25-107): mass-weighted molecule
// centroids (geometric when massless), then x += (s - 1) * centroid.
//
// One thread per molecule; a large molecule's atoms are loaded MOL_RUN at a
// time (their loads in flight together, not one dependent round trip per
// atom) and summed in atom order. (The kernels' duration is the largest
// molecule's chain: the 22-atom peptide of the 2854-atom system.)
constexpr int MOL_RUN = 8;

// Molecule [begin, end)'s mass-weighted centroid in FP64, or its atoms'
// mean where every mass is 0.
__device__ __forceinline__ void moleculeCentroid(const Dev &d, const int *molAtoms, int begin, int end, double w[3]) {
  double total = 0, u[3] = {0, 0, 0};
  w[0] = w[1] = w[2] = 0.0;
  for (int k0 = begin; k0 < end; k0 += MOL_RUN) {
    float4 p[MOL_RUN];
    float im[MOL_RUN];
#pragma unroll
    for (int q = 0; q < MOL_RUN; ++q) {
      const int a = molAtoms[min(k0 + q, end - 1)];
      p[q] = d.pos[a];
      im[q] = d.vel[a].w;
    }
#pragma unroll
    for (int q = 0; q < MOL_RUN; ++q) {
      if (k0 + q >= end) break;
      const double mass = im[q] > 0 ? 1.0 / (double)im[q] : 0.0;
      w[0] += mass * p[q].x;
      w[1] += mass * p[q].y;
      w[2] += mass * p[q].z;
      total += mass;
      u[0] += p[q].x;
      u[1] += p[q].y;
      u[2] += p[q].z;
    }
  }
  if (total > 0) {
    for (int c = 0; c < 3; ++c) w[c] /= total;
  } else {
    const int count = end - begin;
    for (int c = 0; c < 3; ++c) w[c] = u[c] / (count > 0 ? count : 1);
  }
}

// Molecule m's centroid into centroids[m]; returns the molecule's span (its
// atoms' largest distance from the centroid; 0 past the last molecule), which
// the caller's workgroup reduces into Scalars::spanBits with one atomic.
__device__ __forceinline__ float centroid(const Dev &d, const int *molStart, const int *molAtoms, int nMol,
                                          float4 *centroids, int m) {
  if (m >= nMol) return 0.f;
  const int begin = molStart[m], end = molStart[m + 1];
  double w[3];
  moleculeCentroid(d, molAtoms, begin, end, w);
  centroids[m] = make_float4((float)w[0], (float)w[1], (float)w[2], 0.f);
  double span2 = 0.0;
  for (int k0 = begin; k0 < end; k0 += MOL_RUN) {
    float4 p[MOL_RUN];
#pragma unroll
    for (int q = 0; q < MOL_RUN; ++q) p[q] = d.pos[molAtoms[min(k0 + q, end - 1)]];
#pragma unroll
    for (int q = 0; q < MOL_RUN; ++q) {
      const double ex = p[q].x - w[0], ey = p[q].y - w[1], ez = p[q].z - w[2];
      span2 = fmax(span2, ex * ex + ey * ey + ez * ez);
    }
  }
  return (float)sqrt(span2) * 1.0001f;
}

// sum_i F_i . (r_i - c_mol(i)) over molecule m's atoms with the total forces
// (its share of the molecular-scaling correction, E_DSCALE_MOLECULAR), about
// the same mass-weighted centroid as the barostat's, in FP64.
__device__ __forceinline__ double molecularTerm(const Dev &d, const int *molStart, const int *molAtoms, int nMol,
                                                int m) {
  if (m >= nMol) return 0.0;
  const int begin = molStart[m], end = molStart[m + 1];
  double w[3];
  moleculeCentroid(d, molAtoms, begin, end, w);
  double s = 0.0;
  for (int k0 = begin; k0 < end; k0 += MOL_RUN) {
    float4 p[MOL_RUN];
    long long f[MOL_RUN][3];
#pragma unroll
    for (int q = 0; q < MOL_RUN; ++q) {
      const int a = molAtoms[min(k0 + q, end - 1)];
      p[q] = d.pos[a];
      for (int c = 0; c < 3; ++c) f[q][c] = d.force[(size_t)c * d.nPad + a];
    }
#pragma unroll
    for (int q = 0; q < MOL_RUN; ++q) {
      if (k0 + q >= end) break;
      s += (double)f[q][0] * FORCE_UNSCALE * (p[q].x - w[0]) + (double)f[q][1] * FORCE_UNSCALE * (p[q].y - w[1]) +
           (double)f[q][2] * FORCE_UNSCALE * (p[q].z - w[2]);
    }
  }
  return s;
}

__global__ void __launch_bounds__(256) molecularTermKernel(Dev d, const int *molStart, const int *molAtoms, int nMol) {
  const double s = molecularTerm(d, molStart, molAtoms, nMol, blockIdx.x * blockDim.x + threadIdx.x);
  blockSumToSlot(s, &d.slots[E_DSCALE_MOLECULAR * MAX_SLOTS + blockIdx.x]);
}

__global__ void __launch_bounds__(256) molecularTermBatchKernel(const BatchView *__restrict__ v,
                                                                const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  const BatchView &b = v[t.x];
  const double s = molecularTerm(b.d, b.molStart, b.molAtoms, b.nMol, t.y * blockDim.x + threadIdx.x);
  blockSumToSlot(s, &b.d.slots[E_DSCALE_MOLECULAR * MAX_SLOTS + t.y]);
}

__global__ void __launch_bounds__(256) centroidKernel(Dev d, const int *molStart, const int *molAtoms, int nMol,
                                                      float4 *centroids) {
  const float span = centroid(d, molStart, molAtoms, nMol, centroids, blockIdx.x * blockDim.x + threadIdx.x);
  blockMaxToWord(span, &d.sc->spanBits);
}

__device__ __forceinline__ void scaleAtom(const Dev &d, const float4 *centroids, const int *molOfAtom, float factor,
                                          int a) {
  if (a == 0) {
    // A pair's separation in the rebuild frame differs from the scaled
    // frame's by |s - 1| times the molecules' centroid separation, which is
    // at most rl + skin + 2 * span for a pair the list must hold.
    const BoxF &b = *d.box;
    const float rl = sqrtf(b.rl2), skin = rl - b.cutoff;
    const float span = __uint_as_float(d.sc->spanBits);
    d.sc->slack += fabsf(factor - 1.0f) * (rl + skin + 2.0f * span) * 1.01f;
  }
  if (a >= d.n) return;
  const float4 c = centroids[molOfAtom[a]];
  float4 p = d.pos[a];
  p.x += (factor - 1.0f) * c.x;
  p.y += (factor - 1.0f) * c.y;
  p.z += (factor - 1.0f) * c.z;
  d.pos[a] = p;
  float4 r = d.posRebuild[a];
  r.x += (factor - 1.0f) * c.x;
  r.y += (factor - 1.0f) * c.y;
  r.z += (factor - 1.0f) * c.z;
  d.posRebuild[a] = r;
}

__global__ void scaleKernel(Dev d, const float4 *centroids, const int *molOfAtom, float factor) {
  scaleAtom(d, centroids, molOfAtom, factor, blockIdx.x * blockDim.x + threadIdx.x);
}

// Batched barostat scaling: replicas whose scale is 1 are left alone.
__global__ void __launch_bounds__(256) centroidBatchKernel(const BatchView *__restrict__ v,
                                                           const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0 || v[t.x].scale == 1.0f) return;
  const BatchView &b = v[t.x];
  const float span = centroid(b.d, b.molStart, b.molAtoms, b.nMol, b.centroids, t.y * blockDim.x + threadIdx.x);
  blockMaxToWord(span, &b.d.sc->spanBits);
}

__global__ void __launch_bounds__(256) scaleBatchKernel(const BatchView *__restrict__ v,
                                                        const int2 *__restrict__ task) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0 || v[t.x].scale == 1.0f) return;
  const BatchView &b = v[t.x];
  scaleAtom(b.d, b.centroids, b.molOfAtom, b.scale, t.y * blockDim.x + threadIdx.x);
}

// Batched save or restore of positions and forces (and velocities) around a
// block or a barostat trial; see SNAP_*.
__global__ void __launch_bounds__(256) snapshotBatchKernel(const BatchView *__restrict__ v,
                                                           const int2 *__restrict__ task, int mode) {
  const int2 t = task[blockIdx.x];
  if (t.x < 0) return;
  const BatchView &b = v[t.x];
  if ((mode & SNAP_FLAGGED) && !b.restore) return;
  const int a = t.y * blockDim.x + threadIdx.x;
  const int np = b.d.nPad;
  if (a >= np) return;
  if (mode & SNAP_RESTORE) {
    b.d.pos[a] = b.posSaved[a];
    if (mode & SNAP_VELOCITIES) b.d.vel[a] = b.velSaved[a];
    if (mode & SNAP_REBUILD) b.d.posRebuild[a] = b.rebuildSaved[a];
    for (int c = 0; c < 3; ++c) b.d.force[c * np + a] = b.forceSaved[c * np + a];
  } else {
    b.posSaved[a] = b.d.pos[a];
    if (mode & SNAP_VELOCITIES) b.velSaved[a] = b.d.vel[a];
    if (mode & SNAP_REBUILD) b.rebuildSaved[a] = b.d.posRebuild[a];
    for (int c = 0; c < 3; ++c) b.forceSaved[c * np + a] = b.d.force[c * np + a];
  }
}

__global__ void __launch_bounds__(256) snapshotKernel(Dev d, SnapshotBuffers saved, int mode) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int np = d.nPad;
  if (saved.rebuilds && !(mode & SNAP_RESTORE) && a == 0) *saved.rebuilds = d.sc->rebuilds;
  if ((mode & SNAP_SCALARS) && a == 0) {
    if (mode & SNAP_RESTORE) {
      *d.sc = *saved.sc;
    } else {
      d.sc->requiredMax = 0u;
      *saved.sc = *d.sc;
    }
  }
  if (a >= np) return;
  if (mode & SNAP_RESTORE) {
    d.pos[a] = saved.pos[a];
    if (mode & SNAP_VELOCITIES) d.vel[a] = saved.vel[a];
    if (mode & SNAP_REBUILD) d.posRebuild[a] = saved.rebuild[a];
    for (int c = 0; c < 3; ++c) d.force[c * np + a] = saved.force[c * np + a];
  } else {
    saved.pos[a] = d.pos[a];
    if (mode & SNAP_VELOCITIES) saved.vel[a] = d.vel[a];
    if (mode & SNAP_REBUILD) saved.rebuild[a] = d.posRebuild[a];
    for (int c = 0; c < 3; ++c) saved.force[c * np + a] = d.force[c * np + a];
  }
}

void launchSnapshot(const Dev &d, const SnapshotBuffers &saved, int mode, hipStream_t s) {
  snapshotKernel<<<(d.nPad + 255) / 256, 256, 0, s>>>(d, saved, mode);
}

__global__ void forceRebuildFlaggedBatchKernel(const BatchView *__restrict__ v, int replicas) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < replicas && v[r].restore && v[r].rebuild) v[r].d.sc->forceRebuild = 1u;
}

__global__ void clearRequiredBatchKernel(Scalars *sc, int replicas) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < replicas) sc[r].requiredMax = 0u;
}

// Minimiser: the largest per-atom force magnitude, as float bits.
__global__ void maxForceKernel(Dev d, unsigned int *out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.f;
  if (a < d.n) {
    const double fx = (double)d.force[a] * FORCE_UNSCALE, fy = (double)d.force[d.nPad + a] * FORCE_UNSCALE,
                 fz = (double)d.force[2 * d.nPad + a] * FORCE_UNSCALE;
    m = (float)sqrt(fx * fx + fy * fy + fz * fz);
  }
  blockMaxToWord(m, out);
}

// x += clamp(scale * F, cap) per atom (webgpu/kernels/integrate.js:642-665).
__global__ void displaceKernel(Dev d, float scale, float cap) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.n) return;
  const float fx = (float)((double)d.force[a] * FORCE_UNSCALE), fy = (float)((double)d.force[d.nPad + a] * FORCE_UNSCALE),
              fz = (float)((double)d.force[2 * d.nPad + a] * FORCE_UNSCALE);
  float sx = scale * fx, sy = scale * fy, sz = scale * fz;
  const float dist = sqrtf(sx * sx + sy * sy + sz * sz);
  if (dist > cap) {
    const float k = cap / dist;
    sx *= k;
    sy *= k;
    sz *= k;
  }
  float4 p = d.pos[a];
  p.x += sx;
  p.y += sy;
  p.z += sz;
  d.pos[a] = p;
}

// Constraint projection of positions against saved references (the
// minimiser's SHAKE, integrate.js:668-688): the displacement since the
// reference is treated as a drift of unit length.
__global__ void constrainPositionsKernel(Dev d, const float4 *reference, IntegratorParams ip) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= d.nUnits) return;
  const int4 desc = d.units[u];
  if (desc.x == UNIT_FREE) return;
  if (desc.x == UNIT_WATER) {
    const int atoms[3] = {desc.y, desc.z, desc.w};
    double o[9], n[9];
    for (int a = 0; a < 3; ++a) {
      const float4 r = reference[atoms[a]], p = d.pos[atoms[a]];
      o[3 * a] = r.x;
      o[3 * a + 1] = r.y;
      o[3 * a + 2] = r.z;
      n[3 * a] = p.x;
      n[3 * a + 1] = p.y;
      n[3 * a + 2] = p.z;
    }
    const double4 geom = d.waterGeom[u];
    settle(o, n, geom.z, geom.w, geom.x, geom.y);
    for (int a = 0; a < 3; ++a) {
      float4 p = d.pos[atoms[a]];
      p.x = (float)n[3 * a];
      p.y = (float)n[3 * a + 1];
      p.z = (float)n[3 * a + 2];
      d.pos[atoms[a]] = p;
    }
    return;
  }
  if (desc.x != UNIT_CLUSTER && desc.x != UNIT_STAR) return;
  const int cl = desc.y;
  const int ab = d.clusterAtomStart[cl], ae = d.clusterAtomStart[cl + 1];
  double *S = d.clusterScratch + (size_t)CLUSTER_SLOT * ab;
  for (int s = ab; s < ae; ++s) {
    const int a = d.clusterAtoms[s];
    const float4 r = reference[a], p = d.pos[a];
    double *x = S + CLUSTER_SLOT * (s - ab);
    // drift of length one from the reference to the trial position
    x[0] = r.x;
    x[1] = r.y;
    x[2] = r.z;
    x[6] = (double)p.x - r.x;
    x[7] = (double)p.y - r.y;
    x[8] = (double)p.z - r.z;
    x[12] = d.vel[a].w;
  }
  clusterShake(d, ip, cl, S, 1.0);
  for (int s = ab; s < ae; ++s) {
    const int a = d.clusterAtoms[s];
    const double *x = S + CLUSTER_SLOT * (s - ab);
    float4 p = d.pos[a];
    p.x = (float)x[0];
    p.y = (float)x[1];
    p.z = (float)x[2];
    d.pos[a] = p;
  }
}

// ---- launchers ----------------------------------------------------------------

// The water blocks, then the general blocks; returns the in-step kinetic
// slots written.
template <int SM>
int launchIntegrateOpenShaped(const Dev &d, const IntegratorParams &ip, int closing, int early, hipStream_t s) {
  if (!splitWater(d)) {
    integrateOpenKernel<SM><<<unitBlockCount(d), INTEGRATE_BLOCK, ip.clusterLdsBytes, s>>>(d, ip, closing, early, 0);
    return unitBlockCount(d);
  }
  integrateWaterKernel<<<d.waterBlocks, INTEGRATE_BLOCK, 0, s>>>(d, ip, closing, early);
  integrateOpenKernel<SM><<<generalBlockCount(d), INTEGRATE_BLOCK, ip.clusterLdsBytes, s>>>(d, ip, closing, early,
                                                                                           d.waterBlocks);
  return openSlotCount(d);
}

int launchIntegrateOpen(const Dev &d, const IntegratorParams &ip, int closing, int early, hipStream_t s) {
  return ip.starMax <= 3 ? launchIntegrateOpenShaped<3>(d, ip, closing, early, s)
                         : launchIntegrateOpenShaped<4>(d, ip, closing, early, s);
}

int launchIntegrateClose(const Dev &d, const IntegratorParams &ip, int withKinetic, hipStream_t s) {
  const int blocks = unitBlockCount(d);
  if (ip.starMax <= 3)
    integrateCloseKernel<3><<<blocks, INTEGRATE_BLOCK, ip.clusterLdsBytes, s>>>(d, ip, withKinetic);
  else
    integrateCloseKernel<4><<<blocks, INTEGRATE_BLOCK, ip.clusterLdsBytes, s>>>(d, ip, withKinetic);
  return blocks;
}

int launchKinetic(const Dev &d, int term, hipStream_t s) {
  const int blocks = (d.n + 255) / 256;
  kineticKernel<<<blocks, 256, 0, s>>>(d, term);
  return blocks;
}

void launchReduceEnergies(const Dev &d, const SlotCounts &counts, double *out, hipStream_t s) {
  reduceEnergiesKernel<<<E_TERM_COUNT, 256, 0, s>>>(d.slots, counts, out);
}


void launchScaleMolecules(const Dev &d, const int *molStart, const int *molAtoms, int nMol, float4 *centroids,
                          const int *molOfAtom, float scale, hipStream_t s) {
  centroidKernel<<<(nMol + 255) / 256, 256, 0, s>>>(d, molStart, molAtoms, nMol, centroids);
  scaleKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d, centroids, molOfAtom, scale);
}

__global__ void scaleFromKernel(Dev d, const float4 *centroids, const int *molOfAtom, const float *scale) {
  scaleAtom(d, centroids, molOfAtom, *scale, blockIdx.x * blockDim.x + threadIdx.x);
}

void launchScaleMoleculesFrom(const Dev &d, const int *molStart, const int *molAtoms, int nMol, float4 *centroids,
                              const int *molOfAtom, const float *scale, hipStream_t s) {
  centroidKernel<<<(nMol + 255) / 256, 256, 0, s>>>(d, molStart, molAtoms, nMol, centroids);
  scaleFromKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d, centroids, molOfAtom, scale);
}

void launchMaxForce(const Dev &d, unsigned int *out, hipStream_t s) {
  maxForceKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d, out);
}

// The thermostat's draws for every atom at one step: the four uniforms
// (4 per atom, may be null) and the three normals (3 per atom).
__global__ void noiseKernel(int n, unsigned int seed, unsigned int step, float *uniforms, float *normals) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  float u[4], g[3];
  gaussianTriple(seed, a, step, u, g);
  if (uniforms)
    for (int k = 0; k < 4; ++k) uniforms[4 * a + k] = u[k];
  for (int c = 0; c < 3; ++c) normals[3 * a + c] = g[c];
}

void launchNoise(int n, unsigned int seed, unsigned int step, float *uniforms, float *normals, hipStream_t s) {
  noiseKernel<<<(n + 255) / 256, 256, 0, s>>>(n, seed, step, uniforms, normals);
}

void launchDisplace(const Dev &d, float scale, float cap, hipStream_t s) {
  displaceKernel<<<(d.n + 255) / 256, 256, 0, s>>>(d, scale, cap);
}

void launchConstrainPositions(const Dev &d, const float4 *reference, const IntegratorParams &ip, hipStream_t s) {
  constrainPositionsKernel<<<unitBlockCount(d), INTEGRATE_BLOCK, 0, s>>>(d, reference, ip);
}

// ---- batched launchers ---------------------------------------------------------

void launchIntegrateOpenBatch(const BatchView *v, const TaskTable &water, const TaskTable &general, int closing,
                              int early, int lds, int starMax, hipStream_t s) {
  if (water.blocks > 0) integrateWaterBatchKernel<<<water.blocks, INTEGRATE_BLOCK, 0, s>>>(v, water.task, closing, early);
  if (starMax <= 3)
    integrateOpenBatchKernel<3><<<general.blocks, INTEGRATE_BLOCK, lds, s>>>(v, general.task, closing, early);
  else
    integrateOpenBatchKernel<4><<<general.blocks, INTEGRATE_BLOCK, lds, s>>>(v, general.task, closing, early);
}

void launchIntegrateCloseBatch(const BatchView *v, const TaskTable &t, int lds, int starMax, hipStream_t s) {
  if (starMax <= 3)
    integrateCloseBatchKernel<3><<<t.blocks, INTEGRATE_BLOCK, lds, s>>>(v, t.task);
  else
    integrateCloseBatchKernel<4><<<t.blocks, INTEGRATE_BLOCK, lds, s>>>(v, t.task);
}

void launchReduceEnergiesBatch(const BatchView *v, int replicas, int withEarly, hipStream_t s) {
  reduceEnergiesBatchKernel<<<replicas * E_TERM_COUNT, 256, 0, s>>>(v, withEarly);
}

void launchZeroForcesBatch(const BatchView *v, const TaskTable &t, hipStream_t s) {
  zeroForcesBatchKernel<<<t.blocks, 256, 0, s>>>(v, t.task);
}

void launchDecideRebuildBatch(const BatchView *v, int replicas, int force, hipStream_t s) {
  decideRebuildBatchKernel<<<(replicas + 255) / 256, 256, 0, s>>>(v, replicas, force);
}

void launchScaleBatch(const BatchView *v, const TaskTable &molecules, const TaskTable &atoms, hipStream_t s) {
  centroidBatchKernel<<<molecules.blocks, 256, 0, s>>>(v, molecules.task);
  scaleBatchKernel<<<atoms.blocks, 256, 0, s>>>(v, atoms.task);
}

int launchMolecularTerm(const Dev &d, const int *molStart, const int *molAtoms, int nMol, hipStream_t s) {
  const int blocks = (nMol + 255) / 256 > 0 ? (nMol + 255) / 256 : 1;
  molecularTermKernel<<<blocks, 256, 0, s>>>(d, molStart, molAtoms, nMol);
  return blocks;
}

void launchMolecularTermBatch(const BatchView *v, const TaskTable &molecules, hipStream_t s) {
  molecularTermBatchKernel<<<molecules.blocks, 256, 0, s>>>(v, molecules.task);
}

void launchSnapshotBatch(const BatchView *v, const TaskTable &atoms, int mode, hipStream_t s) {
  snapshotBatchKernel<<<atoms.blocks, 256, 0, s>>>(v, atoms.task, mode);
}

void launchForceRebuildFlaggedBatch(const BatchView *v, int replicas, hipStream_t s) {
  forceRebuildFlaggedBatchKernel<<<(replicas + 255) / 256, 256, 0, s>>>(v, replicas);
}

void launchClearRequiredBatch(Scalars *sc, int replicas, hipStream_t s) {
  clearRequiredBatchKernel<<<(replicas + 255) / 256, 256, 0, s>>>(sc, replicas);
}

}  // namespace rmd
