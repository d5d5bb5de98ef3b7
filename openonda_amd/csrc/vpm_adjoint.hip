// Hand-written adjoint (vector-Jacobian product) of the fused pair pass.
//
// Given cotangents lam (of the self-induced velocity u), mu (of the
// stretching S) and xi (of the PSE term D), computes dL/dx and dL/dalpha for
// L = sum_k lam_k.u_k + mu_k.S_k + xi_k.D_k with the GPU-plugin arithmetic
// (sigma_ij, nu_i; ParticlesPhysicsCuda.py:159-191, :396-440, :500-525) and
// sigma held fixed.  One symmetric all-pairs pass per particle m accumulates
// both of its roles -- m as target and m as source -- from its partners' data
// staged in LDS, so no atomics and no second pass are needed.
//
// With r = x_m - x_j, rho^2 = r^2/s^2, t = (rho^2+1)^-1/2, s = sigma_ij, and
//   g = (rho^2+2.5) t^5/s^3,  f = 3(rho^2+3.5) t^7/s^5 = -2 dg/dr^2,
//   h = 7.5(rho^2+4.5) t^9/s^7 = -df/dr^2,  p = t^9/s^2,  q = 4.5 t^11/s^4 = -dp/dr^2,
//   c = r x a_j,  d_mu = mu_m - mu_j,  Q_j = a_j x mu_j,  P_m = a_m x mu_m,
//   phi = d_mu.r,  psi = a_m.c,  z = nu xi,  C = 1/(4 pi):
// velocity  dA_m += C g (lam_j x r)
//           dX_m += C f [lam_m.c - a_m.(lam_j x r)] r - C g [a_j x lam_m - a_m x lam_j]
// stretch   dA_m += C [g (a_j x d_mu) + f phi c]
//           dX_m += C [(f (a_j.P_m + a_m.Q_j) - 2 h psi phi) r + f phi (a_j x a_m) + f psi d_mu]
// PSE       dA_m += 105 C p (z_j - z_m)
//           dX_m += -2 * 105 C q ((a_j - a_m).(z_m - z_j)) r
// Sums over j of g a_j, g lam_j and f phi a_j are accumulated per target and
// crossed with the target's vectors once; the three radial terms share one
// accumulator.
// The same block-centred FP32 offsets, LDS tiles and FP32-partial / FP64-total
// accumulation as the forward pass are used (precision F32C); F64 runs the
// same pass in double.

#include <algorithm>
#include <type_traits>

#include "vpm_common.h"
#include "vpm_prepass.h"

namespace vpm {

namespace {

// Sorted SoA arrays of the adjoint pass (workspace), component c of particle
// k at base[c * n + k].
struct AdjSorted {
  const double* pos;   // [3][n]
  const double* str;   // [3][n]
  const double* lam;   // [3][n]
  const double* mu;    // [3][n]
  const double* z;     // [3][n] nu * xi
  const double* sig;   // [n]
  const double* nu;    // [n]
};

__global__ __launch_bounds__(256) void adj_gather_kernel(vpm_particles_t p, const double* __restrict__ lam,
                                                         const double* __restrict__ mu,
                                                         const double* __restrict__ xi, int64_t ldc,
                                                         int n, const int* __restrict__ perm,
                                                         double* __restrict__ out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int i = perm ? perm[k] : k;
  const int64_t N = n;
  double* pos = out;
  double* str = out + 3 * N;
  double* la = out + 6 * N;
  double* m = out + 9 * N;
  double* z = out + 12 * N;
  double* sg = out + 15 * N;
  double* nu = out + 16 * N;
  const double nui = p.nu ? p.nu[i] : 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pos[c * N + k] = p.pos[c * p.ld + i];
    str[c * N + k] = p.str[c * p.ld + i];
    la[c * N + k] = lam ? lam[c * ldc + i] : 0.0;
    m[c * N + k] = mu ? mu[c * ldc + i] : 0.0;
    z[c * N + k] = xi ? nui * xi[c * ldc + i] : 0.0;
  }
  sg[k] = p.sigma[i];
  nu[k] = nui;
}

template <typename Real>
struct Ar;
template <>
struct Ar<float> {
  static __device__ inline float rsq(float a) { return __builtin_amdgcn_rsqf(a); }
  static __device__ inline float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
};
template <>
struct Ar<double> {
  static __device__ inline double rsq(double a) { return 1.0 / sqrt(a); }
  static __device__ inline double rcp(double a) { return 1.0 / a; }
};

// LDS record of one partner.
template <typename Real>
struct alignas(16) AdjSrc {
  Real x, y, z, ax, ay, az, lx, ly, lz, mx, my, mz, qx, qy, qz, zx, zy, zz, hs, pad;
};

template <typename Real>
struct AdjTgt {
  Real x, y, z, ax, ay, az, lx, ly, lz, mx, my, mz, px, py, pz, zx, zy, zz, hs;
};

// Accumulator slots (per target).  Terms with the same sigma scaling share
// an accumulator; sc = 1/sigma with a uniform radius (offsets pre-scaled),
// 1 otherwise (per-pair 1/sigma_ij powers folded into g, f, h, p, q):
//   aG = sum g a_j, aL = sum g lam_j                      (crossed in the epilogue)
//   aA = sum g (lam_j x r) + 105 p (z_j - z_m)            dA, scale C/s^2
//   aB = sum f phi c - g Q_j                              dA, scale C/s^3
//   aX = sum [f e + sc (f dw - 2 h psi phi) + 945 q dd] r dX, scale C/s^3
//        with dd = (a_j - a_m).(z_j - z_m)
//   aY = sum f phi a_j,  aD = sum f psi d_mu              dX, scale C/s^4
enum : int { aG = 0, aL = 3, aA = 6, aB = 9, aX = 12, aY = 15, aD = 18, aN = 21 };

template <unsigned OUT>
struct AdjSlots {
  static constexpr bool vel = OUT & VPM_OUT_VEL, str = OUT & VPM_OUT_STRETCH,
                        pse = OUT & VPM_OUT_PSE;
  static __host__ __device__ constexpr bool used(int k) {
    return (k < 3 && (vel || str)) || (k >= 3 && k < 6 && vel) ||
           (k >= 6 && k < 9 && (vel || pse)) || (k >= 9 && k < 12 && str) ||
           (k >= 12 && k < 15) || (k >= 15 && str);
  }
  // position of slot k among the used ones (the split partial layout)
  static __host__ __device__ constexpr int packed(int k) {
    int c = 0;
    for (int q = 0; q < k; ++q) c += used(q) ? 1 : 0;
    return c;
  }
  static constexpr int nslot = packed(aN);
};

// Non-inline constants of the pair terms, kept in registers by the caller
// (literal operands halve an instruction's issue rate on gfx950).
template <typename Real>
struct AdjConst {
  Real c15, c3, c75, c2625, c105, c945;
};

template <typename Real, unsigned OUT, bool UNIF>
__device__ __forceinline__ void adj_pair(Real* acc, const AdjTgt<Real>& t, const AdjSrc<Real>& s,
                                         Real sc, const AdjConst<Real>& k) {
  using S = AdjSlots<OUT>;
  using M = Ar<Real>;
  const Real dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  Real A, inv = Real(1), inv2 = Real(1);
  if constexpr (UNIF) {
    A = dz * dz + (dy * dy + (dx * dx + Real(1)));
  } else {
    inv = M::rcp(t.hs + s.hs);
    inv2 = inv * inv;
    A = (dx * dx + dy * dy + dz * dz) * inv2 + Real(1);
  }
  // A t^2 = 1: each factor (rho^2 + c) t^k is written in powers of t alone
  const Real t1 = M::rsq(A);
  const Real t2 = t1 * t1;
  const Real t3 = t2 * t1;
  const Real t5 = t3 * t2;
  const Real t7 = t5 * t2;
  Real xc = Real(0);  // radial coefficient of aX
  if constexpr (S::vel || S::str) {
    Real g = k.c15 * t5 + t3;                        // (rho^2 + 2.5) t^5
    Real f = t5 * (k.c75 * t2 + k.c3);               // 3 (rho^2 + 3.5) t^7
    if constexpr (!UNIF) {
      const Real inv3 = inv2 * inv;
      g *= inv3;
      f *= inv3 * inv2;
    }
    const Real cx = dy * s.az - dz * s.ay;            // c = r x a_j
    const Real cy = dz * s.ax - dx * s.az;
    const Real cz = dx * s.ay - dy * s.ax;
    acc[aG + 0] += g * s.ax;
    acc[aG + 1] += g * s.ay;
    acc[aG + 2] += g * s.az;
    if constexpr (S::vel) {
      const Real kx = s.ly * dz - s.lz * dy;          // lam_j x r
      const Real ky = s.lz * dx - s.lx * dz;
      const Real kz = s.lx * dy - s.ly * dx;
      acc[aA + 0] += g * kx;
      acc[aA + 1] += g * ky;
      acc[aA + 2] += g * kz;
      acc[aL + 0] += g * s.lx;
      acc[aL + 1] += g * s.ly;
      acc[aL + 2] += g * s.lz;
      const Real e = (t.lx * cx + t.ly * cy + t.lz * cz) - (t.ax * kx + t.ay * ky + t.az * kz);
      xc = f * e;
    }
    if constexpr (S::str) {
      Real h = t7 * (k.c2625 * t2 + k.c75);            // 7.5 (rho^2 + 4.5) t^9
      if constexpr (!UNIF) h *= inv2 * inv2 * inv2 * inv;
      const Real dmx = t.mx - s.mx, dmy = t.my - s.my, dmz = t.mz - s.mz;
      const Real phi = dmx * dx + dmy * dy + dmz * dz;
      const Real psi = t.ax * cx + t.ay * cy + t.az * cz;
      const Real dw = (s.ax * t.px + s.ay * t.py + s.az * t.pz) +
                      (t.ax * s.qx + t.ay * s.qy + t.az * s.qz);
      const Real fphi = f * phi;
      acc[aB + 0] += fphi * cx - g * s.qx;
      acc[aB + 1] += fphi * cy - g * s.qy;
      acc[aB + 2] += fphi * cz - g * s.qz;
      acc[aY + 0] += fphi * s.ax;
      acc[aY + 1] += fphi * s.ay;
      acc[aY + 2] += fphi * s.az;
      const Real fpsi = f * psi;
      acc[aD + 0] += fpsi * dmx;
      acc[aD + 1] += fpsi * dmy;
      acc[aD + 2] += fpsi * dmz;
      xc += sc * (f * dw - Real(2) * h * psi * phi);
    }
  }
  if constexpr (S::pse) {
    Real p = t7 * t2, q = p * t2;                     // t^9, t^11
    if constexpr (!UNIF) {
      p *= inv2;
      q *= inv2 * inv2;
    }
    const Real zx = s.zx - t.zx, zy = s.zy - t.zy, zz = s.zz - t.zz;
    const Real p105 = k.c105 * p;
    acc[aA + 0] += p105 * zx;
    acc[aA + 1] += p105 * zy;
    acc[aA + 2] += p105 * zz;
    // (a_j - a_m).(z_m - z_j) = -(a_j - a_m).dz
    const Real dd = (s.ax - t.ax) * zx + (s.ay - t.ay) * zy + (s.az - t.az) * zz;
    xc += k.c945 * q * dd;
  }
  acc[aX + 0] += xc * dx;
  acc[aX + 1] += xc * dy;
  acc[aX + 2] += xc * dz;
}

struct AdjArgs {
  int n;
  AdjSorted s;
  const int* perm;
  double* dpos;
  double* dstr;
  int64_t ld;
  // source split (see vpm_common.h): workgroup (b, k) sums sources
  // [k*chunk, (k+1)*chunk) into part + k*nslot*pld; part == null: one split
  double* part;
  size_t part_cap;
  int64_t pld;
  int chunk;
  int max_splits;  // 0: automatic
};

// FP64 epilogue of one sorted target k from its pair sums T (AdjSlots layout).
template <unsigned OUT, bool UNIF>
__device__ __forceinline__ void adj_store(const AdjArgs& a, int k, const double* T,
                                          double scale) {
  using S = AdjSlots<OUT>;
  const int64_t N = a.n;
  const AdjSorted& A = a.s;
  const double ax = A.str[k], ay = A.str[N + k], az = A.str[2 * N + k];
  double lx = 0, ly = 0, lz = 0, mx = 0, my = 0, mz = 0;
  if constexpr (S::vel) {
    lx = A.lam[k]; ly = A.lam[N + k]; lz = A.lam[2 * N + k];
  }
  if constexpr (S::str) {
    mx = A.mu[k]; my = A.mu[N + k]; mz = A.mu[2 * N + k];
  }
  const double C = kOneOverFourPi;
  double s2 = 1.0, s3 = 1.0, s4 = 1.0;
  if constexpr (UNIF) {
    s2 = scale * scale;
    s3 = s2 * scale;
    s4 = s2 * s2;
  }
  double dA[3], dX[3];
  for (int c = 0; c < 3; ++c) {
    dA[c] = C * s2 * T[aA + c];
    dX[c] = C * s3 * T[aX + c];
  }
  const double G[3] = {T[aG], T[aG + 1], T[aG + 2]};
  if constexpr (S::vel) {  // - C/s^3 (G x lam_m - a_m x Lambda)
    const double Lm[3] = {T[aL], T[aL + 1], T[aL + 2]};
    dX[0] -= C * s3 * ((G[1] * lz - G[2] * ly) - (ay * Lm[2] - az * Lm[1]));
    dX[1] -= C * s3 * ((G[2] * lx - G[0] * lz) - (az * Lm[0] - ax * Lm[2]));
    dX[2] -= C * s3 * ((G[0] * ly - G[1] * lx) - (ax * Lm[1] - ay * Lm[0]));
  }
  if constexpr (S::str) {  // dA += C/s^3 (G x mu_m + B);  dX += C/s^4 (Y x a_m + D)
    dA[0] += C * s3 * ((G[1] * mz - G[2] * my) + T[aB]);
    dA[1] += C * s3 * ((G[2] * mx - G[0] * mz) + T[aB + 1]);
    dA[2] += C * s3 * ((G[0] * my - G[1] * mx) + T[aB + 2]);
    const double Y[3] = {T[aY], T[aY + 1], T[aY + 2]};
    dX[0] += C * s4 * ((Y[1] * az - Y[2] * ay) + T[aD]);
    dX[1] += C * s4 * ((Y[2] * ax - Y[0] * az) + T[aD + 1]);
    dX[2] += C * s4 * ((Y[0] * ay - Y[1] * ax) + T[aD + 2]);
  }
  const int64_t o = a.perm ? a.perm[k] : k;
  for (int c = 0; c < 3; ++c) {
    a.dpos[c * a.ld + o] = dX[c];
    a.dstr[c * a.ld + o] = dA[c];
  }
}

// grid (target blocks, source splits)
template <typename Real, unsigned OUT, bool UNIF, int BLOCK, int TILE, int SUB>
__global__ __launch_bounds__(BLOCK) void adjoint_kernel(AdjArgs a) {
  using S = AdjSlots<OUT>;
  const int lo = static_cast<int>(blockIdx.y) * a.chunk;
  const int hi = a.chunk < a.n - lo ? lo + a.chunk : a.n;
  constexpr bool kF32 = std::is_same<Real, float>::value;
  __shared__ AdjSrc<Real> lds[TILE];
  __shared__ double red[6][BLOCK / 64];
  __shared__ double centre[3];
  const int tid = threadIdx.x;
  const int k = blockIdx.x * BLOCK + tid;
  const bool valid = k < a.n;
  const int64_t N = a.n;
  const AdjSorted& A = a.s;

  double px = 0, py = 0, pz = 0;
  if (valid) {
    px = A.pos[k];
    py = A.pos[N + k];
    pz = A.pos[2 * N + k];
  }
  {
    double mn[3] = {valid ? px : INFINITY, valid ? py : INFINITY, valid ? pz : INFINITY};
    double mx[3] = {valid ? px : -INFINITY, valid ? py : -INFINITY, valid ? pz : -INFINITY};
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double lo = wave_min(mn[c]), hi = wave_max(mx[c]);
      if (lane == 0) {
        red[c][wave] = lo;
        red[3 + c][wave] = hi;
      }
    }
    __syncthreads();
    if (tid < 3) {
      double lo = red[tid][0], hi = red[3 + tid][0];
      for (int w = 1; w < BLOCK / 64; ++w) {
        lo = fmin(lo, red[tid][w]);
        hi = fmax(hi, red[3 + tid][w]);
      }
      centre[tid] = 0.5 * (lo + hi);
    }
    __syncthreads();
  }
  const double cx = centre[0], cy = centre[1], cz = centre[2];
  const double sigma0 = UNIF ? A.sig[0] : 1.0;
  const double scale = UNIF ? 1.0 / sigma0 : 1.0;

  AdjTgt<Real> t{};
  double ax = 0, ay = 0, az = 0, lx = 0, ly = 0, lz = 0, mx = 0, my = 0, mz = 0;
  double zx = 0, zy = 0, zz = 0, sg = 1.0;
  if (valid) {
    ax = A.str[k]; ay = A.str[N + k]; az = A.str[2 * N + k];
    lx = A.lam[k]; ly = A.lam[N + k]; lz = A.lam[2 * N + k];
    mx = A.mu[k]; my = A.mu[N + k]; mz = A.mu[2 * N + k];
    zx = A.z[k]; zy = A.z[N + k]; zz = A.z[2 * N + k];
    sg = A.sig[k];
  }
  t.x = static_cast<Real>((px - cx) * scale);
  t.y = static_cast<Real>((py - cy) * scale);
  t.z = static_cast<Real>((pz - cz) * scale);
  t.ax = static_cast<Real>(ax); t.ay = static_cast<Real>(ay); t.az = static_cast<Real>(az);
  t.lx = static_cast<Real>(lx); t.ly = static_cast<Real>(ly); t.lz = static_cast<Real>(lz);
  t.mx = static_cast<Real>(mx); t.my = static_cast<Real>(my); t.mz = static_cast<Real>(mz);
  t.px = static_cast<Real>(ay * mz - az * my);   // P_m = a_m x mu_m
  t.py = static_cast<Real>(az * mx - ax * mz);
  t.pz = static_cast<Real>(ax * my - ay * mx);
  t.zx = static_cast<Real>(zx); t.zy = static_cast<Real>(zy); t.zz = static_cast<Real>(zz);
  t.hs = static_cast<Real>(0.5 * sg);

  const Real rsc = static_cast<Real>(scale);
  Real part[aN];
  double tot[aN];
#pragma unroll
  for (int q = 0; q < aN; ++q) {
    part[q] = Real(0);
    tot[q] = 0.0;
  }

  AdjConst<Real> kc{Real(1.5), Real(3), Real(7.5), Real(26.25), Real(105), Real(945)};
  if constexpr (kF32)  // keep them in VGPRs
    asm volatile("" : "+v"(kc.c15), "+v"(kc.c3), "+v"(kc.c75), "+v"(kc.c2625), "+v"(kc.c105),
                 "+v"(kc.c945));

  for (int t0 = lo; t0 < hi; t0 += TILE) {
    __syncthreads();
#pragma unroll
    for (int kk = tid; kk < TILE; kk += BLOCK) {
      const int j = t0 + kk;
      AdjSrc<Real> s;
      if (j < hi) {
        const double sax = A.str[j], say = A.str[N + j], saz = A.str[2 * N + j];
        const double smx = A.mu[j], smy = A.mu[N + j], smz = A.mu[2 * N + j];
        s.x = static_cast<Real>((A.pos[j] - cx) * scale);
        s.y = static_cast<Real>((A.pos[N + j] - cy) * scale);
        s.z = static_cast<Real>((A.pos[2 * N + j] - cz) * scale);
        s.ax = static_cast<Real>(sax); s.ay = static_cast<Real>(say); s.az = static_cast<Real>(saz);
        s.lx = static_cast<Real>(A.lam[j]);
        s.ly = static_cast<Real>(A.lam[N + j]);
        s.lz = static_cast<Real>(A.lam[2 * N + j]);
        s.mx = static_cast<Real>(smx); s.my = static_cast<Real>(smy); s.mz = static_cast<Real>(smz);
        s.qx = static_cast<Real>(say * smz - saz * smy);   // Q_j = a_j x mu_j
        s.qy = static_cast<Real>(saz * smx - sax * smz);
        s.qz = static_cast<Real>(sax * smy - say * smx);
        s.zx = static_cast<Real>(A.z[j]);
        s.zy = static_cast<Real>(A.z[N + j]);
        s.zz = static_cast<Real>(A.z[2 * N + j]);
        s.hs = static_cast<Real>(0.5 * A.sig[j]);
      } else {  // far away, all data zero: every term vanishes (PSE's -k_ij
                // lambda_i too: at this distance its powers of t underflow to 0
                // in either precision)
        s = AdjSrc<Real>{};
        s.x = s.y = s.z = static_cast<Real>(std::is_same<Real, float>::value ? 1e18 : 1e40);
        s.hs = Real(0.5);
      }
      s.pad = Real(0);
      lds[kk] = s;
    }
    __syncthreads();
    if (valid) {
      for (int sb = 0; sb < TILE; sb += SUB) {
#pragma unroll 4
        for (int q = 0; q < SUB; ++q) adj_pair<Real, OUT, UNIF>(part, t, lds[sb + q], rsc, kc);
        if constexpr (kF32) {
#pragma unroll
          for (int q = 0; q < aN; ++q) {
            if (!S::used(q)) continue;
            tot[q] += static_cast<double>(part[q]);
            part[q] = Real(0);
          }
        }
      }
    }
  }
  if (!valid) return;
  if constexpr (!kF32) {
#pragma unroll
    for (int q = 0; q < aN; ++q) tot[q] = static_cast<double>(part[q]);
  }

  if (a.part) {  // one split of several: raw sums, coalesced by sorted index
    double* P = a.part + static_cast<int64_t>(blockIdx.y) * S::nslot * a.pld;
#pragma unroll
    for (int q = 0; q < aN; ++q)
      if (S::used(q)) P[S::packed(q) * a.pld + k] = tot[q];
    return;
  }
  adj_store<OUT, UNIF>(a, k, tot, scale);
}

// Adds the splits' sums in split order and runs the epilogue.
template <unsigned OUT, bool UNIF>
__global__ __launch_bounds__(256) void adj_reduce_kernel(AdjArgs a, int nsplit) {
  using S = AdjSlots<OUT>;
  const int k = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (k >= a.n) return;
  double T[aN];
#pragma unroll
  for (int q = 0; q < aN; ++q) T[q] = 0.0;
  for (int sp = 0; sp < nsplit; ++sp) {
    const double* P = a.part + static_cast<int64_t>(sp) * S::nslot * a.pld;
#pragma unroll
    for (int q = 0; q < aN; ++q)
      if (S::used(q)) T[q] += P[S::packed(q) * a.pld + k];
  }
  adj_store<OUT, UNIF>(a, k, T, UNIF ? 1.0 / a.s.sig[0] : 1.0);
}

constexpr int kAdjBlock = 256, kAdjTile = 256, kAdjSub = 32;

template <typename Real, unsigned OUT, bool UNIF>
int launch_adj(const AdjArgs& a0, hipStream_t st) {
  const int blocks = static_cast<int>(ceil_div(a0.n, kAdjBlock));
  auto kern = adjoint_kernel<Real, OUT, UNIF, kAdjBlock, kAdjTile, kAdjSub>;
  AdjArgs a = a0;
  const int ntiles = static_cast<int>(ceil_div(a.n, kAdjTile));
  const int64_t pld = static_cast<int64_t>(blocks) * kAdjBlock;
  const size_t per_split = static_cast<size_t>(AdjSlots<OUT>::nslot) * pld * sizeof(double);
  int ks = 1;
  if (a.part) {
    int kmax = max_splits_for(blocks, ntiles);
    const size_t by_cap = a.part_cap / per_split;
    if (by_cap < static_cast<size_t>(kmax)) kmax = static_cast<int>(by_cap);
    if (a.max_splits > 0 && a.max_splits < kmax) kmax = a.max_splits;
    if (kmax > 1) ks = choose_splits(blocks, ntiles, kmax, resident_blocks(kern, kAdjBlock));
  }
  if (ks > 1) {
    a.chunk = static_cast<int>(ceil_div(ntiles, ks)) * kAdjTile;  // ks is effective
    a.pld = pld;
  } else {
    a.part = nullptr;
    a.chunk = a.n;
    a.pld = 0;
  }
  hipLaunchKernelGGL(kern, dim3(blocks, ks), dim3(kAdjBlock), 0, st, a);
  VPM_LAUNCH_CHECK("adjoint_kernel");
  if (ks > 1) {
    hipLaunchKernelGGL((adj_reduce_kernel<OUT, UNIF>), dim3(static_cast<unsigned>(ceil_div(a.n, 256))),
                       dim3(256), 0, st, a, ks);
    VPM_LAUNCH_CHECK("adj_reduce_kernel");
  }
  return VPM_OK;
}

template <typename Real, bool UNIF>
int adj_mask(const AdjArgs& a, unsigned mask, hipStream_t st) {
  switch (mask) {
    case 1u: return launch_adj<Real, 1u, UNIF>(a, st);
    case 2u: return launch_adj<Real, 2u, UNIF>(a, st);
    case 3u: return launch_adj<Real, 3u, UNIF>(a, st);
    case 4u: return launch_adj<Real, 4u, UNIF>(a, st);
    case 5u: return launch_adj<Real, 5u, UNIF>(a, st);
    case 6u: return launch_adj<Real, 6u, UNIF>(a, st);
    default: return launch_adj<Real, 7u, UNIF>(a, st);
  }
}

struct AdjLayout {
  size_t bbox_part, bbox, keys_in, keys_out, idx_in, perm, sorted, split, cub, total, cub_bytes,
      split_bytes;
};

int adj_layout(int64_t n, int presorted, AdjLayout* L) {
  const size_t N = static_cast<size_t>(n > 0 ? n : 1);
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t at = off;
    off += align256(b);
    return at;
  };
  L->bbox_part = take(kBboxBlocks * 6 * sizeof(double));
  L->bbox = take(8 * sizeof(double));
  L->keys_in = take(N * 4);
  L->keys_out = take(N * 4);
  L->idx_in = take(N * 4);
  L->perm = take(N * 4);
  L->sorted = take(17 * N * sizeof(double));
  // room for the largest source split, sized for all cotangents present
  const int64_t blocks = ceil_div(n, kAdjBlock);
  const int kmax = max_splits_for(blocks, ceil_div(n, kAdjTile));
  L->split_bytes = kmax > 1 ? static_cast<size_t>(kmax) * AdjSlots<7u>::nslot *
                                  static_cast<size_t>(blocks * kAdjBlock) * sizeof(double)
                            : 0;
  L->split = take(L->split_bytes);
  size_t cb = 0;
  if (!presorted) {
    const int rc = sort_temp_bytes(n, &cb);
    if (rc) return rc;
  }
  L->cub_bytes = cb;
  L->cub = take(cb + 16);
  L->total = off;
  return VPM_OK;
}

// ---- velocity at evaluation points: adjoint w.r.t. the sources (FP64) ----
// For u(y_i) = -C sum_j g_ij (r_ij x a_j), r_ij = y_i - x_j, sigma_j,
// r^2 > 1e-8 (ParticlesPhysicsCuda.py:65-97):
//   dA_j = C sum_i g (r x lam_i),  dX_j = C sum_i [g (a_j x lam_i) - f (r.(a_j x lam_i)) r].
constexpr int kPB = 256;

__global__ __launch_bounds__(kPB) void vel_at_adjoint_kernel(vpm_points_t tg, vpm_particles_t p,
                                                             const double* __restrict__ lam,
                                                             int64_t ldl, double* __restrict__ dpos,
                                                             double* __restrict__ dstr, int64_t ldo) {
  __shared__ double sy[3][kPB], sl[3][kPB];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kPB + threadIdx.x;
  const bool valid = j < p.n;
  double xj[3] = {0, 0, 0}, aj[3] = {0, 0, 0}, sig = 1.0;
  if (valid) {
    for (int c = 0; c < 3; ++c) {
      xj[c] = p.pos[c * p.ld + j];
      aj[c] = p.str[c * p.ld + j];
    }
    sig = p.sigma[j];
  }
  const double s2 = sig * sig, is2 = 1.0 / s2, is3 = is2 / sig, is5 = is3 * is2;
  double dA[3] = {0, 0, 0}, dX[3] = {0, 0, 0};
  for (int64_t t0 = 0; t0 < tg.m; t0 += kPB) {
    const int64_t i = t0 + threadIdx.x;
    __syncthreads();
    if (i < tg.m) {
      for (int c = 0; c < 3; ++c) {
        sy[c][threadIdx.x] = tg.pos[c * tg.ld + i];
        sl[c][threadIdx.x] = lam[c * ldl + i];
      }
    }
    __syncthreads();
    const int cnt = static_cast<int>(tg.m - t0 < kPB ? tg.m - t0 : kPB);
    if (valid) {
      for (int q = 0; q < cnt; ++q) {
        const double rx = sy[0][q] - xj[0], ry = sy[1][q] - xj[1], rz = sy[2][q] - xj[2];
        const double r2 = rx * rx + ry * ry + rz * rz;
        if (!(r2 > kEpsPair)) continue;
        const double rho2 = r2 * is2;
        const double t = 1.0 / sqrt(rho2 + 1.0);
        const double t2 = t * t, t5 = t2 * t2 * t, t7 = t5 * t2;
        const double g = (rho2 + 2.5) * t5 * is3;
        const double f = 3.0 * (rho2 + 3.5) * t7 * is5;
        const double lx = sl[0][q], ly = sl[1][q], lz = sl[2][q];
        dA[0] += g * (ry * lz - rz * ly);
        dA[1] += g * (rz * lx - rx * lz);
        dA[2] += g * (rx * ly - ry * lx);
        const double wx = aj[1] * lz - aj[2] * ly, wy = aj[2] * lx - aj[0] * lz,
                     wz = aj[0] * ly - aj[1] * lx;
        const double e = f * (rx * wx + ry * wy + rz * wz);
        dX[0] += g * wx - e * rx;
        dX[1] += g * wy - e * ry;
        dX[2] += g * wz - e * rz;
      }
    }
  }
  if (valid) {
    for (int c = 0; c < 3; ++c) {
      dstr[c * ldo + j] = kOneOverFourPi * dA[c];
      dpos[c * ldo + j] = kOneOverFourPi * dX[c];
    }
  }
}

}  // namespace

}  // namespace vpm

using namespace vpm;

extern "C" {

size_t vpm_adjoint_workspace_bytes(int64_t n, const vpm_forward_opts_t* opts) {
  AdjLayout L{};
  if (adj_layout(n, opts ? opts->presorted : 0, &L) != VPM_OK) return 0;
  return L.total;
}

int vpm_adjoint(const vpm_particles_t* p, const vpm_adjoint_in_t* in, const vpm_forward_opts_t* o,
                vpm_adjoint_out_t* out, void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  if (!in || !o || !out) return fail(VPM_ERR_ARG, "vpm_adjoint: null argument");
  unsigned mask = 0;
  if (in->lam) mask |= VPM_OUT_VEL;
  if (in->mu) mask |= VPM_OUT_STRETCH;
  if (in->xi) mask |= VPM_OUT_PSE;
  if (o->precision != VPM_PREC_F64 && o->precision != VPM_PREC_F32C)
    return fail(VPM_ERR_ARG, "vpm_adjoint: unknown precision");
  if (o->sigma_mode != VPM_SIGMA_GENERAL && o->sigma_mode != VPM_SIGMA_UNIFORM)
    return fail(VPM_ERR_ARG, "vpm_adjoint: unknown sigma mode");
  int rc = check_particles(p, in->xi != nullptr, "vpm_adjoint");
  if (rc) return rc;
  const int64_t n = p->n;
  if (in->ld < n) return fail(VPM_ERR_ARG, "vpm_adjoint: cotangent ld < n");
  if (out->ld < n || (n && (!out->dpos || !out->dstr)))
    return fail(VPM_ERR_ARG, "vpm_adjoint: bad output");
  AdjLayout L{};
  if ((rc = adj_layout(n, o->presorted, &L))) return rc;
  if (!workspace || ws_bytes < L.total) {
    set_error("vpm_adjoint: workspace %zu bytes < required %zu", ws_bytes, L.total);
    return VPM_ERR_WORKSPACE;
  }
  if (n == 0) return VPM_OK;
  hipStream_t st = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  int* perm = nullptr;
  if (!o->presorted) {
    perm = reinterpret_cast<int*>(ws + L.perm);
    rc = morton_sort(p, reinterpret_cast<double*>(ws + L.bbox_part),
                     reinterpret_cast<double*>(ws + L.bbox), reinterpret_cast<uint32_t*>(ws + L.keys_in),
                     reinterpret_cast<uint32_t*>(ws + L.keys_out), reinterpret_cast<int*>(ws + L.idx_in),
                     perm, ws + L.cub, L.cub_bytes, st);
    if (rc) return rc;
  }
  double* sorted = reinterpret_cast<double*>(ws + L.sorted);
  const int ni = static_cast<int>(n);
  hipLaunchKernelGGL(adj_gather_kernel, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0,
                     st, *p, in->lam, in->mu, in->xi, in->ld, ni, static_cast<const int*>(perm),
                     sorted);
  VPM_LAUNCH_CHECK("adj_gather_kernel");
  AdjArgs a{};
  a.n = ni;
  const int64_t N = n;
  a.s = AdjSorted{sorted, sorted + 3 * N, sorted + 6 * N, sorted + 9 * N, sorted + 12 * N,
                  sorted + 15 * N, sorted + 16 * N};
  a.perm = perm;
  a.dpos = out->dpos;
  a.dstr = out->dstr;
  a.ld = out->ld;
  a.part = L.split_bytes ? reinterpret_cast<double*>(ws + L.split) : nullptr;
  a.part_cap = L.split_bytes;
  if (o->max_splits < 0) return fail(VPM_ERR_ARG, "vpm_adjoint: max_splits < 0");
  a.max_splits = o->max_splits;
  if (mask == 0) mask = VPM_OUT_VEL;  // all cotangents null: zero gradients from the vel pass
  const bool unif = o->sigma_mode == VPM_SIGMA_UNIFORM;
  if (o->precision == VPM_PREC_F32C)
    return unif ? adj_mask<float, true>(a, mask, st) : adj_mask<float, false>(a, mask, st);
  return unif ? adj_mask<double, true>(a, mask, st) : adj_mask<double, false>(a, mask, st);
}

int vpm_induced_velocity_at_adjoint(const vpm_points_t* targets, const vpm_particles_t* src,
                                    const double* lam, int64_t ld_lam, double* dpos, double* dstr,
                                    int64_t ld_out, void* stream) {
  clear_error();
  int rc = check_points(targets, "vpm_induced_velocity_at_adjoint");
  if (rc) return rc;
  if ((rc = check_particles(src, false, "vpm_induced_velocity_at_adjoint"))) return rc;
  if (targets->m > 0 && (!lam || ld_lam < targets->m))
    return fail(VPM_ERR_ARG, "vpm_induced_velocity_at_adjoint: bad cotangent");
  if (src->n > 0 && (!dpos || !dstr || ld_out < src->n))
    return fail(VPM_ERR_ARG, "vpm_induced_velocity_at_adjoint: bad output");
  if (src->n == 0) return VPM_OK;
  hipLaunchKernelGGL(vel_at_adjoint_kernel, dim3(static_cast<unsigned>(ceil_div(src->n, kPB))),
                     dim3(kPB), 0, as_stream(stream), *targets, *src, lam, ld_lam, dpos, dstr,
                     ld_out);
  VPM_LAUNCH_CHECK("vel_at_adjoint_kernel");
  return VPM_OK;
}

}  // extern "C"
