// Fused all-pairs forward pass (the fast path).
//
// One pass over all (target, source) pairs computes any subset of the GPU
// plugin's outputs -- self-induced velocity, transpose-scheme stretching, PSE
// diffusion, LES smoothed vorticity and the enstrophy scalar -- sharing r_ij,
// A = rho^2 + 1 and a single rsqrt per pair (A^-5/2 = t^5, A^-7/2 = t^7,
// A^-9/2 = t^9 with t = rsqrt(A)).  Reference arithmetic being fused:
// ParticlesPhysicsCuda.py:159-191 (velocity), :396-440 (stretching), :500-525
// (PSE), :319-346 (vorticity), :886-909 + :777 (enstrophy).
//
// Layout and precision:
//  * Prepass: bounding box -> 30-bit Morton keys -> radix sort (hipCUB) ->
//    gather into a packed 64-byte FP64 record per particle (x,y,z,a,s,nu).
//    Sorting makes each block of targets spatially compact.
//  * A workgroup owns BLOCK*R consecutive sorted targets (R per lane,
//    register blocked).  It computes its bounding-box centre c and stages each
//    source tile into LDS as FP32 offsets (x_j - c)*scale computed in FP64,
//    so FP32 pair arithmetic works on small numbers: r_ij keeps ~24 bits
//    relative to the block size instead of to |x|.  With uniform sigma the
//    offsets are pre-scaled by 1/sigma, so r^2 is rho^2 directly.
//  * Each lane reads the same LDS record (broadcast) and updates R targets.
//  * Per source tile and wave, one of three loops: the far-field moment form
//    (tile far from the workgroup's targets), FP64 per-pair G with an FP64
//    triple product (waves of coherent strengths), or FP32 partials of G.
//  * Accumulation: FP32 partials over SUB sources; partials enter a Kahan
//    (sum, compensation) pair per tile (COMP=0) and the compensated tile sum
//    is folded into FP64 totals; COMP=1 folds partials straight into FP64.
//  * Epilogue scales by the 1/(4 pi) and 1/sigma^k constants in FP64 and
//    scatters to the caller's (unsorted) order.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "vpm_common.h"
#include "vpm_prepass.h"

namespace vpm {

namespace {

// ------------------------------------------------------------------ main pass
struct FusedArgs {
  int n;
  const Src8* src;      // sorted particles
  const int* perm;      // sorted k -> caller index (null: identity)
  double* vel;          // outputs (caller order); unrequested ones point to scratch
  double* str;
  double* pse;
  double* vort;
  int64_t ld_vel, ld_str, ld_pse, ld_vort;
  double* ens_part;     // per-block enstrophy partials
  // Source split: workgroup (b, k) sums sources [k*chunk, (k+1)*chunk) into
  // part + k*nslot*pld (FP64 raw sums per target, by sorted index); a reduce
  // kernel adds the splits in order k = 0, 1, ... and stores.  part == null:
  // one split, stored directly.
  double* part;
  size_t part_cap;      // bytes available at part
  int max_splits;       // 0: automatic
  int g64_mode;         // VPM_G64_*: per-pair FP64 alpha_j sum of the stretching
  int64_t pld;
  int chunk;
  // Far-field records of the FP32 pass (far_records_kernel): per chunk of
  // SUB sorted sources its centre and half-extent in radius units (far_box, 6 doubles)
  // and per source FP32 chunk-centred offsets, strength and moment
  // (far_rec, FarRec<OUT>::kStride floats).  Null where unused.
  const float* far_rec;
  const double* far_box;
};

template <typename Real>
struct Arith;
template <>
struct Arith<float> {
  static __device__ inline float rsq(float a) { return __builtin_amdgcn_rsqf(a); }
  static __device__ inline float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
  static __device__ inline float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};
// FP64: the hardware estimates refined by two Newton steps (each squares the
// relative error of the estimate) instead of the IEEE sqrt + division
// sequences; the FP64 parity tests hold the pass to 1e-12 of the oracle.
// Arguments are finite and >= 1 (A) or > 0 (sigma sums): no special cases.
template <>
struct Arith<double> {
  static __device__ inline double rsq(double a) {
    double y = __builtin_amdgcn_rsq(a);
    double e = ::fma(-a * y, y, 1.0);  // 1 - a y^2
    y = ::fma(0.5 * y, e, y);
    e = ::fma(-a * y, y, 1.0);
    return ::fma(0.5 * y, e, y);
  }
  static __device__ inline double rcp(double a) {
    double y = __builtin_amdgcn_rcp(a);
    double e = ::fma(-a, y, 1.0);      // 1 - a y
    y = ::fma(y, e, y);
    e = ::fma(-a, y, 1.0);
    return ::fma(y, e, y);
  }
  static __device__ inline double fma(double a, double b, double c) { return ::fma(a, b, c); }
};

// LDS source record.  Uniform sigma needs position + strength (p0, p1: x and
// y of the moment a_j x x_j of the far-field form); general sigma adds
// sigma_j/2, 1/sigma_j^2 and 1/sigma_j^3.
template <typename Real, bool UNIF>
struct alignas(16) SrcR;
template <typename Real>
struct alignas(16) SrcR<Real, true> {
  Real x, y, z, ax, ay, az, p0, p1;
};
template <typename Real>
struct alignas(16) SrcR<Real, false> {
  Real x, y, z, ax, ay, az, hs, is2, is3, p0, p1, p2;
};

// (lx, ly, lz): the low parts alpha - fp32(alpha) of an FP32 pass with PSE.
template <typename Real>
struct Tgt {
  Real x, y, z, ax, ay, az, hs, lx, ly, lz;
};

// Accumulator slots; kH is the moment sum of the far-field form (folded into
// kU's FP64 totals per sub-tile, never stored).
// kT9 is the far-field form's sum of the PSE weights (folded into kP's
// partials per sub-tile, never stored).
enum : int { kU = 0, kG = 3, kSd = 6, kP = 9, kW = 12, kE = 15, kH = 18, kT9 = 21, kNAcc = 22 };

template <unsigned OUT>
struct Slots {
  static constexpr bool vel = OUT & VPM_OUT_VEL, str = OUT & VPM_OUT_STRETCH,
                        pse = OUT & VPM_OUT_PSE, vort = OUT & VPM_OUT_VORT,
                        ens = OUT & VPM_OUT_ENSTROPHY;
  // Slots with FP64 totals.  The stretching's alpha_j sum G has FP64
  // totals of its own (gsum): waves whose targets' strengths are coherent
  // (vortex rings, tubes, sheets) add to it in FP64 per pair on near tiles,
  // because there alpha_i x G is a near-exact cancellation and FP32 partials
  // of G cost an order of magnitude in the stretching's relative error
  // (coherent_wave below); other waves, and every wave on far tiles, fold
  // FP32 partials kG..kG+2 into it per sub-tile.
  static __host__ __device__ constexpr bool used(int k) {
    return (k < 3 && vel) || (k >= 6 && k < 9 && str) || (k >= 9 && k < 12 && pse) ||
           (k >= 12 && k < 15 && vort) || (k >= 15 && k < 18 && ens);
  }
  // position of slot k among the used ones (the split partial layout)
  static __host__ __device__ constexpr int packed(int k) {
    int c = 0;
    for (int q = 0; q < k; ++q) c += used(q) ? 1 : 0;
    return c;
  }
  static constexpr int nacc = packed(kNAcc);
  static constexpr int nslot = nacc + (str ? 3 : 0);  // + the FP64 alpha_j sum
};

// Scaling and storing of one target's totals T (FP32-pass slots, FP64) and G
// (the FP64 alpha_j sum of the stretching); returns its enstrophy term.
template <unsigned OUT, bool UNIF>
__device__ __forceinline__ double store_target(const FusedArgs& a, const Src8& ti, int64_t o,
                                               const double* T, const double* G, double scale) {
  using S = Slots<OUT>;
  const double C = kOneOverFourPi;
  double s1 = 1.0, s2 = 1.0, s3 = 1.0;  // 1/sigma^k for uniform sigma
  if constexpr (UNIF) {
    s1 = scale;
    s2 = scale * scale;
    s3 = s2 * scale;
  }
  if constexpr (S::vel) {
    const double f = -C * s2;
    a.vel[o] = f * T[kU + 0];
    a.vel[a.ld_vel + o] = f * T[kU + 1];
    a.vel[2 * a.ld_vel + o] = f * T[kU + 2];
  }
  if constexpr (S::str) {
    const double ax = ti.ax, ay = ti.ay, az = ti.az;
    const double f = C * s3;
    a.str[o] = f * ((ay * G[2] - az * G[1]) + T[kSd + 0]);
    a.str[a.ld_str + o] = f * ((az * G[0] - ax * G[2]) + T[kSd + 1]);
    a.str[2 * a.ld_str + o] = f * ((ax * G[1] - ay * G[0]) + T[kSd + 2]);
  }
  if constexpr (S::pse) {
    const double f = 105.0 * ti.nu * C * s2;
    a.pse[o] = f * T[kP + 0];
    a.pse[a.ld_pse + o] = f * T[kP + 1];
    a.pse[2 * a.ld_pse + o] = f * T[kP + 2];
  }
  if constexpr (S::vort) {
    const double f = 7.5 * C * s3;
    a.vort[o] = f * T[kW + 0];
    a.vort[a.ld_vort + o] = f * T[kW + 1];
    a.vort[2 * a.ld_vort + o] = f * T[kW + 2];
  }
  if constexpr (S::ens) return (ti.ax * T[kE + 0] + ti.ay * T[kE + 1] + ti.az * T[kE + 2]) * s1;
  return 0.0;
}

// Non-inline constants of the pair update, held in registers by the caller:
// as literal operands (v_fmac_f32 v, 0x3fc00000, v) they halve the issue rate
// of their instruction on gfx950.
template <typename Real>
struct PairConst {
  Real c15, c75, c22, c3, c25;
};

// slo: the source's low parts (lx, ly, lz) when the PSE difference is taken
// in two parts (FP32 pass), else null.
// SELF: the pair may be the target's own (self == true then); its term
// g(0) alpha_i = 2.5 alpha_i leaves the FP32 partials of G.  alpha_i x G
// annihilates it in exact arithmetic, but in FP32 partials it sets the size
// of every rounding of its sub-tile, and in sparse sets (radius small
// against the spacing) it dwarfs the other terms of G: a 37-particle cloud
// lost the 1e-6 bar (2.5e-6) with it.  Velocity, stretching's r-term and
// PSE vanish for the self pair; the enstrophy keeps it (the reference sums
// i = j there).
template <typename Real, unsigned OUT, bool UNIF, bool G64, bool SELF = false>
__device__ __forceinline__ void pair_update(Real* acc, double* g64, const Tgt<Real>& t,
                                            const SrcR<Real, UNIF>& s, const double* aj64,
                                            const double* ai64, const Real* slo, Real eps_r2,
                                            const PairConst<Real>& k, bool self = false) {
  using S = Slots<OUT>;
  using M = Arith<Real>;
  const Real dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  // A = rho^2 + 1.  With a uniform radius (offsets pre-scaled by 1/sigma) and
  // no vorticity mask to evaluate, the +1 starts the FMA chain; every factor
  // below is written in A (rho^2 + 2.5 = A + 1.5, 3 rho^2 + 10.5 = 3A + 7.5,
  // rho^2 + 3.2 = A + 2.2) and then in powers of t = A^-1/2.
  Real r2 = Real(0), A, inv = Real(1), inv2 = Real(1);
  if constexpr (UNIF && !S::vort) {
    if constexpr (std::is_same<Real, float>::value)
      A = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, 1.0f)));
    else
      A = fma(dz, dz, fma(dy, dy, fma(dx, dx, 1.0)));
  } else {
    r2 = dx * dx + dy * dy + dz * dz;
    if constexpr (UNIF) {
      A = r2 + Real(1);
    } else {
      const Real sij = t.hs + s.hs;
      inv = M::rcp(sij);
      inv2 = inv * inv;
      A = r2 * inv2 + Real(1);
    }
  }
  // Since A t^2 = 1:  (A + 1.5) t^5 = t^3 + 1.5 t^5,  (3A + 7.5) t^7 = t^5 (3 + 7.5 t^2)
  // and (A + 2.2) t^3 = t + 2.2 t^3 -- one instruction fewer each, and fewer
  // factors of the (1-ulp) hardware rsqrt than the A-weighted forms.
  const Real t1 = M::rsq(A);
  const Real t2 = t1 * t1;
  const Real t3 = t2 * t1;
  const Real t5 = t3 * t2;
  if constexpr (S::vel || S::str) {
    Real g = M::fma(k.c15, t5, t3);
    if constexpr (!UNIF) g *= inv2 * inv;
    Real cx = dy * s.az - dz * s.ay;
    Real cy = dz * s.ax - dx * s.az;
    Real cz = dx * s.ay - dy * s.ax;
    // Coherent waves (G64) of an FP32 pass with stretching evaluate the
    // triple product e = alpha_i . (r x alpha_j) in FP64, from the FP32
    // offsets and the FP64 strengths.  Along a vortex tube alpha_i, r and
    // alpha_j are nearly coplanar and e is a small difference of products of
    // size |alpha|^2 |r|; in FP32 its rounding dominated the stretching (thin
    // ring of 126592 in emulation: 4e-6 with FP32 strengths, 8.9e-7 with
    // their low parts in FP32, 3.4e-7 in FP64).  The velocity's c stays FP32
    // (8.8e-8 there either way).
    constexpr bool kE64 = G64 && S::str && std::is_same<Real, float>::value;
    if constexpr (S::vel) {
      acc[kU + 0] += g * cx;
      acc[kU + 1] += g * cy;
      acc[kU + 2] += g * cz;
    }
    if constexpr (S::str) {
      if constexpr (G64) {
        const double gd = static_cast<double>(g);
        g64[0] = fma(gd, aj64[0], g64[0]);
        g64[1] = fma(gd, aj64[1], g64[1]);
        g64[2] = fma(gd, aj64[2], g64[2]);
      } else {
        const Real gG = (SELF && self) ? Real(0) : g;
        acc[kG + 0] += gG * s.ax;
        acc[kG + 1] += gG * s.ay;
        acc[kG + 2] += gG * s.az;
      }
      Real e;
      if constexpr (kE64) {
        const double rx = dx, ry = dy, rz = dz;
        const double qx = ry * aj64[2] - rz * aj64[1];
        const double qy = rz * aj64[0] - rx * aj64[2];
        const double qz = rx * aj64[1] - ry * aj64[0];
        e = static_cast<Real>(ai64[0] * qx + ai64[1] * qy + ai64[2] * qz);
      } else {
        e = t.ax * cx + t.ay * cy + t.az * cz;
      }
      Real f2 = t5 * M::fma(k.c75, t2, k.c3);
      if constexpr (!UNIF) f2 *= inv2 * inv2 * inv;
      const Real w = f2 * e;
      acc[kSd + 0] += w * dx;
      acc[kSd + 1] += w * dy;
      acc[kSd + 2] += w * dz;
    }
  }
  if constexpr (S::pse) {
    Real t9 = t5 * t2 * t2;
    if constexpr (!UNIF) t9 *= inv2;
    Real dax = s.ax - t.ax, day = s.ay - t.ay, daz = s.az - t.az;
    if constexpr (std::is_same<Real, float>::value) {
      // alpha_j - alpha_i from high and low parts: the high difference is
      // exact when the strengths are close (smooth fields), where a rounded
      // FP32 strength would lose the small difference
      dax += slo[0] - t.lx;
      day += slo[1] - t.ly;
      daz += slo[2] - t.lz;
    }
    acc[kP + 0] += t9 * dax;
    acc[kP + 1] += t9 * day;
    acc[kP + 2] += t9 * daz;
  }
  if constexpr (S::vort) {
    Real wv;
    if constexpr (UNIF) {
      wv = t5 * t2;
    } else {  // sigma_j, not sigma_ij (ParticlesPhysicsCuda.py:323,334-337)
      const Real Aj = r2 * s.is2 + Real(1);
      const Real tj = M::rsq(Aj);
      const Real tj2 = tj * tj;
      wv = tj2 * tj2 * tj2 * tj * s.is3;
    }
    wv = r2 > eps_r2 ? wv : Real(0);
    acc[kW + 0] += wv * s.ax;
    acc[kW + 1] += wv * s.ay;
    acc[kW + 2] += wv * s.az;
  }
  if constexpr (S::ens) {
    Real en = M::fma(k.c22, t3, t1);
    if constexpr (!UNIF) en *= inv;
    acc[kE + 0] += en * s.ax;
    acc[kE + 1] += en * s.ay;
    acc[kE + 2] += en * s.az;
  }
}

// g(0) of the self pair, by the operations pair_update applies to it (d = 0
// gives A = 1 exactly): the product g(0) alpha_i formed from it is the term
// the FP64 sums of G took for the self pair.
template <typename Real, bool UNIF>
__device__ __forceinline__ Real self_g(Real hs, const PairConst<Real>& k) {
  using M = Arith<Real>;
  Real A = Real(1);
  asm volatile("" : "+v"(A));  // evaluated at run time, as in the sweep
  const Real t1 = M::rsq(A);
  const Real t2 = t1 * t1;
  const Real t3 = t2 * t1;
  const Real t5 = t3 * t2;
  Real g = M::fma(k.c15, t5, t3);
  if constexpr (!UNIF) {
    const Real inv = M::rcp(hs + hs);
    const Real inv2 = inv * inv;
    g *= inv2 * inv;
  }
  return g;
}

// Far-field moment form of velocity + stretching (+ PSE), FP32, uniform
// radius.  With d = x_i - x_j and coordinates relative to the workgroup
// centre,
//   sum_j g d x a_j  = x_i x G + sum_j g (a_j x x_j),   G = sum_j g a_j,
//   a_i . (d x a_j)  = a_j . (a_i x x_i) + a_i . (a_j x x_j),
// so with G already summed for the stretching and the source moments
// m_j = a_j x x_j staged with the tile, a pair costs 29 VALU instructions
// instead of 32 (the cross product d x a_j and its three multiply-adds
// into u go).  x_i x G + H is a cancellation of relative size |x_i| / |d|:
// the form is used only for source tiles whose bounding box keeps a gap of
// more than kFarGap radii from the workgroup's targets, where it stays
// small (random cloud of 2^20 in FP32 emulation: rel-L2 unchanged at
// 3.1e-7 / 6.2e-7 with 97% of the pairs in the far form; using it for
// every pair gives 1.2e-6 / 8.7e-7).  Coherent waves (vortex tubes) use it
// on far tiles too, with FP32 partials of G there: far terms of G are small
// beside the near ones that alpha_i x G cancels, and the triple product's
// two moment terms are each small along a tube (thin ring of 131072 in
// emulation: 1.27e-7 / 1.02e-6 either way, 88% of pairs far).
// p = a_i x x_i per target.
constexpr double kFarGap = 4.0;
// The far records are centred on their own chunk, so |x_j| is at most the
// chunk's radius and x_i x G + H is no cancellation at any gap; the gap is
// still 4 radii: the far form's FP32 triple product is what limits it.  At
// 2 radii the ring of 131072 with PSE ran 17.6 -> 16.2 ms, but the batch's
// thin rings (126,592 particles, |S| ~ 2e-3 |alpha||u|) lost the 1e-6 bar
// in the stretching (1.3e-6), as did the single thin ring of 126,592.
constexpr double kFarGapS = 4.0;

template <typename Real>
struct Mom {  // a source's moment a_j x x_j
  Real x, y, z;
};

// Far-field record of one source in the FP32 pass, relative to the centre of
// its chunk of SUB sorted sources and scaled by 1/sigma: offset (3),
// strength (3) and moment a_j x x_j (3; 3 a_j x x_j in the sets with
// stretching and PSE, see pair_moment); the far form of PSE needs no low
// parts of the strengths (pair_moment).  A wave reads the records of a far chunk through
// scalar loads (every lane pairs with the same source), so they reach the
// pair instructions as scalar operands with no LDS staging, no LDS reads and
// no barrier; the target's offset to the chunk centre, and a_i x x_i, are
// formed once per chunk.  Relative to the chunk centre |x_j| is at most the
// chunk's radius, so x_i x G + H is not a cancellation on a far chunk.
constexpr int kMinFarChunk = 64, kMaxFarChunk = 128;  // SUB of the shapes that use them
#ifndef VPM_FAR_GROUP
#define VPM_FAR_GROUP 4
#endif
constexpr int kFarGroup = VPM_FAR_GROUP;  // far records per scalar-load group
// Floats per far record: 9, padded to 12 in the sets with PSE, whose
// kernels ran faster with groups of four records on whole 64-byte lines
// (ring of 131,072: 16.0 against 16.3 ms; the 96-set batch: 723 against
// 743 ms per step); velocity + stretching ran faster unpadded (618 against
// 624 ms at N = 2^20, same machine).
template <unsigned OUT>
struct FarRec {
  static constexpr int kStride = (OUT & VPM_OUT_PSE) ? 12 : 9;
  // Sets with stretching and PSE store 3 m_j (pair_moment, E3), the others m_j.
  static constexpr bool kE3 = (OUT & VPM_OUT_PSE) && (OUT & VPM_OUT_STRETCH);
  static constexpr double kMomScale = kE3 ? 3.0 : 1.0;
};
constexpr int kMaxFarStride = 12;

// Constant-address-space view of the records: a load through it with a
// wave-uniform address becomes a scalar (s_load) load.
typedef __attribute__((address_space(4))) const float far_float_t;

// Builds the far records of chunk c (sources [c*CH, c*CH + CH)) of a set of
// n sorted records: one workgroup of CH lanes.  Uniform radius sigma0 (scale
// = 1/sigma0, as the pass computes it).  Lanes past n write zero-strength
// padding far away, which contributes exactly 0.  rec / box point at the
// chunk's own records and box.
template <int CH>
__device__ __forceinline__ void far_records_chunk(const Src8* __restrict__ src, int n, int c,
                                                  int stride, double mscale, double sigma0,
                                                  float* __restrict__ rec,
                                                  double* __restrict__ box) {
  static_assert(CH % 64 == 0 && CH <= 256, "chunk");
  __shared__ double red[6][CH / 64];
  __shared__ double cen[3];
  const int k = c * CH + static_cast<int>(threadIdx.x);
  const bool ok = k < n;
  const double scale = 1.0 / sigma0;
  Src8 g{};
  if (ok) g = src[k];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  {
    double v[3] = {g.x, g.y, g.z};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double lo = wave_min(ok ? v[q] : INFINITY), hi = wave_max(ok ? v[q] : -INFINITY);
      if (lane == 0) {
        red[q][wave] = lo;
        red[3 + q][wave] = hi;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    double lo = red[q][0], hi = red[3 + q][0];
    for (int w = 1; w < CH / 64; ++w) {
      lo = fmin(lo, red[q][w]);
      hi = fmax(hi, red[3 + q][w]);
    }
    cen[q] = 0.5 * (lo + hi);
    box[q] = cen[q] * scale;  // centre and half-extent in radius units
    box[3 + q] = 0.5 * (hi - lo) * scale;
  }
  __syncthreads();
  float* r = rec + static_cast<int64_t>(threadIdx.x) * stride;
  if (ok) {
    const double x = (g.x - cen[0]) * scale, y = (g.y - cen[1]) * scale, z = (g.z - cen[2]) * scale;
    const float ax = static_cast<float>(g.ax), ay = static_cast<float>(g.ay),
                az = static_cast<float>(g.az);
    r[0] = static_cast<float>(x);
    r[1] = static_cast<float>(y);
    r[2] = static_cast<float>(z);
    r[3] = ax;
    r[4] = ay;
    r[5] = az;
    r[6] = static_cast<float>(mscale * (g.ay * z - g.az * y));
    r[7] = static_cast<float>(mscale * (g.az * x - g.ax * z));
    r[8] = static_cast<float>(mscale * (g.ax * y - g.ay * x));
    for (int q = 9; q < stride; ++q) r[q] = 0.0f;
  } else {
    r[0] = r[1] = r[2] = 1e18f;
    for (int q = 3; q < stride; ++q) r[q] = 0.0f;
  }
}

// The far records of one set: workgroup c builds chunk c.
template <int CH>
__global__ __launch_bounds__(CH) void far_records_kernel(const Src8* __restrict__ src, int n,
                                                         int stride, double mscale,
                                                         float* __restrict__ rec,
                                                         double* __restrict__ box) {
  const int c = blockIdx.x;
  far_records_chunk<CH>(src, n, c, stride, mscale, src[0].s,
                        rec + static_cast<int64_t>(c) * CH * stride,
                        box + static_cast<int64_t>(c) * 6);
}

// E3: p and the moment m come scaled by 3 (the far records of sets with
// stretching and PSE hold 3 m_j; their H is divided by 3 in the FP64 fold),
// so e arrives as 3 e and the weight t5 (3 + 7.5 t2) e is (t5 + 2.5 t7) 3e;
// t7 = t5 t2 is then shared with PSE's t9 = t7 t2 and the vorticity's weight
// (one instruction fewer per pair with PSE, two with PSE and vorticity).
template <typename Real, unsigned OUT, bool UNIF, bool E3 = false>
__device__ __forceinline__ void pair_moment(Real* acc, const Tgt<Real>& t, const Real* p,
                                            const SrcR<Real, UNIF>& s, const Mom<Real>& m,
                                            const Real* slo, const PairConst<Real>& k) {
  using S = Slots<OUT>;
  using M = Arith<Real>;
  const Real dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  const Real A = M::fma(dz, dz, M::fma(dy, dy, M::fma(dx, dx, Real(1))));
  const Real t1 = M::rsq(A);
  const Real t2 = t1 * t1;
  const Real t3 = t2 * t1;
  const Real t5 = t3 * t2;
  const Real g = M::fma(k.c15, t5, t3);
  acc[kG + 0] = M::fma(g, s.ax, acc[kG + 0]);
  acc[kG + 1] = M::fma(g, s.ay, acc[kG + 1]);
  acc[kG + 2] = M::fma(g, s.az, acc[kG + 2]);
  acc[kH + 0] = M::fma(g, m.x, acc[kH + 0]);
  acc[kH + 1] = M::fma(g, m.y, acc[kH + 1]);
  acc[kH + 2] = M::fma(g, m.z, acc[kH + 2]);
  Real e = s.ax * p[0];
  e = M::fma(s.ay, p[1], e);
  e = M::fma(s.az, p[2], e);
  e = M::fma(t.ax, m.x, e);
  e = M::fma(t.ay, m.y, e);
  e = M::fma(t.az, m.z, e);
  const Real t7 = t5 * t2;
  Real w;
  if constexpr (E3)
    w = M::fma(k.c25, t7, t5) * e;
  else
    w = t5 * M::fma(k.c75, t2, k.c3) * e;
  acc[kSd + 0] = M::fma(w, dx, acc[kSd + 0]);
  acc[kSd + 1] = M::fma(w, dy, acc[kSd + 1]);
  acc[kSd + 2] = M::fma(w, dz, acc[kSd + 2]);
  if constexpr (S::vort) {  // uniform radius: t^7; r^2 > 1e-8 always holds on a far tile
    const Real wv = t7;
    acc[kW + 0] = M::fma(wv, s.ax, acc[kW + 0]);
    acc[kW + 1] = M::fma(wv, s.ay, acc[kW + 1]);
    acc[kW + 2] = M::fma(wv, s.az, acc[kW + 2]);
  }
  if constexpr (S::ens) {
    const Real en = M::fma(k.c22, t3, t1);
    acc[kE + 0] = M::fma(en, s.ax, acc[kE + 0]);
    acc[kE + 1] = M::fma(en, s.ay, acc[kE + 1]);
    acc[kE + 2] = M::fma(en, s.az, acc[kE + 2]);
  }
  if constexpr (S::pse) {
    // sum_j t9 (a_j - a_i) = sum_j t9 a_j - a_i sum_j t9 (the second term
    // is applied per sub-tile, far_pse_fold).  On a far tile or chunk every
    // pair is more than 4 radii apart, so t9 = A^-9/2 < 17^-9/2 = 2.9e-6: both
    // sums are that small beside the near tiles' terms of the same target,
    // and so are their roundings -- the near tiles keep the differences.
    const Real t9 = t7 * t2;
    acc[kP + 0] = M::fma(t9, s.ax, acc[kP + 0]);
    acc[kP + 1] = M::fma(t9, s.ay, acc[kP + 1]);
    acc[kP + 2] = M::fma(t9, s.az, acc[kP + 2]);
    acc[kT9] += t9;
  }
}

// The far form's PSE partials: sum t9 a_j - a_i sum t9 (see pair_moment).
template <typename Real, unsigned OUT>
__device__ __forceinline__ void far_pse_fold(Real* acc, const Tgt<Real>& t) {
  if constexpr ((OUT & VPM_OUT_PSE) != 0) {
    acc[kP + 0] = Arith<Real>::fma(-t.ax, acc[kT9], acc[kP + 0]);
    acc[kP + 1] = Arith<Real>::fma(-t.ay, acc[kT9], acc[kP + 1]);
    acc[kP + 2] = Arith<Real>::fma(-t.az, acc[kT9], acc[kP + 2]);
    acc[kT9] = Real(0);
  }
}

// Coherence of 64 consecutive sorted targets (one lane row of a wave),
// |sum alpha_i| / sum |alpha_i|: for unrelated strengths it is ~0.14 and
// stayed below 0.35 over the 16384 rows of a random cloud of 2^20 particles;
// along vortex rings, tubes and sheets (the example cases, leapfrogging
// rings) it stayed above 0.59.  A wave with a row above kCoherent sums G in
// FP64 per pair.
constexpr double kCoherent = 0.45;

template <int R>
__device__ __forceinline__ bool coherent_wave(const Src8* tg, const bool* valid) {
  bool any = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double sx = 0.0, sy = 0.0, sz = 0.0, m = 0.0;
    if (valid[r]) {
      sx = tg[r].ax;
      sy = tg[r].ay;
      sz = tg[r].az;
      m = sqrt(sx * sx + sy * sy + sz * sz);
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    m = wave_sum(m);
    any = any || sqrt(sx * sx + sy * sy + sz * sz) > kCoherent * m;
  }
  return any;
}

// One workgroup's work: BLOCK*R consecutive sorted targets of a particle set
// of n particles (src, perm) starting at `base`, against sources [lo, hi).
// split_out == null: writes the outputs of its targets and its enstrophy
// partial to *ens_slot.  Otherwise writes its raw sums to split_out (Slots::packed layout,
// leading dimension pld) for the split reduce.
// FARS: far chunks of an FP32 pass with the moment form read the far records
// (far_rec, far_box: this set's, indexed by its sorted source index) through
// scalar loads; otherwise far tiles are staged in LDS like near ones.
template <typename Real, unsigned OUT, bool UNIF, int BLOCK, int R, int TILE, int SUB, int COMP,
          bool FARS = false>
__device__ __forceinline__ void fused_block(const FusedArgs& a, const int n, const Src8* src,
                                            const int* perm, const int base, double* ens_slot,
                                            const int lo, const int hi, double* split_out,
                                            const int64_t pld, const float* far_rec = nullptr,
                                            const double* far_box = nullptr) {
  static_assert(TILE % SUB == 0 && (TILE % BLOCK == 0 || BLOCK % TILE == 0), "tile shape");
  using S = Slots<OUT>;
  constexpr bool kComp = std::is_same<Real, float>::value && COMP == 0;
  constexpr bool kFold32 = std::is_same<Real, float>::value;
  // low parts alpha - fp32(alpha) of the strengths (PSE differences)
  constexpr bool kLo = S::pse && std::is_same<Real, float>::value;
  // Output sets with stretching carry two near-tile loops in an FP32 pass:
  // FP64 per-pair G (and triple product) for coherent waves, FP32 partials of
  // G otherwise.  Sets with vorticity or enstrophy but no PSE are the
  // exception: they unroll deeper and a second loop would cost occupancy, so
  // they sum G in FP64 only.  With velocity and a uniform radius every wave
  // runs the far-field moment form (pair_moment) on distant tiles.
  constexpr bool kBothG = S::str && (S::pse || (!S::vort && !S::ens));
  constexpr bool kMoment = S::str && S::vel && UNIF;
  // far chunks from the scalar-loaded far records (kFarS) or far tiles
  // staged in LDS (kFarL)
  constexpr bool kFarS = FARS && kMoment && std::is_same<Real, float>::value;
  constexpr bool kFarL = kMoment && !kFarS;
  __shared__ SrcR<Real, UNIF> lds[TILE];
  __shared__ double lds_a64[S::str ? TILE : 1][3];
  __shared__ float4 lds_lo[kLo ? TILE : 1];
  __shared__ Real lds_mz[kFarL ? TILE : 1];  // a_j x x_j: (p0, p1) of the record, z here
  __shared__ double lds_x64[kMoment ? BLOCK * R : 1][3];  // the targets' x_i in FP64
  __shared__ double red[6][BLOCK / 64];
  __shared__ int wave_coh[BLOCK / 64];
  __shared__ double wbox[kFarS ? BLOCK / 64 : 1][6];
  __shared__ double centre[3];
  __shared__ double half_ext[3];  // the targets' bounding box, scaled: [-half_ext, half_ext]

  const int tid = threadIdx.x;
  // Uniform sigma: every record holds the same radius; offsets are scaled by
  // 1/sigma so that r^2 below is rho^2.  The vorticity mask r^2 > 1e-8
  // (ParticlesPhysicsCuda.py:332) is rescaled alike.
  const double sigma0 = UNIF ? src[0].s : 1.0;
  const double scale = UNIF ? 1.0 / sigma0 : 1.0;

  // ---- targets and the block centre ------------------------------------
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  Src8 tg[R];
  bool valid[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = base + r * BLOCK + tid;
    valid[r] = k < n;
    if (valid[r]) {
      tg[r] = src[k];
      mn[0] = fmin(mn[0], tg[r].x); mx[0] = fmax(mx[0], tg[r].x);
      mn[1] = fmin(mn[1], tg[r].y); mx[1] = fmax(mx[1], tg[r].y);
      mn[2] = fmin(mn[2], tg[r].z); mx[2] = fmax(mx[2], tg[r].z);
    } else {
      tg[r] = Src8{0, 0, 0, 0, 0, 0, 1, 0};
    }
  }
  // G accumulation mode (FP32 passes with stretching): FP64 per pair when a
  // wave of the workgroup holds coherent strengths.  The choice is made for
  // the whole workgroup, so that every wave runs the same tile loop and
  // reaches the same barriers.
  bool coh = false;
  if constexpr (kBothG && std::is_same<Real, float>::value)
    coh = a.g64_mode == VPM_G64_ALWAYS ||
          (a.g64_mode == VPM_G64_AUTO && coherent_wave<R>(tg, valid));
  // each wave's own target box (centre and half-extent in radius units) in wbox: far
  // chunks are decided per wave against it, the staging of a tile per
  // workgroup
  {
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double lo = wave_min(mn[c]), hi = wave_max(mx[c]);
      if (lane == 0) {
        red[c][wave] = lo;
        red[3 + c][wave] = hi;
        if constexpr (kFarS) {  // wbox has one row per wave only then
          wbox[wave][c] = 0.5 * (lo + hi) * scale;
          wbox[wave][3 + c] = 0.5 * (hi - lo) * scale;
        }
      }
    }
    if (lane == 0) wave_coh[wave] = coh ? 1 : 0;
    __syncthreads();
    if (tid < 3) {
      double lo = red[tid][0], hi = red[3 + tid][0];
      for (int w = 1; w < BLOCK / 64; ++w) {
        lo = fmin(lo, red[tid][w]);
        hi = fmax(hi, red[3 + tid][w]);
      }
      centre[tid] = 0.5 * (lo + hi);
      half_ext[tid] = 0.5 * (hi - lo) * scale;
    }
    __syncthreads();
  }
  const double cx = centre[0], cy = centre[1], cz = centre[2];
  const double eps_r2_d = kEpsPair * scale * scale;

  Tgt<Real> t[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    t[r].x = static_cast<Real>((tg[r].x - cx) * scale);
    t[r].y = static_cast<Real>((tg[r].y - cy) * scale);
    t[r].z = static_cast<Real>((tg[r].z - cz) * scale);
    t[r].ax = static_cast<Real>(tg[r].ax);
    t[r].ay = static_cast<Real>(tg[r].ay);
    t[r].az = static_cast<Real>(tg[r].az);
    t[r].hs = static_cast<Real>(0.5 * tg[r].s);
    t[r].lx = t[r].ly = t[r].lz = Real(0);
    if constexpr (kLo) {
      t[r].lx = static_cast<Real>(tg[r].ax - static_cast<double>(t[r].ax));
      t[r].ly = static_cast<Real>(tg[r].ay - static_cast<double>(t[r].ay));
      t[r].lz = static_cast<Real>(tg[r].az - static_cast<double>(t[r].az));
    }
  }
  double a64t[R][3];  // the targets' strengths in FP64 (triple products of coherent waves)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    a64t[r][0] = tg[r].ax;
    a64t[r][1] = tg[r].ay;
    a64t[r][2] = tg[r].az;
  }
  Real pm[R][3];  // a_i x x_i of the moment form
#pragma unroll
  for (int r = 0; r < R; ++r) {
    pm[r][0] = pm[r][1] = pm[r][2] = Real(0);
    if constexpr (kMoment) {
      const double x = (tg[r].x - cx) * scale, y = (tg[r].y - cy) * scale,
                   z = (tg[r].z - cz) * scale;
      // far records: the target's position in radius units (a far chunk's
      // centre is subtracted from it); otherwise its offset to the block
      // centre
      lds_x64[r * BLOCK + tid][0] = kFarS ? tg[r].x * scale : x;
      lds_x64[r * BLOCK + tid][1] = kFarS ? tg[r].y * scale : y;
      lds_x64[r * BLOCK + tid][2] = kFarS ? tg[r].z * scale : z;
      pm[r][0] = static_cast<Real>(tg[r].ay * z - tg[r].az * y);
      pm[r][1] = static_cast<Real>(tg[r].az * x - tg[r].ax * z);
      pm[r][2] = static_cast<Real>(tg[r].ax * y - tg[r].ay * x);
    }
  }
  const Real eps_r2 = static_cast<Real>(eps_r2_d);
  // Workgroup-uniform choice of the G accumulation (FP32 passes).
  bool g64 = S::str && std::is_same<Real, float>::value;
  if constexpr (kBothG && std::is_same<Real, float>::value) {
    int any = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) any |= wave_coh[w];
    g64 = __builtin_amdgcn_readfirstlane(any) != 0;
  }

  Real part[R][kNAcc];
  Real ksum[R][kNAcc];
  Real kcmp[R][kNAcc];
  double tot[R][kNAcc];
  double gsum[R][3];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int q = 0; q < kNAcc; ++q) {
      part[r][q] = Real(0);
      ksum[r][q] = Real(0);
      kcmp[r][q] = Real(0);
      tot[r][q] = 0.0;
    }
#pragma unroll
  for (int r = 0; r < R; ++r) gsum[r][0] = gsum[r][1] = gsum[r][2] = 0.0;

  PairConst<Real> kc{Real(1.5), Real(7.5), Real(2.2), Real(3), Real(2.5)};
  if constexpr (std::is_same<Real, float>::value)  // keep them in VGPRs
    asm volatile("" : "+v"(kc.c15), "+v"(kc.c75), "+v"(kc.c22), "+v"(kc.c3));
  // Far chunks of the sets with stretching and PSE take the weight from t7
  // (pair_moment, E3; their records hold 3 m_j).
  constexpr bool kE3 = kFarS && FarRec<OUT>::kE3;

  // ---- sweep the source tiles ------------------------------------------
  // One instantiation per G mode (G64: FP64 per-pair G, coherent waves), so
  // that each loop keeps only its own state live.  The mode is uniform over
  // the workgroup, so every barrier below is reached by all of its waves.
  auto sweep = [&](auto mode) {
    constexpr bool G64 = decltype(mode)::value;
    constexpr bool kFar = kFarL;
    constexpr bool kStageLo = kLo;
    constexpr bool kStageA64 = S::str;
    // FP32 partials of G: the sub-tiles holding this wave's own targets take
    // the self pair out of G (pair_update, SELF).
    constexpr bool kSelfMask = !G64 && S::str && std::is_same<Real, float>::value;
    // first target of this wave, in a scalar register: the test below is
    // then scalar arithmetic (per sub-tile, far tiles skip it)
    const int wave0 = __builtin_amdgcn_readfirstlane(base + (tid & ~63));
    auto self_sub = [&](int j0) {  // [j0, j0 + SUB) meets a wave's targets
      bool meet = false;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k0 = wave0 + r * BLOCK;
        meet = meet || (j0 < k0 + 64 && k0 < j0 + SUB);
      }
      return meet;
    };
    // Sources per unrolled step: deep unrolling lets the compiler prefetch LDS
    // records for 8 sources, which the heaviest output sets pay for in occupancy.
    constexpr int kUnroll = (S::pse && (S::vort || S::ens)) ? 1 : 8;
    // The FP32 far loop runs 16 deep where the near loop runs 8 (0.5-0.65%
    // faster at N = 2^20 than 8 on two machines, the batch 0.3%; 4 is 0.7%
    // slower; the sets unrolled 1 for occupancy would spill).  FP64 records are twice the registers: the
    // FP64 pass's far loop prefetches fewer of them.
    constexpr int kUnrollFar = std::is_same<Real, float>::value ? (kUnroll == 8 ? 16 : kUnroll) : 2;
    constexpr int kChunks = TILE / SUB;
    constexpr unsigned kAllFar = (1u << kChunks) - 1u;
    static_assert(kChunks <= 32 && 64 % kChunks == 0, "far bits");
    // Chunks of this tile in the far form (kFarS): their box keeps a gap of
    // more than kFarGapS radii from the targets' box.  Returns the
    // workgroup's bits (workgroup-uniform: a tile with a near chunk is
    // staged) and sets the wave's (its own, smaller box: a superset of the
    // workgroup's bits; wave-uniform).  Chunks at or past hi count as far
    // and are skipped.  The tests run for a window of 64 chunks at a time,
    // lane l testing chunk win + l, and two ballots hold the window's bits
    // (tiles never straddle a window: 64 is a multiple of kChunks and the
    // first window starts at the first tile).
    int win = 0;
    bool have_win = false;
    uint64_t wg_far = 0, wv_far = 0;
    auto far_bits = [&](int t0, unsigned* wave_bits) {
      const int c0 = t0 / SUB;
      if (!have_win || c0 >= win + 64) {
        auto gap2 = [&](const double* cb, const double* cen, const double* half) {
          const double gx = fmax(fabs(cb[0] - cen[0]) - half[0] - cb[3], 0.0),
                       gy = fmax(fabs(cb[1] - cen[1]) - half[1] - cb[4], 0.0),
                       gz = fmax(fabs(cb[2] - cen[2]) - half[2] - cb[5], 0.0);
          return gx * gx + gy * gy + gz * gz;
        };
        const double wgc[3] = {cx * scale, cy * scale, cz * scale};  // radius units
        const double wgh[3] = {half_ext[0], half_ext[1], half_ext[2]};
        win = c0;
        have_win = true;
        const int64_t cc = static_cast<int64_t>(win) + (tid & 63);
        bool f = true, fw = true;
        if (cc * SUB < hi) {
          const double* cb = far_box + cc * 6;
          f = gap2(cb, wgc, wgh) > kFarGapS * kFarGapS;
          fw = f || gap2(cb, &wbox[tid >> 6][0], &wbox[tid >> 6][3]) > kFarGapS * kFarGapS;
        }
        wg_far = __ballot(f);
        wv_far = __ballot(fw);
      }
      const int sh = c0 - win;
      *wave_bits = static_cast<unsigned>((wv_far >> sh) & kAllFar);
      return static_cast<unsigned>((wg_far >> sh) & kAllFar);
    };
    // One far chunk [j0, j0 + SUB) from the far records, in scalar loads.
    auto far_chunk = [&](int j0) {
      if constexpr (kFarS) {
        constexpr int RS = FarRec<OUT>::kStride;
        const double* cb = far_box + static_cast<int64_t>(j0 / SUB) * 6;
        const double ox = -cb[0], oy = -cb[1], oz = -cb[2];  // positions in radius units
        Tgt<Real> tc[R];
        Real pc[R][3];
#pragma unroll
        for (int r = 0; r < R; ++r) {  // the targets relative to the chunk centre
          tc[r] = t[r];
          tc[r].x = static_cast<Real>(lds_x64[r * BLOCK + tid][0] + ox);
          tc[r].y = static_cast<Real>(lds_x64[r * BLOCK + tid][1] + oy);
          tc[r].z = static_cast<Real>(lds_x64[r * BLOCK + tid][2] + oz);
          pc[r][0] = tc[r].ay * tc[r].z - tc[r].az * tc[r].y;
          pc[r][1] = tc[r].az * tc[r].x - tc[r].ax * tc[r].z;
          pc[r][2] = tc[r].ax * tc[r].y - tc[r].ay * tc[r].x;
          if constexpr (kE3) {  // with the records' 3 m_j, e arrives as 3 e
            pc[r][0] *= kc.c3;
            pc[r][1] *= kc.c3;
            pc[r][2] *= kc.c3;
          }
        }
        const far_float_t* rec = (const far_float_t*)(far_rec + static_cast<int64_t>(j0) * RS);
        PairConst<Real> kf = kc;  // 2.5 in a register only while the chunk runs
        if constexpr (kE3) asm volatile("" : "+v"(kf.c25));
        // Sources in groups of kFarGroup, two register buffers: the loads of
        // the next group issue after the first pair of the current one (whose
        // wait covers only the current group's loads), so they overlap the
        // rest of the group's pairs.  Scalar loads return out of order, so a
        // wait always covers every load in flight; the scheduling barriers
        // keep the compiler from hoisting the next group's loads above it.
        constexpr int GR = kFarGroup * RS;
        static_assert(SUB % (2 * kFarGroup) == 0, "far group");
        float bA[GR], bB[GR];
        auto load_group = [&](float* b, int q) {
#pragma unroll
          for (int u = 0; u < GR; ++u) b[u] = rec[q * RS + u];
        };
        auto pair_of = [&](const float* f) {
          SrcR<Real, UNIF> sr;
          sr.x = f[0];
          sr.y = f[1];
          sr.z = f[2];
          sr.ax = f[3];
          sr.ay = f[4];
          sr.az = f[5];
          const Mom<Real> m{f[6], f[7], f[8]};
          const Real slo[3] = {Real(0), Real(0), Real(0)};  // unused by the far form
#pragma unroll
          for (int r = 0; r < R; ++r)
            pair_moment<Real, OUT, UNIF, kE3>(part[r], tc[r], pc[r], sr, m, slo, kf);
        };
        auto group = [&](const float* b, float* nb, int qn) {
          pair_of(b);
          __builtin_amdgcn_sched_barrier(0);
          load_group(nb, qn);  // the last group's prefetch reads the padding
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 1; u < kFarGroup; ++u) pair_of(b + u * RS);
        };
        load_group(bA, 0);
#pragma unroll 1
        for (int q = 0; q < SUB; q += 2 * kFarGroup) {
          group(bA, bB, q + kFarGroup);
          group(bB, bA, q + 2 * kFarGroup);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {  // u += x_i x G + H of the chunk, in FP64
          const double X = lds_x64[r * BLOCK + tid][0] + ox, Y = lds_x64[r * BLOCK + tid][1] + oy,
                       Z = lds_x64[r * BLOCK + tid][2] + oz;
          const double gx = part[r][kG + 0], gy = part[r][kG + 1], gz = part[r][kG + 2];
          constexpr double hs = kE3 ? 1.0 / 3.0 : 1.0;  // H of the records' 3 m_j
          tot[r][kU + 0] += (Y * gz - Z * gy) + static_cast<double>(part[r][kH + 0]) * hs;
          tot[r][kU + 1] += (Z * gx - X * gz) + static_cast<double>(part[r][kH + 1]) * hs;
          tot[r][kU + 2] += (X * gy - Y * gx) + static_cast<double>(part[r][kH + 2]) * hs;
          part[r][kH + 0] = part[r][kH + 1] = part[r][kH + 2] = Real(0);
          far_pse_fold<Real, OUT>(part[r], t[r]);
        }
      }
    };
    for (int t0 = lo; t0 < hi; t0 += TILE) {
      unsigned farbits = 0u, wfarbits = 0u;
      if constexpr (kFarS) farbits = far_bits(t0, &wfarbits);
      // A tile whose chunks are all far needs no staging, but its barrier
      // stays: it keeps the workgroup's waves within a tile of each other, so
      // that they read the same far records in the same stretch of time (the
      // scalar cache then serves most of them; random cloud of 2^20: 632 ->
      // 622 ms on one machine).
      bool far = false;
      if (kFarS && farbits == kAllFar) __syncthreads();
      if (!kFarS || farbits != kAllFar) {
        __syncthreads();
        Real bmn[3] = {Real(INFINITY), Real(INFINITY), Real(INFINITY)};
        Real bmx[3] = {Real(-INFINITY), Real(-INFINITY), Real(-INFINITY)};
#pragma unroll
        for (int k = tid; k < TILE; k += BLOCK) {
          const int j = t0 + k;
          SrcR<Real, UNIF> s;
          if (j < hi) {
            const Src8 g = src[j];
            const double x = (g.x - cx) * scale, y = (g.y - cy) * scale, z = (g.z - cz) * scale;
            s.x = static_cast<Real>(x);
            s.y = static_cast<Real>(y);
            s.z = static_cast<Real>(z);
            s.ax = static_cast<Real>(g.ax);
            s.ay = static_cast<Real>(g.ay);
            s.az = static_cast<Real>(g.az);
            if constexpr (kFarL) {
              s.p0 = static_cast<Real>(g.ay * z - g.az * y);
              s.p1 = static_cast<Real>(g.az * x - g.ax * z);
              lds_mz[k] = static_cast<Real>(g.ax * y - g.ay * x);
              bmn[0] = s.x < bmn[0] ? s.x : bmn[0]; bmx[0] = s.x > bmx[0] ? s.x : bmx[0];
              bmn[1] = s.y < bmn[1] ? s.y : bmn[1]; bmx[1] = s.y > bmx[1] ? s.y : bmx[1];
              bmn[2] = s.z < bmn[2] ? s.z : bmn[2]; bmx[2] = s.z > bmx[2] ? s.z : bmx[2];
            }
            if constexpr (kStageA64) {
              lds_a64[k][0] = g.ax;
              lds_a64[k][1] = g.ay;
              lds_a64[k][2] = g.az;
            }
            if constexpr (kStageLo)
              lds_lo[k] = make_float4(static_cast<float>(g.ax - static_cast<double>(s.ax)),
                                      static_cast<float>(g.ay - static_cast<double>(s.ay)),
                                      static_cast<float>(g.az - static_cast<double>(s.az)), 0.0f);
            if constexpr (!UNIF) {
              s.hs = static_cast<Real>(0.5 * g.s);
              s.is2 = static_cast<Real>(1.0 / (g.s * g.s));
              s.is3 = static_cast<Real>(1.0 / (g.s * g.s * g.s));
            }
          } else {  // padding: far away, zero strength -> contributes exactly 0
            // (PSE weights alpha_j - alpha_i = -alpha_i by t^9: at this distance
            // t^9 underflows to 0 in either precision; A stays finite)
            s.x = s.y = s.z = static_cast<Real>(std::is_same<Real, float>::value ? 1e18 : 1e40);
            s.ax = s.ay = s.az = Real(0);
            if constexpr (kStageA64) {
              lds_a64[k][0] = 0.0;
              lds_a64[k][1] = 0.0;
              lds_a64[k][2] = 0.0;
            }
            if constexpr (kStageLo) lds_lo[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if constexpr (kFarL) {
              s.p0 = s.p1 = Real(0);
              lds_mz[k] = Real(0);
            }
            if constexpr (!UNIF) {
              s.hs = Real(0.5);
              s.is2 = Real(1);
              s.is3 = Real(1);
            }
          }
          lds[k] = s;
        }
        if constexpr (kFarL) {  // the tile's bounding box, per wave
          const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const Real l = wave_min(bmn[c]), h = wave_max(bmx[c]);
            if (lane == 0) {
              red[c][wave] = static_cast<double>(l);
              red[3 + c][wave] = static_cast<double>(h);
            }
          }
        }
        __syncthreads();
        // Moment form for this tile: its gap to the targets exceeds kFarGap.
        if constexpr (kFar) {
          double gap2 = 0.0;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double l = red[c][0], h = red[3 + c][0];
#pragma unroll
            for (int w = 1; w < BLOCK / 64; ++w) {
              l = fmin(l, red[c][w]);
              h = fmax(h, red[3 + c][w]);
            }
            const double gp = fmax(fmax(l - half_ext[c], -half_ext[c] - h), 0.0);
            gap2 += gp * gp;
          }
          far = __builtin_amdgcn_readfirstlane(static_cast<int>(gap2 > kFarGap * kFarGap)) != 0;
        }
      }

      // Folds the sub-tile's FP32 partials: G into its FP64 totals, the others
      // into the Kahan pairs (or FP64 totals).  FAR: a far tile adds nothing
      // to u's partials (its velocity went to FP64 directly) and u's Kahan
      // pair is zero throughout such a tile, so its update is skipped.
      auto fold_sub = [&](auto far_tile) {
        constexpr bool FAR = decltype(far_tile)::value;
        // G leaves its partials every sub-tile also in the FP64 pass when the
        // far form is on: a far sub-tile's fold (x_i x G + H) must see only
        // its own G.
        if constexpr (S::str && (kFold32 || kMoment)) {
#pragma unroll
          for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {  // zero in FP64-G waves' near sub-tiles
              gsum[r][c] += static_cast<double>(part[r][kG + c]);
              part[r][kG + c] = Real(0);
            }
        }
        if constexpr (kFold32) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int q = 0; q < kNAcc; ++q) {
              if (!S::used(q) || (FAR && q >= kU && q < kU + 3)) continue;
              if constexpr (kComp) {  // Kahan: ksum + kcmp carries the tile sum
                const Real y = part[r][q] - kcmp[r][q];
                const Real z = ksum[r][q] + y;
                kcmp[r][q] = (z - ksum[r][q]) - y;
                ksum[r][q] = z;
              } else {
                tot[r][q] += static_cast<double>(part[r][q]);
              }
              part[r][q] = Real(0);
            }
          }
        }
      };

      if (far) {  // constant false unless kFar: a far tile sweeps its sub-tiles alone
#pragma unroll 1
        for (int sb = 0; sb < TILE; sb += SUB) {
#pragma unroll kUnrollFar
          for (int q = 0; q < SUB; ++q) {
            const SrcR<Real, UNIF> s = lds[sb + q];
            const Mom<Real> m{s.p0, s.p1, lds_mz[sb + q]};
            Real slo[3] = {Real(0), Real(0), Real(0)};
            if constexpr (S::pse && kStageLo) {
              const float4 l = lds_lo[sb + q];
              slo[0] = l.x;
              slo[1] = l.y;
              slo[2] = l.z;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
              if constexpr (kFar) pair_moment<Real, OUT, UNIF>(part[r], t[r], pm[r], s, m, slo, kc);
            }
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {  // u += x_i x G + H of the sub-tile, in FP64
            if constexpr (kFar) {
              const double X = lds_x64[r * BLOCK + tid][0], Y = lds_x64[r * BLOCK + tid][1],
                           Z = lds_x64[r * BLOCK + tid][2];
              const double gx = part[r][kG + 0], gy = part[r][kG + 1], gz = part[r][kG + 2];
              tot[r][kU + 0] += (Y * gz - Z * gy) + static_cast<double>(part[r][kH + 0]);
              tot[r][kU + 1] += (Z * gx - X * gz) + static_cast<double>(part[r][kH + 1]);
              tot[r][kU + 2] += (X * gy - Y * gx) + static_cast<double>(part[r][kH + 2]);
              part[r][kH + 0] = part[r][kH + 1] = part[r][kH + 2] = Real(0);
              far_pse_fold<Real, OUT>(part[r], t[r]);
            }
          }
          fold_sub(std::integral_constant<bool, kFar>{});
        }
      } else {
#pragma unroll 1
        for (int sb = 0; sb < TILE; sb += SUB) {
          if constexpr (kFarS) {
            if ((wfarbits >> (sb / SUB)) & 1u) {
              if (t0 + sb < hi) {
                far_chunk(t0 + sb);
                fold_sub(std::true_type{});
              }
              continue;
            }
          }
          if constexpr (G64) {
#pragma unroll kUnroll
            for (int q = 0; q < SUB; ++q) {
              const SrcR<Real, UNIF> s = lds[sb + q];
              const double aj64[3] = {lds_a64[sb + q][0], lds_a64[sb + q][1],
                                      lds_a64[sb + q][2]};
              Real slo[3] = {Real(0), Real(0), Real(0)};
              if constexpr (kStageLo) {
                const float4 l = lds_lo[sb + q];
                slo[0] = l.x;
                slo[1] = l.y;
                slo[2] = l.z;
              }
#pragma unroll
              for (int r = 0; r < R; ++r)
                pair_update<Real, OUT, UNIF, true>(part[r], gsum[r], t[r], s, aj64, a64t[r], slo,
                                                   eps_r2, kc);
            }
          } else if (!(kSelfMask && self_sub(t0 + sb))) {
#pragma unroll kUnroll
            for (int q = 0; q < SUB; ++q) {
              const SrcR<Real, UNIF> s = lds[sb + q];
              Real slo[3] = {Real(0), Real(0), Real(0)};
              if constexpr (kStageLo) {
                const float4 l = lds_lo[sb + q];
                slo[0] = l.x;
                slo[1] = l.y;
                slo[2] = l.z;
              }
#pragma unroll
              for (int r = 0; r < R; ++r)
                pair_update<Real, OUT, UNIF, false>(part[r], gsum[r], t[r], s, nullptr, nullptr,
                                                    slo, eps_r2, kc);
            }
          } else {  // the sub-tile holding this wave's targets: once per tile
#pragma unroll 1
            for (int q = 0; q < SUB; ++q) {
              const SrcR<Real, UNIF> s = lds[sb + q];
              Real slo[3] = {Real(0), Real(0), Real(0)};
              if constexpr (kStageLo) {
                const float4 l = lds_lo[sb + q];
                slo[0] = l.x;
                slo[1] = l.y;
                slo[2] = l.z;
              }
              const int j = t0 + sb + q;
#pragma unroll
              for (int r = 0; r < R; ++r)
                pair_update<Real, OUT, UNIF, false, kSelfMask>(
                    part[r], gsum[r], t[r], s, nullptr, nullptr, slo, eps_r2, kc,
                    j == base + r * BLOCK + tid);
            }
          }
          fold_sub(std::false_type{});
        }
      }
    }
    // The Kahan pairs carry the whole sweep: (sum, compensation) keeps the
    // sum of the sub-tile partials to about twice FP32's precision, so they
    // reach the FP64 totals once, here.  (Folded per tile, three totals the
    // allocator keeps in scratch were reloaded and stored for every tile:
    // 3e10 bytes written per launch at N = 2^20.)
    if constexpr (kComp) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < kNAcc; ++q) {
          if (!S::used(q)) continue;
          tot[r][q] += static_cast<double>(ksum[r][q]) - static_cast<double>(kcmp[r][q]);
          ksum[r][q] = Real(0);
          kcmp[r][q] = Real(0);
        }
    }
  };
  if (g64)  // constant unless kBothG
    sweep(std::true_type{});
  else
    sweep(std::false_type{});
  if constexpr (!kFold32) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int q = 0; q < kNAcc; ++q) tot[r][q] += static_cast<double>(part[r][q]);
      if constexpr (S::str) {
#pragma unroll
        for (int c = 0; c < 3; ++c) gsum[r][c] += static_cast<double>(part[r][kG + c]);
      }
    }
  }
  // The FP64 sums of G (coherent waves of an FP32 pass, the FP64 pass) hold
  // the self pair's g(0) alpha_i; the FP32 partials leave it out (pair_update,
  // SELF).  The reference sums the stretching over j != i, and alpha_i x G
  // cancels the term only to a rounding of g(0) |alpha_i|^2 (a lone particle
  // came out at 5e-15, not 0), so it is subtracted here, as the same rounded
  // product the sweep added, for the targets whose index lies in [lo, hi).
  if constexpr (S::str) {
    if (!kFold32 || g64) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k = base + r * BLOCK + tid;
        if (!valid[r] || k < lo || k >= hi) continue;
        const double g0 = static_cast<double>(self_g<Real, UNIF>(t[r].hs, kc));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double term = g0 * a64t[r][c];
          asm volatile("" : "+v"(term));  // rounded, not fused into the subtraction
          gsum[r][c] -= term;
        }
      }
    }
  }

  // ---- epilogue ----------------------------------------------------------
  if (split_out) {  // one split of several: raw sums, coalesced by sorted index
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!valid[r]) continue;
      const int k = base + r * BLOCK + tid;
#pragma unroll
      for (int q = 0; q < kNAcc; ++q)
        if (S::used(q)) split_out[S::packed(q) * pld + k] = tot[r][q];
      if constexpr (S::str) {
#pragma unroll
        for (int c = 0; c < 3; ++c) split_out[(S::nacc + c) * pld + k] = gsum[r][c];
      }
    }
    return;
  }
  double ens_local = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (!valid[r]) continue;
    const int k = base + r * BLOCK + tid;
    ens_local += store_target<OUT, UNIF>(a, src[k], perm ? perm[k] : k, tot[r], gsum[r], scale);
  }
  if constexpr (S::ens) {
    __syncthreads();
    const double bs = block_sum<BLOCK>(ens_local, &red[0][0]);
    if (tid == 0) *ens_slot = bs;
  }
}

// Minimum waves per SIMD the register allocation must allow.  The
// velocity/stretching FP32 pass with 512 lanes x one target holds three
// loops (FP64 G, FP32 G near, moment form far); left alone the allocator
// settles just above 128 VGPRs, which would halve the workgroups per CU.
// With the far form in scalar-loaded records (FARS) the single-set kernel
// fits 80 VGPRs: six waves per SIMD, three workgroups of 512 lanes per CU
// (random cloud of 2^20: 638 ms against 663 ms at four waves and 659 ms at
// five; the near FP64-G loop then keeps a few values in scratch, rings
// unchanged).
#ifndef VPM_W256
#define VPM_W256 3
#endif
template <typename Real, unsigned OUT, int BLOCK, int R, int TILE, bool FARS = false>
constexpr int min_waves() {
  using S = Slots<OUT>;
  if (!std::is_same<Real, float>::value || R != 1 || TILE > 512 || !S::str) return 1;
  // 512 lanes with far records: six waves per SIMD (three workgroups fit the
  // LDS of 512-source tiles), eight with 256-source tiles
  if (BLOCK == 512) return (!S::vort && !S::ens) ? (FARS ? (TILE <= 256 ? 8 : 6) : 4) : 1;
  // velocity + stretching + PSE with 256 lanes: five waves per SIMD (96
  // VGPRs, a few spills outside the pair loops; ring of 131,072: 17.6 ->
  // 16.9 ms against four); the LES sets of 256 lanes: three workgroups per CU
  if (BLOCK == 256 && S::pse && !S::vort && !S::ens) return 5;
  if (BLOCK == 256) return (S::pse || S::vort || S::ens) ? VPM_W256 : 1;
  return 1;
}

// Single-set FP32 passes with the moment form and chunks of 64 or more
// sources read far records (shapes with 32-source sub-tiles stage far tiles
// in LDS).
template <typename Real, unsigned OUT, bool UNIF, int SUB>
constexpr bool uses_far_records() {
  return std::is_same<Real, float>::value && UNIF && (OUT & VPM_OUT_VEL) &&
         (OUT & VPM_OUT_STRETCH) && SUB >= 64;
}

// grid (target blocks, splits)
template <typename Real, unsigned OUT, bool UNIF, int BLOCK, int R, int TILE, int SUB, int COMP>
__global__ __launch_bounds__(BLOCK, (min_waves<Real, OUT, BLOCK, R, TILE,
                                               uses_far_records<Real, OUT, UNIF, SUB>()>()))
void fused_kernel(
    FusedArgs a) {
  const int k = static_cast<int>(blockIdx.y);
  const int lo = k * a.chunk;
  const int hi = a.chunk < a.n - lo ? lo + a.chunk : a.n;
  double* part = a.part ? a.part + static_cast<int64_t>(k) * Slots<OUT>::nslot * a.pld : nullptr;
  fused_block<Real, OUT, UNIF, BLOCK, R, TILE, SUB, COMP,
              uses_far_records<Real, OUT, UNIF, SUB>()>(
      a, a.n, a.src, a.perm, static_cast<int>(blockIdx.x) * (BLOCK * R), a.ens_part + blockIdx.x,
      lo, hi, part, a.pld, a.far_rec, a.far_box);
}

// Adds the splits' raw sums in split order (deterministic) and stores; one
// enstrophy partial per workgroup.
template <unsigned OUT, bool UNIF>
__global__ __launch_bounds__(256) void split_reduce_kernel(FusedArgs a, int nsplit) {
  using S = Slots<OUT>;
  __shared__ double red[4];
  const int k = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  double ens_local = 0.0;
  if (k < a.n) {
    double T[kNAcc], G[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kNAcc; ++q) T[q] = 0.0;
    for (int sp = 0; sp < nsplit; ++sp) {
      const double* P = a.part + static_cast<int64_t>(sp) * S::nslot * a.pld;
#pragma unroll
      for (int q = 0; q < kNAcc; ++q)
        if (S::used(q)) T[q] += P[S::packed(q) * a.pld + k];
      if constexpr (S::str) {
#pragma unroll
        for (int c = 0; c < 3; ++c) G[c] += P[(S::nacc + c) * a.pld + k];
      }
    }
    const double scale = UNIF ? 1.0 / a.src[0].s : 1.0;
    ens_local = store_target<OUT, UNIF>(a, a.src[k], a.perm ? a.perm[k] : k, T, G, scale);
  }
  if constexpr (S::ens) {
    const double bs = block_sum<256>(ens_local, red);
    if (threadIdx.x == 0) a.ens_part[blockIdx.x] = bs;
  }
}

// ------------------------------------------------------------------ configs
struct TileCfg {
  int block, r, tile, sub, comp;
};
// Configs 0 and 1 exist for every output set; configs 2.. for the hot
// output sets (VEL|STRETCH, VEL|STRETCH|PSE, VEL|STRETCH|VORT|ENSTROPHY) with
// uniform sigma.
constexpr TileCfg kCfgs[] = {
    {512, 1, 512, 64, 0},  {256, 1, 512, 64, 0},  {256, 2, 256, 32, 1}, {256, 1, 256, 64, 0},
    {256, 1, 256, 64, 1},  {512, 1, 512, 128, 0}, {512, 1, 1024, 64, 0}, {128, 1, 256, 64, 0},
    {256, 2, 512, 64, 0},  {256, 2, 256, 32, 0},  {512, 1, 256, 128, 0}, {256, 1, 256, 128, 1},
};

constexpr bool hot_mask(unsigned m) { return m == 3u || m == 7u || m == 27u; }

// Shapes for tile_config = -1, the autotuner's winners on one MI355X:
//  * velocity + stretching, uniform sigma: config 5 (512 lanes x one target,
//    128-source FP32 partials: half the folds of config 0; random cloud of
//    2^20: 685 ms against 693 ms, rel-L2 5.0e-7 against 4.5e-7);
//  * velocity + stretching + PSE, uniform sigma: config 11 (256 lanes,
//    256-source tiles, 128-source chunks, partials folded straight into
//    FP64; ring of 131072 with far records and five waves per SIMD: 16.1 ms
//    against 16.4-16.5 for config 4, its 64-source-chunk twin, 17.0 with
//    512-source tiles, 17.2 for config 3 and 18.6 for config 0);
//  * other sets: config 1 (256 lanes) with PSE, vorticity or enstrophy --
//    their kernels need more than 128 VGPRs, where a 512-lane workgroup would
//    run alone on its CU -- and config 0 without.
constexpr unsigned kHeavyOutputs = VPM_OUT_PSE | VPM_OUT_VORT | VPM_OUT_ENSTROPHY;
constexpr int default_config(unsigned kernel_mask, bool unif) {
  return (unif && kernel_mask == 3u)      ? 5
         : (unif && kernel_mask == 7u)    ? 11
         : (kernel_mask & kHeavyOutputs) ? 1
                                          : 0;
}

// The batched launch: config 4 with PSE or the LES outputs (96 sets, 3.3e6
// particles, RK4 with PSE and LES, far tiles in the moment form: 882 ms per
// step against 893 for config 1, 896 for config 3 and 956 for config 0 on
// one box), config 0 for velocity + stretching alone.
constexpr int default_batch_config(unsigned batch_mask) {
  return (batch_mask & kHeavyOutputs) ? 4 : 0;
}

// The configuration the single-set launch really uses for a request.
constexpr int effective_config(unsigned kernel_mask, bool unif, int requested) {
  return requested < 0 ? default_config(kernel_mask, unif)
         : (requested <= 1 || (unif && hot_mask(kernel_mask))) ? requested
                                                                : default_config(kernel_mask, unif);
}
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);
constexpr TileCfg kCfgF64 = {256, 1, 256, 256, 1};


template <typename Real, unsigned OUT, bool UNIF, int BLOCK, int R, int TILE, int SUB, int COMP>
int launch_cfg(const FusedArgs& a0, hipStream_t st, int* nblocks) {
  const int blocks = static_cast<int>(ceil_div(a0.n, BLOCK * R));
  *nblocks = blocks;
  if (blocks == 0) return VPM_OK;
  auto kern = fused_kernel<Real, OUT, UNIF, BLOCK, R, TILE, SUB, COMP>;
  FusedArgs a = a0;
  const int ntiles = static_cast<int>(ceil_div(a.n, TILE));
  const int64_t pld = static_cast<int64_t>(blocks) * BLOCK * R;
  const size_t per_split = static_cast<size_t>(Slots<OUT>::nslot) * pld * sizeof(double);
  int ks = 1;
  if (a.part) {
    const size_t by_cap = a.part_cap / per_split;
    int kmax = max_splits_for(blocks, ntiles);
    if (by_cap < static_cast<size_t>(kmax)) kmax = static_cast<int>(by_cap);
    if (a.max_splits > 0 && a.max_splits < kmax) kmax = a.max_splits;
    if (kmax > 1) ks = choose_splits(blocks, ntiles, kmax, resident_blocks(kern, BLOCK));
  }
  if (ks > 1) {
    const int per = (ntiles + ks - 1) / ks;  // tiles per split; ks is effective
    a.chunk = per * TILE;
    a.pld = pld;
  } else {
    a.part = nullptr;
    a.chunk = a.n;
    a.pld = 0;
  }
  if constexpr (uses_far_records<Real, OUT, UNIF, SUB>()) {
    if (!a.far_rec || !a.far_box) return fail(VPM_ERR_WORKSPACE, "fused pass: no far records");
    static_assert(SUB == kMinFarChunk || SUB == kMaxFarChunk, "far chunk");
    hipLaunchKernelGGL(far_records_kernel<SUB>, dim3(static_cast<unsigned>(ceil_div(a.n, SUB))),
                       dim3(SUB), 0, st, a.src, a.n, FarRec<OUT>::kStride,
                       FarRec<OUT>::kMomScale, const_cast<float*>(a.far_rec),
                       const_cast<double*>(a.far_box));
    VPM_LAUNCH_CHECK("far_records_kernel");
  }
  hipLaunchKernelGGL(kern, dim3(blocks, ks), dim3(BLOCK), 0, st, a);
  VPM_LAUNCH_CHECK("fused_kernel");
  if (ks > 1) {
    const int rb = static_cast<int>(ceil_div(a.n, 256));
    hipLaunchKernelGGL((split_reduce_kernel<OUT, UNIF>), dim3(rb), dim3(256), 0, st, a, ks);
    VPM_LAUNCH_CHECK("split_reduce_kernel");
    *nblocks = rb;
  }
  return VPM_OK;
}

template <unsigned OUT, bool UNIF>
int launch_f32(const FusedArgs& a, int cfg, hipStream_t st, int* nb) {
  switch (cfg) {
#define VPM_CFG(I)                                                                          \
  case I:                                                                                    \
    return launch_cfg<float, OUT, UNIF, kCfgs[I].block, kCfgs[I].r, kCfgs[I].tile,          \
                      kCfgs[I].sub, kCfgs[I].comp>(a, st, nb);
    VPM_CFG(0) VPM_CFG(1)
    default:
      break;
  }
  if constexpr (UNIF && hot_mask(OUT)) {
    switch (cfg) {
      VPM_CFG(2) VPM_CFG(3) VPM_CFG(4) VPM_CFG(5) VPM_CFG(6) VPM_CFG(7) VPM_CFG(8) VPM_CFG(9)
      VPM_CFG(10) VPM_CFG(11)
      default:
        break;
    }
#undef VPM_CFG
  }
  return VPM_ERR_ARG;  // callers pass effective_config()
}

template <unsigned OUT>
int launch_mask(const FusedArgs& a, int prec, bool unif, int cfg, hipStream_t st, int* nb) {
  if (prec == VPM_PREC_F32C) return unif ? launch_f32<OUT, true>(a, cfg, st, nb)
                                         : launch_f32<OUT, false>(a, cfg, st, nb);
  if (unif)
    return launch_cfg<double, OUT, true, kCfgF64.block, kCfgF64.r, kCfgF64.tile, kCfgF64.sub,
                      kCfgF64.comp>(a, st, nb);
  return launch_cfg<double, OUT, false, kCfgF64.block, kCfgF64.r, kCfgF64.tile, kCfgF64.sub,
                    kCfgF64.comp>(a, st, nb);
}

// Output sets with their own instantiation; any other non-empty set runs the
// smallest superset and the extra outputs are not stored.
constexpr unsigned kMasks[] = {1u, 2u, 3u, 4u, 7u, 8u, 16u, 24u, 27u, 31u};

unsigned kernel_mask(unsigned want) {
  for (unsigned m : kMasks)
    if ((m & want) == want) return m;
  return 31u;
}

int dispatch(const FusedArgs& a, unsigned mask, int prec, bool unif, int cfg, hipStream_t st,
             int* nb) {
  switch (mask) {
    case 1u: return launch_mask<1u>(a, prec, unif, cfg, st, nb);
    case 2u: return launch_mask<2u>(a, prec, unif, cfg, st, nb);
    case 3u: return launch_mask<3u>(a, prec, unif, cfg, st, nb);
    case 4u: return launch_mask<4u>(a, prec, unif, cfg, st, nb);
    case 7u: return launch_mask<7u>(a, prec, unif, cfg, st, nb);
    case 8u: return launch_mask<8u>(a, prec, unif, cfg, st, nb);
    case 16u: return launch_mask<16u>(a, prec, unif, cfg, st, nb);
    case 24u: return launch_mask<24u>(a, prec, unif, cfg, st, nb);
    case 27u: return launch_mask<27u>(a, prec, unif, cfg, st, nb);
    default: return launch_mask<31u>(a, prec, unif, cfg, st, nb);
  }
}

// ------------------------------------------------------------------ workspace
struct WsLayout {
  size_t bbox_part, bbox, keys_in, keys_out, idx_in, perm, src8, ens_part, scratch_vec, split,
      far_rec, far_box, cub, total;
  size_t cub_bytes, split_bytes;
};

// Far records of an FP32 pass with a uniform radius and velocity +
// stretching: records for whole chunks of the largest size (widest record),
// boxes for chunks of the smallest.
bool wants_far_records(const vpm_forward_opts_t* o) {
  if (!o || o->precision != VPM_PREC_F32C || o->sigma_mode != VPM_SIGMA_UNIFORM) return false;
  const unsigned m = kernel_mask(o->outputs);  // the instantiation that runs
  return (m & VPM_OUT_VEL) && (m & VPM_OUT_STRETCH);
}

// FP64 partial slots per target of a source split (see Slots::nslot)
int split_slots(unsigned mask) {
  int c = 0;
  for (unsigned b = 1; b <= VPM_OUT_ENSTROPHY; b <<= 1) c += (mask & b) ? 3 : 0;
  return c + ((mask & VPM_OUT_STRETCH) ? 3 : 0);
}

// Room for the largest source split launch_cfg may choose for this call.
size_t split_capacity(int64_t n, const vpm_forward_opts_t* o) {
  if (!o || n <= 0) return 0;
  int br = kCfgF64.block * kCfgF64.r, tile = kCfgF64.tile;
  if (o->precision == VPM_PREC_F32C) {
    const int c = effective_config(kernel_mask(o->outputs), o->sigma_mode == VPM_SIGMA_UNIFORM,
                                   o->tile_config < kNumCfgs ? o->tile_config : -1);
    br = kCfgs[c].block * kCfgs[c].r;
    tile = kCfgs[c].tile;
  }
  const int64_t blocks = ceil_div(n, br);
  const int kmax = max_splits_for(blocks, ceil_div(n, tile));
  if (kmax < 2) return 0;
  return static_cast<size_t>(kmax) * split_slots(kernel_mask(o->outputs)) *
         static_cast<size_t>(blocks * br) * sizeof(double);
}

// The hipCUB scratch comes last, so the other offsets do not depend on it;
// calls that do not sort pass need_cub = false and skip the size query.
// split_cap: at most this many bytes for the source-split partials (the
// launch then picks fewer splits).
int layout(int64_t n, const vpm_forward_opts_t* o, WsLayout* L, bool need_cub = true,
           size_t split_cap = SIZE_MAX) {
  const size_t N = static_cast<size_t>(n > 0 ? n : 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align256(bytes);
    return at;
  };
  L->bbox_part = take(kBboxBlocks * 6 * sizeof(double));
  L->bbox = take(8 * sizeof(double));
  L->keys_in = take(N * sizeof(uint32_t));
  L->keys_out = take(N * sizeof(uint32_t));
  L->idx_in = take(N * sizeof(int));
  L->perm = take(N * sizeof(int));
  L->src8 = take(N * sizeof(Src8));
  L->ens_part = take((N / 64 + 2) * sizeof(double));
  // Unrequested outputs of a superset kernel land here: 3 vectors of N.
  L->scratch_vec = take(3 * N * sizeof(double));
  L->split_bytes = std::min(split_capacity(n, o), split_cap);
  L->split = take(L->split_bytes);
  const bool far = wants_far_records(o);
  const int64_t nn = n > 0 ? n : 1;
  // + one group of records read ahead past the last chunk
  L->far_rec = take(far ? (static_cast<size_t>(ceil_div(nn, kMaxFarChunk)) * kMaxFarChunk +
                           kFarGroup) * kMaxFarStride * sizeof(float)
                        : 0);
  L->far_box = take(far ? static_cast<size_t>(ceil_div(nn, kMinFarChunk)) * 6 * sizeof(double) : 0);
  size_t cub_bytes = 0;
  if (need_cub && (!o || !o->presorted)) {
    const int rc = sort_temp_bytes(n, &cub_bytes);
    if (rc) return rc;
  }
  L->cub_bytes = cub_bytes;
  L->cub = take(cub_bytes + 16);
  L->total = off;
  return VPM_OK;
}

// ------------------------------------------------------------------ batched
// Many particle sets ("instances") of uneven size in one persistent launch.
// The sets are concatenated (set s owns caller indices [off_s, off_s + n_s))
// and sorted per set (set id in the high bits of a 64-bit Morton key), so a
// set stays contiguous after sorting.  Work items are (set, target block);
// the host lists them largest set first and workgroups pull them from a
// counter, so the long items start first and the tail is short.
struct BatchArgs {
  FusedArgs f;            // outputs in caller order; f.ens_part: one slot per item;
                          // f.far_rec / f.far_box: every set's far records
  const Src8* src;        // sorted, concatenated
  const int* perm;        // sorted position -> caller index
  const int* inst_off;    // [S+1]
  const int* item_inst;   // [items]
  const int* item_block;  // [items]
  const int* set_chunk0;  // [S+1] first far-record chunk of each set (SUB-source chunks)
  int nsets;
  int n_total;            // particles over all sets
  int n_items;
  int* counter;
};

template <typename Real, unsigned OUT, bool UNIF, int BLOCK, int R, int TILE, int SUB, int COMP>
__global__ __launch_bounds__(BLOCK, (min_waves<Real, OUT, BLOCK, R, TILE,
                                               uses_far_records<Real, OUT, UNIF, SUB>()>()))
void fused_batch_kernel(BatchArgs b) {
  constexpr bool kFarS = uses_far_records<Real, OUT, UNIF, SUB>();
  __shared__ int item_s;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0)
      item_s = __hip_atomic_fetch_add(b.counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int item = item_s;
    // every workgroup ends once the list is drained (or on a counter that is
    // not a valid item: its workspace was overwritten)
    if (item < 0 || item >= b.n_items) break;
    const int inst = b.item_inst[item];
    const int off = b.inst_off[inst];
    const int n = b.inst_off[inst + 1] - off;
    const int64_t c0 = kFarS ? b.set_chunk0[inst] : 0;
    fused_block<Real, OUT, UNIF, BLOCK, R, TILE, SUB, COMP, kFarS>(
        b.f, n, b.src + off, b.perm + off, b.item_block[item] * (BLOCK * R), b.f.ens_part + item,
        0, n, nullptr, 0, kFarS ? b.f.far_rec + c0 * SUB * FarRec<OUT>::kStride : nullptr,
        kFarS ? b.f.far_box + c0 * 6 : nullptr);
  }
}

// set_chunk0[s] = sum over sets t < s of ceil(n_t / ch); set_chunk0[S] = the
// total.  One workgroup: each thread sums a contiguous range of sets, then a
// scan of the per-thread sums in LDS.
__global__ __launch_bounds__(1024) void set_chunks_kernel(const int* __restrict__ off, int nsets,
                                                          int ch, int* __restrict__ chunk0) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nsets + 1023) / 1024;
  const int b = t * per, e = min(nsets, b + per);
  int sum = 0;
  for (int s = b; s < e; ++s) sum += (off[s + 1] - off[s] + ch - 1) / ch;
  part[t] = sum;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int k = 0; k < 1024; ++k) {
      const int v = part[k];
      part[k] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int acc = part[t];
  for (int s = b; s < e; ++s) {
    chunk0[s] = acc;
    acc += (off[s + 1] - off[s] + ch - 1) / ch;
  }
  if (t == 1023) chunk0[nsets] = acc;
}

// The far records of every set: workgroup g builds chunk g (of all sets'
// chunks; the owning set by binary search in set_chunk0), like
// far_records_kernel for one set.
template <int CH>
__global__ __launch_bounds__(CH) void set_far_records_kernel(const Src8* __restrict__ src,
                                                             const int* __restrict__ off,
                                                             const int* __restrict__ chunk0,
                                                             int nsets, int stride,
                                                             double mscale,
                                                             float* __restrict__ rec,
                                                             double* __restrict__ box) {
  const int g = blockIdx.x;
  if (g >= chunk0[nsets]) return;  // workgroup-uniform: no barrier is skipped by part of it
  int lo = 0, hi = nsets - 1;      // the set s with chunk0[s] <= g < chunk0[s+1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (chunk0[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const int s = lo;
  const int n = off[s + 1] - off[s];
  const int c = g - chunk0[s];
  far_records_chunk<CH>(src + off[s], n, c, stride, mscale, src[off[s]].s,
                        rec + (int64_t)g * CH * stride, box + (int64_t)g * 6);
}

__global__ void counter_reset_kernel(int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = 0;
}

// per-set bounding boxes: one block per set
__global__ __launch_bounds__(256) void set_bbox_kernel(const double* __restrict__ pos, int64_t ld,
                                                       const int* __restrict__ off,
                                                       double* __restrict__ bbox) {
  __shared__ double red[6][4];
  const int s = blockIdx.x;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += 256)
    for (int c = 0; c < 3; ++c) {
      const double v = pos[c * ld + i];
      mn[c] = fmin(mn[c], v);
      mx[c] = fmax(mx[c], v);
    }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = 0; c < 3; ++c) {
    const double lo = wave_min(mn[c]), hi = wave_max(mx[c]);
    if (lane == 0) {
      red[c][wave] = lo;
      red[3 + c][wave] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = red[threadIdx.x][0];
    for (int w = 1; w < 4; ++w)
      v = threadIdx.x < 3 ? fmin(v, red[threadIdx.x][w]) : fmax(v, red[threadIdx.x][w]);
    bbox[s * 6 + threadIdx.x] = v;
  }
}

__global__ __launch_bounds__(256) void set_keys_kernel(const double* __restrict__ pos, int64_t ld,
                                                       int n, const int* __restrict__ part_inst,
                                                       const double* __restrict__ bbox,
                                                       uint64_t* __restrict__ keys,
                                                       int* __restrict__ idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int s = part_inst[i];
  uint32_t q[3];
  for (int c = 0; c < 3; ++c) {
    const double lo = bbox[s * 6 + c], ext = bbox[s * 6 + 3 + c] - lo;
    const double f = ext > 0.0 ? (pos[c * ld + i] - lo) / ext : 0.0;
    int v = static_cast<int>(f * 1023.0);
    v = v < 0 ? 0 : (v > 1023 ? 1023 : v);
    q[c] = static_cast<uint32_t>(v);
  }
  const uint32_t m = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  keys[i] = (static_cast<uint64_t>(s) << 30) | m;
  idx[i] = i;
}

// per-set enstrophy: one block per set sums the partials of its items
__global__ __launch_bounds__(256) void set_enstrophy_kernel(const double* __restrict__ part,
                                                            const int* __restrict__ first,
                                                            const int* __restrict__ count,
                                                            double scale,
                                                            double* __restrict__ out) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  double v = 0.0;
  for (int k = threadIdx.x; k < count[s]; k += 256) v += part[first[s] + k];
  v = block_sum<256>(v, red);
  if (threadIdx.x == 0) out[s] = v * scale;
}

// Dynamic LES filter per set, vpmModule.py:887-962, one block per set:
// C_nu recursion from the set's enstrophy, mean vorticity, nu_t candidates,
// the max > 1e-9 test and the commit; also writes nu_eff = nu + nu_t.
__global__ __launch_bounds__(256) void set_les_kernel(const int* __restrict__ off, int64_t ld,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ radii,
                                                      const double* __restrict__ ens,
                                                      double* __restrict__ state,
                                                      const double* __restrict__ nu,
                                                      double* __restrict__ nu_t,
                                                      double* __restrict__ nu_eff,
                                                      double* __restrict__ cand) {
  __shared__ double red[4];
  __shared__ double sh[4];
  const int s = blockIdx.x;
  const int b = off[s], e = off[s + 1], n = e - b;
  double* st = state + 3 * s;
  if (threadIdx.x == 0) {
    const double en = sqrt(ens[s]);
    double ref = st[2] != 0.0 ? st[1] : en;
    ref = 0.1 * en + (1 - 0.1) * ref;
    double c = st[0] * en / (ref + 1e-9);
    c = fmax(1e-9, fmin(c, 0.0761));
    st[0] = c;
    st[1] = ref;
    st[2] = 1.0;
    sh[3] = c;
  }
  double m[3] = {0, 0, 0};
  for (int i = b + threadIdx.x; i < e; i += 256)
    for (int c = 0; c < 3; ++c) m[c] += w[c * ld + i];
  for (int c = 0; c < 3; ++c) {
    const double v = block_sum<256>(m[c], red);
    if (threadIdx.x == 0) sh[c] = v / n;
  }
  __syncthreads();
  const double mx = sh[0], my = sh[1], mz = sh[2], cnu = sh[3];
  const double mm = sqrt(mx * mx + my * my + mz * mz);
  const double ux = mx / mm, uy = my / mm, uz = mz / mm;
  const double beta0 = kPi / 4.0;
  double vmax = 0.0;
  for (int i = b + threadIdx.x; i < e; i += 256) {
    const double wx = w[i], wy = w[ld + i], wz = w[2 * ld + i];
    const double wm = sqrt(wx * wx + wy * wy + wz * wz);
    const double c = (wx / wm) * ux + (wy / wm) * uy + (wz / wm) * uz;
    const double beta = acos(fmin(fmax(c, -1.0), 1.0));
    const double psi = (beta0 <= beta && beta <= kPi - beta0) ? 1.0 : 0.0;
    const double cr = cnu * radii[i];
    const double v = psi * (cr * cr) * wm;
    cand[i] = v;
    vmax = fmax(vmax, v);
  }
  vmax = wave_max(vmax);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmax;
  __syncthreads();
  const double top = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  for (int i = b + threadIdx.x; i < e; i += 256) {
    if (top > 1e-9) nu_t[i] = cand[i];
    nu_eff[i] = nu[i] + nu_t[i];
  }
}

template <typename Real, unsigned OUT, int BLOCK, int R, int TILE, int SUB, int COMP>
int launch_batch_cfg(const BatchArgs& b, int grid, hipStream_t st) {
  hipLaunchKernelGGL(counter_reset_kernel, dim3(1), dim3(64), 0, st, b.counter);
  VPM_LAUNCH_CHECK("counter_reset_kernel");
  if (b.n_items == 0) return VPM_OK;
  if (grid <= 0) {
    int per_cu = 0, dev = 0, cus = 0;
    VPM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, fused_batch_kernel<Real, OUT, true, BLOCK, R, TILE, SUB, COMP>, BLOCK, 0));
    VPM_HIP_TRY(hipGetDevice(&dev));
    VPM_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::max(1, per_cu) * std::max(1, cus);
  }
  grid = std::min(grid, b.n_items);
  if constexpr (uses_far_records<Real, OUT, true, SUB>()) {
    static_assert(SUB == kMinFarChunk || SUB == kMaxFarChunk, "far chunk");
    if (!b.f.far_rec || !b.f.far_box || !b.set_chunk0)
      return fail(VPM_ERR_WORKSPACE, "batch pass: no far records");
    int* chunk0 = const_cast<int*>(b.set_chunk0);
    hipLaunchKernelGGL(set_chunks_kernel, dim3(1), dim3(1024), 0, st, b.inst_off, b.nsets, SUB,
                       chunk0);
    VPM_LAUNCH_CHECK("set_chunks_kernel");
    const int64_t max_chunks = ceil_div(b.n_total, SUB) + b.nsets;  // bound; extra groups exit
    hipLaunchKernelGGL(set_far_records_kernel<SUB>, dim3(static_cast<unsigned>(max_chunks)),
                       dim3(SUB), 0, st, b.src, b.inst_off, b.set_chunk0, b.nsets,
                       FarRec<OUT>::kStride, FarRec<OUT>::kMomScale,
                       const_cast<float*>(b.f.far_rec), const_cast<double*>(b.f.far_box));
    VPM_LAUNCH_CHECK("set_far_records_kernel");
  }
  hipLaunchKernelGGL((fused_batch_kernel<Real, OUT, true, BLOCK, R, TILE, SUB, COMP>), dim3(grid),
                     dim3(BLOCK), 0, st, b);
  VPM_LAUNCH_CHECK("fused_batch_kernel");
  return VPM_OK;
}

template <unsigned OUT>
int launch_batch_mask(const BatchArgs& b, int prec, int cfg, int grid, hipStream_t st) {
  if (prec == VPM_PREC_F64)
    return launch_batch_cfg<double, OUT, kCfgF64.block, kCfgF64.r, kCfgF64.tile, kCfgF64.sub,
                            kCfgF64.comp>(b, grid, st);
#define VPM_BCFG(I)                                                                            \
  case I:                                                                                      \
    return launch_batch_cfg<float, OUT, kCfgs[I].block, kCfgs[I].r, kCfgs[I].tile,             \
                            kCfgs[I].sub, kCfgs[I].comp>(b, grid, st);
  switch (cfg) {
    VPM_BCFG(1) VPM_BCFG(2) VPM_BCFG(3) VPM_BCFG(4) VPM_BCFG(5) VPM_BCFG(6) VPM_BCFG(7)
    VPM_BCFG(8) VPM_BCFG(9) VPM_BCFG(10) VPM_BCFG(11)
    default:
      break;
  }
#undef VPM_BCFG
  return launch_batch_cfg<float, OUT, kCfgs[0].block, kCfgs[0].r, kCfgs[0].tile, kCfgs[0].sub,
                          kCfgs[0].comp>(b, grid, st);
}

int dispatch_batch(const BatchArgs& b, unsigned mask, int prec, int cfg, int grid,
                   hipStream_t st) {
  switch (mask) {
    case 3u: return launch_batch_mask<3u>(b, prec, cfg, grid, st);
    case 7u: return launch_batch_mask<7u>(b, prec, cfg, grid, st);
    case 27u: return launch_batch_mask<27u>(b, prec, cfg, grid, st);
    default: return launch_batch_mask<31u>(b, prec, cfg, grid, st);
  }
}

unsigned batch_mask(unsigned want) {
  for (unsigned m : {3u, 7u, 27u, 31u})
    if ((m & want) == want) return m;
  return 31u;
}

struct BatchLayout {
  size_t bbox, keys_in, keys_out, idx_in, perm, src8, ens_part, cand, scratch_vec, counter,
      set_chunk0, far_rec, far_box, cub, total, cub_bytes;
};

int batch_layout(int64_t n, int nsets, int n_items, BatchLayout* L, bool need_cub = true) {
  const size_t N = static_cast<size_t>(n > 0 ? n : 1);
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t at = off;
    off += align256(b);
    return at;
  };
  L->bbox = take(static_cast<size_t>(nsets > 0 ? nsets : 1) * 6 * sizeof(double));
  L->keys_in = take(N * 8);
  L->keys_out = take(N * 8);
  L->idx_in = take(N * 4);
  L->perm = take(N * 4);
  L->src8 = take(N * sizeof(Src8));
  L->ens_part = take(static_cast<size_t>(n_items > 0 ? n_items : 1) * sizeof(double));
  L->cand = take(N * sizeof(double));
  L->scratch_vec = take(3 * N * sizeof(double));
  L->counter = take(64);
  // far records: every set's chunks of 64 or 128 sources, at most one partial
  // chunk per set, plus one group read ahead past the last
  const size_t S = static_cast<size_t>(nsets > 0 ? nsets : 1);
  L->set_chunk0 = take((S + 1) * sizeof(int));
  L->far_rec = take((N + kMaxFarChunk * S + kFarGroup) * kMaxFarStride * sizeof(float));
  L->far_box = take((N / kMinFarChunk + S + 1) * 6 * sizeof(double));
  size_t cb = 0;
  if (need_cub) {
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(
        nullptr, cb, static_cast<const uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
        static_cast<const int*>(nullptr), static_cast<int*>(nullptr), static_cast<int>(N), 0, 62,
        hipStream_t(0));
    if (e != hipSuccess) {
      set_error("hipcub SortPairs size query failed: %s", hipGetErrorString(e));
      return VPM_ERR_LAUNCH;
    }
  }
  L->cub_bytes = cb;
  L->cub = take(cb + 16);
  L->total = off;
  return VPM_OK;
}

int check_desc(const vpm_batch_desc_t* d, const vpm_particles_t* p, const char* who) {
  if (!d || d->nsets < 0 || d->n_items < 0) {
    set_error("%s: bad batch descriptor", who);
    return VPM_ERR_ARG;
  }
  if (d->nsets > (1 << 24)) {
    set_error("%s: too many sets", who);
    return VPM_ERR_ARG;
  }
  if (d->nsets > 0 && (!d->set_off || !d->part_set || !d->item_set || !d->item_block ||
                       !d->set_item_first || !d->set_item_count)) {
    set_error("%s: null descriptor array", who);
    return VPM_ERR_ARG;
  }
  return p ? check_particles(p, false, who) : VPM_OK;
}

}  // namespace

}  // namespace vpm

using namespace vpm;

extern "C" {

int vpm_tile_config_count(void) { return kNumCfgs; }

int vpm_tile_config_default(uint32_t outputs) {
  return default_config(kernel_mask(outputs == 0 ? VPM_OUT_VEL : outputs), true);
}

int vpm_batch_tile_config_default(uint32_t outputs) {
  return default_batch_config(batch_mask(outputs == 0 ? VPM_OUT_VEL : outputs));
}

int vpm_tile_config_info(int index, int* block, int* r, int* tile, int* sub, int* comp) {
  clear_error();
  if (index < 0 || index >= kNumCfgs) return fail(VPM_ERR_ARG, "tile config index out of range");
  if (block) *block = kCfgs[index].block;
  if (r) *r = kCfgs[index].r;
  if (tile) *tile = kCfgs[index].tile;
  if (sub) *sub = kCfgs[index].sub;
  if (comp) *comp = kCfgs[index].comp;
  return VPM_OK;
}

size_t vpm_forward_workspace_bytes(int64_t n, const vpm_forward_opts_t* opts) {
  WsLayout L{};
  if (layout(n, opts, &L) != VPM_OK) return 0;
  return L.total;
}

namespace {

// Validation shared by the three forward entry points.
int check_forward(const vpm_particles_t* p, const vpm_forward_opts_t* o, const char* who) {
  if (!o) {
    set_error("%s: null opts", who);
    return VPM_ERR_ARG;
  }
  const unsigned want = o->outputs;
  if (want == 0 || (want & ~31u)) {
    set_error("%s: bad output mask %u", who, want);
    return VPM_ERR_ARG;
  }
  if (o->precision != VPM_PREC_F64 && o->precision != VPM_PREC_F32C) {
    set_error("%s: unknown precision", who);
    return VPM_ERR_ARG;
  }
  if (o->sigma_mode != VPM_SIGMA_GENERAL && o->sigma_mode != VPM_SIGMA_UNIFORM) {
    set_error("%s: unknown sigma mode", who);
    return VPM_ERR_ARG;
  }
  if (o->tile_config < -1 || o->tile_config >= kNumCfgs) {
    set_error("%s: tile_config out of range", who);
    return VPM_ERR_ARG;
  }
  if (o->max_splits < 0) {
    set_error("%s: max_splits < 0", who);
    return VPM_ERR_ARG;
  }
  if (o->g64_mode < VPM_G64_AUTO || o->g64_mode > VPM_G64_NEVER) {
    set_error("%s: unknown g64_mode", who);
    return VPM_ERR_ARG;
  }
  return check_particles(p, (want & VPM_OUT_PSE) != 0, who);
}

int check_ws(int64_t n, const vpm_forward_opts_t* o, void* ws, size_t ws_bytes, WsLayout* L,
             const char* who, bool need_cub) {
  int rc = layout(n, o, L, need_cub);
  if (rc) return rc;
  if (ws_bytes < L->total || !ws) {
    set_error("%s: workspace %zu bytes < required %zu", who, ws_bytes, L->total);
    return VPM_ERR_WORKSPACE;
  }
  return VPM_OK;
}

}  // namespace

namespace {

// The sort and gather of a forward pass into a workspace laid out as L.
int forward_prepare_impl(const vpm_particles_t* p, const vpm_forward_opts_t* o, char* ws,
                         const WsLayout& L, hipStream_t st) {
  const int64_t n = p->n;
  if (n == 0) return VPM_OK;
  const int ni = static_cast<int>(n);
  int* perm = nullptr;
  if (!o->presorted) {
    perm = reinterpret_cast<int*>(ws + L.perm);
    const int rc = morton_sort(p, reinterpret_cast<double*>(ws + L.bbox_part),
                               reinterpret_cast<double*>(ws + L.bbox),
                               reinterpret_cast<uint32_t*>(ws + L.keys_in),
                               reinterpret_cast<uint32_t*>(ws + L.keys_out),
                               reinterpret_cast<int*>(ws + L.idx_in), perm, ws + L.cub,
                               L.cub_bytes, st);
    if (rc) return rc;
  }
  Src8* src8 = reinterpret_cast<Src8*>(ws + L.src8);
  hipLaunchKernelGGL(gather_kernel, dim3(static_cast<unsigned>(ceil_div(n, kPreBlock))),
                     dim3(kPreBlock), 0, st, *p, ni, static_cast<const int*>(perm), src8);
  VPM_LAUNCH_CHECK("gather_kernel");
  return VPM_OK;
}

int check_forward_out(const vpm_particles_t* p, const vpm_forward_opts_t* o,
                      const vpm_forward_out_t* out, const char* who) {
  if (!out) {
    set_error("%s: null out", who);
    return VPM_ERR_ARG;
  }
  const unsigned want = o->outputs;
  const int64_t n = p->n;
  const char* miss = nullptr;
  if (out->ld < n) miss = "out->ld < n";
  else if ((want & VPM_OUT_VEL) && n && !out->vel) miss = "vel output missing";
  else if ((want & VPM_OUT_STRETCH) && n && !out->stretch) miss = "stretch output missing";
  else if ((want & VPM_OUT_PSE) && n && !out->pse) miss = "pse output missing";
  else if ((want & VPM_OUT_VORT) && n && !out->vort) miss = "vort output missing";
  else if ((want & VPM_OUT_ENSTROPHY) && !out->enstrophy) miss = "enstrophy output missing";
  if (miss) {
    set_error("%s: %s", who, miss);
    return VPM_ERR_ARG;
  }
  return VPM_OK;
}

// The pass itself, from the sorted records of forward_prepare_impl.
int forward_run_impl(const vpm_particles_t* p, const vpm_forward_opts_t* o,
                     vpm_forward_out_t* out, char* ws, const WsLayout& L, hipStream_t st) {
  const unsigned want = o->outputs;
  const int64_t n = p->n;
  if (n == 0) {
    if (want & VPM_OUT_ENSTROPHY) {
      hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, st, out->enstrophy, int64_t(1), 0.0);
      VPM_LAUNCH_CHECK("fill_kernel");
    }
    return VPM_OK;
  }
  // Unrequested outputs of a superset kernel share one scratch region; a
  // thread writes only its own target's rows there, so they cannot race.
  const unsigned mask = kernel_mask(want);
  double* scratch = reinterpret_cast<double*>(ws + L.scratch_vec);
  FusedArgs a{};
  a.n = static_cast<int>(n);
  a.src = reinterpret_cast<const Src8*>(ws + L.src8);
  a.perm = o->presorted ? nullptr : reinterpret_cast<const int*>(ws + L.perm);
  auto route = [&](unsigned bit, double* dst, double** ptr, int64_t* ld) {
    if (want & bit) {
      *ptr = dst;
      *ld = out->ld;
    } else {
      *ptr = scratch;
      *ld = n;
    }
  };
  route(VPM_OUT_VEL, out->vel, &a.vel, &a.ld_vel);
  route(VPM_OUT_STRETCH, out->stretch, &a.str, &a.ld_str);
  route(VPM_OUT_PSE, out->pse, &a.pse, &a.ld_pse);
  route(VPM_OUT_VORT, out->vort, &a.vort, &a.ld_vort);
  a.ens_part = reinterpret_cast<double*>(ws + L.ens_part);
  a.part = L.split_bytes ? reinterpret_cast<double*>(ws + L.split) : nullptr;
  a.part_cap = L.split_bytes;
  a.max_splits = o->max_splits;
  a.g64_mode = o->g64_mode;
  if (wants_far_records(o)) {
    a.far_rec = reinterpret_cast<const float*>(ws + L.far_rec);
    a.far_box = reinterpret_cast<const double*>(ws + L.far_box);
  }
  int nblocks = 0;
  const int cfg = effective_config(mask, o->sigma_mode == VPM_SIGMA_UNIFORM, o->tile_config);
  int rc = dispatch(a, mask, o->precision, o->sigma_mode == VPM_SIGMA_UNIFORM, cfg, st, &nblocks);
  if (rc) return rc;
  if (want & VPM_OUT_ENSTROPHY) {
    const double k = (1.0 / (8.0 * kPi)) * kOneOverFourPi;
    hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(1024), 0, st, a.ens_part,
                       static_cast<int64_t>(nblocks), k, out->enstrophy);
    VPM_LAUNCH_CHECK("final_sum_kernel");
  }
  return VPM_OK;
}

}  // namespace

int vpm_forward_prepare(const vpm_particles_t* p, const vpm_forward_opts_t* o, void* workspace,
                        size_t ws_bytes, void* stream) {
  clear_error();
  int rc = check_forward(p, o, "vpm_forward_prepare");
  if (rc) return rc;
  WsLayout L{};
  if ((rc = check_ws(p->n, o, workspace, ws_bytes, &L, "vpm_forward_prepare", true))) return rc;
  return forward_prepare_impl(p, o, static_cast<char*>(workspace), L, as_stream(stream));
}

int vpm_forward_run(const vpm_particles_t* p, const vpm_forward_opts_t* o, vpm_forward_out_t* out,
                    void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  int rc = check_forward(p, o, "vpm_forward_run");
  if (rc) return rc;
  if ((rc = check_forward_out(p, o, out, "vpm_forward_run"))) return rc;
  WsLayout L{};
  if ((rc = check_ws(p->n, o, workspace, ws_bytes, &L, "vpm_forward_run", false))) return rc;
  return forward_run_impl(p, o, out, static_cast<char*>(workspace), L, as_stream(stream));
}

// Workspace of vpm_forward_noalloc: a file-scope device array, one copy per
// GPU (the code object's globals), held until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): its device address is looked up once per device
// with hipGetSymbolAddress and kept in scratch_dev[].
constexpr size_t kScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
void* scratch_dev[kMaxDevices];
hipError_t scratch_err[kMaxDevices];
std::once_flag scratch_once[kMaxDevices];

// The current device's copy of `scratch`; the first call on a device looks
// its address up (that device is current).
int scratch_ptr(char** ptr) {
  int dev = -1;
  VPM_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(VPM_ERR_DEVICE, "vpm_forward_noalloc: device index out of range");
  std::call_once(scratch_once[dev], [dev] {
    scratch_err[dev] = hipGetSymbolAddress(&scratch_dev[dev], HIP_SYMBOL(scratch));
  });
  if (scratch_err[dev] != hipSuccess || !scratch_dev[dev]) {
    set_error("vpm_forward_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
              hipGetErrorString(scratch_err[dev]));
    return VPM_ERR_LAUNCH;
  }
  *ptr = static_cast<char*>(scratch_dev[dev]);
  return VPM_OK;
}

}  // namespace

size_t vpm_forward_noalloc_capacity(void) { return kScratchBytes; }

int vpm_forward_noalloc(const vpm_particles_t* p, const vpm_forward_opts_t* o,
                        vpm_forward_out_t* out, void* stream) {
  clear_error();
  int rc = check_forward(p, o, "vpm_forward_noalloc");
  if (rc) return rc;
  if ((rc = check_forward_out(p, o, out, "vpm_forward_noalloc"))) return rc;
  // the layout without source-split partials must fit; the split partials
  // then get what is left (fewer splits than vpm_forward may choose)
  WsLayout L{};
  if ((rc = layout(p->n, o, &L, true, 0))) return rc;
  if (L.total > kScratchBytes) {
    set_error("vpm_forward_noalloc: n = %lld needs %zu workspace bytes, more than the static "
              "%zu (use vpm_forward with a caller workspace)",
              (long long)p->n, L.total, kScratchBytes);
    return VPM_ERR_WORKSPACE;
  }
  if ((rc = layout(p->n, o, &L, true, kScratchBytes - L.total))) return rc;
  while (L.total > kScratchBytes && L.split_bytes > 0) {  // alignment slack
    if ((rc = layout(p->n, o, &L, true, L.split_bytes - std::min<size_t>(L.split_bytes, 4096))))
      return rc;
  }
  char* ws = nullptr;
  if ((rc = scratch_ptr(&ws))) return rc;
  hipStream_t st = as_stream(stream);
  if ((rc = forward_prepare_impl(p, o, ws, L, st))) return rc;
  return forward_run_impl(p, o, out, ws, L, st);
}

int vpm_forward(const vpm_particles_t* p, const vpm_forward_opts_t* o, vpm_forward_out_t* out,
                void* workspace, size_t ws_bytes, void* stream) {
  int rc = vpm_forward_prepare(p, o, workspace, ws_bytes, stream);
  if (rc) return rc;
  return vpm_forward_run(p, o, out, workspace, ws_bytes, stream);
}

size_t vpm_autotune_workspace_bytes(int64_t n, const vpm_forward_opts_t* opts) {
  if (!opts) return 0;
  vpm_forward_opts_t o = *opts;
  size_t most = 0;
  for (int c = 0; c < kNumCfgs; ++c) {
    o.tile_config = c;
    WsLayout L{};
    if (layout(n, &o, &L) != VPM_OK) return 0;
    most = std::max(most, L.total);
  }
  return most;
}

int vpm_autotune(const vpm_particles_t* p, const vpm_forward_opts_t* opts, vpm_forward_out_t* out,
                 void* workspace, size_t ws_bytes, void* stream, int reps, int* best,
                 float* ms) {
  clear_error();
  if (!opts || !best) return fail(VPM_ERR_ARG, "vpm_autotune: null opts/best");
  if (reps < 1) return fail(VPM_ERR_ARG, "vpm_autotune: reps < 1");
  if (opts->precision != VPM_PREC_F32C)
    return fail(VPM_ERR_ARG, "vpm_autotune: tile configurations apply to VPM_PREC_F32C");
  if (ws_bytes < vpm_autotune_workspace_bytes(p ? p->n : 0, opts))
    return fail(VPM_ERR_WORKSPACE, "vpm_autotune: workspace < vpm_autotune_workspace_bytes");
  vpm_forward_opts_t o = *opts;
  o.tile_config = -1;
  int rc = vpm_forward_prepare(p, &o, workspace, ws_bytes, stream);  // sorts once
  if (rc) return rc;
  hipStream_t st = as_stream(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VPM_HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(VPM_ERR_LAUNCH, "vpm_autotune: hipEventCreate failed");
  }
  float best_ms = 0.0f;
  *best = -1;
  for (int c = 0; c < kNumCfgs && rc == VPM_OK; ++c) {
    o.tile_config = c;
    rc = vpm_forward_run(p, &o, out, workspace, ws_bytes, stream);  // warm
    if (rc) break;
    if (hipEventRecord(e0, st) != hipSuccess) rc = VPM_ERR_LAUNCH;
    for (int k = 0; k < reps && rc == VPM_OK; ++k)
      rc = vpm_forward_run(p, &o, out, workspace, ws_bytes, stream);
    float t = 0.0f;
    if (rc == VPM_OK && (hipEventRecord(e1, st) != hipSuccess ||
                         hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = fail(VPM_ERR_LAUNCH, "vpm_autotune: event timing failed");
    if (rc) break;
    t /= static_cast<float>(reps);
    if (ms) ms[c] = t;
    if (*best < 0 || t < best_ms) {
      *best = c;
      best_ms = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

size_t vpm_batch_workspace_bytes(int64_t n_total, int nsets, int n_items) {
  BatchLayout L{};
  if (batch_layout(n_total, nsets, n_items, &L) != VPM_OK) return 0;
  return L.total;
}

int vpm_batch_prepare(const vpm_particles_t* p, const vpm_batch_desc_t* d, int resort,
                      void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  int rc = check_desc(d, p, "vpm_batch_prepare");
  if (rc) return rc;
  BatchLayout L{};
  if ((rc = batch_layout(p->n, d->nsets, d->n_items, &L, resort != 0))) return rc;
  if (!workspace || ws_bytes < L.total) return fail(VPM_ERR_WORKSPACE, "vpm_batch_prepare: workspace too small");
  if (p->n == 0) return VPM_OK;
  hipStream_t st = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  const int ni = static_cast<int>(p->n);
  int* perm = reinterpret_cast<int*>(ws + L.perm);
  if (resort) {
    double* bbox = reinterpret_cast<double*>(ws + L.bbox);
    hipLaunchKernelGGL(set_bbox_kernel, dim3(d->nsets), dim3(256), 0, st, p->pos, p->ld,
                       d->set_off, bbox);
    VPM_LAUNCH_CHECK("set_bbox_kernel");
    uint64_t* kin = reinterpret_cast<uint64_t*>(ws + L.keys_in);
    uint64_t* kout = reinterpret_cast<uint64_t*>(ws + L.keys_out);
    int* iin = reinterpret_cast<int*>(ws + L.idx_in);
    hipLaunchKernelGGL(set_keys_kernel, dim3(static_cast<unsigned>(ceil_div(p->n, 256))), dim3(256),
                       0, st, p->pos, p->ld, ni, d->part_set, bbox, kin, iin);
    VPM_LAUNCH_CHECK("set_keys_kernel");
    int bits = 30;
    while ((int64_t(1) << (bits - 30)) < d->nsets) ++bits;
    size_t cb = L.cub_bytes;
    VPM_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(ws + L.cub, cb, kin, kout, iin, perm, ni, 0,
                                                   bits, st));
  }
  hipLaunchKernelGGL(gather_kernel, dim3(static_cast<unsigned>(ceil_div(p->n, kPreBlock))),
                     dim3(kPreBlock), 0, st, *p, ni, static_cast<const int*>(perm),
                     reinterpret_cast<Src8*>(ws + L.src8));
  VPM_LAUNCH_CHECK("gather_kernel");
  return VPM_OK;
}

int vpm_batch_run(const vpm_particles_t* p, const vpm_batch_desc_t* d, const vpm_forward_opts_t* o,
                  vpm_forward_out_t* out, void* workspace, size_t ws_bytes, void* stream) {
  clear_error();
  int rc = check_desc(d, p, "vpm_batch_run");
  if (rc) return rc;
  if (!o || !out) return fail(VPM_ERR_ARG, "vpm_batch_run: null opts/out");
  const unsigned want = o->outputs;
  if (want == 0 || (want & ~31u)) return fail(VPM_ERR_ARG, "vpm_batch_run: bad output mask");
  if (o->sigma_mode != VPM_SIGMA_UNIFORM)
    return fail(VPM_ERR_UNSUPPORTED, "vpm_batch_run: each set needs a uniform radius (sigma_mode UNIFORM)");
  if (o->precision != VPM_PREC_F64 && o->precision != VPM_PREC_F32C)
    return fail(VPM_ERR_ARG, "vpm_batch_run: unknown precision");
  if (o->tile_config < -1 || o->tile_config >= kNumCfgs)
    return fail(VPM_ERR_ARG, "vpm_batch_run: tile_config out of range");
  if (o->g64_mode < VPM_G64_AUTO || o->g64_mode > VPM_G64_NEVER)
    return fail(VPM_ERR_ARG, "vpm_batch_run: unknown g64_mode");
  if ((want & VPM_OUT_PSE) && p->n && !p->nu) return fail(VPM_ERR_ARG, "vpm_batch_run: PSE needs nu");
  const int64_t n = p->n;
  if (out->ld < n) return fail(VPM_ERR_ARG, "vpm_batch_run: out->ld < n");
  if ((want & VPM_OUT_VEL) && n && !out->vel) return fail(VPM_ERR_ARG, "vpm_batch_run: vel missing");
  if ((want & VPM_OUT_STRETCH) && n && !out->stretch) return fail(VPM_ERR_ARG, "vpm_batch_run: stretch missing");
  if ((want & VPM_OUT_PSE) && n && !out->pse) return fail(VPM_ERR_ARG, "vpm_batch_run: pse missing");
  if ((want & VPM_OUT_VORT) && n && !out->vort) return fail(VPM_ERR_ARG, "vpm_batch_run: vort missing");
  if ((want & VPM_OUT_ENSTROPHY) && !out->enstrophy) return fail(VPM_ERR_ARG, "vpm_batch_run: enstrophy missing");
  BatchLayout L{};
  if ((rc = batch_layout(n, d->nsets, d->n_items, &L, false))) return rc;
  if (!workspace || ws_bytes < L.total) return fail(VPM_ERR_WORKSPACE, "vpm_batch_run: workspace too small");
  hipStream_t st = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  double* scratch = reinterpret_cast<double*>(ws + L.scratch_vec);
  BatchArgs b{};
  auto route = [&](unsigned bit, double* dst, double** ptr, int64_t* ld) {
    if (want & bit) {
      *ptr = dst;
      *ld = out->ld;
    } else {
      *ptr = scratch;
      *ld = n > 0 ? n : 1;
    }
  };
  route(VPM_OUT_VEL, out->vel, &b.f.vel, &b.f.ld_vel);
  route(VPM_OUT_STRETCH, out->stretch, &b.f.str, &b.f.ld_str);
  route(VPM_OUT_PSE, out->pse, &b.f.pse, &b.f.ld_pse);
  route(VPM_OUT_VORT, out->vort, &b.f.vort, &b.f.ld_vort);
  b.f.ens_part = reinterpret_cast<double*>(ws + L.ens_part);
  b.f.g64_mode = o->g64_mode;
  b.src = reinterpret_cast<const Src8*>(ws + L.src8);
  b.perm = reinterpret_cast<const int*>(ws + L.perm);
  b.inst_off = d->set_off;
  b.item_inst = d->item_set;
  b.item_block = d->item_block;
  b.set_chunk0 = reinterpret_cast<const int*>(ws + L.set_chunk0);
  b.f.far_rec = reinterpret_cast<const float*>(ws + L.far_rec);
  b.f.far_box = reinterpret_cast<const double*>(ws + L.far_box);
  b.nsets = d->nsets;
  b.n_total = static_cast<int>(n);
  b.n_items = d->n_items;
  b.counter = reinterpret_cast<int*>(ws + L.counter);
  if ((rc = dispatch_batch(b, batch_mask(want), o->precision,
                           o->tile_config < 0 ? default_batch_config(batch_mask(want))
                                              : o->tile_config,
                           d->grid, st)))
    return rc;
  if ((want & VPM_OUT_ENSTROPHY) && d->nsets > 0) {
    const double k = (1.0 / (8.0 * kPi)) * kOneOverFourPi;
    hipLaunchKernelGGL(set_enstrophy_kernel, dim3(d->nsets), dim3(256), 0, st, b.f.ens_part,
                       d->set_item_first, d->set_item_count, k, out->enstrophy);
    VPM_LAUNCH_CHECK("set_enstrophy_kernel");
  }
  return VPM_OK;
}

int vpm_batch_les_update(const vpm_batch_desc_t* d, int64_t ld, const double* vorticity,
                         const double* radii, const double* enstrophy, double* state,
                         const double* nu, double* nu_t, double* nu_eff, void* workspace,
                         size_t ws_bytes, int64_t n_total, void* stream) {
  clear_error();
  int rc = check_desc(d, nullptr, "vpm_batch_les_update");
  if (rc) return rc;
  if (n_total < 0 || ld < n_total) return fail(VPM_ERR_ARG, "vpm_batch_les_update: bad n/ld");
  BatchLayout L{};
  if ((rc = batch_layout(n_total, d->nsets, d->n_items, &L, false))) return rc;
  if (!workspace || ws_bytes < L.total) return fail(VPM_ERR_WORKSPACE, "vpm_batch_les_update: workspace too small");
  if (d->nsets == 0 || n_total == 0) return VPM_OK;
  if (!vorticity || !radii || !enstrophy || !state || !nu || !nu_t || !nu_eff)
    return fail(VPM_ERR_ARG, "vpm_batch_les_update: null pointer");
  hipLaunchKernelGGL(set_les_kernel, dim3(d->nsets), dim3(256), 0, as_stream(stream), d->set_off, ld,
                     vorticity, radii, enstrophy, state, nu, nu_t, nu_eff,
                     reinterpret_cast<double*>(static_cast<char*>(workspace) + L.cand));
  VPM_LAUNCH_CHECK("set_les_kernel");
  return VPM_OK;
}

}  // extern "C"
