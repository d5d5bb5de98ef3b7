/*
 * vpm_hip.h -- C ABI of the MI355X (gfx950) all-pairs regularised Biot-Savart
 * library (libvpm_hip.so).
 *
 * The reference solver's hot path is a duck-typed Python "physics module"
 * chosen by importlib (OpenONDA/solvers/VPM/vpmModule.py:141-154) whose eight
 * functions take host float64 arrays of shape (N,3)/(N,) and return fresh
 * host arrays (OpenONDA/solvers/VPM/ParticlesPhysicsCuda.py:11,103,198,351,
 * 452,537,647,756).  This header is the native boundary that replaces the
 * Numba-CUDA kernels behind those functions.  A ctypes binding
 * (openonda_amd/_lib.py) drives it with PyTorch-ROCm device buffers.
 *
 * Conventions
 *  - Every pointer is a device pointer unless stated otherwise.  The caller
 *    owns every buffer; the library allocates nothing persistent except the
 *    HIP graph objects of vpm_rk4_graph_* and kernel-resident constants.
 *  - Vector fields are SoA: component c of particle i lives at p[c*ld + i]
 *    (ld >= n).  Scalars per particle are dense arrays of length n.
 *  - All work is enqueued on the given hipStream_t (passed as void*, NULL =
 *    the default stream).  No call synchronises the device except where noted.
 *  - Return value 0 = success, < 0 = error (see VPM_ERR_*).  The message of
 *    the last error of the calling thread is returned by vpm_last_error().
 *  - No C++ exception crosses this boundary.
 */
#ifndef VPM_HIP_H
#define VPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPM_ABI_VERSION 3

#define VPM_OK 0
#define VPM_ERR_ARG (-1)        /* bad shape, stride, option or null pointer */
#define VPM_ERR_DEVICE (-2)     /* no HIP device / pointer not on the device */
#define VPM_ERR_LAUNCH (-3)     /* HIP runtime or launch failure */
#define VPM_ERR_WORKSPACE (-4)  /* caller workspace too small */
#define VPM_ERR_UNSUPPORTED (-5)
#define VPM_ERR_NONFINITE (-6)  /* input holds NaN/Inf, or a radius <= 0 (vpm_scan_particles) */

/* Which reference plugin's arithmetic to reproduce.
 * CUDA: OpenONDA/solvers/VPM/ParticlesPhysicsCuda.py (the GPU plugin).
 * CPU : OpenONDA/solvers/VPM/ParticlesPhysics.py (the Numba CPU plugin:
 *       sigma_i in the self-induced velocity, full-curl vorticity, PSE with
 *       nu_ij and A+1e-8, KE with A+1e-8, single 1/(4pi) or 1/(8pi) on the
 *       helicity and enstrophy). */
#define VPM_VARIANT_CUDA 0
#define VPM_VARIANT_CPU 1

/* A particle set (sources, and targets for self-evaluations). */
typedef struct vpm_particles {
  int64_t n;           /* number of particles, >= 0 */
  int64_t ld;          /* component stride of pos/str, >= n */
  const double* pos;   /* [3][ld] positions x_i */
  const double* str;   /* [3][ld] strengths alpha_i = omega_i vol_i */
  const double* sigma; /* [n] smoothing radii */
  const double* nu;    /* [n] effective viscosities, may be NULL unless PSE */
} vpm_particles_t;

/* Evaluation points for the *_at functions. */
typedef struct vpm_points {
  int64_t m;
  int64_t ld;          /* component stride, >= m */
  const double* pos;   /* [3][ld] */
} vpm_points_t;

/* ---- library ---------------------------------------------------------- */

int vpm_abi_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* vpm_last_error(void);
/* Number of visible HIP devices (0 when none); never fails. */
int vpm_device_count(void);

/* ---- plain FP64 drop-in: one entry point per reference function --------
 * Faithful FP64 restatements of the reference arithmetic (IEEE division and
 * pow), with an LDS-tiled SoA memory design.  Results agree with the FP64
 * oracle to a few ulp of the sum. */

/* ParticlesPhysicsCuda.py:11-97 (get_induced_velocity_at; same arithmetic in
 * ParticlesPhysics.py:9-74).  out[3][ld_out] at the m targets. */
int vpm_get_induced_velocity_at(const vpm_points_t* targets, const vpm_particles_t* src,
                                double* out, int64_t ld_out, void* stream);

/* ParticlesPhysicsCuda.py:103-191 (variant CUDA: sigma_ij) and
 * ParticlesPhysics.py:81-143 (variant CPU: sigma_i). */
int vpm_get_selfinduced_velocity(const vpm_particles_t* p, int variant,
                                 double* out, int64_t ld_out, void* stream);

/* ParticlesPhysicsCuda.py:198-219,291-346 (CUDA: simplified zeta form) and
 * ParticlesPhysics.py:211-262 (CPU: full curl form).  Output has m rows. */
int vpm_get_vorticity_field_at(const vpm_points_t* targets, const vpm_particles_t* src,
                               int variant, double* out, int64_t ld_out, void* stream);

/* ParticlesPhysicsCuda.py:351-445 and ParticlesPhysics.py:268-312 (same). */
int vpm_get_strength_gradients(const vpm_particles_t* p, double* out, int64_t ld_out,
                               void* stream);

/* ParticlesPhysicsCuda.py:452-530 (CUDA: nu_i) and ParticlesPhysics.py:318-382
 * (CPU: nu_ij, A + 1e-8).  p->nu must be non-NULL. */
int vpm_get_strength_gradients_pse(const vpm_particles_t* p, int variant,
                                   double* out, int64_t ld_out, void* stream);

/* Scalar pair reductions.  *result (device, 1 double) is overwritten; the sum
 * is deterministic (fixed-order two-pass).  workspace must hold
 * vpm_reduction_workspace_bytes(n) bytes.
 * ParticlesPhysicsCuda.py:537-642 / ParticlesPhysics.py:388-429 (energy),
 * ParticlesPhysicsCuda.py:647-750 / ParticlesPhysics.py:432-468 (helicity),
 * ParticlesPhysicsCuda.py:756-777,867-920 / ParticlesPhysics.py:517-550
 * (enstrophy). */
size_t vpm_reduction_workspace_bytes(int64_t n);
int vpm_get_total_kinetic_energy(const vpm_particles_t* p, int variant, double* result,
                                 void* workspace, size_t ws_bytes, void* stream);
int vpm_get_total_helicity(const vpm_particles_t* p, int variant, double* result,
                           void* workspace, size_t ws_bytes, void* stream);
int vpm_get_total_enstrophy(const vpm_particles_t* p, int variant, double* result,
                            void* workspace, size_t ws_bytes, void* stream);

/* ---- fused forward (the fast path) -------------------------------------
 * One all-pairs pass computing any subset of the GPU-variant outputs
 * (ParticlesPhysicsCuda.py semantics):
 *   VEL      self-induced velocity      (:133-191)
 *   STRETCH  transpose-scheme stretching (:375-445)
 *   PSE      PSE diffusion with nu_i     (:478-530)
 *   VORT     LES smoothed vorticity      (:291-346, targets = particles)
 *   ENSTROPHY total enstrophy scalar     (:756-777,867-920)
 * precision F32C: FP32 pair arithmetic on block-centred offsets of
 * Morton-sorted particles, FP32 sub-tile partials combined with Kahan
 * compensation per source tile and folded into FP64 totals.
 * precision F64: the same pass in FP64 throughout. */
#define VPM_OUT_VEL 1u
#define VPM_OUT_STRETCH 2u
#define VPM_OUT_PSE 4u
#define VPM_OUT_VORT 8u
#define VPM_OUT_ENSTROPHY 16u

#define VPM_PREC_F64 0
#define VPM_PREC_F32C 1

#define VPM_SIGMA_GENERAL 0 /* sigma_ij = (sigma_i+sigma_j)/2 per pair */
#define VPM_SIGMA_UNIFORM 1 /* caller asserts sigma_i == sigma[0] for all i */

typedef struct vpm_forward_opts {
  uint32_t outputs;     /* OR of VPM_OUT_* (non-zero) */
  int32_t precision;    /* VPM_PREC_* */
  int32_t sigma_mode;   /* VPM_SIGMA_* */
  int32_t presorted;    /* 1: particles already spatially coherent, skip the Morton sort */
  int32_t tile_config;  /* -1: default; else index < vpm_tile_config_count() */
  int32_t max_splits;   /* 0: automatic; k >= 1: at most k source splits.  When the
                         * target blocks alone cannot fill the device, the sources
                         * are split among workgroups and the FP64 partial sums
                         * are added in split order (deterministic). */
  int32_t g64_mode;     /* VPM_G64_*: how the FP32 pass accumulates the stretching's
                         * sum_j g alpha_j on source tiles near the targets, for
                         * output sets with STRETCH (sets with VORT or ENSTROPHY
                         * and without PSE: always FP64; far tiles of a uniform
                         * radius with VEL: always FP32 partials; ignored by the
                         * FP64 pass) */
} vpm_forward_opts_t;

/* VPM_G64_AUTO: FP64 per pair in waves with a row of 64 consecutive sorted
 * targets whose strengths are coherent (|sum a_i| > 0.45 sum |a_i|: rings,
 * tubes, sheets, where a_i x G nearly cancels), FP32 partials with FP64
 * folds elsewhere. */
#define VPM_G64_AUTO 0
#define VPM_G64_ALWAYS 1
#define VPM_G64_NEVER 2

typedef struct vpm_forward_out {
  int64_t ld;           /* component stride of the vector outputs, >= n */
  double* vel;          /* [3][ld] or NULL (required iff VEL requested) */
  double* stretch;      /* [3][ld] or NULL */
  double* pse;          /* [3][ld] or NULL */
  double* vort;         /* [3][ld] or NULL */
  double* enstrophy;    /* [1] or NULL */
} vpm_forward_out_t;

size_t vpm_forward_workspace_bytes(int64_t n, const vpm_forward_opts_t* opts);
/* vpm_forward = vpm_forward_prepare + vpm_forward_run on the same workspace.
 * prepare sorts (unless presorted) and packs the particles into the
 * workspace; run executes the pair pass and the reductions.  run may be
 * repeated on one prepared workspace while the particles are unchanged. */
int vpm_forward(const vpm_particles_t* p, const vpm_forward_opts_t* opts,
                vpm_forward_out_t* out, void* workspace, size_t ws_bytes, void* stream);

int vpm_forward_prepare(const vpm_particles_t* p, const vpm_forward_opts_t* opts,
                        void* workspace, size_t ws_bytes, void* stream);
int vpm_forward_run(const vpm_particles_t* p, const vpm_forward_opts_t* opts,
                    vpm_forward_out_t* out, void* workspace, size_t ws_bytes, void* stream);

/* vpm_forward without a caller workspace and without any allocation: the
 * sort, the records and the split partials live in a static device array of
 * vpm_forward_noalloc_capacity() bytes (256 MiB) per GPU, whose address is
 * looked up once per device.  Returns VPM_ERR_WORKSPACE when n needs more
 * than that (vpm_forward_workspace_bytes without the split partials); the
 * source-split partials get what is left, so small sets may split less than
 * vpm_forward.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (different streams or threads) race.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
int vpm_forward_noalloc(const vpm_particles_t* p, const vpm_forward_opts_t* opts,
                        vpm_forward_out_t* out, void* stream);
size_t vpm_forward_noalloc_capacity(void);


/* Tile configurations the fused F32C kernel is instantiated for (used by the
 * autotuner).  Writes block threads, targets per lane, source tile, sub-tile
 * length and compensation mode (0 Kahan+FP64, 1 direct FP64 fold). */
int vpm_tile_config_count(void);
/* index used for tile_config = -1 with these outputs: one set with a uniform
 * radius / a batch (FP32) */
int vpm_tile_config_default(uint32_t outputs);
int vpm_batch_tile_config_default(uint32_t outputs);
int vpm_tile_config_info(int index, int* block, int* targets_per_lane, int* tile,
                         int* subtile, int* comp);

/* Autotuning (no counterpart in the reference): sorts once, then times every
 * tile configuration of the VPM_PREC_F32C pass `reps` times on the caller's
 * particles with HIP events, and returns the fastest index in *best and, if
 * ms is not NULL, each configuration's mean time in ms[0..count).  The
 * workspace must hold vpm_autotune_workspace_bytes (the largest layout).
 * Synchronises the stream.  (openonda_amd/autotune.py adds an accuracy check
 * against the FP64 pass and a cache keyed by N.) */
size_t vpm_autotune_workspace_bytes(int64_t n, const vpm_forward_opts_t* opts);
int vpm_autotune(const vpm_particles_t* p, const vpm_forward_opts_t* opts, vpm_forward_out_t* out,
                 void* workspace, size_t ws_bytes, void* stream, int reps, int* best, float* ms);

/* ---- batched persistent launch ------------------------------------------
 * No counterpart in the reference (one particle set per call there).  Many
 * particle sets of uneven size are concatenated into one vpm_particles_t
 * (set s owns indices [set_off[s], set_off[s+1])), each with a uniform radius
 * of its own, and evaluated by one persistent launch whose workgroups pull
 * (set, target block) work items from a device counter.  Work lists are
 * built by the caller: items of one set are consecutive, sets ordered by
 * decreasing size, item_block[k] counts target blocks of
 * block*targets_per_lane particles of the tile config used.  Device arrays: */
typedef struct vpm_batch_desc {
  int32_t nsets;
  int32_t n_items;
  int32_t grid;                   /* persistent workgroups; 0 = occupancy x CUs */
  const int32_t* set_off;         /* [nsets+1] */
  const int32_t* part_set;        /* [n] set of each particle */
  const int32_t* item_set;        /* [n_items] */
  const int32_t* item_block;      /* [n_items] */
  const int32_t* set_item_first;  /* [nsets] */
  const int32_t* set_item_count;  /* [nsets] */
} vpm_batch_desc_t;

size_t vpm_batch_workspace_bytes(int64_t n_total, int nsets, int n_items);
/* resort=1: per-set Morton sort, then pack; resort=0: pack with the
 * permutation stored in the workspace by the last resort=1 call (no sort
 * kernels, so the call can be captured in a HIP graph). */
int vpm_batch_prepare(const vpm_particles_t* p, const vpm_batch_desc_t* d, int resort,
                      void* workspace, size_t ws_bytes, void* stream);
/* One persistent launch of the fused pass over every set (opts.sigma_mode
 * must be UNIFORM: uniform per set).  out->enstrophy, if requested, receives
 * nsets values. */
int vpm_batch_run(const vpm_particles_t* p, const vpm_batch_desc_t* d,
                  const vpm_forward_opts_t* opts, vpm_forward_out_t* out, void* workspace,
                  size_t ws_bytes, void* stream);
/* Per-set dynamic LES filter (vpmModule.py:887-962), one workgroup per set:
 * state[3*s..] = {C_nu, enstrophy_ref, has_ref}; writes nu_t (if the set's
 * max exceeds 1e-9) and nu_eff = nu + nu_t. */
int vpm_batch_les_update(const vpm_batch_desc_t* d, int64_t ld, const double* vorticity,
                         const double* radii, const double* enstrophy, double* state,
                         const double* nu, double* nu_t, double* nu_eff, void* workspace,
                         size_t ws_bytes, int64_t n_total, void* stream);

/* ---- adjoint (vector-Jacobian product) ----------------------------------
 * No counterpart in the reference (which has no gradients).  For
 * L = sum_k lam_k.u_k + mu_k.S_k + xi_k.D_k over the fused outputs (GPU
 * arithmetic: sigma_ij, nu_i; sigma held fixed) writes dL/dx and dL/dalpha.
 * A NULL cotangent means zero.  opts: precision, sigma_mode, presorted are
 * used as in vpm_forward (outputs and tile_config are ignored). */
typedef struct vpm_adjoint_in {
  int64_t ld;           /* component stride of the cotangents, >= n */
  const double* lam;    /* [3][ld] cotangent of the velocity, or NULL */
  const double* mu;     /* [3][ld] cotangent of the stretching, or NULL */
  const double* xi;     /* [3][ld] cotangent of the PSE term, or NULL (needs p->nu) */
} vpm_adjoint_in_t;

typedef struct vpm_adjoint_out {
  int64_t ld;           /* component stride, >= n */
  double* dpos;         /* [3][ld] dL/dx, overwritten */
  double* dstr;         /* [3][ld] dL/dalpha, overwritten */
} vpm_adjoint_out_t;

size_t vpm_adjoint_workspace_bytes(int64_t n, const vpm_forward_opts_t* opts);
int vpm_adjoint(const vpm_particles_t* p, const vpm_adjoint_in_t* in,
                const vpm_forward_opts_t* opts, vpm_adjoint_out_t* out, void* workspace,
                size_t ws_bytes, void* stream);

/* Adjoint of vpm_get_induced_velocity_at w.r.t. the sources (targets fixed):
 * given lam[3][ld_lam] at the m targets, writes dL/dx and dL/dalpha of the
 * n sources (FP64).  No reference counterpart. */
int vpm_induced_velocity_at_adjoint(const vpm_points_t* targets, const vpm_particles_t* src,
                                    const double* lam, int64_t ld_lam, double* dpos, double* dstr,
                                    int64_t ld_out, void* stream);

/* ---- device-resident time integration (all FP64, no host sync) ---------
 * The O(N) parts of a reference solver step, so a step can run (and be
 * captured in a HIP graph) without leaving the device. */

/* out = x + c*y for [3][ld] vectors (RK stage positions, vpmModule.py:604,637,676). */
int vpm_axpy3(int64_t n, int64_t ld, const double* x, const double* y, double c, double* out,
              void* stream);
/* out = x0 + dt/6 (k1 + 2 k2 + 2 k3 + k4), vpmModule.py:688. */
int vpm_rk4_combine(int64_t n, int64_t ld, const double* x0, const double* k1, const double* k2,
                    const double* k3, const double* k4, double dt, double* out, void* stream);
/* Step-indexed strength update (Euler, then the BDF-shaped explicit
 * combinations of steps 1, 2, >=3) and history push, in place:
 * vpmModule.py:749-791 (stretching) and :797-845 (PSE). */
int vpm_strength_update(int64_t n, int64_t ld, int step, double dt, double* strengths,
                        double* prev1, double* prev2, double* prev3, const double* rhs,
                        void* stream);
/* sigma <- sqrt(sigma^2 + 5.34 (nu + nu_t) dt), vpmModule.py:851-862 (nu_t may be NULL). */
int vpm_core_spreading(int64_t n, double* sigma, const double* nu, const double* nu_t, double dt,
                       void* stream);
/* alpha <- 0.95 alpha + 0.05 (omega/|omega|)|alpha|, vpmModule.py:867-883. */
int vpm_relax_strengths(int64_t n, int64_t ld, double* strengths, const double* vorticity,
                        void* stream);
/* Particle-set scan for inputs the pair sums cannot take: ORs VPM_BAD_*
 * bits into *flags (device memory; the caller zeroes it and reads it after
 * the stream reaches this point -- a nonzero value maps to
 * VPM_ERR_NONFINITE).  No reference counterpart: the reference plugins
 * propagate NaN and Inf silently (e.g. the energy's division by |r| at
 * coincident particles, ParticlesPhysicsCuda.py:595-600, and the LES
 * filter's division by |mean vorticity|, vpmModule.py:936-941). */
#define VPM_BAD_POSITION 1u  /* a position is NaN/Inf */
#define VPM_BAD_STRENGTH 2u  /* a strength is NaN/Inf */
#define VPM_BAD_RADIUS 4u    /* a radius is NaN/Inf or <= 0 */
#define VPM_BAD_VISCOSITY 8u /* a viscosity (if given) is NaN/Inf or < 0 */
int vpm_scan_particles(const vpm_particles_t* p, uint32_t* flags, void* stream);

/* Workspace of vpm_diagnostics / vpm_les_update. */
size_t vpm_state_workspace_bytes(int64_t n);
/* out10 = [linear impulse (3), angular impulse (3), sum alpha (3), sum |alpha|]
 * (vpmModule.py:264-284), deterministic. */
int vpm_diagnostics(int64_t n, int64_t ld, const double* pos, const double* str, double* out10,
                    void* workspace, size_t ws_bytes, void* stream);
/* Dynamic LES filter, vpmModule.py:887-962: if update_constant, updates
 * state3 = {C_nu, enstrophy_ref, has_ref} from the device enstrophy scalar
 * (:887-909); then writes nu_t (only if its maximum exceeds 1e-9) from the
 * particle vorticity, radii and state3[0] = C_nu (:916-962). */
int vpm_les_update(int64_t n, int64_t ld, const double* vorticity, const double* radii,
                   const double* enstrophy, double* state3, double* nu_t, int update_constant,
                   void* workspace, size_t ws_bytes, void* stream);

/* ---- HIP graphs ----------------------------------------------------------
 * Capture every library call issued on `stream` between begin and end into
 * one executable graph (stream must not be the default stream).  The graph
 * replays the same kernels on the same buffers. */
int vpm_graph_begin(void* stream);
int vpm_graph_end(void* stream, void** graph_exec);
int vpm_graph_launch(void* graph_exec, void* stream);
int vpm_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* VPM_HIP_H */
