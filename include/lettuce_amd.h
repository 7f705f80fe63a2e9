/*
 * lettuce_amd C-ABI: the MI355X (gfx950) lattice-Boltzmann time step.
 *
 * This is the drop-in boundary for lettuce's native step.  lettuce binds its
 * native path as one Python callable per simulation (`invoke(simulation)`,
 * lettuce/cuda_native/_template.py:35-40) that calls one pybind11 function
 * `lettuce(f, f_next, ncm, nsm, tau_inv, ...)` taking at::Tensor arguments
 * (_template.py:47-53, :72-81) and then swaps `flow.f, flow.f_next`
 * (_template.py:39).  Here the same job is done by plain-C entry points that
 * take raw device pointers, sizes and a hipStream_t (passed as void*), so any
 * FFI (ctypes, cffi, a C++ host) can bind them.  No torch types appear here.
 *
 * Conventions
 *  - Population storage is SoA, `f[q][i0][i1][i2]`, the last index contiguous,
 *    exactly lettuce's `flow.f` of shape [Q, X, Y(, Z)] (lettuce/_flow.py:92).
 *    2D lattices use i0 == 0 (n[0] == 1) with (i1, i2) = (x, y).
 *  - Every entry point returns 0 on success or a negative LBM_E_* code;
 *    lbm_last_error() returns a thread-local message for the last failure.
 *  - Nothing here allocates or frees caller memory; launches are
 *    stream-ordered and nothing synchronises the device.
 */
#ifndef LETTUCE_AMD_H
#define LETTUCE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 3

/* error codes */
#define LBM_OK 0
#define LBM_E_ARG (-1)      /* invalid argument (shape, stride, enum, NULL) */
#define LBM_E_VERSION (-2)  /* args->abi_version != LBM_ABI_VERSION          */
#define LBM_E_UNSUPPORTED (-3) /* combination not available natively         */
#define LBM_E_HIP (-4)      /* HIP runtime error reported at launch           */

/* lattices: lettuce/ext/_stencil/{d2q9,d3q19,d3q27}.py (same direction order) */
enum lbm_stencil { LBM_D2Q9 = 0, LBM_D3Q19 = 1, LBM_D3Q27 = 2 };

/* population storage.  The *_SHIFTED formats hold f_q - w_q * rho0_shift;
 * all arithmetic is done in FP32 (FP64 for LBM_F64). */
enum lbm_storage {
  LBM_F32 = 0,
  LBM_F64 = 1,
  LBM_F16_SHIFTED = 2,
  LBM_BF16_SHIFTED = 3
};

/* collision operators on cells of op LBM_OP_COLLIDE:
 *  NONE    lettuce/ext/_collision/no_collision.py:10-11
 *  BGK     lettuce/ext/_collision/bgk_collision.py:17-22
 *  TRT     lettuce/ext/_collision/trt_collision.py:16-27
 *  KBC     lettuce/ext/_collision/kbc_collision.py:96-160 (D2Q9, D3Q27)
 *  LEARNED KBC decomposition whose higher-order rates come from a per-cell
 *          MLP (no lettuce counterpart; nearest hook MRTCollision,
 *          lettuce/ext/_collision/mrt_collision.py:8-35), D2Q9 and D3Q27.
 *  BGK_GUO BGK with Guo forcing by a uniform acceleration a
 *          (bgk_collision.py:17-22 with ext/_force/guo.py:15-36):
 *          u* = j / rho + a / (2 rho),
 *          f' = f - (f - feq(rho, u*)) / tau
 *               + (1 - 1 / (2 force_tau)) w (3 (e - u*) + 9 (e.u*) e) . a */
enum lbm_collision {
  LBM_COLL_NONE = 0,
  LBM_COLL_BGK = 1,
  LBM_COLL_TRT = 2,
  LBM_COLL_KBC = 3,
  LBM_COLL_LEARNED = 4,
  LBM_COLL_BGK_GUO = 5
};

/* streaming strategy, same bit values as lettuce's StreamingStrategy
 * (lettuce/cuda_native/_default_code_gen.py:13-25):
 * bit 1 = stream before collision (pull), bit 0 = stream after (push). */
enum lbm_streaming {
  LBM_NO_STREAMING = 0,
  LBM_POST_STREAMING = 1,
  LBM_PRE_STREAMING = 2,
  LBM_DOUBLE_STREAMING = 3
};

/* what a cell does in place of collision (one op per cell, the transformer
 * selected by lettuce's no_collision_mask, lettuce/_simulation.py:258-305) */
enum lbm_cell_op {
  LBM_OP_COLLIDE = 0,     /* the simulation's collision                     */
  LBM_OP_BOUNCE_BACK = 1, /* f_q <- f_opp(q)      bounce_back_boundary.py:17 */
  LBM_OP_EQUILIBRIUM = 2, /* f <- feq(rho, u)     equilibrium_boundary_pu.py:79-84 */
  LBM_OP_OUTLET = 3,      /* f <- feq(rho, u(x - dir)) equilibrium_outlet_p.py:63-73 */
  LBM_OP_KEEP = 4,        /* f unchanged (no transformer, halfway-BB solid) */
  LBM_OP_ANTI_BOUNCE_BACK = 5 /* f_opp(v) <- -f_v + w_v rho (2 + 9 (e_v.u_w)^2
                                 - 3 |u_w|^2), u_w = u + (u - u(x - dir)) / 2,
                                 for e_v.dir = 1; anti_bounce_back_outlet.py:79-92 */
};

/* cell-class flags */
#define LBM_CLASS_SOLID 0x01      /* halfway bounce-back solid cell          */
#define LBM_CLASS_NEAR_SOLID 0x02 /* some neighbour x - e_q is SOLID         */
#define LBM_CLASS_BC_FIELD 0x04   /* EQUILIBRIUM: rho,u read per cell from bc_field */

/* One entry of the cell-class table (48 bytes).  A class is one distinct
 * (no_collision_mask value, no_streaming_mask column, solid-neighbour) triple. */
typedef struct lbm_cell_class {
  uint8_t op;         /* enum lbm_cell_op                                      */
  uint8_t flags;      /* LBM_CLASS_*                                           */
  int8_t dir[3];      /* OUTLET, ANTI_BOUNCE_BACK: outward normal (i0,i1,i2);
                         the neighbour velocity is taken at x - dir          */
  uint8_t pad[3];
  uint32_t no_stream; /* bit q set: population q is not streamed into the cell */
  uint32_t pad2;
  double rho;         /* EQUILIBRIUM / OUTLET density (lattice units)          */
  double u[3];        /* EQUILIBRIUM velocity, lattice components (x, y, z)    */
} lbm_cell_class;

/* Learned collision.  With df = f - feq, ds the KBC shear part of df and
 * dh = df - ds (D2Q9: dh holds only the raw moments xxy, xyy, xxyy):
 *   x   = (T, N, Pxy, xxy, xyy, xxyy)(df) / rho * in_scale        (bf16)
 *         (D3Q27: T, Nxz, Nyz, Pxy, Pxz, Pyz of df)
 *   o   = W3 silu(W2 silu(W1 x + b1) + b2) + b3    (bf16 MFMA, FP32 sums)
 *   r_g = 1/tau + o_g, clamped into [rate_lo, rate_hi]
 *   f'  = f - ds/tau - sum_g r_g m_g(dh) phi_g
 * (D2Q9: the groups are the moments xxy, xyy, xxyy; D3Q27: the raw moments
 * x^a y^b z^c of order 3, of order 4, and of orders 5 and 6).
 * A cell whose raw r_g is NaN or outside [stable_lo, stable_hi] takes the
 * KBC update instead.  `packed` is the device copy of the
 * LBM_LEARNED_PACKED_BYTES buffer written by lbm_learned_pack.
 * The kernel evaluates the MLP with every pre-activation z scaled to
 * -log2(e) z (silu becomes s / (1 + 2^s)), so the bf16 roundings of x and
 * of the activations apply to the scaled values. */
#define LBM_LEARNED_HIDDEN 32
#define LBM_LEARNED_PACKED_BYTES 17408
typedef struct lbm_learned {
  const void* packed; /* device pointer, LBM_LEARNED_PACKED_BYTES, 16-byte aligned */
  float in_scale;
  float rate_lo;
  float rate_hi;
  float stable_lo;
  float stable_hi;
  int32_t pad;
} lbm_learned;

typedef struct lbm_step_args {
  int32_t abi_version; /* LBM_ABI_VERSION                                     */
  int32_t stencil;     /* enum lbm_stencil                                    */
  int32_t storage;     /* enum lbm_storage                                    */
  int32_t collision;   /* enum lbm_collision                                  */
  int32_t streaming;   /* enum lbm_streaming                                  */
  int32_t n_classes;   /* entries in `classes` (0 when cell_class == NULL)    */
  int64_t n[3];        /* extents (i0, i1, i2); 2D: n[0] == 1                 */
  int64_t stride[4];   /* element strides of q, i0, i1, i2; stride[3] == 1.
                        * One i0 plane ((n1-1)*stride[2] + n2 elements) must
                        * stay below 2^31 elements and, for formats of up to
                        * 4 bytes, 2^32 bytes; with cell_class the grid below
                        * 2^31 cells.  Otherwise LBM_E_ARG.                   */
  const void* f_in;    /* populations before the step                         */
  void* f_out;         /* populations after the step (must not alias f_in)    */
  const uint8_t* cell_class;      /* [n0*n1*n2] class ids, or NULL (all collide, periodic) */
  const lbm_cell_class* classes;  /* device table of n_classes entries        */
  const double* bc_field; /* [4][cells] (rho, ux, uy, uz) for LBM_CLASS_BC_FIELD, else NULL */
  double tau;          /* BGK / TRT tau_plus / KBC tau                        */
  double tau_minus;    /* TRT tau_minus                                       */
  double rho0_shift;   /* shifted storage reference density (usually 1)       */
  const lbm_learned* learned; /* host pointer to parameters, LEARNED only      */
  double force_tau;    /* BGK_GUO: tau of the source term (Guo's own tau)      */
  double accel[3];     /* BGK_GUO: acceleration in lattice units (x, y, z)     */
  /* Stochastic rounding of LBM_BF16_SHIFTED storage (no lettuce counterpart):
   * sr_seed == 0 rounds every stored value to nearest even; otherwise the 16
   * bits the conversion drops are replaced by 16 random bits added before
   * truncation, so a value rounds away from zero with a probability equal to
   * the dropped fraction (unbiased on average).  The bits are a hash of
   * sr_seed, the step index (sr_step for the first step of an lbm_run call,
   * sr_step + k for step k), the stored element's offset and the population,
   * so two-step launches equal two single steps bit for bit.  Other storage
   * formats and x-slab steps take sr_seed == 0 only (LBM_E_UNSUPPORTED). */
  uint32_t sr_seed;
  uint32_t sr_step;
} lbm_step_args;

/* ABI version of the loaded library (LBM_ABI_VERSION it was built with). */
int lbm_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char* lbm_last_error(void);

/* sizeof(lbm_step_args), sizeof(lbm_cell_class), sizeof(lbm_learned): lets a
 * binding check its struct mirror. */
int64_t lbm_struct_size(int which);

/* One time step: f_out = step(f_in).  Replaces the generated
 * `lettuce(f, f_next, ...)` kernel host function
 * (lettuce/cuda_native/_template.py:72-81, launched by invoke() at :35-40);
 * the caller swaps f_in / f_out afterwards as invoke() does (:39). */
int lbm_step(const lbm_step_args* args, void* stream);

/* `nsteps` time steps, alternating f_in and f_out (step k reads the buffer
 * written by step k-1).  The result is in f_out when nsteps is odd and in
 * f_in when it is even; f_in is overwritten when nsteps >= 2.  Replaces the
 * per-step Python loop of Simulation.__call__ (lettuce/_simulation.py:317-320)
 * when no reporter runs in between.  A post-streaming run (lettuce's default
 * order, collide then stream, _simulation.py:54, :154-157) of nsteps >= 2
 * without the learned collision runs as (S C S) (C S)^(nsteps-2) C: one
 * collide-only launch, nsteps - 2 pull steps, one double-streaming launch;
 * the composition is the same, so every cell takes the same values as with
 * nsteps push steps. */
int lbm_run(const lbm_step_args* args, int64_t nsteps, void* stream);

/* How lbm_run would execute `nsteps` steps, without launching anything:
 * plan[0] = launches of the two-step kernel (two fused steps each: periodic
 * D3Q19 / D3Q27 without boundaries, pull streaming (or the pull steps of a
 * split post-streaming run), contiguous extent a multiple of 64, i1 extent a
 * multiple of 8; by default for D3Q19 in every storage format and for D3Q27
 * in FP32), plan[1] = single-step launches.  No lettuce counterpart; for
 * reporting. */
int lbm_run_plan(const lbm_step_args* args, int64_t nsteps, int64_t* plan);

/* Launch tuning for measurements and tests (no lettuce counterpart): set the
 * process-wide knob `name` to `value`, or back to the value its environment
 * variable gave at load when value < 0.  Knobs: "two_step" (0 never, 1
 * default selection, 2 wherever the two-step kernel applies),
 * "two_step_planes" (x chunk, 0 = per grid), "two_step_barriers" (1 or 2,
 * 0 = per format), "wide_offsets" (1 forces 64-bit plane offsets),
 * "post_split" (0 runs post-streaming with the push kernel only).  Not
 * synchronised with launches on other threads. */
int lbm_set_tuning(const char* name, int32_t value);

/* f[q][cell] = feq_q(rho[cell], u[:, cell]) written in the given storage:
 * QuadraticEquilibrium (lettuce/ext/_equilibrium/quadratic_equilibrium.py:11-24)
 * as used by Flow.initialize (lettuce/_flow.py:127-143).
 * rho: [cells], u: [D][cells] in FP32 (storage F32/F16/BF16) or FP64 (F64),
 * contiguous; f uses `stride` as in lbm_step_args. */
int lbm_equilibrium(int32_t stencil, int32_t storage, const int64_t n[3],
                    const int64_t stride[4], const void* rho, const void* u,
                    double rho0_shift, void* f, void* stream);

/* Decode any storage into FP32 or FP64 populations (out_f64 selects), i.e.
 * the true f_q = stored_q + w_q * rho0_shift for the shifted formats.
 * `out` is contiguous [Q][cells]. */
int lbm_decode(int32_t stencil, int32_t storage, const int64_t n[3],
               const int64_t stride[4], const void* f, double rho0_shift,
               void* out, int32_t out_f64, void* stream);

/* Encode contiguous FP32/FP64 populations `in` ([Q][cells]; in_f64 selects)
 * into `f` in the given storage (inverse of lbm_decode). */
int lbm_encode(int32_t stencil, int32_t storage, const int64_t n[3],
               const int64_t stride[4], const void* in, int32_t in_f64,
               double rho0_shift, void* f, void* stream);

/* Moments in FP64: rho[cell] = sum_q f_q, u[d][cell] = j_d / rho
 * (Flow.rho / Flow.j / Flow.u, lettuce/_flow.py:157-193).  rho and u are
 * contiguous FP64 arrays of [cells] and [D][cells]. */
int lbm_moments_f64(int32_t stencil, int32_t storage, const int64_t n[3],
                    const int64_t stride[4], const void* f, double rho0_shift,
                    double* rho, double* u, void* stream);

/* FP64 global reductions over the cells whose `mask` byte is 0 (mask may be
 * NULL = all cells).  Writes out[0..4] = { sum rho, sum j_x, sum j_y, sum j_z,
 * sum |u|^2 / 2 }.  The kinetic-energy sum is IncompressibleKineticEnergy in
 * lattice units before the PU conversion (observable_reporter.py:34-42).
 * `workspace` holds at least lbm_reduce_workspace_bytes(n) bytes. */
int64_t lbm_reduce_workspace_bytes(const int64_t n[3]);
int lbm_reduce_f64(int32_t stencil, int32_t storage, const int64_t n[3],
                   const int64_t stride[4], const void* f, double rho0_shift,
                   const uint8_t* mask, double* out, void* workspace,
                   void* stream);

/* lbm_reduce_f64 without the workspace argument; it allocates nothing.  No
 * lettuce counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-block partial sums go to a static device array of
 * the library, LBM_REDUCE_STATIC_BYTES bytes on each device it runs on,
 * held until the process ends; a grid whose reduction needs more partials
 * is rejected with LBM_E_ARG (none does: the partials are capped at 2048
 * blocks, lbm_reduce_workspace_bytes).  It runs on the device current on
 * the calling thread, which must own f, mask, out and stream.  Calls on one
 * device share the array: calls in one stream are ordered, calls that
 * overlap on one device (other streams, other threads) race. */
#define LBM_REDUCE_STATIC_BYTES 81920
int lbm_reduce_f64_static(int32_t stencil, int32_t storage, const int64_t n[3],
                          const int64_t stride[4], const void* f, double rho0_shift,
                          const uint8_t* mask, double* out, void* stream);

/* Failure scan over all cells, the checks of NaNReporter.is_failed
 * (torch.isnan(flow.f).any(), lettuce/ext/_reporter/failure_reporter.py:144-145)
 * and HighMaReporter.is_failed (|u| / cs > threshold, :168-171) in one pass:
 * out[0..2] = { number of NaN populations, number of cells whose |u|^2 is
 * not finite, max |u|^2 over the other cells } (FP64, deterministic).
 * `workspace` holds at least lbm_reduce_workspace_bytes(n) bytes. */
int lbm_scan_f64(int32_t stencil, int32_t storage, const int64_t n[3],
                 const int64_t stride[4], const void* f, double rho0_shift,
                 double* out, void* workspace, void* stream);

/* FP64 enstrophy sum, sum over cells of |curl u|^2 with u in lattice units
 * and 6th-order periodic central differences (dx = 1), the sum computed by
 * Enstrophy (observable_reporter.py:45-68) via torch_gradient
 * (lettuce/util/utility.py:37-99) before the PU scaling.  u is FP64
 * [D][cells] as written by lbm_moments_f64. */
int lbm_enstrophy_f64(int32_t stencil, const int64_t n[3], const double* u,
                      double* out, void* workspace, void* stream);

/* Momentum-exchange force on halfway-bounce-back solids from the stored
 * post-collision populations: F = 2 sum_links f_d(x) e_d over links
 * (x fluid, x + e_d solid), the force of
 * examples/advanced_projects/efficient_bounce_back_obstacle/boundary/
 * halfway_bounce_back_boundary.py:211-218.  links: int32 [L][2] = (cell, d).
 * out[0..2] FP64 force components. */
int lbm_link_force_f64(int32_t stencil, int32_t storage, const int64_t n[3],
                       const int64_t stride[4], const void* f,
                       double rho0_shift, const int32_t* links, int64_t n_links,
                       double* out, void* workspace, void* stream);

/* Bandwidth probe with no lettuce counterpart: dst = src for `bytes` bytes
 * (a multiple of 16, both pointers 16-byte aligned), 16 bytes per lane,
 * optionally with non-temporal stores.  It measures the read + write HBM
 * ceiling that the step kernels are compared with. */
int lbm_stream_copy(const void* src, void* dst, int64_t bytes,
                    int32_t nontemporal, void* stream);

/* Host function: pack the MLP 6 -> 32 -> 32 -> 3 (row-major FP32 weights
 * w1[32][6], w2[32][32], w3[3][32] and biases b1[32], b2[32], b3[3]) into the
 * per-lane MFMA operand layout used by the learned-collision kernel; the
 * weights are rounded to bf16 (round to nearest even) and the biases are
 * stored multiplied by -log2(e), for the kernel's scaled evaluation.  `out`
 * is host memory of LBM_LEARNED_PACKED_BYTES bytes; copy it to the device. */
int lbm_learned_pack(const float* w1, const float* b1, const float* w2,
                     const float* b2, const float* w3, const float* b3,
                     void* out);

/* ---------------------------------------------------------------------------
 * Training of the learned collision (no lettuce counterpart: lettuce has no
 * trainable collision; these are the forward and adjoint kernels behind
 * lettuce_amd.ext.LearnedRollout, a torch.autograd.Function).
 *
 * Periodic pull streaming in F32 or F64 (storage), the MLP as one vector
 * `params` of the same type:
 *   [w1 (32x6), b1 (32), w2 (32x32), b2 (32), w3 (3x32), b3 (3)]
 *   stencil LBM_D2Q9   a batch of n[0] independent n[1] x n[2] fields,
 *                      populations contiguous as [9][n0][n1][n2];
 *   stencil LBM_D3Q27  one n[0] x n[1] x n[2] field, [27][n0][n1][n2].
 * The forward step is the learned collision's formula with the MLP in the
 * arithmetic type (no bf16 rounding) and without the KBC fallback:
 *   f' = g - ds/tau - sum_k clamp(1/tau + o_k, rate_lo, rate_hi) m_k(dh) phi_k
 * with g the pulled populations (D3Q27: the rates of the raw moments of dh
 * of orders 3, 4 and 5-6).  The clamp passes the gradient for raw rates in
 * [rate_lo, rate_hi] and blocks it outside.
 * ------------------------------------------------------------------------- */
#define LBM_LEARNED_PARAMS 1379
/* per-cell record of the backward kernel: x (6), a1 (32), a2 (32),
 * dL/dz1 (32), dL/dz2 (32), dL/do (3), stored [LBM_LEARNED_RECORD][cells];
 * the weight gradients are sums over cells of outer products of its rows */
#define LBM_LEARNED_RECORD 137

/* f_out = learned collide(pull(f_in)) */
int lbm_learned_train_forward(int32_t stencil, int32_t storage, const int64_t n[3],
                              double tau, const void* params, double in_scale,
                              double rate_lo, double rate_hi,
                              const void* f_in, void* f_out, void* stream);

/* The adjoint of one step taken from f_in.  lam is the adjoint of that
 * step's output: read at the cell itself (lam_gathered == 0) or, when lam is
 * the previous backward call's mu, as lam_q(x) = mu_q(x + e_q)
 * (lam_gathered != 0), the transpose of pull streaming.  Writes
 * mu = (d collide / d g)^T lam and, if record is not null, the per-cell
 * record (LBM_LEARNED_RECORD x cells values of the storage type). */
int lbm_learned_train_backward(int32_t stencil, int32_t storage, const int64_t n[3],
                               double tau, const void* params, double in_scale,
                               double rate_lo, double rate_hi,
                               const void* f_in, const void* lam,
                               int32_t lam_gathered, void* mu, void* record,
                               void* stream);

/* lam_q(x) = mu_q(x + e_q): the adjoint of periodic pull streaming (the
 * gradient with respect to the populations before the step) */
int lbm_unpull(int32_t stencil, int32_t storage, const int64_t n[3], const void* mu,
               void* lam, void* stream);

/* Parameter gradients from a record of lbm_learned_train_backward (no
 * lettuce counterpart), accumulated into grad [LBM_LEARNED_PARAMS] of the
 * storage type (LBM_F32 or LBM_F64) in the parameter layout above:
 *   dW1[i][j] += sum_c dz1_i(c) x_j(c)    db1[i] += sum_c dz1_i(c)
 *   dW2[i][j] += sum_c dz2_i(c) a1_j(c)   db2[i] += sum_c dz2_i(c)
 *   dW3[i][j] += sum_c do_i(c) a2_j(c)    db3[i] += sum_c do_i(c)
 * over the record's `cells` columns.  workspace holds
 * lbm_learned_grad_workspace(storage, cells) bytes of per-block partial
 * sums, added in a fixed order (deterministic).  lbm_learned_grad_workspace
 * returns -1 for an unsupported storage or cells < 1. */
int64_t lbm_learned_grad_workspace(int32_t storage, int64_t cells);
int lbm_learned_grad_reduce(int32_t storage, int64_t cells, const void* record,
                            void* workspace, void* grad, void* stream);

/* ---------------------------------------------------------------------------
 * x-slabs: one device's share of a lattice split along i0 (no lettuce
 * counterpart: lettuce steps the whole lattice on one device,
 * lettuce/_simulation.py:241-256 rolls every axis periodically).
 *
 * A slab holds planes 0 .. n[0]-1 of a 3D lattice (D3Q19 / D3Q27) that is
 * periodic along i1 and i2; the planes i0 = -1 and i0 = n[0] live on the
 * neighbouring slabs.  Per pull step the populations that cross an i0 face
 * travel as packed halo buffers, each [k][n[1]][n[2]] contiguous elements of
 * the storage type, k = lbm_halo_planes(stencil) populations (those with
 * e_x = +1, resp. -1, in stencil order):
 *   send_hi  the slab's plane n[0]-1, e_x = +1  -> the upper neighbour's recv_lo
 *   send_lo  the slab's plane 0,      e_x = -1  -> the lower neighbour's recv_hi
 * The transfer between devices is the caller's (RCCL send/recv, a peer copy).
 * One slab whose recv_lo / recv_hi are copies of its own send_hi / send_lo
 * steps exactly like the periodic lattice.
 * ------------------------------------------------------------------------- */
enum lbm_slab_part {
  LBM_SLAB_ALL = 0,      /* interior planes, then the two face planes       */
  LBM_SLAB_INTERIOR = 1, /* planes 1 .. n[0]-2 (needs no halo: overlaps the transfer) */
  LBM_SLAB_FACES = 2     /* planes 0 and n[0]-1 (reads recv_*, writes send_*) */
};

/* Populations that cross one i0 face: 5 (D3Q19), 9 (D3Q27); -1 otherwise. */
int64_t lbm_halo_planes(int32_t stencil);

/* send_lo / send_hi <- the leaving populations of planes 0 / n[0]-1 of f (raw
 * storage bits), for a state that no lbm_slab_step produced (the initial
 * condition, a restart).  f uses `stride` as in lbm_step_args. */
int lbm_halo_pack(int32_t stencil, int32_t storage, const int64_t n[3],
                  const int64_t stride[4], const void* f, void* send_lo,
                  void* send_hi, void* stream);

/* One pull step (args->streaming == LBM_PRE_STREAMING, no cell_class, any
 * collision but LEARNED) of the slab f_in -> f_out, or the part of it that
 * `part` selects.  The face planes pull the entering populations from
 * recv_lo / recv_hi and write the leaving ones of f_out to send_lo / send_hi
 * (the next step's halos), so no pack pass runs between steps.  The four
 * halo buffers must be distinct; they are ignored for LBM_SLAB_INTERIOR.  The
 * arithmetic is that of lbm_step; the caller swaps f_in / f_out after the
 * face part, and must not overwrite recv_* before the face part has run. */
int lbm_slab_step(const lbm_step_args* args, int32_t part,
                  const void* recv_lo, const void* recv_hi,
                  void* send_lo, void* send_hi, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LETTUCE_AMD_H */
