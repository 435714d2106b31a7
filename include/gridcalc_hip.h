/*
 * gridcalc_hip.h -- C ABI of libgridcalc_hip.so, the MI355X (gfx950) engine behind
 * gridcalc's periodic-grid operators, reductions and explicit time stepping.
 *
 * The reference (gridcalc 0.15.1) is a header-only C++17 library whose boundary is
 * its template API; the C++ headers shipped in grid-calculus_amd/include/gridcalc/
 * keep that API and call these functions.  Each entry point below names the
 * reference function it replaces (path:line in the reference tree).
 *
 * ABI rules
 *  - Plain C types only; no exceptions cross this boundary.  Every call returns a
 *    gc_status; on failure gc_last_error() returns a message for the calling thread.
 *  - The caller owns every device buffer passed in (the library never frees them).
 *    Buffers come from gc_malloc (or any hipMalloc'ed memory on the current device).
 *  - Scratch (RK4 stage fields, reduction partials) lives in a gc_context, created
 *    per stream.  Calls on one context are ordered on its stream; contexts are
 *    independent, so distinct contexts may be used from distinct threads.
 *  - Field layout: i-fastest, site (i,j,k) at i + Nx*(j + Ny*k)  (core/Grid.hpp:132-137).
 *    Vec3 fields are 3 values per site (x,y,z) contiguous (Field<Vec3d>, AoS);
 *    Mat3 fields are 9 values per site, column-major, M(r,c) at r + 3c
 *    (Field<Mat3d>, Eigen's default storage order).
 *  - GC_F64 computes in double precision in the reference's arithmetic order without
 *    contracted multiply-adds (bit-identical to the CPU path compiled the same way).
 *    GC_F32 stores and computes fields in single precision; reductions always
 *    accumulate in double precision.
 *  - Grids are limited to 2^31-1 sites per field.
 */
#ifndef GRIDCALC_HIP_H
#define GRIDCALC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gc_status {
  GC_OK = 0,
  GC_ERR_INVALID = 1,    /* bad argument (null pointer, bad order/dtype, negative dt or n_steps) */
  GC_ERR_EXTENT = 2,     /* an axis extent aliases the stencil (2 <= N <= 2r) */
  GC_ERR_CFL = 3,        /* explicit-diffusion CFL bound violated */
  GC_ERR_NONPOS_D = 4,   /* diffusivity field has a non-positive value */
  GC_ERR_DEVICE = 5,     /* HIP runtime error */
  GC_ERR_NOMEM = 6,      /* device allocation failed */
  GC_ERR_UNSUPPORTED = 7 /* valid request this build does not implement */
} gc_status;

typedef enum gc_dtype { GC_F64 = 0, GC_F32 = 1 } gc_dtype;
typedef enum gc_reduce_method { GC_PAIRWISE = 0, GC_KAHAN = 1 } gc_reduce_method;
typedef enum gc_integrator { GC_EULER = 0, GC_RK4 = 1 } gc_integrator;

/* core::Grid (core/Grid.hpp:28-148): extents and cell size. */
typedef struct gc_grid {
  int32_t nx, ny, nz;
  double hx, hy, hz;
} gc_grid;

/* Parameters of the Cahn-Hilliard right-hand side M*lap(-c + c^3 - kappa*lap c)
 * (examples/common/cahn_hilliard.hpp:38-59). */
typedef struct gc_ch_params {
  double M;
  double kappa;
} gc_ch_params;

typedef struct gc_context gc_context;

/* ---- library, errors, memory ------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char* gc_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int gc_version(void);
/* Number of visible HIP devices (0 when there is no GPU). */
gc_status gc_device_count(int* count);

/* A context binds a HIP stream (stream == NULL creates a private one) and owns
 * scratch memory that grows on demand and is freed by gc_context_destroy. */
gc_status gc_context_create(gc_context** ctx, void* stream);
gc_status gc_context_destroy(gc_context* ctx);
gc_status gc_context_sync(gc_context* ctx);
void* gc_context_stream(gc_context* ctx);
/* Bytes of scratch currently held by the context. */
size_t gc_context_scratch_bytes(gc_context* ctx);

gc_status gc_malloc(void** ptr, size_t bytes);
gc_status gc_free(void* ptr);
/* Page-locked host memory (hipHostMalloc) for host copies of fields: transfers to and from
 * it run as DMA without staging.  gc_host_malloc fails with GC_ERR_NOMEM when the system
 * cannot pin more memory; gc_host_free releases what it returned. */
gc_status gc_host_malloc(void** ptr, size_t bytes);
gc_status gc_host_free(void* ptr);
/* Copies ordered on the context's stream; host-side copies complete before return. */
gc_status gc_memcpy_htod(gc_context* ctx, void* dst, const void* src, size_t bytes);
gc_status gc_memcpy_dtoh(gc_context* ctx, void* dst, const void* src, size_t bytes);
gc_status gc_memcpy_dtod(gc_context* ctx, void* dst, const void* src, size_t bytes);

/* Stream-ordered timer: gc_timer_start records an event on the context's stream,
 * gc_timer_stop records a second one, waits for it and returns the elapsed ms. */
gc_status gc_timer_start(gc_context* ctx);
gc_status gc_timer_stop(gc_context* ctx, float* ms);

/* ---- stencil operators (out is a distinct buffer from every input) ----------- */

/* diff::laplacian<Order> (diff/Laplacian.hpp:44-75); order 2 or 4. */
gc_status gc_laplacian(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                       const void* in, void* out);
/* diff::gradient<Order>(Field<double>) -> Field<Vec3d> (diff/Gradient.hpp:43-74). */
gc_status gc_gradient(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                      const void* in, void* out_vec3);
/* diff::gradient<Order>(Field<Vec3d>) -> Field<Mat3d> (diff/Gradient.hpp:100-137). */
gc_status gc_gradient_vec(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                          const void* in_vec3, void* out_mat3);
/* diff::divergence<Order>(Field<Vec3d>) -> Field<double> (diff/Divergence.hpp:43-74). */
gc_status gc_divergence(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                        const void* in_vec3, void* out);
/* diff::divergence<Order>(Field<Mat3d>) -> Field<Vec3d> (diff/Divergence.hpp:100-135). */
gc_status gc_divergence_mat(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                            const void* in_mat3, void* out_vec3);
/* diff::biharmonic<Order> / diff::d4 (diff/Biharmonic.hpp:48-119). */
gc_status gc_biharmonic(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                        const void* in, void* out);
/* diff::detail::mixedDerivative<Nx, Ny, Nz, Order> (diff/detail/MultiIndexDerivative.hpp:
 * 96-136): d^(nx+ny+nz) psi / dx^nx dy^ny dz^nz with the tensor product of one axis stencil
 * per axis (first to fourth central difference, Order 2 or 4), 1 <= nx+ny+nz <= 4; behind the
 * named partials of diff/MixedPartial.hpp (dxx .. dyz), diff/ThirdOrder.hpp (d3dx3 ..
 * d3dxdydz) and diff/FourthOrder.hpp (d4dx4 .. d4dxdydz2).  Extent check per axis with that
 * axis's stencil radius. */
gc_status gc_mixed_partial(gc_context* ctx, const gc_grid* g, gc_dtype dt, int nx, int ny, int nz,
                           int order, const void* in, void* out);
/* fvm::cellLaplacian(psi, D) (fvm/CellLaplacian.hpp:91-143).  Checks D > 0
 * (CellLaplacian.hpp:99-109) inside the operator's kernel: on a violation returns
 * GC_ERR_NONPOS_D (out then holds no meaningful values) and, when first_bad is not NULL,
 * stores the smallest violating flat index there. */
gc_status gc_cell_laplacian(gc_context* ctx, const gc_grid* g, gc_dtype dt,
                            const void* psi, const void* D, void* out, int64_t* first_bad);
/* Smallest flat index with !(D > 0), or -1 (the scan of CellLaplacian.hpp:99-109). */
gc_status gc_first_nonpositive(gc_context* ctx, const gc_grid* g, gc_dtype dt,
                               const void* D, int64_t* first_bad);
/* Largest harmonic-mean face diffusivity over the +x/+y/+z faces
 * (solve/Diffusion.hpp:187-207). */
gc_status gc_max_face_d(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D,
                        double* dmax);

/* ---- reductions (synchronous: the result is on the host at return) ----------- */

/* func::integrate(Field, Pairwise|Kahan) (func/Integrate.hpp:116-135):
 * cellVolume * sum.  GC_PAIRWISE reproduces pairwiseSum's recursion tree
 * (Integrate.hpp:65-74) exactly; GC_KAHAN is a compensated (Neumaier/TwoSum)
 * parallel sum in double precision. */
gc_status gc_integrate(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                       const void* in, double* result);
/* No counterpart in the reference: func::integrate for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and the same result bits as gc_integrate, written as one double
 * (cellVolume * sum) to the device address result_dev, asynchronously on the context's
 * stream: no allocation, no host copy, no synchronisation.  Partial sums live in a fixed
 * per-device static array of about 1 MiB, enough for every grid the library accepts
 * (sizes beyond it fail with GC_ERR_UNSUPPORTED).  All calls on one device share that array:
 * calls that overlap on one GPU (from different streams or contexts) race; calls ordered on
 * one stream do not.  The context must have been created beforehand (gc_context_create
 * allocates). */
gc_status gc_integrate_noalloc(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                               const void* in, double* result_dev);
/* ---- asynchronous snapshots ------------------------------------------------------ */

/* A snapshot slot for writing fields out while stepping continues (the VTK snapshots of
 * examples/common/vtk_writer.hpp:34-67 and examples/cahn_hilliard.cpp:130-135).  It owns a
 * device staging buffer, a page-locked host buffer of `bytes`, and a copy stream of its own
 * (default priority).  gc_snapshot_capture copies the field into the staging buffer on the
 * context's stream (ordered after the work that produced it) and queues the device-to-host
 * copy on the copy stream, so later kernels on the context's stream do not wait for it;
 * gc_snapshot_wait blocks until that copy is done and returns the host buffer.  A slot
 * holds one capture at a time (capture waits for the previous one). */
typedef struct gc_snapshot gc_snapshot;
gc_status gc_snapshot_create(gc_context* ctx, size_t bytes, gc_snapshot** out);
gc_status gc_snapshot_capture(gc_snapshot* s, const void* src_dev, size_t bytes);
gc_status gc_snapshot_wait(gc_snapshot* s, const void** host);
gc_status gc_snapshot_destroy(gc_snapshot* s);

/* ---- tensor contraction expressions ------------------------------------------ */

/* One instruction of a per-site register program (tensor::expr of tensor/Expressions.hpp,
 * lowered by gridcalc/tensor/Expressions.hpp).  Slots 0..7 hold nc = 1 (scalar), 3 (Vec3) or
 * 9 (Mat3, column-major) doubles.  dst = op(a, b); LOAD reads nc doubles per site from the FP64
 * field at ptr; SCALE multiplies by c. */
typedef enum gc_expr_op {
  GC_EXPR_LOAD = 0,     /* dst = field(ptr)                              */
  GC_EXPR_IDENTITY = 1, /* dst = Mat3 identity                           */
  GC_EXPR_SCALE = 2,    /* dst = c * a                                   */
  GC_EXPR_ADD = 3,      /* dst = a + b                                   */
  GC_EXPR_SUB = 4,      /* dst = a - b                                   */
  GC_EXPR_MUL = 5,      /* dst = a * b (scalars)                         */
  GC_EXPR_TRACE = 6,    /* dst = (a00 + a11) + a22                       */
  GC_EXPR_SYM = 7,      /* dst = 0.5 * (a + a^T)                         */
  GC_EXPR_MATMUL = 8,   /* dst = a * b (3x3 matrix product)              */
  GC_EXPR_DDOT = 9      /* dst = sum of a_i * b_i over the 9 entries     */
} gc_expr_op;

typedef struct gc_expr_insn {
  int32_t op, dst, a, b, nc;
  double c;
  const void* ptr;
} gc_expr_insn;

/* Runs a program of n_insn <= 32 instructions at every site and writes slot out_slot
 * (out_nc doubles per site) to the FP64 field out (tensor::expr::materialize,
 * tensor/Expressions.hpp:628-651; the eager tensor::trace / singleContract / doubleContract
 * of tensor/Contraction.hpp:33-101 are one-node programs). */
gc_status gc_expr_eval(gc_context* ctx, const gc_grid* g, const gc_expr_insn* prog, int n_insn,
                       int out_slot, int out_nc, void* out);

/* Densities func::evaluate integrates on the device (func/Functional.hpp:88-148, the scalar
 * overload, and :183-220, the Vec3 overload), each a formula of the Order-2 derivatives the
 * reference materialises.  The C++ structs in gridcalc/func/Densities.hpp evaluate the same
 * formulas on the host operation for operation (so a caller's lambda of the same expression
 * gives the same bits).  a and b are the density's parameters. */
typedef enum gc_density_kind {
  GC_DENSITY_VALUE = 0,           /* f(p) = p                                    arity 1 */
  GC_DENSITY_GRADIENT_ENERGY = 1, /* f(p, g) = |g|^2                             arity 2 */
  GC_DENSITY_CAHN_HILLIARD = 2,   /* -p^2/2 + p^4/4 + a/2 |g|^2 (a = kappa)       arity 2 */
  GC_DENSITY_GINZBURG_LANDAU = 3, /* |g|^2/2 + a (p^2 - 1)^2 (a = W)              arity 3 */
  GC_DENSITY_PFC = 4,             /* (q0^4 - eps)/2 p^2 + q0^2 p lap p + p bih p / 2 + p^4/4
                                     (a = q0, b = eps)                           arity 4 */
  GC_DENSITY_LINEAR_ELASTIC = 5   /* Vec3 field u: e = sym(grad u),
                                     (lambda tr(e)^2 + 2 mu e:e) / 2 (a = lambda, b = mu) */
} gc_density_kind;

typedef struct gc_density {
  int32_t kind; /* gc_density_kind */
  double a, b;
} gc_density;

/* func::evaluate(field, f, Pairwise|Kahan) for a recognised density f, fused: the derivatives
 * each site needs, the density and the reduction in one pass (func/Functional.hpp:88-220).
 * field is scalar, or Vec3 (3 values per site) for GC_DENSITY_LINEAR_ELASTIC.  Extent checks
 * as the materialised operators: radius 1 for any gradient or Laplacian, radius 2 for the
 * biharmonic. */
gc_status gc_evaluate(gc_context* ctx, const gc_grid* g, gc_dtype dt, const gc_density* f,
                      int method, const void* field, double* result);
/* Cahn-Hilliard free energy func::evaluate(psi, f(p, grad p)) with
 * f = -p^2/2 + p^4/4 + kappa/2 |grad p|^2 (func/Functional.hpp:122-130,
 * examples/common/cahn_hilliard.hpp:64-68), fused: gradient + density + reduction. */
gc_status gc_ch_free_energy(gc_context* ctx, const gc_grid* g, gc_dtype dt, double kappa,
                            int method, const void* psi, double* F);
/* Gradient energy func::evaluate(psi, |grad p|^2) (examples/common/cahn_hilliard.hpp:73-77). */
gc_status gc_gradient_energy(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                             const void* psi, double* E);

/* ---- pointwise integrator steps (generic right-hand-side path) --------------- */

/* out = a + scale*b (solve::detail::axpyFresh, solve/RK4.hpp:51-64). */
gc_status gc_axpy(gc_context* ctx, gc_dtype dt, size_t n, const void* a, double scale,
                  const void* b, void* out);
/* psi += dt*r (solve/ExplicitEuler.hpp:106-110). */
gc_status gc_euler_update(gc_context* ctx, gc_dtype dt, size_t n, void* psi, double step,
                          const void* r);
/* psi += (dt/6)*(k1+k4) + (dt/3)*(k2+k3) (solve/RK4.hpp:145-152). */
gc_status gc_rk4_combine(gc_context* ctx, gc_dtype dt, size_t n, void* psi, double step,
                         const void* k1, const void* k2, const void* k3, const void* k4);

/* ---- fused time stepping --------------------------------------------------- */

/* Cahn-Hilliard right-hand side M*lap(-c + c^3 - kappa*lap c), Order 2
 * (examples/common/cahn_hilliard.hpp:38-59), fused into one kernel. */
gc_status gc_ch_rhs(gc_context* ctx, const gc_grid* g, gc_dtype dt, const gc_ch_params* p,
                    const void* psi, void* out);
/* solve::integrate(psi, computeRhs, dt, n_steps, RK4|ExplicitEuler) for the
 * Cahn-Hilliard right-hand side (solve/RK4.hpp:99-154, solve/ExplicitEuler.hpp:79-112):
 * every RK4 stage is one kernel that computes mu, lap(mu) and the stage update with
 * one pass over HBM.  psi is updated in place. */
gc_status gc_ch_integrate(gc_context* ctx, const gc_grid* g, gc_dtype dt, const gc_ch_params* p,
                          void* psi, double step, int n_steps, int integrator);
/* solve::diffuse(psi, D, dt, n_steps, tag) with constant D (solve/Diffusion.hpp:89-112). */
gc_status gc_diffuse_const(gc_context* ctx, const gc_grid* g, gc_dtype dt, void* psi, double D,
                           double step, int n_steps, int integrator);
/* solve::diffuse(psi, Field D, dt, n_steps, tag) (solve/Diffusion.hpp:161-216). */
gc_status gc_diffuse_field(gc_context* ctx, const gc_grid* g, gc_dtype dt, void* psi,
                           const void* D, double step, int n_steps, int integrator,
                           int64_t* first_bad);

/* ---- tile-shape autotuner of the fused Cahn-Hilliard stage kernel ------------------
 * No counterpart in the reference (which has no GPU kernels).  The stage kernel tiles
 * the grid in (OX x OY) columns marched through chunks of ZC planes (2-D: OX x OY
 * tiles).  Candidate shapes are numbered 0..gc_ch_config_count()-1 per kernel family
 * (2-D when Nz == 1, 3-D when Nz >= 8; grids with 1 < Nz < 8 are untiled and have no
 * candidates).  The choice is remembered per (extents, dtype) for the process. */

/* Number of candidate shapes for this grid's kernel family (0 if untiled). */
int gc_ch_config_count(const gc_grid* g);
/* Tile shape of candidate `id` (as run in FP32; two-stage candidates whose FP32 tile needs
 * more LDS than a workgroup has in FP64 run 64 x 16 x zc in FP64). */
gc_status gc_ch_config_shape(const gc_grid* g, int id, int* ox, int* oy, int* zc);
/* Nonzero if candidate `id` fuses RK4 stages (same results bit for bit): 3-D candidates run
 * RK4 as two kernels of two stages each (S1+S2, S3+S4: 9 words of HBM traffic per site and
 * step instead of 17), 1 for the kernel that keeps three planes of every intermediate in
 * LDS, 2 for the kernel that keeps the z direction in registers; 3 for a 2-D candidate that
 * runs all four stages of a step in one kernel (the tile and an 8-site halo in LDS, 2 words
 * per site and step).  0 for one kernel per stage, -1 for an invalid id.  Euler and single
 * right-hand-side evaluations always use one-stage kernels. */
int gc_ch_config_fused(const gc_grid* g, int id);
/* Forces candidate `id` for (extents, dtype); -1 restores the built-in default. */
gc_status gc_ch_set_config(const gc_grid* g, gc_dtype dt, int id);
/* Candidate currently used for (extents, dtype). */
gc_status gc_ch_get_config(const gc_grid* g, gc_dtype dt, int* id);
/* Times one RK4 step (4 or 2 kernels) of every candidate on scratch data, keeps the
 * fastest for (extents, dtype) and returns its id; if ms is not NULL it receives the
 * measured ms per step of each candidate (gc_ch_config_count entries). */
gc_status gc_ch_tune(gc_context* ctx, const gc_grid* g, gc_dtype dt, int* best, float* ms);

/* ---- batched ensembles of 2-D Cahn-Hilliard grids ---------------------------------
 * No counterpart in the reference: many independent periodic 2-D grids (uneven extents,
 * their own M, kappa and cell size, one dtype, one time step) advanced by explicit RK4
 * with the right-hand side of examples/common/cahn_hilliard.hpp:38-59.  Each RK4 step is
 * ONE kernel: a block runs all four stages of one (member, tile) work item in LDS
 * (temporal blocking with an 8-site halo), so a step reads each state once (plus halos)
 * and writes it once.  Results equal the per-member solve::integrate(..., RK4) arithmetic
 * up to the sign of an exact zero.  States live in the ensemble's own device pool. */

typedef struct gc_ensemble gc_ensemble;

typedef struct gc_ensemble_member {
  int32_t nx, ny;
  double hx, hy;
  double M, kappa;
} gc_ensemble_member;

/* Creates an ensemble of n members with zero initial states. */
gc_status gc_ensemble_create(gc_context* ctx, gc_dtype dt, int n,
                             const gc_ensemble_member* members, double step,
                             gc_ensemble** out);
gc_status gc_ensemble_destroy(gc_ensemble* e);
/* Copies member i's state (nx*ny values of the ensemble's dtype) from / to the host. */
gc_status gc_ensemble_set_state(gc_ensemble* e, int i, const void* host);
gc_status gc_ensemble_get_state(gc_ensemble* e, int i, void* host);
/* Device pointer of member i's current state (valid until the next step). */
gc_status gc_ensemble_state_ptr(gc_ensemble* e, int i, void** dev);
/* Total sites over all members. */
gc_status gc_ensemble_sites(gc_ensemble* e, size_t* sites);
/* Advances every member n_steps RK4 steps; use_graph != 0 replays a captured hipGraph of
 * 20 steps (the first call for a shape captures it). */
gc_status gc_ensemble_step(gc_ensemble* e, int n_steps, int use_graph);
/* Tile shapes (OX x OY) of the step kernel: query, force, or time them all (states are
 * preserved) and keep the fastest; ms (optional) receives ms per step of each shape. */
gc_status gc_ensemble_config_shape(gc_ensemble* e, int id, int* ox, int* oy);
gc_status gc_ensemble_set_config(gc_ensemble* e, int id);
gc_status gc_ensemble_tune(gc_ensemble* e, int* best, float* ms);
/* Per-member total mass (integrate) and free energy (evaluate with the Cahn-Hilliard
 * density), FP64 reductions; either output may be NULL. */
gc_status gc_ensemble_diagnostics(gc_ensemble* e, int method, double* mass, double* F);

/* ---- learned corrector for 2-D Cahn-Hilliard ---------------------------------------
 * No counterpart in the reference.  A 6-layer 3x3 periodic CNN (1 -> 32 x5 -> 1 channels,
 * GELU, residual output) that maps a 2-D state c(t) to c(t + 16 dt):
 *   c_new = c + scale * (out - mean(out))      (the mean removal conserves mass)
 * The 32 -> 32 layers run on the matrix cores with BF16 weights and activations; the
 * 1 -> 32 and 32 -> 1 layers take the state and their weights in FP32; every accumulation
 * is FP32.  Parameters: one FP32 vector of gc_cnn_param_count() values in the order
 * W1[co][tap] b1 {W_l[co][tap][ci] b_l}x4 W6[tap][ci] b6 (tap = 3(dy+1) + (dx+1)).
 * Images are H x W with H a multiple of 8 and W a multiple of 32. */

typedef struct gc_cnn gc_cnn;

int gc_cnn_param_count(void);
/* The network runs on ctx's stream; destroy it before ctx. */
gc_status gc_cnn_create(gc_context* ctx, int H, int W, int batch_max, gc_cnn** out);
gc_status gc_cnn_destroy(gc_cnn* n);
/* Host FP32 parameters and the output scale; resets the optimiser state. */
gc_status gc_cnn_set_params(gc_cnn* n, const float* host, double scale);
gc_status gc_cnn_get_params(gc_cnn* n, float* host);
/* Gradient of the last gc_cnn_loss_grad or gc_cnn_loss_grad_chain call (host FP32,
 * param-count values). */
gc_status gc_cnn_get_grads(gc_cnn* n, float* host);
/* Network output `out` (before scale and mean removal) for B device images of dtype dt;
 * out_dev (device FP32, B*H*W) may be NULL.  One kernel runs all six layers per 8 x 32
 * tile (inference); gc_cnn_forward_layers runs the per-layer kernels of the training path.
 * Both give the same bits. */
gc_status gc_cnn_forward(gc_cnn* n, gc_dtype dt, const void* c, int B, float* out_dev);
gc_status gc_cnn_forward_layers(gc_cnn* n, gc_dtype dt, const void* c, int B, float* out_dev);
/* One training evaluation: images c[idx[b]] and targets target[idx[b]] (device arrays of
 * dtype dt; idx is a host array of B indices, or NULL for 0..B-1).  Loss is the mean of
 * (out - (target - c)/scale)^2; its gradient goes to the parameter gradient buffer.  If
 * loss is not NULL it receives the loss (synchronous). */
gc_status gc_cnn_loss_grad(gc_cnn* n, gc_dtype dt, const void* c, const void* target,
                           const int* idx, int B, double* loss);
/* Unrolled-chain training evaluation (back-propagation through K residual applications).
 * data is a device array of n_images images of dtype dt; idx (host, B entries) names the
 * starting image of each chain, whose targets after k = 1..K applications are images
 * idx[b] + k (so idx[b] + K < n_images).  With c_0 = data[idx[b]] and
 * c_k = c_{k-1} + scale (out(c_{k-1}) - mean(out(c_{k-1}))) (per image, FP32 states), the loss
 * is sum_k weights[k-1] mean(((c_k - data[idx + k]) / scale)^2) (weights NULL: 1/K each);
 * its gradient with respect to the parameters, through every application, goes to the
 * parameter gradient buffer.  1 <= K <= 8.  If loss is not NULL it receives the loss
 * (synchronous).  The network's chain buffers are allocated on the first call (and again
 * for a larger K). */
gc_status gc_cnn_loss_grad_chain(gc_cnn* n, gc_dtype dt, const void* data, size_t n_images,
                                 const int* idx, int B, int K, const double* weights,
                                 double* loss);
/* Adam update of the parameters from the current gradient. */
gc_status gc_cnn_adam(gc_cnn* n, double lr, double beta1, double beta2, double eps);
/* c_new = c + scale*(out - mean(out)) for B device images (c_new may not alias c). */
gc_status gc_cnn_predict(gc_cnn* n, gc_dtype dt, const void* c, void* c_new, int B);

/* Advances a 2-D Cahn-Hilliard state (grid H x W x 1, matching the network) by n_macro
 * macro steps of `substeps` RK4 steps of size `step`.  Each macro step takes the
 * network's proposal unless the state's mean drifts from its initial value by more than
 * mass_tol or the free energy increases; then it recomputes the macro step with
 * gc_ch_integrate (RK4).  n_fallback (may be NULL) receives the number of recomputed
 * macro steps.  psi is updated in place.  Proposals are made up to 16 macro steps ahead
 * of the host's decisions (one synchronisation per 16 macro steps); the result is that of
 * deciding after every macro step. */
gc_status gc_learned_integrate(gc_cnn* n, const gc_grid* g, gc_dtype dt, const gc_ch_params* p,
                               void* psi, double step, int substeps, int n_macro,
                               double mass_tol, int* n_fallback);
/* gc_learned_integrate with one network per segment of macro steps (no counterpart in the
 * reference; the phase-specialised corrector): macro step m is proposed by the first
 * network i with first[i] <= m < end[i], and runs as plain RK4 (gc_ch_integrate) where no
 * network covers it.  The networks must share one context and be built for the grid.  The
 * proposal pipeline continues across segment boundaries, and the mass drift is measured
 * from psi's mean at the call.  gc_learned_integrate(n, ...) is this with the one segment
 * [0, n_macro). */
gc_status gc_learned_integrate_phases(gc_cnn* const* nets, const int* first, const int* end,
                                      int n_nets, const gc_grid* g, gc_dtype dt,
                                      const gc_ch_params* p, void* psi, double step,
                                      int substeps, int n_macro, double mass_tol,
                                      int* n_fallback);

/* Element-wise conversion between GC_F64 and GC_F32 device arrays. */
gc_status gc_cast(gc_context* ctx, gc_dtype from, gc_dtype to, size_t n, const void* src,
                  void* dst);

#ifdef __cplusplus
}
#endif

#endif /* GRIDCALC_HIP_H */
