/*
 * chmy_mi355x.h — C ABI of the MI355X (gfx950) staggered-grid stencil library.
 *
 * Drop-in boundary for Chmy.jl's hot path: the Launcher-launched kernels that
 * apply GridOperators (∂x/∂y/∂z, lerp, divg) to Fields in the explicit
 * diffusion and pseudo-transient (PT) viscoelastic Stokes + temperature
 * updates, and the bc! boundary batches.  Every entry point takes plain
 * pointers, POD descriptors and a hipStream_t passed as void*.
 *
 * Conventions
 *  - Layout: Chmy's Field storage (src/Fields/field.jl:6-18,56-62): data has
 *    size n .+ 4H per dim, x fastest; logical index I (1-based) lives at
 *    data[I .+ 2H].  chmy_field_t.n is the logical size (size(grid, loc)).
 *  - Iteration box: chmy_box_t gives logical start and length per dim.  The
 *    Launcher's box is lo = 0, len = ncells + 2 (src/KernelLaunch.jl:41,
 *    105-126: worksize n .+ 2 at Offset(-1)).  Boxes must lie in
 *    [0, ncells+1]; halos must be >= 1 so every ±1 neighbour is in storage.
 *    A box with no points along some dim launches nothing and returns 0.
 *  - Ownership: the caller owns all device memory; descriptors are read
 *    during the call only.
 *  - Status: 0 ok; > 0 a hipError_t; < 0 misuse (CHMY_E_*).  The message of
 *    the last failure on the calling thread is chmy_last_error().
 *  - Asynchrony: kernels are enqueued on `stream` and return without
 *    synchronizing, except the reductions returning host scalars.
 *  - Arithmetic: IEEE f64/f32 with no contraction; fma only where Chmy
 *    writes muladd (lerp, first-order BCs), so f64 results match Chmy's CPU
 *    semantics to rounding of the same expressions.
 */
#ifndef CHMY_MI355X_H
#define CHMY_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CHMY_BF16 (build extension): bfloat16 storage, accepted only for the
 * forcing/tau_old fields of FP32 correction sweeps and for chmy_maxabs*. */
enum { CHMY_F64 = 0, CHMY_F32 = 1, CHMY_BF16 = 2 };
enum { CHMY_CENTER = 0, CHMY_VERTEX = 1 };
enum { CHMY_DIRICHLET = 0, CHMY_NEUMANN = 1 };
enum { CHMY_VALUE_ZERO = 0, CHMY_VALUE_SCALAR = 1, CHMY_VALUE_FIELD = 2 };
enum {
  CHMY_OK = 0,
  CHMY_E_ARG = -1,      /* null / malformed argument */
  CHMY_E_DTYPE = -2,    /* dtype mismatch between fields */
  CHMY_E_LOC = -3,      /* field location or size does not match the grid */
  CHMY_E_BOX = -4,      /* iteration box outside [0, ncells+1] */
  CHMY_E_HALO = -5,     /* halo too small for the stencil */
  CHMY_E_NODEV = -6,    /* no HIP device */
  CHMY_E_SIZE = -7      /* size above a stated maximum */
};

/* Field{T,N,L,H} descriptor (src/Fields/field.jl:6-9). */
typedef struct {
  void*   data;      /* device pointer to the padded storage */
  int32_t nd;        /* 1, 2 or 3 */
  int32_t dtype;     /* CHMY_F64 / CHMY_F32 */
  int64_t n[3];      /* logical size per dim = size(grid, loc) */
  int32_t halo[3];   /* H per dim */
  int32_t loc[3];    /* CHMY_CENTER / CHMY_VERTEX per dim */
} chmy_field_t;

/* StructuredGrid (src/Grids/structured_grid.jl:5-9).  A dim with axis[d] ==
 * NULL is a UniformAxis (uniform_axis.jl:1-11, UniformGrid
 * structured_grid.jl:27-39): spacing = extent/n, inv_spacing = 1/spacing, as
 * the host computed them.  A dim with axis[d] != NULL is a FunctionAxis
 * (function_axis.jl:1-9): axis[d] is a device table of CHMY_AXIS_ROWS rows of
 * n[d] + 2 + 2*CHMY_AXIS_PAD doubles; entry (row, i) of logical index i in
 * [-CHMY_AXIS_PAD, n[d] + 1 + CHMY_AXIS_PAD] sits at
 * axis[d][row * (n[d] + 2 + 2*CHMY_AXIS_PAD) + i + CHMY_AXIS_PAD]:
 *   CHMY_AX_VERTEX   vertex(i) = f(i)
 *   CHMY_AX_CENTER   center(i) = 0.5 * (vertex(i) + vertex(i+1))   (abstract_axis.jl:18)
 *   CHMY_AX_DC       spacing(Center, i) = vertex(i+1) - vertex(i)  (abstract_axis.jl:26)
 *   CHMY_AX_DV       spacing(Vertex, i) = center(i) - center(i-1)  (abstract_axis.jl:27)
 *   CHMY_AX_IDC, CHMY_AX_IDV   inv(...) of the two                 (abstract_axis.jl:29)
 *   CHMY_AX_WCV      the Center -> Vertex lerp weight at vertex i,
 *                    0.5 * spacing(Center, i-1) * inv_spacing(Vertex, i) (interpolation.jl:29)
 * Grids with a FunctionAxis are accepted by chmy_op, chmy_op_masked,
 * chmy_bc_side, chmy_bc_batchset and chmy_set_program; every other entry
 * point returns CHMY_E_ARG for them. */
enum { CHMY_AXIS_PAD = 4, CHMY_AXIS_ROWS = 7 };
enum { CHMY_AX_VERTEX = 0, CHMY_AX_CENTER = 1, CHMY_AX_DC = 2, CHMY_AX_DV = 3, CHMY_AX_IDC = 4, CHMY_AX_IDV = 5,
       CHMY_AX_WCV = 6 };
typedef struct {
  int32_t nd;
  int32_t reserved;
  int64_t n[3];            /* cells per dim */
  double  origin[3];
  double  extent[3];
  double  spacing[3];
  double  inv_spacing[3];
  const double* axis[3];   /* FunctionAxis metric tables (device), NULL = uniform */
} chmy_grid_t;

typedef struct { int64_t lo[3]; int64_t len[3]; } chmy_box_t;

/* FirstOrderBC{T,Kind} (first_order_boundary_condition.jl:9-40): value is
 * nothing (0), a number, or a field indexed by the tangential indices. */
typedef struct {
  int32_t kind;                 /* CHMY_DIRICHLET / CHMY_NEUMANN */
  int32_t vkind;                /* CHMY_VALUE_* */
  double  value;
  const chmy_field_t* vfield;   /* device field, vkind == CHMY_VALUE_FIELD */
} chmy_bc_t;

/* ---------------------------------------------------------------- runtime */
/* Thread-local message of the last failing call ("" if none). */
const char* chmy_last_error(void);
/* Architectures.jl:46-49 Arch(backend; device_id) / set_device!: make `device`
 * current on this thread (0-based).  Also used by get_device. */
int chmy_init(int device);
int chmy_device_count(int* count);
/* KernelAbstractions.zeros(backend, T, dims) (field.jl:59): zero-filled. */
int chmy_malloc_zeros(void** ptr, size_t bytes, void* stream);
int chmy_free(void* ptr);
/* Zero `bytes` bytes at device pointer `ptr` on `stream` with a fill kernel
 * (KernelAbstractions.zeros' fill, field.jl:59), so it may sit in a captured
 * HIP graph. */
int chmy_memset_zero(void* ptr, size_t bytes, void* stream);
/* A non-blocking stream for a Launcher worker task (KernelLaunch.jl:44,
 * activate!(...; priority)); only the default priority (0) is supported, so
 * the binding maps :high to 0. */
int chmy_stream_create(int priority, void** stream);
int chmy_stream_destroy(void* stream);
int chmy_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int chmy_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
/* KernelAbstractions.synchronize(backend) (src/KernelLaunch.jl:117). */
int chmy_stream_sync(void* stream);

/* UniformAxis coord (uniform_axis.jl:18-19): out[q] = coord(ax, loc, first+q),
 * vertex(i) = fma(i-1, Δ, o), center(i) = fma(i-1, Δ, fma(0.5, Δ, o)).  Host;
 * uniform dims only (a FunctionAxis's coordinates are its table's rows). */
int chmy_coords(const chmy_grid_t* grid, int dim, int loc, int64_t first, int64_t count, double* out);

/* ---------------------------------------------------------------- bc! */
/* bc_kernel! for one (dim, side) of a FieldBatch of K fields
 * (src/BoundaryConditions/batch.jl:163-184): tangential indices 0..n_t+2,
 * fields applied in batch order.  dim is 0-based, side 1 or 2. */
int chmy_bc_side(const chmy_grid_t* grid, int dim, int side, int K,
                 const chmy_field_t* fields, const chmy_bc_t* bcs, void* stream);
/* bc!(arch, grid, batch::BatchSet) (batch.jl:20-29): dims N..1, side 1 then
 * side 2.  counts[2*dim+side-1] fields per batch, fields/bcs concatenated in
 * (dim, side) order.  One launch for the whole set (each point's final value
 * evaluated directly, bitwise equal to the per-dim sequence) when it holds at
 * most 8 distinct fields; otherwise one launch per dim. */
int chmy_bc_batchset(const chmy_grid_t* grid, const int32_t* counts,
                     const chmy_field_t* fields, const chmy_bc_t* bcs, void* stream);

/* ---------------------------------------------------------------- operators
 * Generic single-output operator over a box, the GPU form of Chmy's
 * GridOperators as used by the reference's operator tests
 * (test/test_grid_operators.jl, test_interpolations.jl):
 *   CHMY_OP_PARTIAL  dst[I] = ∂_dim src[I]           (partial_derivatives.jl:2-5)
 *   CHMY_OP_LERP     dst[I] = lerp(src, loc(dst))[I]  (interpolation.jl:14-94)
 *   CHMY_OP_DIVG     dst[I] = divg(src[0..nd-1])[I]  (field_operators.jl:50-59)
 *   CHMY_OP_VMAG     dst[I] = vmag(src[0..nd-1])[I]  (field_operators.jl:116-125)
 *   CHMY_OP_LEFT/RIGHT  leftx/rightx... (cartesian_field_operators.jl:17-28) */
enum { CHMY_OP_PARTIAL = 0, CHMY_OP_LERP = 1, CHMY_OP_DIVG = 2, CHMY_OP_VMAG = 3,
       CHMY_OP_D2 = 4, CHMY_OP_DKD = 5, CHMY_OP_LAPL = 6, CHMY_OP_DIVG_GRAD = 7,
       CHMY_OP_LEFT = 8, CHMY_OP_RIGHT = 9, CHMY_OP_HLERP = 10, CHMY_OP_DELTA = 11 };
int chmy_op(int op, const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* dst,
            int nsrc, const chmy_field_t* src, int dim, void* stream);
/* Every operator, plain or masked (src/GridOperators/masked_operators.jl:11-39,
 * masked_field_operators.jl:1-127, field_mask.jl:1-117), and the second-order
 * and harmonic ones (partial_derivatives.jl:7-21, field_operators.jl:72-81,95-104,
 * interpolation.jl:15,101):
 *   CHMY_OP_DELTA      dst[I] = δ_dim src[I]            (partial_derivatives.jl:2)
 *   CHMY_OP_D2         dst[I] = ∂²_dim src[I]           (at src's location)
 *   CHMY_OP_DKD        dst[I] = ∂k∂_dim(src, k)[I]      (k lerped, never masked)
 *   CHMY_OP_LAPL       dst[I] = lapl(src)[I]            (Σ_D ∂²_D, left-associated)
 *   CHMY_OP_DIVG_GRAD  dst[I] = divg_grad(src, k)[I]    (Σ_D ∂k∂_D)
 *   CHMY_OP_HLERP      dst[I] = hlerp(src, loc(dst))[I]
 * and the chmy_op operators.  masks (NULL = unmasked): one field per source,
 * the FieldMask component at that source's location, at(ω, loc(src[q]));
 * a masked left/right reads f[Il]·ω[Il], so masked δ, ∂, ∂², ∂k∂, divg,
 * lapl, divg_grad are built on f·ω at the knots.  LERP, HLERP, VMAG have no
 * masked form.  k: the coefficient of DKD/DIVG_GRAD (any location), NULL
 * otherwise.  Masks and k have the sources' dtype. */
int chmy_op_masked(int op, const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* dst,
                   int nsrc, const chmy_field_t* src, const chmy_field_t* masks, const chmy_field_t* k,
                   int dim, void* stream);

/* ---------------------------------------------------------------- diffusion */
/* compute_q! (examples/diffusion_2d.jl:8-13): q_D[I] = -χ ∂_D C[I]. */
int chmy_diffusion_flux(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* q,
                        const chmy_field_t* C, double chi, void* stream);
/* update_C! (examples/diffusion_2d.jl:15-19): C[I] -= Δt divg(q)[I]. */
int chmy_diffusion_update(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* C,
                          const chmy_field_t* q, double dt, void* stream);
/* compute_q! then update_C! (examples/diffusion_2d.jl:8-19, the loop body at
 * :44-48) in one launch over the same box, writing C_out (a second buffer,
 * distinct from C) instead of updating C in place; q is stored as compute_q!
 * stores it.  C_out and q equal the two launches' C and q bitwise (build
 * extension: the reference launches the two kernels). */
int chmy_diffusion_step(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* C,
                        const chmy_field_t* C_out, const chmy_field_t* q, double chi, double dt, void* stream);
/* niter steps of the diffusion loop (examples/diffusion_2d.jl:44-48) in one
 * call: chmy_diffusion_step over the Launcher box from buffer p =
 * *parity_inout into 1-p, then the C batch (bc_counts[2*dim+side-1] <= 1
 * entries, the conditions in bcs, as chmy_bc_batchset) on 1-p.  When every
 * (dim, side) has one rule with a scalar (or no) value, the batch runs inside
 * the step's launch (one launch per step, same bits).  On return
 * *parity_inout names the buffer holding C (build extension). */
int chmy_diffusion_iterate(const chmy_grid_t* grid, const chmy_field_t* C0, const chmy_field_t* C1,
                           const chmy_field_t* q, const int32_t* bc_counts, const chmy_bc_t* bcs, double chi,
                           double dt, int niter, int* parity_inout, void* stream);

/* ---------------------------------------------------------------- Stokes + T
 * tau arrays: 2D (xx, yy, xy), 3D (xx, yy, zz, xy, xz, yz) at the locations of
 * TensorField (field.jl:182-206); V arrays: VectorField (field.jl:148).
 * Viscosity: eta == NULL is the reference's scalar form (eta_s, eta_ve_s).
 * eta != NULL (build extension for variable viscosity): eta at cell centres on
 * the full padded storage; at a tau location eta = lerp(eta); per point
 * eta_ve = 1/(1/eta + 1/(G dt)); the velocity update uses
 * eta_ve(max(eta_ml[I-e_D], eta_ml[I])) with eta_ml from chmy_maxloc. */
/* update_old! (examples/stokes_3d_inc_ve_T.jl:11-21); T/T_old may be NULL. */
int chmy_stokes_update_old(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* T,
                           const chmy_field_t* tau, const chmy_field_t* T_old, const chmy_field_t* tau_old,
                           void* stream);
/* update_stress! (examples/stokes_3d_inc_ve_T.jl:23-46; 2D stokes_2d_inc_ve_T.jl:20-34). */
int chmy_stokes_stress(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* tau,
                       const chmy_field_t* Pr, const chmy_field_t* divV, const chmy_field_t* V,
                       const chmy_field_t* tau_old, const chmy_field_t* eta, double eta_s, double eta_ve_s,
                       double G, double dt, double dtau_Pr, double dtau_r, void* stream);
/* update_velocity! (examples/stokes_3d_inc_ve_T.jl:48-57; 2D :36-43); rhog
 * (z component in 3D, y in 2D, at the V_last location) may be NULL. */
int chmy_stokes_velocity(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* V,
                         const chmy_field_t* r_V, const chmy_field_t* Pr, const chmy_field_t* tau,
                         const chmy_field_t* rhog, const chmy_field_t* eta_ml, double eta_ve_s, double G,
                         double dt, double nudtau, void* stream);
/* dst[I] = max of src over the 3^N cell neighbourhood of I (build extension). */
int chmy_maxloc(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* dst,
                const chmy_field_t* src, void* stream);
/* update_thermal_flux! (examples/stokes_3d_inc_ve_T.jl:59-71). */
int chmy_thermal_flux(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* qT,
                      const chmy_field_t* T, const chmy_field_t* V, double lambda_rhoCp, void* stream);
/* update_thermal! (examples/stokes_3d_inc_ve_T.jl:73-77). */
int chmy_thermal_update(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* T,
                        const chmy_field_t* T_old, const chmy_field_t* qT, double dt, void* stream);
/* update_thermal_flux! then update_thermal! (stokes_3d_inc_ve_T.jl:167-168)
 * in one launch over the same box, writing T_out (a second buffer,
 * distinct from T) instead of T; qT is stored as update_thermal_flux! stores
 * it.  T_out and qT equal the two launches' T and qT bitwise (build
 * extension). */
int chmy_thermal_step(const chmy_grid_t* grid, const chmy_box_t* box, const chmy_field_t* T,
                      const chmy_field_t* T_out, const chmy_field_t* T_old, const chmy_field_t* qT,
                      const chmy_field_t* V, double lambda_rhoCp, double dt, void* stream);

/* One fused PT iteration = update_stress! then update_velocity! over the
 * Launcher box (examples/stokes_3d_inc_ve_T.jl:164-165), double-buffered:
 * reads (V, tau, Pr) and writes (V_out, tau_out, Pr_out), so no block reads a
 * value another block has already advanced.  tau/Pr/V outside the box keep
 * their stored values, so V_out/tau_out/Pr_out must equal the inputs there
 * (e.g. copies made once before the first iteration).  Every field has halo 1.
 * flags: CHMY_ITER_RESIDUAL also writes r_V and divV (needed only at
 * convergence checks); CHMY_ITER_EXACT evaluates the reference's expressions
 * (results bitwise those of chmy_stokes_stress + chmy_stokes_velocity);
 * without it the algebraically equal form with 7 instead of 31 divisions per
 * point is used (a few ulp per iteration).  The V boundary batch is applied
 * by the caller afterwards (chmy_bc_batchset on V_out), as Launcher's bc= does.
 * Defect correction (build extension): with f_V/f_P (vector / Center fields)
 * the iteration solves for a correction δ driven by a residual R of an f64
 * state: r_V += f_V, and Pr uses (divV - f_P); the stress forcing enters as
 * tau_old = G dt R_tau (see chmy_stokes_residual).  NULL f_V/f_P: plain PT. */
enum { CHMY_ITER_RESIDUAL = 1, CHMY_ITER_EXACT = 2, CHMY_ITER_PAIRS = 4, CHMY_ITER_SINGLE = 8 };
int chmy_stokes_iter(const chmy_grid_t* grid, const chmy_field_t* V, const chmy_field_t* tau,
                     const chmy_field_t* Pr, const chmy_field_t* V_out, const chmy_field_t* tau_out,
                     const chmy_field_t* Pr_out, const chmy_field_t* tau_old, const chmy_field_t* eta,
                     const chmy_field_t* eta_ml, const chmy_field_t* rhog, const chmy_field_t* f_V,
                     const chmy_field_t* f_P, const chmy_field_t* divV, const chmy_field_t* r_V, int flags,
                     double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                     double nudtau, void* stream);

/* niter PT iterations in one call (the inner loop of stokes_3d_inc_ve_T.jl:163-169):
 * chmy_stokes_iter from buffer set p = *parity_inout into set 1-p, then the V
 * batch on set 1-p, alternating.  The V batch: bc_counts[2*dim+side-1]
 * entries per (dim, side) in bc! order, entry q applies bcs[q] to V component
 * bc_comp[q] (at most 64 entries).  `flags` for every iteration, `flags_last`
 * for the last (e.g. CHMY_ITER_RESIDUAL for a convergence check).  On return
 * *parity_inout names the set holding the new state.
 * Build extension: with a free-slip V batch (each component Dirichlet on its
 * own dim, Neumann on the others, scalar values) and no residual fields,
 * iterations run in pairs, one pass each with the batch applied in-kernel
 * between the two (as chmy_stokes_iter2; in 3D not with rho g and forcing
 * together); CHMY_ITER_SINGLE (in `flags`) keeps single iterations, and
 * CHMY_ITER_PAIRS is accepted and ignored.  A pair writes set 1-p like one
 * iteration, so the parity flips once per pair; results are bitwise those of
 * single iterations. */
int chmy_stokes_iterate(const chmy_grid_t* grid, const chmy_field_t* V0, const chmy_field_t* tau0,
                        const chmy_field_t* Pr0, const chmy_field_t* V1, const chmy_field_t* tau1,
                        const chmy_field_t* Pr1, const chmy_field_t* tau_old, const chmy_field_t* eta,
                        const chmy_field_t* eta_ml, const chmy_field_t* rhog, const chmy_field_t* f_V,
                        const chmy_field_t* f_P, const chmy_field_t* divV, const chmy_field_t* r_V,
                        const int32_t* bc_counts, const int32_t* bc_comp, const chmy_bc_t* bcs, int niter,
                        int flags, int flags_last, double eta_s, double eta_ve_s, double G, double dt,
                        double dtau_Pr, double dtau_r, double nudtau, int* parity_inout, void* stream);

/* Two PT iterations in one pass (build extension; the reference runs
 * update_stress!, update_velocity! and the V batch per iteration,
 * examples/stokes_3d_inc_ve_T.jl:164-165): chmy_stokes_iter from (V, tau, Pr),
 * the V batch, and chmy_stokes_iter again into (V_out, tau_out, Pr_out), with
 * the first iteration's state never stored: the batch between the two is
 * evaluated in-kernel.  The V batch (given as in chmy_stokes_iterate) must be
 * the free-slip batch: each component Dirichlet on its own dim, Neumann on the
 * others, scalar values, at least 2 cells per dim; else CHMY_E_ARG.  The
 * caller applies the batch to V_out afterwards, as after chmy_stokes_iter.
 * flags: CHMY_ITER_EXACT only.  In 3D, rho g and f_V/f_P not together.
 * Results are bitwise those of the two chmy_stokes_iter calls with
 * chmy_bc_batchset between them. */
int chmy_stokes_iter2(const chmy_grid_t* grid, const chmy_field_t* V, const chmy_field_t* tau,
                      const chmy_field_t* Pr, const chmy_field_t* V_out, const chmy_field_t* tau_out,
                      const chmy_field_t* Pr_out, const chmy_field_t* tau_old, const chmy_field_t* eta,
                      const chmy_field_t* eta_ml, const chmy_field_t* rhog, const chmy_field_t* f_V,
                      const chmy_field_t* f_P, const int32_t* bc_counts, const int32_t* bc_comp,
                      const chmy_bc_t* bcs, int flags, double eta_s, double eta_ve_s, double G, double dt,
                      double dtau_Pr, double dtau_r, double nudtau, void* stream);

/* chmy_stokes_iter2 that also stores the first iteration's V, before its V
 * batch, into V_first (build extension).  In the physical time loop the
 * reference runs update_thermal_flux! and update_thermal! after every PT
 * iteration, reading V (examples/stokes_3d_inc_ve_T.jl:164-169,
 * stokes_2d_inc_ve_T.jl:145-151); with V_first
 * two PT iterations can run as one pass and both thermal updates still see
 * their own V.  V_first is written at exactly the points the pass writes
 * V_out; the caller applies the V batch to both.  Then V_first is bitwise
 * V after the first chmy_stokes_iter + batch.  2D or 3D, FP64, no f_V/f_P
 * (else CHMY_E_ARG or CHMY_E_DTYPE); V_first distinct from V and V_out; other
 * arguments as in chmy_stokes_iter2. */
int chmy_stokes_iter2_first(const chmy_grid_t* grid, const chmy_field_t* V, const chmy_field_t* tau,
                            const chmy_field_t* Pr, const chmy_field_t* V_out, const chmy_field_t* tau_out,
                            const chmy_field_t* Pr_out, const chmy_field_t* V_first, const chmy_field_t* tau_old,
                            const chmy_field_t* eta, const chmy_field_t* eta_ml, const chmy_field_t* rhog,
                            const int32_t* bc_counts, const int32_t* bc_comp, const chmy_bc_t* bcs, int flags,
                            double eta_s, double eta_ve_s, double G, double dt, double dtau_Pr, double dtau_r,
                            double nudtau, void* stream);

/* Defect correction (build extension; the reference has no mixed-precision
 * path).  Residuals of an f64 state in the reference's expression order
 * (update_stress!/update_velocity!, examples/stokes_3d_inc_ve_T.jl:23-57):
 * R_tau = -(tau - tau_old)/(G dt) - tau/eta + 2 dev(eps(V)), R_P = -divg(V),
 * R_V = -grad(Pr) + divg(tau) - rhog (+ f_ext, an optional f64 vector field:
 * the forcing of an adjoint solve), over the Launcher box.  Writes the
 * correction forcing in the dtype of the output fields: tau_old_c = G dt R_tau,
 * f_P = R_P, f_V = R_V; divV/r_V (f64, optional) receive divg(V) and R_V for
 * the residual norms.  Viscosity as in chmy_stokes_stress. */
int chmy_stokes_residual(const chmy_grid_t* grid, const chmy_field_t* V, const chmy_field_t* tau,
                         const chmy_field_t* Pr, const chmy_field_t* tau_old, const chmy_field_t* eta,
                         const chmy_field_t* rhog, const chmy_field_t* f_ext, double eta_s, double G, double dt,
                         const chmy_field_t* tau_old_c, const chmy_field_t* f_V, const chmy_field_t* f_P,
                         const chmy_field_t* divV, const chmy_field_t* r_V, void* stream);
/* x += d over the Launcher box for (V, tau, Pr) in f64, d in f32 or f64. */
int chmy_stokes_accumulate(const chmy_grid_t* grid, const chmy_field_t* V, const chmy_field_t* tau,
                           const chmy_field_t* Pr, const chmy_field_t* V_c, const chmy_field_t* tau_c,
                           const chmy_field_t* Pr_c, void* stream);
/* Gradient of a functional J of the converged steady Stokes state with respect
 * to the Center viscosity field (build extension).  tau: the forward solve's
 * stresses; tau_a: the stresses of the adjoint solve, i.e. the same PT problem
 * with forcing f_V = -dJ/dV, homogeneous boundary values and tau_old = 0
 * (the steady operator is symmetric up to Pr -> -lambda_P).  Writes
 * grad = dJ/deta (logscale = 0) or dJ/dlog(eta) (logscale != 0) at centres
 * 0..n+1:  [interior] sum_d tau_a_dd tau_dd / (2 eta^2) + 1/4 sum over the
 * active edges that lerp eta from the centre of tau_a_e tau_e / eta_e^2. */
int chmy_stokes_grad_eta(const chmy_grid_t* grid, const chmy_field_t* tau, const chmy_field_t* tau_a,
                         const chmy_field_t* eta, const chmy_field_t* grad, int logscale, void* stream);

/* ---------------------------------------------------------------- SIA (2D, f64)
 * Shallow-ice nonlinear diffusion of examples/sia.jl with the lazy fields
 * S = B + H, ∇S, D = (A H^(n+2) + As H^n) |∇S|^(n-1), M = min(mb_a (S - mb_b), mb_c)
 * evaluated at every use.  A (per-cell rate factor) may be NULL (A = 1, the
 * reference's D, sia.jl:46-49).  H, B, A are Center fields with halos (H's
 * halo from its Neumann batch, B's filled by the caller). */
/* update_q! (sia.jl:12-16): q at vertices 1..n in both dims. */
int chmy_sia_flux(const chmy_grid_t* grid, const chmy_field_t* q, const chmy_field_t* H, const chmy_field_t* B,
                  const chmy_field_t* A, double As, int npow, void* stream);
/* update_H! (sia.jl:7-10) over cells 1..n (H_out may alias H). */
int chmy_sia_update(const chmy_grid_t* grid, const chmy_field_t* H, const chmy_field_t* H_out,
                    const chmy_field_t* q, const chmy_field_t* B, double mb_a, double mb_b, double mb_c,
                    double dt, void* stream);
/* D at cells 1..n (sia.jl:46-49), for the time-step rule Δt = Δ²/max(D)/4.1 (:76). */
int chmy_sia_diffusivity(const chmy_grid_t* grid, const chmy_field_t* D, const chmy_field_t* H,
                         const chmy_field_t* B, const chmy_field_t* A, double As, int npow, void* stream);
/* One explicit step = update_q!, the q Dirichlet batch and update_H! fused,
 * H -> H_out (distinct); apply H's Neumann batch to H_out afterwards. */
int chmy_sia_step(const chmy_grid_t* grid, const chmy_field_t* H, const chmy_field_t* H_out,
                  const chmy_field_t* B, const chmy_field_t* A, double As, int npow, double mb_a, double mb_b,
                  double mb_c, double dt, void* stream);
/* Adjoint of chmy_sia_step + Neumann batch (build extension): given the step's
 * input state H (with halos) and lam_in = dJ/dH_out (cells), writes
 * lam_out = dJ/dH (cells) and adds dJ/dlog(A) into grad_logA (if A != NULL).
 * work == NULL: one tiled pass (LDS-staged D, mu, abar).  Otherwise five
 * gather passes through 8 scratch fields {mu: Center, Dvx, gdx: (Vertex,Center),
 * Dvy, gdy: (Center,Vertex), ax, ay, HD: Center}; both give the same bits. */
int chmy_sia_step_adjoint(const chmy_grid_t* grid, const chmy_field_t* H, const chmy_field_t* lam_in,
                          const chmy_field_t* lam_out, const chmy_field_t* grad_logA, const chmy_field_t* B,
                          const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c,
                          double dt, const chmy_field_t* work, void* stream);
/* The time loop of sia.jl:71-96 with a given Δt sequence (the reference's
 * adaptive rule evaluated beforehand, or fixed): states[q] -> states[q+1]
 * by one step followed by the H Neumann batch (sia.jl:81), both in one tiled
 * launch, for q = 0..nstates-2.  Consecutive states must not share storage;
 * a plain run alternates two fields, a checkpointed run passes its
 * checkpoint fields at the stored steps (build extension: no copies). */
int chmy_sia_run(const chmy_grid_t* grid, const chmy_field_t* states, int nstates, const chmy_field_t* B,
                 const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c,
                 const double* dts, void* stream);
/* Reverse sweep of a checkpointed run (build extension): ckpt[s] holds the
 * state before step s*stride (nckpt = ceil(nsteps/stride)); each segment is
 * recomputed once into seg (2*(stride-1) Center fields, two alternating
 * halves), the recompute of segment s-1 riding in the same launches as the
 * adjoint steps of segment s.  The recompute steps also store the
 * diffusivity D of their starting state, which the adjoint steps read, in
 * device memory the library holds (2*stride Center fields per device, grown
 * on demand, kept until the process ends; calls that overlap on one device
 * race).  lam_a holds dJ/dH_N on entry; on return the field named by
 * *lam_slot (0 = lam_a, 1 = lam_b) holds dJ/dH_0, and grad_logA has
 * dJ/dlog(A) added (A != NULL). */
int chmy_sia_reverse(const chmy_grid_t* grid, const chmy_field_t* ckpt, int nckpt, int stride, int nsteps,
                     const chmy_field_t* seg, const chmy_field_t* lam_a, const chmy_field_t* lam_b,
                     const chmy_field_t* grad_logA, const chmy_field_t* B, const chmy_field_t* A, double As,
                     int npow, double mb_a, double mb_b, double mb_c, const double* dts, int* lam_slot,
                     void* stream);
/* chmy_sia_reverse without workspace arguments, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * segment states and the second adjoint buffer live in the library's
 * file-scope device scratch of 48 MiB (one per GPU, held until the process
 * ends): 2 (stride - 1) + 1 Center storages, each rounded up to 256 B, must
 * fit, else CHMY_E_SIZE and nothing runs (512 x 512 cells: stride <= 12).
 * The call allocates nothing, clears the part of the scratch it uses, and
 * leaves dJ/dH_0 in lam (which holds dJ/dH_N on entry); grad_logA as in
 * chmy_sia_reverse.  Results are bitwise those of chmy_sia_reverse with
 * zero-filled seg and lam_b.  Every call on a GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams or host threads) race. */
int chmy_sia_reverse_noalloc(const chmy_grid_t* grid, const chmy_field_t* ckpt, int nckpt, int stride, int nsteps,
                             const chmy_field_t* lam, const chmy_field_t* grad_logA, const chmy_field_t* B,
                             const chmy_field_t* A, double As, int npow, double mb_a, double mb_b, double mb_c,
                             const double* dts, void* stream);

/* ---------------------------------------------------------------- set! / FunctionField */
/* Device evaluation of set!(f, grid, fun; discrete, parameters) and of
 * FunctionField values (src/Fields/field.jl:119-142,
 * src/Fields/function_field.jl:51-61).  The function is given as a point
 * program: nprog operations in SSA form, operation q writing register q from
 * registers < q; the last register is stored at each point of the box
 * (logical indices of dst; NULL = its interior 1..n, as set! does; any box
 * inside its storage, e.g. the whole storage for a FunctionField read in the
 * halo), converted to dst's dtype.  Registers are f64; each operation is one
 * IEEE operation (no contraction).  Operands:
 *   CONST a            consts[a]
 *   INDEX a, b         logical index along dim a, plus b
 *   COORD a, b, c      coord(grid, loc b, index + c) along dim a: vertex
 *                      fma(i-1, Δ, o), centre fma(i-1, Δ, fma(0.5, Δ, o))
 *                      (uniform_axis.jl:18-19)
 *   SPACING a, b, c    spacing(grid, loc b, index + c) along dim a (uniform: Δ);
 *   INV_SPACING a, b, c    its inverse.  On a FunctionAxis COORD and SPACING
 *                      read the axis table and index + c must stay in its range
 *   LOAD a, b          srcs[a] at the point's logical index plus the offsets
 *                      packed in b (one signed byte per dim, x lowest); every
 *                      point of the box plus the offsets must lie in the
 *                      source's storage (else CHMY_E_BOX)
 *   unary a            NEG ABS SQRT EXP LOG SIN COS TANH FLOOR NOT
 *   binary a, b        ADD SUB MUL DIV POW MIN MAX (Julia's min/max: NaN wins,
 *                      -0.0 < 0.0) LT LE GT GE EQ NE AND OR (results 1.0/0.0)
 *   POWI a, b          r[a]^b by repeated products left to right (Julia's
 *                      literal powers; b < 0: the same on 1/r[a]), |b| <= 64
 *   SELECT a, b, c     ifelse(r[a] != 0, r[b], r[c]);  FMA a, b, c  fma(r[a], r[b], r[c])
 * Limits: 128 operations, 64 constants, 8 source fields. */
typedef struct { int32_t op, a, b, c; } chmy_pop_t;
enum {
  CHMY_P_CONST = 0, CHMY_P_INDEX = 1, CHMY_P_COORD = 2, CHMY_P_SPACING = 3, CHMY_P_INV_SPACING = 4,
  CHMY_P_LOAD = 5,
  CHMY_P_NEG = 8, CHMY_P_ABS = 9, CHMY_P_SQRT = 10, CHMY_P_EXP = 11, CHMY_P_LOG = 12, CHMY_P_SIN = 13,
  CHMY_P_COS = 14, CHMY_P_TANH = 15, CHMY_P_FLOOR = 16, CHMY_P_NOT = 17,
  CHMY_P_ADD = 24, CHMY_P_SUB = 25, CHMY_P_MUL = 26, CHMY_P_DIV = 27, CHMY_P_POW = 28, CHMY_P_MIN = 29,
  CHMY_P_MAX = 30, CHMY_P_LT = 31, CHMY_P_LE = 32, CHMY_P_GT = 33, CHMY_P_GE = 34, CHMY_P_EQ = 35,
  CHMY_P_NE = 36, CHMY_P_AND = 37, CHMY_P_OR = 38,
  CHMY_P_POWI = 40, CHMY_P_SELECT = 48, CHMY_P_FMA = 49
};
int chmy_set_program(const chmy_grid_t* grid, const chmy_field_t* dst, const chmy_box_t* box,
                     const chmy_pop_t* prog, int nprog, const double* consts, int nconst,
                     const chmy_field_t* srcs, int nsrc, void* stream);

/* ---------------------------------------------------------------- reductions */
/* maximum(abs.(view)) over a box of a field (stokes_3d_inc_ve_T.jl:172-175),
 * exact (order-independent); result in f64 written to *out_dev (device).
 * NaN in the box yields NaN. */
int chmy_maxabs_dev(const chmy_field_t* f, const chmy_box_t* box, double* out_dev, void* stream);
/* Same, synchronizing and returning the scalar on the host. */
int chmy_maxabs(const chmy_field_t* f, const chmy_box_t* box, double* out_host, void* stream);
/* Σ f² over a box in f64, deterministic order (run-to-run identical); host result. */
int chmy_sumsq(const chmy_field_t* f, const chmy_box_t* box, double* out_host, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CHMY_MI355X_H */
