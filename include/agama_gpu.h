/*
 * agama_gpu.h -- C ABI of the MI355X (gfx950) Agama-compatible potential
 * evaluator.  Plain pointers and sizes only: every pointer argument named
 * `d_*` or documented as "device" is a device (HBM) address, e.g. a torch
 * tensor's data_ptr(); `stream` is a hipStream_t (NULL = default stream).
 * Calls are asynchronous on `stream`; the library never allocates on the
 * evaluation path.
 *
 * Output convention (Agama's, reference gpu_potential.py:8-16):
 *   phi   [N]      potential
 *   force [N][3]   -grad Phi
 *   deriv [N][6]   force derivatives [dFx/dx, dFy/dy, dFz/dz, dFx/dy, dFy/dz, dFz/dx]
 * Input positions are AoS [N][3] FP64 (the reference splits them into SoA
 * copies first, gpu_potential.py:812-814; reading AoS directly saves those
 * three copy kernels).
 *
 * Return value: 0 on success, a negative AGX_E* code otherwise
 * (agx_error_string() describes it).
 */
#ifndef AGAMA_GPU_H
#define AGAMA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AGX_OK = 0,
    AGX_E_NULL = -1,        /* required pointer is NULL                        */
    AGX_E_SIZE = -2,        /* N < 0 or a table dimension out of range         */
    AGX_E_LMAX = -3,        /* lmax > 32                                       */
    AGX_E_LAUNCH = -4,      /* kernel launch failed (hipGetLastError)          */
    AGX_E_WHAT = -5,        /* no output requested / unknown flag              */
    AGX_E_TYPE = -6,        /* unknown component type in a composite           */
    AGX_E_WORKSPACE = -7,   /* workspace too small                             */
    AGX_E_SCRATCH = -8      /* N beyond the static scratch of a no-alloc call  */
};

/* output selection flags */
#define AGX_PHI   1u
#define AGX_FORCE 2u
#define AGX_DERIV 4u

/* ------------------------------------------------------------------------
 * Multipole (spherical-harmonic BFE with quintic log-r splines).
 *
 * Tables are built on the host (agama_gpu_amd/multipole_prep.py, restating
 * the reference's _build_multipole_data, gpu_potential.py:408-626) and
 * uploaded once; this struct only carries their device addresses and the
 * scalar parameters.
 *   poly_km : [n_int][n_lm][6] Horner coefficients a0..a5 of the scaled
 *             radial functions on interval k, s in [0,1]
 *   program : the Legendre recurrence schedule, 32-byte records (device,
 *             16-byte aligned), n_groups + n_steps + 2 of them:
 *       [0, n_groups)   one per active |m| group, in increasing |m|:
 *                       int32 m, step_begin, step_end, stride s (1 or 2);
 *                       double qmm (mul_m x normalised P_mm / sin^m theta), m
 *       n_groups        all zero
 *       [n_groups + 1, n_groups + 1 + n_steps)   one per recurrence step
 *                       l = m, m + s, .. l_last of each group, in order:
 *                       int32 l, int32 offsets (int16 cos_off in the low
 *                       half, int16 sin_off in the high half), double gamma;
 *                       double alpha, beta
 *                       (P_l = alpha (c + beta) P_{l-s} + gamma P_{l-2s},
 *                       c = cos(theta) for s = 1, cos^2(theta) for s = 2;
 *                       beta = 0 where s = 1; s = 2 only where every
 *                       active term of the group has l - m even; the l = m
 *                       record's coefficients are 0;
 *                       cos_off / sin_off = 6 x column of the (l, +m) /
 *                       (l, -m) term in a poly_km slab, -1 if not active)
 *       last            all zero
 *             step_begin/step_end index the step records from 0.
 *   prog_flags : an OR of
 *             AGX_PROG_STRIDE1  every group of `program` with more than one
 *                               step has stride 1;
 *             AGX_PROG_STRIDE2  every group of `program` with more than one
 *                               step has stride 2;
 *             AGX_PROG_COS_ONLY no step of `program` has a sin term (every
 *                               sin_off is -1: tables with m >= 0 only);
 *             the kernels then run variants without the per-group stride
 *             arithmetic or the sin sums;
 *             plus L << AGX_PROG_TRIAXIAL_SHIFT (bits 8-15, even L >= 2)
 *             when `program` is the triaxial program of degree L: groups
 *             m = 0, 2, .. L, group m with the steps l = m, m + 2, .. L
 *             (stride 2), cos terms only, cos_off = 6 j for step j >= 1 and
 *             -1 for step 0 (l = m = 0) -- the program of a table with
 *             every harmonic of even l <= L and even m in 0..l, columns in
 *             program order (multipole_prep.triaxial_degree).  Requires
 *             n_groups = L/2 + 1 and n_lm = n_steps = (L/2 + 1)(L/2 + 2)/2
 *             (checked: AGX_E_WHAT otherwise); the LDS-staged path then
 *             runs a kernel with that program unrolled at compile time
 *             (degrees 2..12, the potential alone 2..10, tables up to
 *             80 KB; AGX_OPT_MP_TRIAXIAL);
 *             or L << AGX_PROG_FULL_SHIFT (bits 16-23, L >= 1) when
 *             `program` is the full program of degree L: groups m = 0, 1,
 *             .. L, group m with the steps l = m, m + 1, .. L (stride 1),
 *             cos_off = 6 j for 1 <= j <= L, -1 for step 0, and for a
 *             later step j cos_off = 6 (2j - L - 1), sin_off = 6 (2j - L)
 *             -- the program of a table with every harmonic up to l = L,
 *             columns in program order (multipole_prep.full_degree).
 *             Requires n_groups = L + 1, n_steps = (L + 1)(L + 2)/2 and
 *             n_lm = (L + 1)^2, and no triaxial degree (checked); the
 *             LDS-staged path then runs the unrolled program (even degrees
 *             2..10, tables up to 160 KB).
 *             0 is always correct; a flag set for a program that does not
 *             satisfy it gives wrong results.  No counterpart in the
 *             reference (its kernel runs every table through the same loop).
 * ---------------------------------------------------------------------- */
#define AGX_PROG_STRIDE1 1u
#define AGX_PROG_COS_ONLY 2u
#define AGX_PROG_STRIDE2 4u
#define AGX_PROG_TRIAXIAL_SHIFT 8
#define AGX_PROG_TRIAXIAL_MASK 0xFF00u
#define AGX_PROG_FULL_SHIFT 16
#define AGX_PROG_FULL_MASK 0xFF0000u

typedef struct agx_multipole {
    const double *poly_km;
    const void *program;
    int32_t n_int, n_lm, n_groups, n_steps, lmax;
    int32_t log_scaling, inner_log, outer_log;
    uint32_t prog_flags;
    int32_t pad;
    double logr_min, dlogr, inv_dlogr, invPhi0;
    double inner_s, inner_U, inner_W;
    double outer_s, outer_U, outer_W;
    /* NULL, or device memory holding [invPhi0, inner_U, inner_W, outer_U,
     * outer_W], read by the kernels in place of the five fields above: lets a
     * caller whose table is built on the device (torch.autograd) launch
     * without copying them to the host first. */
    const double *law;
    /* NULL, or device memory holding a centre [x0, y0, z0] (1.3.0): the
     * agx_multipole_* evaluation, density, sorted, no-allocation and adjoint
     * entry points then evaluate at x - centre (the adjoint's dL/dx is still
     * with respect to x).  A potential shifted along a fitted trajectory
     * (torch.autograd) so needs no shifted copy of the positions.
     * agx_multipole_pack rejects it (AGX_E_WHAT): a composite leaf's centre
     * is agx_component.center. */
    const double *origin;
} agx_multipole;

/* Potential / force / force derivatives at N points.
 * Replaces multipole_potential_kernel, multipole_force_kernel and
 * multipole_hess_kernel (reference _multipole_potential_kernel.cu:415-455,
 * 466-800) as launched by MultipolePotentialGPU._launch_eval/_launch_hess
 * (gpu_potential.py:757-782).  `what` is any OR of AGX_PHI|AGX_FORCE|AGX_DERIV;
 * unrequested outputs may be NULL. */
int agx_multipole_eval(const agx_multipole *mp, const double *d_xyz, int64_t n,
                       double *d_phi, double *d_force, double *d_deriv,
                       uint32_t what, void *stream);

/* As agx_multipole_eval, through the interval-sorted path: the points are
 * first sorted by radial interval into d_work (agx_multipole_work_size(N)
 * bytes, overwritten, no clearing needed), then every wavefront evaluates 64
 * points of one interval with the interval's coefficients in scalar
 * registers.  For tables too large for LDS and large N.  Needs n_int <= 127
 * and N < 2^31 - 8192. */
int agx_multipole_eval_sorted(const agx_multipole *mp, const double *d_xyz, int64_t n,
                              double *d_phi, double *d_force, double *d_deriv, uint32_t what,
                              void *d_work, int64_t work_bytes, void *stream);

/* As agx_multipole_eval_sorted, with the sort in a static device array of the
 * library (AGX_NOALLOC_SCRATCH_BYTES per GPU, reserved when the library's code
 * is loaded on that GPU and held until the process ends) instead of a caller
 * workspace: the call allocates nothing (no hipMalloc, no allocating library
 * routine), for callers that cannot allocate -- code under HIP graph
 * capture, serving loops, harnesses that forbid allocation.  No counterpart
 * in the reference.  N <= AGX_NOALLOC_MAX_N (AGX_E_SCRATCH otherwise).  Every
 * call on one GPU uses that GPU's one array: calls that overlap on a GPU (two
 * streams, two host threads) race and must be serialised by the caller; the
 * array keeps the last call's sort, and each call rewrites what it reads. */
#define AGX_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define AGX_NOALLOC_MAX_N 11000000ll
int agx_multipole_eval_sorted_noalloc(const agx_multipole *mp, const double *d_xyz, int64_t n,
                                      double *d_phi, double *d_force, double *d_deriv,
                                      uint32_t what, void *stream);

/* Density rho = Laplacian(Phi) / (4 pi G) at N points.
 * Replaces multipole_density_kernel (_multipole_potential_kernel.cu:809-958)
 * as launched by _launch_density (gpu_potential.py:784-793). */
int agx_multipole_density(const agx_multipole *mp, const double *d_xyz, int64_t n,
                          double inv_4piG, double *d_rho, void *stream);

/* ------------------------------------------------------------------------
 * CylSpline (azimuthal harmonics x bicubic spline in asinh-scaled R, z).
 * Tables are built on the host (agama_gpu_amd/cylspline_prep.py, restating
 * _build_cylspline_data, gpu_potential.py:1143-1253):
 *   lR_grid [nR], lz_grid [nz]     asinh(R/Rscale), asinh(z/Rscale) (device)
 *   nodes   [n_harm][nR][nz][4]    f, df/dlR, df/dlz, d2f/dlR dlz (device)
 *   m_arr   [n_harm]               azimuthal number of each harmonic, sorted
 *                                  by (|m|, cos before sin) (device)
 *   W_outer [(lmax_outer+1)^2]     outer PowerLaw coefficients, l(l+1)+m (device)
 * ---------------------------------------------------------------------- */
typedef struct agx_cylspline {
    const double *lR_grid, *lz_grid, *nodes;
    const int32_t *m_arr;
    const double *W_outer;
    int32_t nR, nz, n_harm, log_scaling, lmax_outer, mmax_outer;
    double Rscale, r0_outer;
} agx_cylspline;

/* Replaces cylspl_potential_kernel, cylspl_force_kernel, cylspl_hess_kernel
 * (_cylspl_potential_kernel.cu:871-972) as launched by CylSplinePotentialGPU
 * (gpu_potential.py:1345-1376). */
int agx_cylspline_eval(const agx_cylspline *cs, const double *d_xyz, int64_t n,
                       double *d_phi, double *d_force, double *d_deriv, uint32_t what,
                       void *stream);
/* Density = Laplacian / (4 pi G); the reference derives it from the Hessian
 * trace (gpu_potential.py:1389-1395). */
int agx_cylspline_density(const agx_cylspline *cs, const double *d_xyz, int64_t n,
                          double *d_rho, void *stream);
int64_t agx_cylspline_pack_size(void);
int agx_cylspline_pack(const agx_cylspline *cs, void *dst, int64_t size);

/* ------------------------------------------------------------------------
 * Fused composite: a sum of components, each evaluated at
 *     x_leaf = scale * (x - center)
 * and accumulated as  Phi += ampl * Phi_leaf,  force += ampl * scale * F_leaf,
 * deriv += ampl * scale^2 * D_leaf,  rho += ampl * scale^2 * rho_leaf,
 * in one launch.  Composite, Shifted, Scaled and Evolving wrappers flatten into
 * such leaves (Evolving lerp weights go into ampl).
 * Replaces the reference's per-component launches and elementwise glue of
 * CompositePotentialGPU (gpu_potential.py:1439-1529), EvolvingPotentialGPU
 * (1536-1644), ShiftedPotentialGPU (1651-1764) and ScaledPotentialGPU
 * (1767-1918), and the analytic CuPy kernels (_analytic_potentials.py:176-1165).
 *
 * Analytic parameters p[] per type:
 *   NFW            {G M, r_s, M}           Plummer   {G M, b, M}
 *   Hernquist      {G M, a, M}             Dehnen    {G M, a, gamma, M}
 *   Isochrone      {G M, b, M}             MiyamotoNagai {G M, a, b, M}
 *   LogHalo        {v0^2, r_c^2, p^2, q^2} DiskAnsatz {4 pi G Sigma, R_d, h, r_in}
 *   UniformAcceleration {a_x, a_y, a_z}
 * Expansion types carry `table`: the device address of a descriptor image
 * written by agx_multipole_pack / agx_cylspline_pack and copied to device
 * memory.
 * ---------------------------------------------------------------------- */
enum {
    AGX_T_NFW = 1, AGX_T_PLUMMER = 2, AGX_T_HERNQUIST = 3, AGX_T_DEHNEN = 4,
    AGX_T_ISOCHRONE = 5, AGX_T_MIYAMOTO_NAGAI = 6, AGX_T_LOGHALO = 7,
    AGX_T_DISK_ANSATZ = 8, AGX_T_UNIFORM_ACCEL = 9,
    AGX_T_MULTIPOLE = 16, AGX_T_CYLSPLINE = 17
};
#define AGX_MAX_COMPONENTS 16

typedef struct agx_component {
    int32_t type;
    uint32_t prog_flags;     /* AGX_T_MULTIPOLE: the table's agx_multipole.prog_flags
                                (0 is always correct); 0 for the other types.
                                Padding before version 1.2.0: callers written
                                for 1.1 must zero it (or set it as above), since
                                the kernel now trusts it */
    double p[8];
    const void *table;
    double ampl;
    double scale;
    double center[3];
} agx_component;

int agx_composite_eval(const agx_component *comps, int n_comps, const double *d_xyz, int64_t n,
                       double *d_phi, double *d_force, double *d_deriv, uint32_t what,
                       void *stream);
int agx_composite_density(const agx_component *comps, int n_comps, const double *d_xyz,
                          int64_t n, double *d_rho, void *stream);

/* Device descriptor image of a Multipole table for agx_component.table:
 * agx_multipole_pack writes agx_multipole_pack_size() bytes into the host
 * buffer `dst`; the caller copies them to device memory. */
int64_t agx_multipole_pack_size(void);
int agx_multipole_pack(const agx_multipole *mp, void *dst, int64_t size);

/* Reverse mode of agx_multipole_eval's potential+force outputs.  Given
 * dL/dPhi (d_gphi, [N], may be NULL) and dL/dF (d_gforce, [N][3], may be
 * NULL) writes dL/dx into d_gxyz [N][3] (may be NULL; a Hessian-vector
 * product) and ADDS dL/d(poly_km) into d_gpoly [n_int][n_lm][6] and
 * dL/d(invPhi0, inner_U, inner_W, outer_U, outer_W) into d_gscal [5]; the
 * caller zeroes those first.  d_work is device scratch of at least
 * agx_multipole_work_size(N) bytes (about 6 N; the particles are
 * sorted by radial interval there); it is overwritten and needs no clearing.
 * flags: AGX_WORK_SORTED if d_work already holds the sort of these d_xyz
 * for this radial grid (left by agx_multipole_eval_sorted or an earlier
 * adjoint call with the same points and grid): the sort is skipped;
 * AGX_ADJ_CENTER_GRAD: also ADDS dL/dc into d_gscal[5..7] (which then has 8
 * doubles) for positions that are x - c for a centre c: minus the sum over
 * particles of dL/dx (d_gxyz may still be NULL); other bits give
 * AGX_E_WHAT.
 * Table gradients are summed per wavefront and interval before FP64 atomic
 * adds, so their last bits depend on the order of the atomics.  Needs
 * n_int <= 127 and N < 2^31 - 8192.  No reference counterpart (the
 * reference has no derivatives with respect to positions or tables); it
 * serves torch.autograd for fitting potentials to tracer forces. */
int agx_multipole_adjoint(const agx_multipole *mp, const double *d_xyz, int64_t n,
                          const double *d_gphi, const double *d_gforce, double *d_gxyz,
                          double *d_gpoly, double *d_gscal, void *d_work, int64_t work_bytes,
                          uint32_t flags, void *stream);
#define AGX_WORK_SORTED 1u
#define AGX_ADJ_CENTER_GRAD 2u
/* Device scratch (bytes) of the interval-sorted entry points for N points. */
int64_t agx_multipole_work_size(int64_t n);

/* The raw coefficient table of a Multipole on the device and the frozen
 * choices of its host build (multipole_prep.build_multipole_data, restating
 * _build_multipole_data, gpu_potential.py:408-626): term columns in kernel
 * order, law modes, and the C2-quintic system as two dense maps
 * d2 = Mf f + Md df (nR x nR, row-major).  Used to rebuild poly_km and the law
 * scalars on the device, differentiably. */
typedef struct agx_multipole_source {
    const double *phi;      /* [nR][n_cols] Phi_lm(r_i), the coefficient file's columns */
    const double *dphi;     /* [nR][n_cols] dPhi_lm/dr(r_i) */
    const double *r;        /* [nR] radii (log-uniform) */
    const int32_t *cols;    /* [n_lm] file column of each kernel term, (0,0) first */
    const double *Mf, *Md;  /* [nR][nR] */
    int32_t nR, n_cols, n_lm;
    int32_t log_scaling;
    int32_t inner_mode;     /* 0 constant, 1 logarithmic, 2 s = 2 core */
    int32_t inner_reset;    /* s = 2 core with invPhi0 = 0 */
    int32_t outer_mode;     /* 0 Keplerian, 1 logarithmic */
    int32_t pad;
    double dlogr;
} agx_multipole_source;

/* poly_km [n_int][n_lm][6] and law [invPhi0, inner_U, inner_W, outer_U,
 * outer_W] (the array agx_multipole.law points at) from the raw table, one
 * workgroup per term; nR <= 128.  No reference counterpart: the device-side
 * counterpart of the host build, for fitting loops that change the table
 * every step. */
int agx_multipole_build(const agx_multipole_source *src, double *d_poly, double *d_law,
                        void *stream);
/* Reverse of agx_multipole_build: ADDS dL/dphi, dL/ddphi [nR][n_cols] (the
 * caller zeroes them) given dL/dpoly_km and dL/dlaw; d_work is one double of
 * device scratch.  The (0,0) column's gradient is summed with FP64 atomics. */
int agx_multipole_build_adjoint(const agx_multipole_source *src, const double *d_gpoly,
                                const double *d_glaw, double *d_gphi, double *d_gdphi,
                                double *d_work, void *stream);

/* Reverse mode of agx_cylspline_eval's potential+force outputs.  Given
 * dL/dPhi (d_gphi, [N], may be NULL) and dL/dF (d_gforce, [N][3], may be
 * NULL) writes dL/dx into d_gxyz [N][3] (may be NULL; a Hessian-vector
 * product) and ADDS dL/d(nodes) into d_gnodes [n_harm][nR][nz][4] and
 * dL/d(W_outer) into d_gW [(lmax_outer+1)^2]; the caller zeroes those first.
 * d_work is device scratch of agx_cylspline_adjoint_work_size(cs, N) bytes
 * (about 6 N + 256 per grid cell: the particles are sorted by grid cell
 * there); overwritten, needs no clearing.  A size of 0 means the grid has
 * more than 8191 cells: then d_work may be NULL and every particle adds its
 * own node gradients.  Gradients are summed per wavefront before FP64 atomic
 * adds, so their last bits depend on the order of the atomics.  N < 2^31 -
 * 2^19.  No reference counterpart (the reference has no derivatives with
 * respect to positions or tables); it serves torch.autograd for fitting disc
 * potentials to tracer forces. */
int agx_cylspline_adjoint(const agx_cylspline *cs, const double *d_xyz, int64_t n,
                          const double *d_gphi, const double *d_gforce, double *d_gxyz,
                          double *d_gnodes, double *d_gW, void *d_work, int64_t work_bytes,
                          void *stream);
int64_t agx_cylspline_adjoint_work_size(const agx_cylspline *cs, int64_t n);

/* ---------------------------------------------------------------------------
 * Ragged ensembles: many independent Multipole potentials, each with its own
 * particle range of the concatenated arrays, in one persistent launch.
 *
 * agx_batch_plan (host only) cuts potential p's counts[p] rows (ranges are
 * consecutive, in potential order) into items of at most chunk[p] rows, with
 * path[p] (AGX_BATCH_LANE, or AGX_BATCH_LDS plus an even triaxial degree up
 * to 32 in bits 8-15; AGX_E_WHAT otherwise), and orders them by
 * cost[p] * rows, largest first.  With items == NULL it only returns the
 * item count in *n_items.
 *
 * agx_multipole_batch_eval evaluates the items: d_desc holds n_pot descriptor
 * images (agx_multipole_pack) at desc_stride bytes apart; d_items the plan;
 * d_xyz [n_rows][3]; outputs as agx_multipole_eval, rows of the concatenation.
 * d_queue is one int32 of device memory the call resets (by a kernel, not a
 * memset) and uses as the work counter; concurrent calls need separate
 * queues.  Items outside [0, n_rows) or naming no potential are skipped; an
 * AGX_BATCH_LDS item whose table is larger than
 * agx_batch_lds_table_doubles() doubles is evaluated as AGX_BATCH_LANE.
 * prog_flags: the AGX_PROG_* flags that hold for every potential's program
 * (the AND of their agx_multipole.prog_flags; 0 is always correct).  No
 * reference counterpart (the reference launches once per potential,
 * gpu_potential.py:757-793).
 * ------------------------------------------------------------------------- */
#define AGX_BATCH_MAX_ITEM 16384
#define AGX_BATCH_LANE 0     /* one particle per lane, table reads from L2 */
#define AGX_BATCH_LDS 1      /* table staged in LDS per item */
typedef struct {
    int64_t start;           /* first row */
    int32_t count;           /* rows, 1..AGX_BATCH_MAX_ITEM */
    int32_t pot;             /* potential index */
    int32_t path;            /* AGX_BATCH_*; for AGX_BATCH_LDS, bits 8-15 may
                                hold the potential's triaxial degree (its
                                prog_flags bits 8-15): degrees 2..10 run the
                                unrolled program for the potential and for
                                pot+force where the descriptor's sizes agree */
    int32_t pad;
} agx_batch_item;

int agx_batch_plan(int32_t n_pot, const int64_t *counts, const int32_t *chunk,
                   const int32_t *path, const double *cost, agx_batch_item *items,
                   int64_t capacity, int64_t *n_items);
int64_t agx_batch_lds_table_doubles(void);
int agx_multipole_batch_eval(const void *d_desc, int64_t desc_stride, int32_t n_pot,
                             const agx_batch_item *d_items, int64_t n_items,
                             const double *d_xyz, int64_t n_rows, double *d_phi,
                             double *d_force, double *d_deriv, uint32_t what,
                             uint32_t prog_flags, int32_t *d_queue, void *stream);

/* ---------------------------------------------------------------------------
 * Leapfrog (kick-drift-kick) of N test particles through a time-dependent
 * potential, one launch per step, for capture in a HIP graph.
 *
 * The potential at step s (time t0 + (s+1) dt) is a schedule row: entries
 * d_sched[s * max_active + e], e < min(d_count[s], max_active), each naming a
 * leaf of d_leaves (type, parameters and table of an agx_component; its own
 * ampl/scale/center fields are ignored) and giving that leaf's weight, scale
 * and centre at that time (the flattening described for agx_component).
 * d_step is a device int32 step counter: the j-th of a call's n_steps steps
 * runs step *d_step + j, and after the last one a kernel advances *d_step by
 * n_steps, so replaying a captured graph of agx_leapfrog launches continues
 * the integration.  Steps outside [0, n_sched) leave the particles
 * untouched.
 * d_x, d_v, d_a [N][3]: positions, velocities and the acceleration at the
 * current positions and time (the caller evaluates it before the first
 * step); all three are updated in place.  No reference counterpart (the
 * reference offers evaluation only; integrators are its callers'). */
typedef struct {
    int32_t leaf, pad;
    double ampl, scale, center[3];
} agx_sched_entry;

int agx_leapfrog(const agx_component *d_leaves, int32_t n_leaves,
                 const agx_sched_entry *d_sched, const int32_t *d_count,
                 int32_t max_active, int64_t n_sched, int32_t *d_step,
                 double *d_x, double *d_v, double *d_a, int64_t n, double dt,
                 int32_t n_steps, void *stream);

/* Tuning options (process-wide; set before launches, not thread-safe):
 *   AGX_OPT_MP_PATH      Multipole evaluation path: 0 automatic (LDS-staged
 *                        when the table fits and N >= 4096, else binned for
 *                        many terms and large N, else per-lane), 1 one particle
 *                        per lane with per-lane table reads, 2 radially binned
 *                        (wave-uniform interval, scalar table reads; needs
 *                        n_int <= 127), 3 one particle per lane with the
 *                        table staged in LDS (tables up to 160 KiB; larger
 *                        ones take path 1), 4 interval-sorted where the caller
 *                        gives a workspace (agx_multipole_eval_sorted; the
 *                        Python layer then always takes it), automatic in
 *                        agx_multipole_eval
 *   AGX_OPT_MP_TILE      particles per binned tile (0 = automatic, <= 8192)
 *   AGX_OPT_MP_BIN_MIN_N smallest N for which the automatic path bins
 *   AGX_OPT_MP_BIN_MIN_TERMS smallest number of active harmonics for which
 *                        the automatic path bins
 * No reference counterpart: the reference has a single launch shape
 * (gpu_potential.py:160, 764). */
#define AGX_OPT_MP_PATH 1
#define AGX_OPT_MP_TILE 2
#define AGX_OPT_MP_BIN_MIN_N 3
#define AGX_OPT_MP_BIN_MIN_TERMS 4
#define AGX_OPT_MP_LDS_GRID 6 /* LDS path grid: this many times the workgroups that
                                 fit at once (1..64, default 4) */
#define AGX_OPT_MP_BIN_GRID 7 /* binned path grid: this many times the workgroups
                                 that fit at once (1..16, default 2) */
#define AGX_OPT_GRID_PER_CU 8 /* workgroups per CU of the grid-stride kernels (lane
                                 Multipole, CylSpline, composite, leapfrog;
                                 1..64, default 32) */
#define AGX_OPT_SORT_GRID 9   /* workgroups of the interval / cell sort's count and
                                 scatter launches (1..4096; default 0: about one
                                 per 4096 points, 64 to 1024) */
#define AGX_OPT_CYL_PATH 10   /* CylSpline evaluation: 0 (default) points outside
                                 the grid box deferred to a workgroup queue and
                                 evaluated in full wavefronts, the grids read
                                 from a copy in LDS (tables with nR + nz <= 512;
                                 larger ones take path 1), 1 one point per
                                 lane whichever branch it takes */
#define AGX_OPT_ADJ_STAGE 11  /* Multipole adjoint's table-gradient sums: 1
                                 (default) terms staged in LDS eight at a time
                                 and summed in one pass where the staging
                                 blocks and the slabs fit in 53 KiB per
                                 workgroup, 0 six cross-lane sums per term */
#define AGX_OPT_MP_TRIAXIAL 12 /* LDS-staged Multipole path: 1 (default) the
                                 unrolled kernels of triaxial and full
                                 programs where prog_flags names a degree
                                 (bits 8-15 / 16-23) that has one, 0 the
                                 generic kernels for every program */
int agx_set_option(int key, int64_t value);

/* Library identification and error text.  agx_version: major * 10000 +
 * minor * 100 + patch (10500: this header).  1.2.0: agx_component.prog_flags
 * (formerly padding) is read; a position with a NaN coordinate gives NaN
 * outputs, and one whose r^2 overflows (|x| above about 1.3e154) gives the
 * outer-law limit 0 for Multipole tables.  1.3.0: agx_multipole.origin (a new
 * last field: callers built against 1.2 pass a shorter struct and must be
 * rebuilt) and agx_multipole_adjoint's AGX_ADJ_CENTER_GRAD flag.  1.4.0: the
 * AGX_OPT_ADJ_STAGE option; CylSpline coordinates and values near the z axis
 * and the midplane, and DiskAnsatz profiles near the midplane, to full
 * relative precision (results change at the rounding level elsewhere).
 * 1.5.0: prog_flags bits 8-15 (the triaxial degree; validated against the
 * program's size) and AGX_OPT_MP_TRIAXIAL; the LDS-staged table's slabs are
 * padded to an odd number of 16-byte words (fewer LDS bank conflicts for
 * tables with an even number of terms; the LDS path admits tables up to
 * 160 KB including the padding); prog_flags bits 16-23 (the full degree,
 * validated likewise); agx_batch_item.path bits 8-15 (the triaxial degree of
 * an LDS item's potential). */
int agx_version(void);
const char *agx_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif /* AGAMA_GPU_H */
