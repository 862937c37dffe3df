/*
 * jetshow_amd.h -- C ABI of the MI355X per-pixel radiative-transfer solver.
 *
 * This is the drop-in boundary for jetshow's Observation::run.  The reference
 * binds that path through C++ classes and one pybind11 function; every entry
 * point below names the reference interface it replaces.  All arguments are
 * plain old data, pointers and sizes; no exception crosses the boundary; every
 * function returns a jsx_status (0 = success) and leaves a message for
 * jsx_last_error().
 *
 * Units follow the reference: CGS throughout (cm, G, cm^-3, Hz), angles in
 * radians, grid coordinates in parsec (as SimulationInterpolater uses them).
 */
#ifndef JETSHOW_AMD_H
#define JETSHOW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JSX_ABI_VERSION 2

typedef enum jsx_status {
    JSX_OK = 0,
    JSX_EINVAL = -1,       /* bad argument or descriptor                        */
    JSX_ENOMEM = -2,       /* device allocation failed                          */
    JSX_EHIP = -3,         /* HIP runtime error (message in jsx_last_error)     */
    JSX_EUNSUPPORTED = -4, /* descriptor kind the kernels do not implement      */
    JSX_EBUDGET = -5,      /* request does not fit the context's HBM budget     */
    JSX_EBUSY = -6,        /* queue full / request still running               */
    JSX_ENOTFOUND = -7,    /* unknown handle or request id                      */
    JSX_ENODEV = -8        /* no GPU / HIP code object could not be loaded      */
} jsx_status;

/* ---- jet description: src/Jet.cpp:5-11 composition of Geometry, VField,
 *      BField, NField, flattened into one tagged POD ------------------------ */

typedef enum jsx_geometry_kind {
    JSX_GEOM_CONE = 1,        /* src/Cone.cpp:16-118 (double-nappe cone)          */
    JSX_GEOM_PARABOLOID = 2,  /* src/Parabaloid.cpp:11-60                          */
    JSX_GEOM_CYLINDER = 3,    /* src/Cylinder.cpp:9-99 (r^2 fixed; see flags)      */
    JSX_GEOM_CONVEX_HULL = 4, /* SimulationGeometry (src/Geometry.cpp:92-132) over
                                 the convex hull of the simulation's points
                                 (src/main.cpp:873-897), as face planes uploaded
                                 with jsx_hull_upload (handle in geo_table)      */
    JSX_GEOM_PARABOLOID_CONE = 5 /* ParabaloidCone (src/ParabaloidCone.cpp:9-94):
                                 paraboloid r0 = geo_param below z = geo_z0, the
                                 cone tangent to it there above                   */
} jsx_geometry_kind;

typedef enum jsx_bfield_kind {
    JSX_B_RANDOM_SCALAR = 1,   /* RandomScalarBField  b0 (|p|/pc)^-e   BField.cpp:35-41 */
    JSX_B_RANDOM_SCALAR_Z = 2, /* RandomScalarBFieldZ b0 (|z|/pc)^-e   BField.cpp:44-50 */
    JSX_B_TOROIDAL = 3,        /* ToroidalBField (-sin fi, cos fi, 0) b0 (z/pc)^-e,
                                  fi = atan(y/x)                       BField.cpp:96-106 */
    JSX_B_TOROIDAL_R = 4,      /* toroidal, b0 (|p|/pc)^-e, azimuth (-y, x)/rho  (ext.)  */
    JSX_B_HELICAL_R = 5,       /* |B| = b0 (|p|/pc)^-e, B_phi = |B| sin(pitch),
                                  B_z = |B| cos(pitch)                          (ext.)  */
    JSX_B_GRID_SCALAR = 6,     /* tangled |B| interpolated from grid b         (ext.)  */
    JSX_B_GRID_VECTOR = 7,     /* SimulationBField (-sin fi Bphi, cos fi Bphi, Bp),
                                  Bp from grid b, Bphi from grid b2    BField.cpp:196-207 */
    JSX_B_CONST_CYLINDER_Z = 8, /* ConstCylinderBFieldZ (0, 0, b0 (|z|/pc)^-e) BField.cpp:76-82 */
    JSX_B_CONST_CYLINDER = 9,  /* ConstCylinderBField (0, 0, b0 (|p|/pc)^-e)   BField.cpp:66-73 */
    JSX_B_RADIAL_CONICAL = 10, /* RadialConicalBField b0 (|p|/pc)^-e p/|p|     BField.cpp:85-94 */
    JSX_B_HELICAL_CYLINDER = 11, /* HelicalCylinderBField (b0 tan(pitch y/r),
                                  -b0 tan(pitch x/r), b0), r = hypot(x, y)     BField.cpp:109-117 */
    JSX_B_SPIRAL_CONICAL = 12, /* SpiralConicalBField, b_z = b0 / (z/pc)^2     BField.cpp:120-131 */
    JSX_B_FORCE_FREE_CYLINDRICAL = 13, /* ForceFreeCylindricalBField, I0/I1 of
                                  b_mu (both transverse components use sin fi,
                                  as the reference writes them)               BField.cpp:134-146 */
    JSX_B_COMPOSITE_RANDOM_SCALAR_Z = 14 /* CompositeRandomScalarBFieldZ: b0 (|z|/pc)^-b_exp
                                  for |z| <= b_z0, ^-b_exp2 beyond             BField.cpp:53-63 */
} jsx_bfield_kind;

typedef enum jsx_nfield_kind {
    JSX_N_CONST = 1,           /* ConstNField            NField.cpp:17-21 */
    JSX_N_BK = 2,              /* BKNField  n0 (|p|/pc)^-e NField.cpp:24-30 */
    JSX_N_BK_Z = 3,            /* BKNFieldZ n0 (|z|/pc)^-e NField.cpp:33-39 */
    JSX_N_GRID = 4,            /* SimulationNField       NField.cpp:56-59 */
    JSX_N_COMPOSITE_BK = 5     /* CompositeBKNField: BKNField(n0, n_exp) for |z| <= n_z0,
                                  BKNField(n0, n_exp2) beyond NField.cpp:42-53 */
} jsx_nfield_kind;

typedef enum jsx_vfield_kind {
    JSX_V_CONST_FLAT = 1,      /* ConstFlatVField        VField.cpp:25-29 */
    JSX_V_CONST_CENTRAL = 2,   /* ConstCentralVField     VField.cpp:61-72 */
    JSX_V_LINEAR_GAMMA_Z = 3,  /* Gamma linear in z between (z0, g0) and (z1, g1),
                                  clamped outside, v along +z            (ext.) */
    JSX_V_GRID = 4,            /* SimulationVField       VField.cpp:269-274 */
    /* Lorentz factors: gamma0 on the axis / spine, gamma1 at the border / sheath */
    JSX_V_SHEARED_FLAT = 5,    /* ShearedFlatVField (gamma0, gamma1, v_r)      VField.cpp:32-42 */
    JSX_V_SHEATH_FLAT = 6,     /* SheathFlatVField (gamma0, gamma1, v_r)       VField.cpp:45-59 */
    JSX_V_SHEARED_CENTRAL = 7, /* ShearedCentralVField (gamma0, gamma1, v_theta, v_origin)
                                  VField.cpp:75-86; theta = acos((z + o_z)/r) unless
                                  JSX_OPT_SHEARED_CENTRAL_REF_BUG              */
    JSX_V_SHEATH_CENTRAL = 8,  /* SheathCentralVField (gamma0, gamma1, v_theta, v_origin)
                                  VField.cpp:89-106                            */
    JSX_V_CONST_PARABOLIC = 9, /* ConstParabolicVField (gamma0, v_Rz0)        VField.cpp:109-129 */
    JSX_V_ACC_PARABOLIC = 10,  /* AccParabolicVField (gamma0, v_R0, v_Rz0)     VField.cpp:132-157 */
    JSX_V_SHEARED_ACC_PARABOLIC = 11, /* ShearedAccParabolicVField (gamma0, gamma1,
                                  v_R0, v_Rz0)                                 VField.cpp:164-204 */
    JSX_V_ACC_PARABOLIC_CONST_CONE = 12, /* AccParabolicConstConeVField (gamma0,
                                  v_Rz0, v_z0)                                 VField.cpp:222-240 */
    JSX_V_SHEARED_ACC_PARABOLIC_CONST_CONE = 13 /* ShearedAccParabolicConstConeVField
                                  (gamma0, gamma1, v_Rz0, v_z0)                VField.cpp:248-268 */
} jsx_vfield_kind;

typedef struct jsx_jet_desc {
    int32_t geometry, bfield, nfield, vfield;
    /* geometry */
    double geo_origin[3];
    double geo_axis[3];           /* normalised by the library, as the ctors do  */
    double geo_param;             /* cone half-angle [rad] | paraboloid r0 [cm] |
                                     cylinder radius [cm]                          */
    double geo_big_scale;         /* [cm] length of infinite path pieces          */
    double z_clip_min, z_clip_max;/* [cm] keep z in [min, max] when max > min (ext.) */
    /* B field */
    int32_t b_in_plasma_frame;    /* VectorBField::in_plasma_frame_ (BField.h:36-44) */
    int32_t grid_b, grid_b2;      /* grid handles for JSX_B_GRID_*                 */
    int32_t geo_table;            /* jsx_hull_upload handle for JSX_GEOM_CONVEX_HULL */
    double b0, b_exp, b_pitch;    /* [G], exponent, helical pitch [rad]            */
    /* electron density */
    int32_t n_in_plasma_frame;    /* NField::in_plasma_frame_ (NField.cpp:7-14)    */
    int32_t grid_n;
    double n0, n_exp;             /* [cm^-3], exponent                             */
    double thermal_fraction;      /* cold e- density / relativistic density (ext.) */
    /* bulk velocity */
    int32_t grid_v, pad1;
    double gamma0, gamma1;        /* Lorentz factors                               */
    double gamma_z0, gamma_z1;    /* [cm] for JSX_V_LINEAR_GAMMA_Z                 */
    double v_origin[3];           /* [cm] ConstCentralVField origin                */
    /* emission */
    double s;                     /* electron energy index (reference hard-codes 3.42,
                                     src/Jet.cpp:32,234)                           */
    double gamma_min;             /* k_F_r / k_C_r low-energy cutoff (utils.h:86)  */
    /* ABI 2: parameters of the further field and geometry classes */
    double geo_z0;                /* [cm] JSX_GEOM_PARABOLOID_CONE switch height   */
    double b_exp2, b_z0;          /* composite B: outer exponent, |z| switch [cm]  */
    double b_mu;                  /* ForceFreeCylindricalBField mu                 */
    double n_exp2, n_z0;          /* CompositeBKNField: outer exponent, |z| switch [cm] */
    double v_r;                   /* [cm] sheared / sheath flat V: border radius   */
    double v_theta;               /* [rad] sheared / sheath central V: angle       */
    double v_R0, v_Rz0, v_z0;     /* [cm] parabolic V fields: R0, Rz0 (paraboloid
                                     radius at z = 1 pc), z0 (cone above it)        */
} jsx_jet_desc;

/* ---- image plane: ImagePlane(pair<ulong,ulong> size, lg_start, lg_stop, los)
 *      src/ImagePlane.cpp:11-33, src/Image.cpp:44-185 ---------------------- */
typedef struct jsx_image_desc {
    int64_t n_across;             /* image_size.first  (must be even, Image.cpp:107) */
    int64_t n_along;              /* image_size.second                            */
    double lg_pixel_size_start_cm;
    double lg_pixel_size_stop_cm;
    double los_angle;             /* [rad]                                        */
} jsx_image_desc;

/* ---- solver parameters: Observation::run(n, tau_max, dt_max, tau_min, ...)
 *      src/Observation.cpp:20-41 plus the constants it hard-codes ------------ */
typedef enum jsx_stokes { JSX_STOKES_I = 1, JSX_STOKES_FULL = 4 } jsx_stokes;
typedef enum jsx_precision { JSX_PREC_FP64 = 0, JSX_PREC_MIXED = 1 } jsx_precision;

/* option bits for jsx_solver_params.flags */
#define JSX_OPT_CYLINDER_REF_BUG 0x1u  /* c0 = |w|^2 - r/pc as Cylinder.cpp:53 */
#define JSX_OPT_SHEARED_CENTRAL_REF_BUG 0x2u /* theta = acos(z + o_z / r) as
                                                VField.cpp:81 writes it (NaN for
                                                |z| > 1 cm)                    */
/* mixed precision only: run the Stokes pass over the whole path to tau_max,
 * as Observation.cpp:225-250 does, instead of from the optical-depth horizon
 * (JSX_PIX_HORIZON) */
#define JSX_OPT_FULL_DEPTH 0x4u

typedef struct jsx_solver_params {
    int32_t stokes;               /* jsx_stokes                                   */
    int32_t precision;            /* jsx_precision                                */
    int32_t n;                    /* initial step divisor (pyjetshow.cpp:150: 100) */
    int32_t max_failed_tries;     /* odeint failed_step_checker limit (500)       */
    double nu;                    /* [Hz] source-frame frequency                   */
    double tau_max;               /* optical-depth truncation                      */
    double tau_min;               /* I skipped when tau <= tau_min                 */
    double dt_max;                /* [cm] step bound handed to the dense stepper;
                                     Observation::run multiplies its argument by pc
                                     first (Observation.cpp:23)                    */
    double tau_abs_tol, tau_rel_tol;       /* Observation.cpp:156: 1e-7, 1e-14    */
    double stokes_abs_tol, stokes_rel_tol; /* Observation.cpp:269: 1e-9, 1e-9     */
    double tau_precision;         /* bisection target |tau - tau_max| (1.0)        */
    double tau_n_min, tau_n_max;  /* steps_schedule thresholds (0.1, 100)          */
    int32_t n_max_factor;         /* steps_schedule n_max = factor * n (10)        */
    uint32_t flags;               /* JSX_OPT_*                                     */
} jsx_solver_params;

/* per-pixel flag word */
#define JSX_PIX_HIT            0x01u /* Geometry::hit returned >= 1 segment        */
#define JSX_PIX_NSEG2          0x02u /* two (half-infinite) segments               */
#define JSX_PIX_TAU_CROSSED    0x04u /* tau reached tau_max, path truncated        */
#define JSX_PIX_I_SKIPPED      0x08u /* tau <= tau_min, Stokes pass not run        */
#define JSX_PIX_FAIL_MAXTRIALS 0x10u /* step-size search gave up (odeint throws)   */
#define JSX_PIX_NAN            0x20u /* a NaN coefficient or state was produced    */
#define JSX_PIX_CLIPPED        0x40u /* z clip removed part or all of the path     */
#define JSX_PIX_FAIL_BISECT    0x80u /* tau bisection hit its iteration cap        */
#define JSX_PIX_HORIZON        0x100u /* mixed: Stokes pass began at the optical-depth
                                        horizon, not at the tau_max point: past
                                        tau = 50 (Stokes I), past int (k_I - |eta|)
                                        ds = 50 (full Stokes, eta = (k_Q, k_U, k_V):
                                        the transfer operator's norm decay), so what
                                        lies deeper reaches the observer attenuated
                                        by e^-50; tau and l still run to tau_max  */

/* Output planes, [n_across][n_along] row-major (Image::getImage order,
 * src/Image.cpp:16-28).  Pointers are device pointers for jsx_observe and
 * host pointers for jsx_observe_host.  Any plane may be NULL except I. */
typedef struct jsx_outputs {
    double *tau, *I, *Q, *U, *V, *l;
    uint32_t *flags;
    uint32_t *n_rhs;              /* right-hand-side evaluations per ray           */
    uint32_t *n_steps;            /* accepted + rejected Dormand-Prince tries      */
} jsx_outputs;

/* grid: values[iz * nr + ir] at (z0 + iz dz, r0 + ir dr) in parsec */
typedef struct jsx_grid_desc {
    int32_t nz, nr;
    double z0, dz, r0, dr;
    double nan_value;             /* SimulationInterpolater nan_value (0, Gamma 1) */
} jsx_grid_desc;

typedef struct jsx_ctx jsx_ctx;

/* ---- context ------------------------------------------------------------- */
/* Creates a solver context on `device` with an HBM admission budget in bytes
 * (0 = 90% of free memory).  Replaces the implicit global state of
 * Observation (src/Observation.cpp:12-17). */
int jsx_ctx_create(int device, size_t hbm_budget, jsx_ctx **out);
int jsx_ctx_destroy(jsx_ctx *ctx);
const char *jsx_last_error(jsx_ctx *ctx);
int jsx_abi_version(void);

/* Launch-shape and scheduling knobs of the batched solver (set by the
 * autotuner, jetshow_amd/tune.py), and counters of the last batched launch.
 * No reference counterpart (the reference has no GPU launch shape).
 *   "blocks_per_cu"   persistent workgroups per CU (1..16)
 *   "queue_block"     ray ids per queue block (>= 1)
 *   "steal"           1: per-XCD queues with stealing, 0: one global queue
 *   "defer_rhs"       right-hand sides after which a mixed full-Stokes ray is
 *                     split into time-parallel pieces (0 disables; default
 *                     512; on the 512^2 paraboloid image 256 lets one
 *                     pixel's Stokes I drift 1.6e-5 from FP64, 384 and up
 *                     keep 5e-7)
 *   "split_w_min"     smallest stiffness per piece
 *   "piece_layout"    how split pieces are integrated: 0 by Dormand-Prince
 *                     in a group of 5 lanes (P and the 4 propagator columns);
 *                     1 by Dormand-Prince in one lane holding all 20
 *                     components; 2 by exponential (4th-order Magnus)
 *                     propagators, one lane per piece; 3 (default) 1 or 2 per
 *                     ray, from cost estimates sampled along its rest
 *   "piece_target"    exponential pieces the planner aims at in the first
 *                     round (0: one per resident lane)
 *   "piece_budget"    right-hand sides after which an exponential piece hands
 *                     the rest of its interval to new pieces in the next of 4
 *                     rounds (0: one round, no budget)
 *   "mg_cost"         layout 3: ratio of estimated costs below which a ray's
 *                     pieces are exponential
 *   "mg_profile"      cuts of exponential rays placed by 0 (default) the
 *                     stiffness profile, 1 the sampled exponential-step profile
 *   "mg_rel"          smallest relative tolerance of an exponential step's
 *                     error estimate (default 5e-6; the estimate, the 2nd-order
 *                     step's error, overstates the 4th-order step's error)
 *   "mg_err"          that estimate: 0 (default) the first-order difference of
 *                     the 2nd- and 4th-order steps (their commutator terms); 1
 *                     the 2nd-order exponential formed and compared
 *   "triage"          1: a first pass writes the rays that need no integration
 *                     and queues only the others, costliest first (see
 *                     "lpt"); 0: the queues deal every ray id in image order;
 *                     2 (default): 1 for mixed-precision launches, 0 for FP64
 *                     (the FP64 grid-field workload ran 24% slower with it)
 *   "lpt"             1 (default): the triage list is ordered by estimated
 *                     cost, costliest first; 0: in triage order
 *   "park_bucket"     mixed full Stokes with "lpt": rays whose estimated cost
 *                     bucket (log2 of the sampled stiffness) is at least this
 *                     are split from their first Stokes step, without a
 *                     sequential prefix (0, the default: off)
 *   "park_wave"       long rays may also be split once at most this many
 *                     lanes of their own wavefront still hold a ray (0: off)
 *   "park_lanes"      rays are split only once the queues are dry and at most
 *                     this many lanes of the launch still hold a ray (0: one
 *                     wavefront per SIMD of the device)
 *   "split_floor_max" cap of the propagator-column error loosening (1 = off;
 *                     >1 trades Q/U/V accuracy for fewer piece steps)
 *   "timing"          1: record HIP events around the batched kernels
 *   "vis_trig"        mixed-precision visibilities: 1 (default) the hardware
 *                     sine and cosine; 0 FP32 polynomials after an exact
 *                     quarter-turn reduction (same accuracy, 2x slower)
 * Read-only (wait for the device): "stat_parked" rays split by the last
 * batched launch ("stat_parked_exp" of them with exponential pieces),
 * "stat_pieces" its pieces, "stat_piece_steps_max" / "stat_piece_steps_mean"
 * right-hand sides per piece record (after the chain-jumping rounds of
 * exponential pieces a record holds the composition of up to 16 pieces), "stat_rhs_batch" / "stat_rhs_pieces" /
 * "stat_rhs_exp" right-hand sides evaluated by its batch / Dormand-Prince
 * piece / exponential piece kernel, "stat_trips" / "stat_trips_uniform" /
 * "stat_trips_tail" the batch kernel's wavefront loop trips (all / with every
 * lane on one request: scalar constant loads / after the queues ran dry), and, with "timing", "stat_triage_ms",
 * "stat_batch_ms", "stat_piece_ms" (both piece kernels: "stat_piece_dp_ms" +
 * "stat_piece_exp_ms"), "stat_split_ms" (all split kernels) means over the
 * launches since timing was enabled ("stat_launches"). */
int jsx_ctx_set_option(jsx_ctx *ctx, const char *name, double value);
int jsx_ctx_get_option(jsx_ctx *ctx, const char *name, double *value);

/* ---- grids: SimulationInterpolater(Delaunay_triangulation*, nan_value)
 *      src/SimulationInterpolater.cpp:30-117; values are copied to HBM once
 *      and referenced by handle from jsx_jet_desc ---------------------------- */
int jsx_grid_upload(jsx_ctx *ctx, const jsx_grid_desc *desc, const double *values,
                    int32_t *handle_out);
int jsx_grid_release(jsx_ctx *ctx, int32_t handle);

/* Face planes of a convex polyhedron for JSX_GEOM_CONVEX_HULL, the GPU form of
 * SimulationGeometry's AABB tree over the convex hull of the simulation points
 * (src/Geometry.cpp:92-132, src/main.cpp:873-897): planes[4 i .. 4 i + 3] =
 * (n_x, n_y, n_z, h) in cm, a point p inside when n . p <= h for every plane
 * (n need not be unit; n_planes >= 4).  The hit is the line's entry and exit
 * points, ordered along the ray direction (the reference orders its two
 * intersection points by the sign of x, which is the same order for the rays of
 * an image plane).  Released with jsx_grid_release. */
int jsx_hull_upload(jsx_ctx *ctx, const double *planes, int32_t n_planes, int32_t *handle_out);

/* ---- device buffers for callers without their own allocator ------------- */
int jsx_alloc_outputs(jsx_ctx *ctx, int64_t n_pixels, int32_t stokes, jsx_outputs *out);
int jsx_free_outputs(jsx_ctx *ctx, jsx_outputs *out);

/* ---- the hot path -------------------------------------------------------- */
/* Observation::run + observe_single_pixel for every pixel of the image:
 * src/Observation.cpp:20-41, 349-395 (ray generation ImagePlane.cpp:11-33,
 * Geometry::hit, integrate_tau_adaptive :126-214, integrate_i_adaptive
 * :242-274 or the restored full-Stokes pass :277-308).  One kernel pass per
 * image; outputs are device pointers; runs on `stream` (NULL = the context's
 * stream) and returns without synchronising. */
int jsx_observe(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                const jsx_solver_params *params, const jsx_outputs *out_device,
                void *stream);

/* Same, with host output pointers; copies back and synchronises. */
int jsx_observe_host(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                     const jsx_solver_params *params, const jsx_outputs *out_host);

/* Geometry::hit for explicit rays (src/Geometry.h:40-43, Cone.cpp:56-118,
 * Parabaloid.cpp:39-60, Cylinder.cpp:35-95).  rays: n x 6 doubles
 * (origin, direction; direction is normalised as Ray::Ray does).  Writes the
 * segment count (0..2) per ray and segments as n x 2 x 6 doubles (point_in,
 * point_out).  Host pointers; synchronous. */
int jsx_hit_host(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_solver_params *params,
                 const double *rays, int64_t n, int32_t *n_seg, double *segments);

/* Field and geometry values at points: what the class API's virtuals
 * return -- VField::vf (include/VField.h:13), VectorBField / ScalarBField::bf
 * and bf_plasma_frame (include/BField.h:38-39,51-52), NField::nf and
 * nf_plasma_frame (include/NField.h:12-13), Geometry::is_within
 * (include/Geometry.h:50), Jet::getV / getKI / getEtaI (include/Jet.h:21,38,50)
 * and the restored Q/U/V/F/C getters (src/Jet.cpp:114-288) -- evaluated by
 * the kernels' own field code.  points: n x 3 (cm); n_los: n x 3 unit
 * vectors (lab frame); nu in Hz.  Writes JSX_FIELDS_STRIDE doubles per point
 * to out (host pointers; synchronous):
 *   [0..2]   v = vf(point)                [3]      Gamma = getG(v)
 *   [4..6]   bf(point) (a scalar field: magnitude in [4])
 *   [7]      1 for a vector field, 0 for a scalar one
 *   [8..10]  bf_plasma_frame(point, v)    [11]     nf(point)
 *   [12]     nf_plasma_frame(point, Gamma)
 *   [13]     D = getD(n_los, v)          [14..16] n_los' = get_n_los_prime
 *   [17]     nu / D
 *   [18..28] k_I, k_Q, k_U, k_V, k_F, k_C, h_Q, eta_I, eta_Q, eta_U, eta_V
 *            in the lab frame (getKI = [18], getEtaI = [25])
 *   [29]     is_within(point)             [30..31] 0
 * params may be NULL (only its option flags are read). */
#define JSX_FIELDS_STRIDE 32
int jsx_fields_host(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_solver_params *params,
                    const double *points, const double *n_los, int64_t n, double nu, double *out);

/* Observation::run_stripe + getStripe("tau") (src/Observation.cpp:310-347,
 * src/Image.cpp:31-40): optical depth of the middle row's pixels
 * n_along/2 .. n_along-1 with the fixed-step RK4 integrate_tau
 * (src/Observation.cpp:58-124; params->n steps per segment, params->tau_max).
 * Writes n_along - n_along/2 doubles to tau_host; synchronous.  Used to size
 * images before the full run (example_simple.py:105-237). */
int jsx_tau_stripe(jsx_ctx *ctx, const jsx_jet_desc *jet, const jsx_image_desc *image,
                   const jsx_solver_params *params, double *tau_host);

/* Pixel sizes [n_across][n_along] in cm, Image::getPixelSizes
 * (src/Image.cpp:56-67, 202-204). Host only, no GPU needed. */
int jsx_image_pixel_sizes(const jsx_image_desc *image, double *sizes);

/* Angular-size distance factor, mas_to_pc (src/utils.cpp:186-208). Host only. */
double jsx_mas_to_pc(double redshift);

/* ---- the reference's text formats (host only, no GPU) ----------------------
 * jsx_write_map: write_2dvector (src/utils.cpp:211-224) of a row-major plane
 * [n_rows][n_cols] into `path`, appended (append != 0, as src/main.cpp:
 * 1447-1462 opens the map files) or truncated: value / scale in the stream's
 * default format (printf %g) followed by a space, a newline per row.
 * jsx_write_stripe: write_vector (src/utils.cpp:227-236), "%g \n" per value. */
int jsx_write_map(const char *path, const double *plane, int64_t n_rows, int64_t n_cols, double scale,
                  int32_t append);
int jsx_write_stripe(const char *path, const double *values, int64_t n, double scale, int32_t append);

/* read_from_txt (src/utils.cpp:381-388): rows of three numbers (z [cm],
 * r [cm], value) until the first that does not parse; *n_points = rows read,
 * their values in zrv[3 * i + k] up to `cap` rows (cap 0: count only;
 * JSX_EINVAL when cap > 0 is too small).  JSX_ENOTFOUND: no such file. */
int jsx_read_field_text(const char *path, double *zrv, int64_t cap, int64_t *n_points);

/* create_triangulation (src/SimulationInterpolater.cpp:105-117) for points
 * that form a complete regular grid: coordinates divided by pc, distinct z and
 * r values uniformly spaced (1e-9), every node present once.  Fills `desc`
 * (and, when values != NULL, values[i * nr + j] at (z0 + i dz, r0 + j dr), cap
 * >= nz * nr) for jsx_grid_upload.  JSX_EUNSUPPORTED: the points are not such a
 * grid (jsx_grid_upload_points takes any point set). */
int jsx_grid_from_points(const double *zrv, int64_t n_points, double nan_value, jsx_grid_desc *desc,
                         double *values, int64_t cap);

/* create_triangulation (src/SimulationInterpolater.cpp:105-117) and the
 * SimulationInterpolater over it (:24-102) for any point set: rows (z [cm],
 * r [cm], value) as jsx_read_field_text returns them.  Points that form a
 * complete regular grid are laid out as by jsx_grid_from_points; any other
 * set is Delaunay-triangulated in (z/pc, r/pc) (a point repeated at the same
 * coordinates keeps its first value) and uploaded with a bucket index.
 * Lookups: the value interpolated linearly in the triangle holding the point
 * (the lowest-numbered one, for a point on a shared edge), nan_value outside
 * the points' convex hull or for a NaN result.  The handle is used like
 * jsx_grid_upload's (JetDesc grid_b / grid_b2 / grid_n / grid_v) and freed by
 * jsx_grid_release.  JSX_EUNSUPPORTED: fewer than three distinct points, or
 * all on one line. */
int jsx_grid_upload_points(jsx_ctx *ctx, const double *zrv, int64_t n_points, double nan_value,
                           int32_t *handle_out);

/* The Delaunay triangulation jsx_grid_upload_points builds, for inspection:
 * zr = (z, r) pairs in any unit; *n_tri triangles of three point indices,
 * counter-clockwise, each starting at its smallest index, sorted (cap 0:
 * count only; JSX_EINVAL if cap is too small).  No reference counterpart
 * (CGAL's Delaunay_triangulation_2 in create_triangulation, which the
 * reference does not expose). */
int jsx_triangulate(const double *zr, int64_t n_points, int32_t *tri, int64_t cap, int64_t *n_tri);

/* ---- visibilities of an image (the step after the path in the fitting loop) ----
 * Replaces FINUFFT_NUNU.forward (fourier.py:39-94: finufft nufft2d3, isign -1,
 * eps 1e-6) and FINUFFT.forward (fourier.py:11-34: nufft2d2, eps 1e-8) as
 * JetModelZoom.ft calls them (jetmodel.py:114-132), with the exact direct sum
 *
 *     vis[k] = sum_j (c_re[j] + i c_im[j]) exp(s 2 pi i (u[k] x[j] + v[k] y[j])),
 *
 * s = +1 if isign >= 0 else -1; u, v in cycles per unit of x, y (wavelengths
 * and radians), so the phase is in revolutions (finufft's frequencies are
 * 2 pi u).  c_im may be NULL (a real image).  vis: n_vis complex doubles,
 * interleaved (re, im).  precision JSX_PREC_FP64: FP64 sine and cosine;
 * JSX_PREC_MIXED: FP64 phase and reduction, FP32 sine, cosine and partial
 * sums over 64 pixels (error ~1e-7 of sum |c|).  n_src == 0 gives zeros.
 * Host buffers, synchronous. */
int jsx_visibilities(jsx_ctx *ctx, const double *x, const double *y, const double *c_re, const double *c_im,
                     int64_t n_src, const double *u, const double *v, int64_t n_vis, int32_t isign,
                     int32_t precision, double *vis);

/* The same on device buffers, asynchronous on `stream` (NULL: the context's
 * stream) -- e.g. the I plane of jsx_observe's device outputs as c_re.
 * Calls on one context share a partial-sum buffer; a launch on another
 * stream waits for the previous launch's use of it. */
int jsx_visibilities_device(jsx_ctx *ctx, const double *x, const double *y, const double *c_re,
                            const double *c_im, int64_t n_src, const double *u, const double *v, int64_t n_vis,
                            int32_t isign, int32_t precision, double *vis, void *stream);

/* The same without allocating: no reference counterpart.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The chunk partial sums go to a
 * fixed device array of JSX_VIS_SCRATCH_MAX double2 entries (64 MiB) that each
 * GPU holds until the process ends; the call issues kernel launches only (no
 * hipMalloc, hipFree or event).  A launch whose partial sums would exceed it
 * (n_chunks x n_vis, n_chunks chosen as by jsx_visibilities_device) fails
 * with JSX_EINVAL before anything runs.  Results equal
 * jsx_visibilities_device's bit for bit.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (other streams, other contexts, other
 * threads) race. */
#define JSX_VIS_SCRATCH_MAX (4 * 1024 * 1024)
int jsx_visibilities_device_noalloc(jsx_ctx *ctx, const double *x, const double *y, const double *c_re,
                                    const double *c_im, int64_t n_src, const double *u, const double *v,
                                    int64_t n_vis, int32_t isign, int32_t precision, double *vis, void *stream);

/* ---- batched and served requests (no reference counterpart beyond
 *      get_i_image called in a loop by JetModelZoom, jetmodel.py:92-99) ----- */
typedef struct jsx_request {
    jsx_jet_desc jet;
    jsx_image_desc image;
    jsx_solver_params params;
    jsx_outputs out_host;         /* host planes to copy results into           */
} jsx_request;

/* One persistent launch over many images (the ABC sweep):
 * all rays of all requests go through one work-stealing ray queue.  Requests
 * are grouped by (precision, stokes) and by the context's HBM budget; host
 * planes are filled on return. */
int jsx_observe_batch(jsx_ctx *ctx, const jsx_request *reqs, int32_t n_reqs);

/* Same with device output planes: reqs[i].out_host holds DEVICE pointers the
 * caller owns.  One persistent launch per run of consecutive requests that
 * share (precision, stokes), enqueued on `stream` (NULL = the context's
 * stream); returns without synchronising. */
int jsx_observe_batch_device(jsx_ctx *ctx, const jsx_request *reqs, int32_t n_reqs, void *stream);

/* ---- request server: continuous batching of images from many callers.
 * Replaces JetModelZoom's loop of blocking get_i_image calls
 * (jetmodel.py:92-99, called per ABC proposal from abc_sampler.py / the
 * emcee samplers) with asynchronous submission.  A server owns one kernel
 * variant (precision x stokes); requests must match it. --------------------- */
typedef struct jsx_server jsx_server;

/* field slot a grid payload fills in the request's jet descriptor */
typedef enum jsx_field_slot { JSX_FIELD_B = 1, JSX_FIELD_B2 = 2, JSX_FIELD_N = 3, JSX_FIELD_V = 4 } jsx_field_slot;

/* a grid that lives in HBM only while its request is admitted */
typedef struct jsx_grid_payload {
    int32_t field;                /* jsx_field_slot                               */
    int32_t pad;
    jsx_grid_desc desc;
    const double *values;         /* host; must stay valid until the request is done */
} jsx_grid_payload;

typedef struct jsx_request_timing {
    double t_submit, t_admit, t_done;  /* seconds, steady clock                 */
    int64_t n_pixels;
    int32_t status;                    /* jsx_status of the request             */
    int32_t pad;
} jsx_request_timing;

/* Starts the scheduler thread.  hbm_budget bounds the HBM held by admitted
 * requests (planes + grids; 0 = half the context budget); max_active bounds
 * requests in flight; epoch_budget is the number of integrator loop trips a
 * lane runs per epoch before its ray state is parked (0 = default). */
int jsx_server_create(jsx_ctx *ctx, int32_t precision, int32_t stokes, int32_t max_active, size_t hbm_budget,
                      int32_t epoch_budget, jsx_server **out);
/* Queues one image; req->out_host planes (host) are filled when it is done.
 * Grid payloads, if any, are uploaded on admission and released when done;
 * with n_grids = 0 the jet descriptor's grid handles are used as they are. */
int jsx_server_submit(jsx_server *srv, const jsx_request *req, int32_t n_grids, const jsx_grid_payload *grids,
                      uint64_t *id_out);
/* *done = 1 once the request's planes are in host memory. */
int jsx_server_poll(jsx_server *srv, uint64_t id, int32_t *done, jsx_request_timing *timing);
/* Blocks until done, fills timing, forgets the id; returns the request's status. */
int jsx_server_wait(jsx_server *srv, uint64_t id, jsx_request_timing *timing);
int jsx_server_stats(jsx_server *srv, int64_t *epochs, size_t *hbm_used);
/* Finishes every queued request, then stops the scheduler. */
int jsx_server_destroy(jsx_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* JETSHOW_AMD_H */
