/*
 * at3d_hip.h -- C ABI of the MI355X SHDOM engine (libat3d_hip.so).
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns 0 on success and a nonzero status on failure, with a NUL-terminated
 * message in the caller's `err` buffer; nothing aborts the process.  Host
 * arrays are C-ordered as documented per argument.
 *
 * The entry points replace the reference's f2py boundary (`at3d.core`, generated
 * from CloudTomography/AT3D tools/generate_pyf.py:51-149) for the solve / render
 * path; the reference routine each one stands in for is cited on it.
 */
#ifndef AT3D_HIP_H
#define AT3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AT3D_OK 0
#define AT3D_ERR_ARG 1     /* invalid argument / unsupported option      */
#define AT3D_ERR_MEMORY 2  /* device or host allocation failed           */
#define AT3D_ERR_DEVICE 3  /* HIP runtime error                          */
#define AT3D_ERR_NUMERIC 4 /* the reference would STOP here (bad trace)  */

/* Scene and numerical parameters.  Mirrors what the reference's RTE object
 * passes to the Fortran core: the property grid (at3d/solver.py:2036-2168,
 * 2324-2528), source (source.py:14-79), Lambertian surface (surface.py:28-105)
 * and numerical parameters (default_config.json, solver.py:1960-2034). */
typedef struct at3d_scene_desc {
  int32_t npx, npy, npz;   /* property grid = base RTE grid                    */
  float delx, dely;        /* km                                               */
  const float* zlevels;    /* [npz] km, increasing                             */
  int32_t npart;           /* number of scattering species                     */
  const float* extinctp;   /* [npart][npx][npy][npz]  1/km                      */
  const float* albedop;    /* [npart][npx][npy][npz]                            */
  const int32_t* iphasep;  /* [npart][npx][npy][npz][nmicro] 1-based into legenp */
  const float* phasewtp;   /* [npart][npx][npy][npz][nmicro] phase weights       */
  int32_t numphase;        /* number of tabulated phase functions              */
  int32_t nlegp;           /* highest Legendre index in legenp                 */
  const float* legenp;     /* [numphase][nlegp+1], coefficients incl. (2l+1);
                              with nstokes 3: [numphase][nlegp+1][6], the six
                              phase-matrix elements a1 a2 a3 a4 b1 b2 of each
                              order (the reference's LEGENP(NSTLEG,0:NLEG,*)) */
  /* source / surface */
  float solarflux;         /* flux on a horizontal surface                     */
  float solarmu;           /* cosine of solar zenith, direction of propagation (<0) */
  float solaraz;           /* radians                                          */
  float skyrad;            /* isotropic downwelling radiance at the domain top */
  float gndalbedo;         /* fixed Lambertian surface albedo                  */
  float wavelength;        /* microns (informational for a solar source)      */
  /* numerical parameters */
  int32_t nmu, nphi;       /* num_mu_bins, num_phi_bins                         */
  int32_t open_bc;         /* 1: open x/y boundaries, 0: periodic              */
  int32_t ip_flag;         /* independent pixel bits (0 = 3D)                  */
  int32_t deltam;
  int32_t accelflag;
  int32_t high_order_radiance;
  int32_t iterfixsh;
  int32_t angle_set;       /* 1 gaussian, 2 reduced gaussian                    */
  float shacc;             /* spherical_harmonics_accuracy                     */
  float solacc;            /* solution_accuracy                                */
  float splitacc;          /* split_accuracy; > 0 refines the grid by adaptive cell splitting */
  float transmin;          /* must be 1.0 (one-cell characteristics)           */
  double tautol;           /* render sub-step optical depth                   */
  double transcut;         /* render transmission cutoff                      */
  /* array capacities of the adaptive grid (RTE._setup_memory, solver.py:2170-2322):
   * grid points, cells, source SH terms and discrete-ordinate work words; the
   * splitting stops where the reference's would run out of them.  0 or less
   * than the base grid's: the base grid's sizes (no room to split). */
  int32_t maxig, maxic, maxiv, maxido;
  /* 1: the per-point medium (property transfer, delta-M, source and
   * single-scatter constants, the two-stream start) is computed on the
   * device by at3d_state_create instead of on the host by at3d_plan_create;
   * the plan's per-point arrays are then made on the host only when asked
   * for.  Same values either way. */
  int32_t device_medium;
  /* Stokes components (num_stokes, solver.py:150): 1 (0 is taken as 1), or 3
   * for the polarized solve and render (I, Q, U; six-element phase matrices,
   * NSTLEG = 6).  A polarized solve runs on the base grid (splitacc 0), with
   * the per-point medium made on the host, and has no Levis gradient. */
  int32_t nstokes;
  /* Phase tables mixed at each property point (num_micro, the reference's
   * MAXNMICRO; 0 is taken as 1, at most 8): iphasep / phasewtp hold nmicro
   * (table, weight) pairs per point and species, and every grid point mixes
   * 8 x nmicro of them (TRILIN_INTERP_PROP, shdom90.f90:17-346).  With more
   * than one, the solve runs on the base grid and has no Levis gradient. */
  int32_t nmicro;
  /* INTERPMETHOD (TRILIN_INTERP_PROP, shdom90.f90:161-235): 0 'ON' mixes the
   * eight corners' tables by scattering weight; 1 'OO' takes the table of the
   * corner with the most scattering (a corner at the point itself wins), as
   * the reference's verification against property files does
   * (tests/test_shdom.py:62-170).  'OO' needs nmicro = 1 and has no Levis
   * gradient. */
  int32_t interpmethod;
} at3d_scene_desc;

typedef struct at3d_plan at3d_plan;   /* host-side setup (no GPU needed)     */
typedef struct at3d_state at3d_state; /* device-resident solver state       */

/* Library version string. */
const char* at3d_version(void);

/* Device blocks of destroyed states are kept by size and reused by the next
 * state of that shape; this returns them to the runtime and gives the bytes
 * it held.  No counterpart in the reference (its arrays are host memory). */
int64_t at3d_device_trim(void);

/* Host setup of one solve: base grid and cells (NEW_GRIDS / INIT_CELL_STRUCTURE,
 * shdomsub2.f:16-309), property transfer with delta-M (TRANSFER_PA_TO_GRID,
 * shdomsub1.f:19-112), discrete ordinates and SH transform tables
 * (MAKE_ANGLE_SET / MAKE_SH_DO_COEF, shdomsub2.f:1061-1220), the Eddington
 * initial radiance (INIT_RADIANCE, shdomsub2.f:614-760), boundary points
 * (BOUNDARY_PNTS, shdomsub1.f:2173) and the per-octant sweep order
 * (SWEEPING_ORDER, shdomsub1.f:3261) with its dependency levels.  Runs on the
 * CPU only.  The direct beam (MAKE_DIRECT, shdomsub2.f:393-473) is traced per
 * state on the device by at3d_state_create, which reports its
 * "DIRECT_BEAM_PROP: ray left the property grid" error; the plan's "dirflux"
 * array is the same trace on the host, made when first asked for. */
int at3d_plan_create(const at3d_scene_desc* desc, at3d_plan** out, char* err, size_t errlen);
void at3d_plan_destroy(at3d_plan* plan);

/* Integer sizes of a plan: keys "npts","ncells","nx","ny","nx1","ny1","nz","nlm","ml","mm",
 * "nmu","nang","nphi0max","bcflag","ipflag","ntop","nbot","nlevels_max",
 * "npart","numphase","nleg","nlegp","nscatangle"; "footprint_bytes": the
 * device bytes a state of this plan will allocate (admission by memory
 * footprint; no counterpart in the reference).  Returns -1 for an unknown key. */
int64_t at3d_plan_size(const at3d_plan* plan, const char* key);

/* Copy a named plan array into `dst` (capacity `nbytes`).  Names and layouts:
 *  "gridpos" f32[npts][3]      "gridptr" i32[ncells][8]   "neighptr" i32[ncells][6]
 *  "cellflags" i8[ncells]      "xgrid" f32[nx1] "ygrid" f32[ny1] "zgrid" f32[nz]
 *  "mu" f32[nmu] "wtmu" f32[nmu] "nphi0" i32[nmu] "phi" f32[nmu][nphi0max]
 *  "wtdo" f32[nmu][nphi0max]   "ymat" f32[nang][nlm]  "wmat" f32[nang][nlm]
 *  "ylmsun" f32[nlm]           "total_ext" f32[npts]  "extinct" f32[npart][npts]
 *  "albedo" f32[npart][npts]   "dirflux" f32[npts]    "radiance0" f32[npts][4]
 *  "chi" f32[npts][ml+1]       "alb" f32[npts]        "sweepord" i32[8][npts]
 *  "level" i32[8][npts]        "bcptr_top" i32[ntop]  "bcptr_bot" i32[nbot]
 * Sweep order entries are (cell<<3)|corner, 0-based, indexed by the octant
 * number IOCT-1 of the reference (not its JOCT slot).
 * Polarized plans (nstokes 3) also give "ypol" / "wpol" f32[6][nang][nlm]
 * (the six-element synthesis and analysis tables of the ordinates),
 * "phasetab" / "phasetab2" f32[numphase][nscatangle] (I-I and I-Q phase
 * values), "legen" and "legen2" / "legen3" / "legen5" f32[numphase][nleg+1]
 * (the prepared a1, a2, a3, b1 tables), "chip" f32[npts][3][ml+1] (the
 * a2, a3, b1 source terms) and "tms5" f32[npts][ml+1]. */
int at3d_plan_get(const at3d_plan* plan, const char* name, void* dst, size_t nbytes,
                  char* err, size_t errlen);

/* Device state for a plan on the current HIP device (allocates HBM). */
int at3d_state_create(const at3d_plan* plan, at3d_state** out, char* err, size_t errlen);
void at3d_state_destroy(at3d_state* st);

/* (Re)initialise the solution: initial radiance, first source function
 * (INIT_SOLUTION, shdomsub1.f:113-442). */
int at3d_init_solution(at3d_state* st, char* err, size_t errlen);

/* Warm start: initialise the solution from the spherical-harmonic radiance of
 * another state on the same grid and angular resolution (typically the
 * converged solution of a slightly different medium), then compute the first
 * source function of this state's medium from it.  It stands where the
 * reference continues from a previous or loaded solution (RTE.solve with
 * init_solution=False, RTE.load_solution, at3d/solver.py:1621-1686). */
int at3d_init_solution_from(at3d_state* st, at3d_state* src, char* err, size_t errlen);

/* Source-function iterations until SOLCRIT <= solacc or `maxiter` total
 * iterations (SOLUTION_ITERATIONS, shdomsub1.f:445-818).  Continues from the
 * current state; writes the final criterion and total iteration count. */
int at3d_solve(at3d_state* st, int32_t maxiter, int32_t verbose, float* solcrit_out,
               int32_t* iters_out, char* err, size_t errlen);

/* Copy a named state array to host.  Names: "source" f32[npts][nlm],
 * "radiance" f32[npts][nlm], "delsource" f32[npts][nlm] (with nstokes 3
 * f32[3][npts][nlm]: the I, Q and U terms in the polarized real generalised
 * spherical harmonics of YLMALL, shdomsub2.f:4244), "fluxes" f32[2][npts],
 * "dirflux" f32[npts], "ns" i32[npts], "nr" i32[npts], "do_radiance" f32[nang][npts],
 * "do_source" f32[nang][npts] (max(0,S)*ext); and the state's current grid,
 * which adaptive cell splitting refines during at3d_solve (the base grid
 * before): "gridpos" f32[npts][3], "gridptr" i32[ncells][8], "neighptr"
 * i32[ncells][6] (0-based: c an end cell, -1 the domain boundary, -(c+2) the
 * reference's negative pointer to a split cell c), "treeptr" i32[ncells][2]
 * (parent, first child; -1 none), "cellflags" i8[ncells] (bits 2-3: split
 * direction), "total_ext" f32[npts], "extinct" / "albedo" f32[npart][npts];
 * and the state's own per-point medium as the solve uses it: "point_ext"
 * f32[npts], "point_alb" f32[npts], "point_chi" / "point_tc" /
 * "point_tms" f32[npts][ml+1] (tms times the direct flux),
 * "point_radiance0" f32[npts][4], "point_ss_off" i32[npts+1],
 * "point_ss_iph" i32[n], "point_ss_coef" f32[n] (n = point_ss_off[npts]). */
int at3d_state_get(at3d_state* st, const char* name, void* dst, size_t nbytes, char* err,
                   size_t errlen);

/* Sizes of a state's current grid: keys "npts", "ncells", "nbpts", "nbcells"
 * (the base grid's), "ntop", "nbot", "refined" (1 once cells were split).
 * Returns -1 for an unknown key. */
int64_t at3d_state_size(const at3d_state* st, const char* key);

/* Radiances of `nrays` rays from the current solution (RENDER,
 * shdomsub4.f:93-286, with INTEGRATE_1RAY shdomsub2.f:2311-2743 and the TMS
 * single-scatter correction of COMPUTE_SOURCE_1CELL_UNPOL :3046-3192).
 * x,y,z: ray positions (km); mu, phi: cosine zenith and azimuth (radians) of the
 * direction TOWARDS the sensor (at3d/sensor.py conventions); out: I per ray,
 * or with nstokes 3 the Stokes vector (I, Q, U) per ray, f32[nrays][3] (the
 * reference's STOKES(NSTOKES, NPIX); Q and U in the meridian plane of the
 * ray, the solar single scatter rotated by ROTATE_POL_PLANE shdomsub2.f:3277).
 * Fails with the reference's errors: negative tabulated phase function
 * (PRECOMPUTE_PHASE_CHECK), ray below the domain. */
int at3d_render(at3d_state* st, int64_t nrays, const float* x, const float* y, const float* z,
                const double* mu, const double* phi, int32_t single_scatter, int32_t nosurface,
                float* out, char* err, size_t errlen);

/* The grid of a saved solution, 0-based as at3d_state_get gives it (npts,
 * ncells, gridpos [npts][3], gridptr [ncells][8], neighptr [ncells][6],
 * treeptr [ncells][2], cellflags [ncells]): the base grid, or one refined by
 * adaptive cell splitting, whose optical properties then come from the
 * property grid at every point as after the reference's load_solution
 * (solver.py:1552-1605, _prepare_optical_properties).  Checked for range and
 * tree consistency; more points or cells than the memory parameters allow is
 * the reference's "Cannot load solution" error.  The state holds no
 * solution afterwards: at3d_state_load (or an initialisation) follows. */
int at3d_state_set_grid(at3d_state* st, int32_t npts, int32_t ncells, const float* gridpos,
                        const int32_t* gridptr, const int32_t* neighptr, const int32_t* treeptr,
                        const int8_t* cellflags, char* err, size_t errlen);

/* Solution transfer in the reference's saved-solution format
 * (RTE.save_solution / RTE.load_solution, at3d/solver.py:1519-1686; SAVE_STATE /
 * RESTORE_STATE of SHDOM): ragged spherical-harmonic rows, 0-based offsets.
 *   shptr[npts+1], source[shptr[npts]]      SOURCE terms of each point
 *   rshptr[npts+2], radiance[rshptr[npts]]  RADIANCE terms (rshptr[npts+1] = rshptr[npts])
 *   fluxes[2][npts]                         hemispherical fluxes (down, up)
 * on the state's current grid (set a saved grid first with at3d_state_set_grid).
 * at3d_state_load makes them the state's solution as INIT_SOLUTION does with
 * INRADFLAG false (shdomsub1.f:356, reached from RTE._init_solution after
 * load_solution, solver.py:2654-2664): no two-stream start and no first
 * source pass, DELSOURCE zero; the next at3d_solve continues from it.  Term
 * counts must be 0..nlm (source) and 1..nlm (radiance) per point.
 * at3d_state_save writes shptr and rshptr always, and source / radiance /
 * fluxes when non-null (capacities in floats; too small is AT3D_ERR_ARG).
 * A polarized state (NSTOKES 3) holds the I, Q and U terms one component
 * after another: source[3][shptr[npts]], radiance[3][rshptr[npts]] (the
 * reference's SOURCE(NSTOKES, *) transposed). */
int at3d_state_load(at3d_state* st, const int32_t* shptr, const float* source,
                    const int32_t* rshptr, const float* radiance, const float* fluxes, char* err,
                    size_t errlen);
int at3d_state_save(at3d_state* st, int32_t* shptr, int32_t* rshptr, float* source,
                    int64_t source_cap, float* radiance, int64_t radiance_cap, float* fluxes,
                    char* err, size_t errlen);

/* Diagnostics of the current solution on the device: COMPUTE_SH
 * (shdomsub2.f:1770-1845; RTE.spherical_harmonics, solver.py:1037) into
 * shterms[nshout][npts] (mean radiance, Fx, Fy, Fz, and for nshout 5 the rms
 * of the higher-order radiance terms), and COMPUTE_NETFLUXDIV
 * (shdomsub2.f:1725-1765; RTE.net_flux_divergence, solver.py:1207) into
 * netfluxdiv[npts].  Either output may be null.  Solar source. */
int at3d_state_sh_terms(at3d_state* st, int32_t nshout, float* shterms, float* netfluxdiv,
                        char* err, size_t errlen);

/* at3d_render without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as at3d_render, except
 * that the call's device tables (the direction source of every grid point
 * for every distinct ray direction, the rays and their radiances) live in
 * one library-wide device scratch array of at3d_render_scratch_bytes()
 * bytes per GPU (64 MiB), held from the library's load until the process
 * ends; a call that needs more fails with AT3D_ERR_ARG before any device
 * work.  The per-state render buffer of at3d_render is not touched.  Calls
 * on one GPU share that array: calls that overlap on one GPU (from several
 * host threads, or on several states of that GPU) race and must be
 * serialised by the caller. */
int at3d_render_noalloc(at3d_state* st, int64_t nrays, const float* x, const float* y,
                        const float* z, const double* mu, const double* phi, int32_t single_scatter,
                        int32_t nosurface, float* out, char* err, size_t errlen);
int64_t at3d_render_scratch_bytes(void);

/* Unknowns of a gradient: NUMDER partial derivatives of the optical properties
 * of one species each, on the property grid (what
 * RTE.calculate_microphysical_partial_derivatives, at3d/solver.py:1327-1516,
 * hands to PREPARE_DERIV_INTERPS, shdomsub4.f:2917-3079).  One phase weight per
 * point (num_micro = 1). */
typedef struct at3d_deriv_desc {
  int32_t numder;
  const int32_t* partder;   /* [numder] 1-based species index                     */
  const float* dext;        /* [numder][npx][npy][npz] d extinction / d unknown   */
  const float* dalb;        /* [numder][npx][npy][npz] d albedo / d unknown       */
  const int32_t* doexact;   /* [numder] 1: phase derivative from the dleg table,
                               0: from dphasewtp                                   */
  int32_t dnumphase;        /* rows of dleg (>= 1)                                */
  const float* dleg;        /* [dnumphase][nlegp+1] d Legendre coefficients incl. (2l+1) */
  const int32_t* diphasep;  /* [numder][npx][npy][npz] 1-based row of dleg        */
  const float* dphasewtp;   /* [numder][npx][npy][npz] d phase weight             */
} at3d_deriv_desc;

/* Cost and gradient by the Levis approximation (LEVISAPPROX_GRADIENT adjoint
 * branch, shdomsub4.f:288-809: forward radiances by INTEGRATE_1RAY, adjoint
 * weights by COMPUTE_ADJOINT_WEIGHTS :3969-4034, ADJOINT_INTEGRATE_1RAY
 * :3223-3967 with COMPUTE_SOURCE_GRAD_1CELL :1546-2042 and
 * COMPUTE_RADIANCE_DERIVATIVE_ADJOINT :4037-4114, direct-beam term
 * COMPUTE_DIRECT_BEAM_DERIV_ADJOINT :4117-4143 over the sun paths of
 * MAKE_DIRECT_DERIVATIVE, shdomsub5.f:1553-2003).
 * Rays as for at3d_render, grouped by pixel: rays_per_pixel[npix] (sum = nrays),
 * ray_weights[nrays], stokes_weights[npix], measurements[npix],
 * uncertainties[npix] (the 1x1 inverse error covariance of each pixel);
 * costfunc "L2" or "LL".  Outputs: grad[numder][npx][npy][npz] (float64),
 * *cost, pixels[npix] (the forward radiances). */
int at3d_levis_gradient(at3d_state* st, const at3d_deriv_desc* deriv, int64_t nrays,
                        const float* x, const float* y, const float* z, const double* mu,
                        const double* phi, int64_t npix, const int32_t* rays_per_pixel,
                        const double* ray_weights, const double* stokes_weights,
                        const float* measurements, const double* uncertainties,
                        const char* costfunc, int32_t exact_single_scatter, double* grad,
                        double* cost, float* pixels, char* err, size_t errlen);

/* The same, plus the Jacobian branch of LEVISAPPROX_GRADIENT (MAKEJACOBIAN,
 * shdomsub4.f:378-467, rays by GRAD_INTEGRATE_1RAY :811-1545 with
 * COMPUTE_RADIANCE_DERIVATIVE :2588-2697 and COMPUTE_DIRECT_BEAM_DERIV
 * :2778-2835): jacobian[npix][njac][numder] (float32) receives, per pixel, the
 * ray- and Stokes-weighted derivative of its radiance with respect to each
 * unknown at the property points jacobian_ptr[njac] (0-based flat
 * [npx][npy][npz] indices; the reference's JACOBIANPTR minus one, as
 * gradient.py:248-252 builds it).  grad and cost are those of the cost
 * function, as from at3d_levis_gradient.  njac = 0 computes no Jacobian. */
int at3d_levis_gradient_jacobian(at3d_state* st, const at3d_deriv_desc* deriv, int64_t nrays,
                                 const float* x, const float* y, const float* z, const double* mu,
                                 const double* phi, int64_t npix, const int32_t* rays_per_pixel,
                                 const double* ray_weights, const double* stokes_weights,
                                 const float* measurements, const double* uncertainties,
                                 const char* costfunc, int32_t exact_single_scatter, int32_t njac,
                                 const int32_t* jacobian_ptr, float* jacobian, double* grad,
                                 double* cost, float* pixels, char* err, size_t errlen);

/* Engine options of a state.  "sweep_general" (0/1): route every grid point
 * through the general multi-cell characteristic trace of the point-ordered
 * sweep (bit-identical to its analytic one-cell path; for tests).
 * "sweep_ring" (0/1/2, default 1): use the level-ordered sweep with per-solve
 * step records and an LDS ring when the ring fits (else the point-ordered
 * sweep); 1 leaves grids whose dependency levels are chains (under four
 * points a level, as periodic grids have) to the point-ordered sweep, which
 * is faster there; 2 takes the ring wherever it fits.  "sweep_ring4" (0/1, default 0): lay the ring out with 4-byte
 * slots (radiances alone; the sources are read from the state's arrays) even
 * where (radiance, source) pairs fit; grids whose pairs exceed LDS take it by
 * themselves (for tests; rebuilds the layout and drops the captured graphs).
 * "sweep_graph" (0/1): replay the sweep launches from a HIP graph.
 * "render_correct_interpolate" (0/1, default 1): RENDER's CORRECTINTERPOLATE
 * (shdomsub2.f:2562); 0 starts each cell after a ray's first with the source
 * and extinction it left the previous cell with.
 * "iteration_graph" (0/1, default 1): run a solve's iterations from captured
 * HIP graphs of four iterations each under device control (the convergence
 * test, FIXSH and the acceleration on the device; the host reads the control
 * block once per four iterations) when the level-ordered sweep is in use; 0,
 * verbose, or the point-ordered sweep runs them one by one with the host
 * deciding after each.
 * "max_subgrid_intervals": sub-step intervals one gradient ray may record.
 * No counterpart in the reference. */
int at3d_state_set_option(at3d_state* st, const char* key, int32_t value, char* err,
                          size_t errlen);

/* Time (ms, HIP events on the state's stream) of each kernel class over the
 * last at3d_solve call: keys "sh_to_do","sweep","do_to_sh","source","total";
 * also "iters", "splitcrit" (SPLITCRIT, the largest adaptive splitting
 * criterion of the last splitting test) and "device_bytes" (the state's
 * device allocations); of the
 * last render call "render" (direction source and rays) and "render_rays"
 * (the ray kernel alone); of the last gradient call "gradient" and
 * "adjoint_rays" (the adjoint ray kernel alone).  The sweep's layout: "sweep_ring"
 * (1 when the level-ordered sweep runs), "sweep_ring_lds" (its LDS bytes),
 * "sweep_ring_slots", "sweep_ring4" (1 for 4-byte slots), "sweep_ring_full" (1 when
 * every column takes the full instantiation), "sweep_sectors".
 * Returns a negative value for unknown keys. */
double at3d_state_timing(const at3d_state* st, const char* key);

#ifdef __cplusplus
}
#endif
#endif /* AT3D_HIP_H */
