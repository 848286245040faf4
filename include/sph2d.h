/*
 * sph2d.h -- C-ABI of the MI355X WCSPH time step (drop-in for the reference's
 * hot-path subroutines of mini-walker/SPH-2D, Taylor_Green_Vortex/).
 *
 * Conventions
 *  - Every entry point returns int: SPH2D_OK (0) or an SPH2D_ERR_* code;
 *    sph2d_last_error() gives the message of the calling thread's last failure.
 *  - Host arrays use the reference's Fortran layout: 1-D arrays of length n,
 *    2-D arrays (ld,2) column-major (x block, then y block), REAL(8) = double,
 *    INTEGER = int32. The library never keeps host pointers after a call.
 *  - A context owns device (HBM) copies of the reference's global particle_*
 *    arrays and of the subdomain_* arrays; state stays on the GPU between
 *    calls and the host synchronises only in sph2d_download_* and sph2d_sample.
 *  - One host thread per context; all work is queued on the context's stream.
 *  - The scalar entry points keep the reference's by-reference argument lists,
 *    so a Fortran bind(C) interface without VALUE can call them unchanged.
 */
#ifndef SPH2D_H
#define SPH2D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPH2D_OK = 0,
  SPH2D_ERR_ARGUMENT = 1,        /* bad parameter or size */
  SPH2D_ERR_CAPACITY = 2,        /* particle, ghost or cell capacity exceeded */
  SPH2D_ERR_HIP = 3,             /* HIP runtime failure */
  SPH2D_ERR_STATE = 4,           /* call out of order (e.g. no state uploaded) */
  SPH2D_ERR_UNSUPPORTED = 5,     /* option not built */
  SPH2D_ERR_NONFINITE = 6        /* non-finite value in the state */
};

/* Solver variants (Solver_Normalized_*_Riemann_Delta_SPS of the reference). */
enum {
  SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN = 0, /* Solver_Normalized_Modified_Linear_Riemann_Delta_SPS.f90 (default, single_step_compute.f90:137) */
  SPH2D_SOLVER_LINEAR_RIEMANN = 1,          /* Solver_Normalized_Linear_Riemann_Delta_SPS.f90 */
  SPH2D_SOLVER_HLLC = 2,                    /* Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90 (single_step_compute.f90:143) */
  SPH2D_SOLVER_MUSCL = 3,                   /* Solver_Normalized_MUSCL_Linear_Riemann_Delta_SPS.f90 (single_step_compute.f90:146) */
  /* the moving-square program (ms = 1; Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/):
   * SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN with sps_on = 0 is its default
   * Solver_Normalized_Modified_Linear_Riemann_Delta.f90 (Single_step_compute.f90:131) */
  SPH2D_SOLVER_MS_MUSCL_HLLC = 4,           /* Solver_Normalized_MUSCL_HLLC_Riemann_Delta.f90; sps_on = 1 adds the SPS
                                               stress of Solver_Normalized_HLLC_Riemann_Delta_SPS.f90 */
  SPH2D_SOLVER_MS_HLLC = 5                  /* Solver_Normalized_HLLC_Riemann_Delta_SPS.f90 (sps_on = 0:
                                               Solver_Normalized_HLLC_Riemann_Delta.f90) */
};

/* Integrators of time_iteration.f90 (leapforg_or_eular, information_module.f90:303). */
enum { SPH2D_INTEGRATOR_PC = 2, SPH2D_INTEGRATOR_RK = 3 };

/* End damping zone of single_step_compute.f90:174-260 on the fluid particles with
 * box_x - x <= damping_length: acceleration and XSPH velocity scaled by
 * 1 - 0.5 A (1 - cos(pi (x + L - Lx) / L)) (COSINE), by
 * (exp(((Lx - x) / L)^3.5) - 1) / (e - 1) (y / depth) (OPENFOAM) or by
 * 1 - 6 ((x + L - Lx) / L)^2 ((depth - y) / depth) (DUALSPH). LINEAR has no reference
 * counterpart: the Cosin_Type zone with the ramp 1 - A (x + L - Lx) / L, a linear
 * damping zone over the last damping_length of the tank. */
enum { SPH2D_DAMPING_OFF = 0, SPH2D_DAMPING_COSINE = 1, SPH2D_DAMPING_OPENFOAM = 2, SPH2D_DAMPING_DUALSPH = 3,
       SPH2D_DAMPING_LINEAR = 4 };

/* Arithmetic forms. */
enum { SPH2D_FP64 = 0, SPH2D_FP32_MIXED = 1 };

/* Re-use of the previous step's rates (time_iteration.f90:542-555, 942-957):
 * REFERENCE reads them by the new step's local index, as the reference does;
 * BY_PARTICLE carries them with the particle (gid). */
enum { SPH2D_RATE_REUSE_REFERENCE = 0, SPH2D_RATE_REUSE_BY_PARTICLE = 1 };

/* Runtime form of information_module.f90's compile-time parameters. */
typedef struct {
  int32_t struct_size;       /* sizeof(sph2d_params), checked */
  int32_t nx, ny;            /* interior_particle_number_x/_y (:75-76): lattice for sph2d_init_tgv */
  double box_x, box_y;       /* periodic box [0,box_x]x[0,box_y] (:54-57, :94-103) */
  double rho0;               /* water_rho_0 (:31) */
  double c0sq;               /* square_c_0 (:145) */
  double nu;                 /* water_kinematic_viscosity (:37); REAL(4) literal in the reference */
  double U;                  /* U_Module (:52) */
  double dx;                 /* interior_dx (:79) */
  double h;                  /* smooth_length = smooth_length_factor*dx (:135) */
  double dt;                 /* dt (:309) */
  int32_t integrator;        /* SPH2D_INTEGRATOR_* */
  int32_t solver;            /* SPH2D_SOLVER_* */
  int32_t precision;         /* SPH2D_FP64 or SPH2D_FP32_MIXED */
  int32_t rate_reuse;        /* SPH2D_RATE_REUSE_* */
  int32_t shift_on;          /* ShiftParticle_Switch (:349) */
  int32_t mls_on;            /* RefreshDensityByMLS_Switch (:324) */
  int32_t mls_every;         /* MLS_time_step (:332) */
  int32_t mls_order;         /* RefreshDensity_MLS_Order (:330): 1 or 6 */
  int32_t sample_every;      /* sampling_time_step (:336) */
  int32_t periodic;          /* DistributePeriodBoundaryParticleOrNot (:92) */
  double alpha, beta, fai;   /* Monaghan artificial viscosity (:210-212) */
  double delta, delta_alpha; /* delta-SPH (:217-218) */
  double epsilon_xsph;       /* XSPH epsilon (:207) */
  double shift_beta, shift_R;/* shift_belta, shift_R (:350-351) */
  double max_shift;          /* max_shift = 0.25*dx (:353) */
  double mls_alpha;          /* slighted_modified_alpha (:359) */
  int32_t damping;           /* end damping zone, SPH2D_DAMPING_* (EndDampingZoneType :286-288 with
                                EndDampingZone_Switch 'on' and make_wave_or_not 1) */
  double damping_length;     /* EndZoneDampingLengh (:291) */
  double damping_amplitude;  /* EndZoneDampingAmplitude (:294), Cosin_Type only; REAL(4) 0.3 */
  double water_depth;        /* water_depth (:194) */
  int32_t walls;             /* fix_ghost_layer (:90) layers of fixed ghost particles around the box
                                (DistributeFixedBoundaryParticleOrNot = 1, :86); needs periodic = 0 */
  double gravity;            /* g (:32) when Gravity_Switch (:33) is 'on', else 0: -g on the fluid's
                                y acceleration, hydrostatic term in the wall pressure */
  int32_t wavemaker_layers;  /* wave_maker_layer (:214) columns of piston wavemaker particles left of
                                x = 0 (fix_wave_maker_or_not, make_wave_or_not = 1); needs walls */
  int32_t sps_on;            /* 1: the solvers' _SPS forms (single_step_compute.f90:137, the default);
                                0: their Delta forms without the SPS stress, e.g.
                                Solver_Normalized_Modified_Linear_Riemann_Delta.f90, whose XSPH term
                                is epsilon m_j v_ij / (rho_i + rho_j) W instead of 2 epsilon (...) */
  int32_t renormalization;   /* 0: the reference's identity in place of the dyadic tensor's inverse
                                (Define_freesurface_and_Normailized_derivates.f90:168-188); 1: the
                                2x2 inverse itself, T_i = sum_j (x_j - x_i) (x) grad_i W V_j, used
                                for every normalised kernel gradient (FP64 form; singular tensors
                                keep the identity and are counted in n_singular). No reference
                                counterpart: the reference computes the inverse and discards it. */
  /* The moving-square program (Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/), ms = 1,
   * with fixed walls (walls layers around [0,box_x]x[0,box_y], the same number inside the
   * square body; sph2d_init_moving_square), the predictor-corrector and MLS order 3 or off:
   * linear EOS p = c0sq (rho - rho0) + background_p (Status_Equation.f90:56-90), 4 dx cells
   * (Information_Module.f90:189), pairs down to 1.0E-10 (Find_particle_pair.f90:73-75), no
   * free surface, the fixed-ghost MLS with mass and the characteristic correction
   * (Assign_boundary_particles.f90:611-760), XSPH inside the continuity equation, the MS
   * predictor-corrector (SPH_Time_Iteration.f90:313-347), the MS density reinitialisation,
   * the body moved at u(t) = body_speed min(t / body_ramp, 1) in x (Body_Motion_Solver.f90
   * with this closed form instead of the Motion_Body.dat spline) and its force
   * (Calculate_Body_Force.f90) on sampling steps. */
  int32_t ms;
  double background_p;       /* Background_Pressure (Information_Module.f90:487) */
  double release_time;       /* Shock_wave_release_time (:715) */
  double body_cx, body_cy;   /* Block_Center_X/_Y at t = 0 (:152-153) */
  double body_len;           /* Block_Length (:154) */
  double body_speed;         /* U of the prescribed motion */
  double body_ramp;          /* its ramp time */
} sph2d_params;

typedef struct sph2d_ctx sph2d_ctx;

/* --- the moving-square program (ms = 1) ------------------------------------------
 * The tank and body layout (no GPU needed): fluid on the nx x ny lattice of the box
 * without the points of the square (rows in y; MS/SPH_Subroutine_Module.f90:871-925),
 * then `walls` layers of fixed ghosts around the tank (bottom, right, top, left) and
 * `walls` layers inside the square's edges (MS/Initial_Particle_From_First_Step.f90
 * :178-620). Fills at most cap entries of x, y (positions), btype (-6, -7, -8; 0 for the
 * fluid) and init (particle_initial_type: 2 fluid, -3 wall, -4 body); *n_fluid and
 * *n_total receive the counts. Any output pointer may be NULL (counts only). */
int sph2d_moving_square_layout(const sph2d_params* p, int32_t cap, double* x, double* y, int32_t* btype,
                               int32_t* init, int32_t* n_fluid, int32_t* n_total);
/* The layout above on the GPU: fluid at rest, rho = rho0, p = background_p, m = rho0 dx dy;
 * ghosts at rest with rho0, m = rho0 dx dy, p = background_p; time 0. sph2d_time_step then
 * runs the MS step (SPH_Time_Iteration.f90:58-540, Iteration_Method = 2). */
int sph2d_init_moving_square(sph2d_ctx* ctx);
/* Body_Motion_Solver.f90:62-273 at physical time t (the prescribed ramp): moves the body
 * particles to the centre x(t) and sets the body velocity and acceleration used by the
 * next evaluations. sph2d_time_step calls it itself at t = i_time_step * dt (summed). */
int sph2d_body_motion(sph2d_ctx* ctx, double t);
/* The body force of the last evaluation of a sampling step (Single_step_compute.f90
 * :219-249, Calculate_Body_Force.f90:51-142): out6 = pressure x, y, viscous x, y,
 * total x, y, the force on the water (Body_Force.dat divides by 0.5 rho0 U^2 L);
 * *step receives that step (0: none yet). */
int sph2d_body_force(sph2d_ctx* ctx, double* out6, int32_t* step);

/* Counters of the last step (sph2d_get_diagnostics). */
typedef struct {
  int32_t n_fluid;           /* particle_ture_number */
  int32_t n_ghost;           /* Period_Boundary_Particle_Number */
  int32_t n_local;           /* total_particle_number_in_subdomain */
  int32_t n_ill;             /* particles typed 120 */
  int32_t n_surface;         /* subdomain_surface_particle_number after the geometric test */
  int32_t n_singular;        /* singular dyadic tensors (matrix_inversion_less_three check_error=0) of
                                fluid particles; counted when renormalization = 1 forms the tensor */
  int32_t n_mls_fallback;    /* MLS points that fell back to Shepard */
  int32_t max_cell_count;    /* largest background-cell population */
  int64_t n_pairs;           /* pair_number of the last force evaluation (counted on request) */
  int32_t rebin_needed;      /* a particle moved > h/2 since binning (stencil no longer covers 3h) */
  int32_t error_flags;       /* device-side capacity flags */
} sph2d_diag;

/* --- library --------------------------------------------------------------- */
const char* sph2d_last_error(void);
/* Reference defaults of Taylor_Green_Vortex/information_module.f90 (100x100 TGV). */
int sph2d_default_params(sph2d_params* p);

/* --- context (no reference counterpart: owns the HBM-resident arrays) ------ */
int sph2d_ctx_create(const sph2d_params* p, int32_t n_fluid_capacity, sph2d_ctx** out);
int sph2d_ctx_destroy(sph2d_ctx* ctx);

/* Distribute_Initial_Particle.f90:204-249: TGV lattice, rho from inverse Tait,
 * mass = dx*dy*rho, c = c0, e = e_0, type 2. Computed on the GPU. */
int sph2d_init_tgv(sph2d_ctx* ctx);

/* Upload the fluid particles in global (reference) order: pos/vel are (ld,2). */
int sph2d_upload_state(sph2d_ctx* ctx, int32_t n, const double* pos, const double* vel,
                       const double* rho, const double* mass, const double* energy,
                       const double* h, const int32_t* type, int32_t ld);
/* Download global arrays (particle_*; ghosts excluded): pos/vel (ld,2). Any may be NULL. */
int sph2d_download_state(sph2d_ctx* ctx, int32_t n, double* pos, double* vel, double* rho,
                         double* press, double* c, double* mass, int32_t* type, int32_t ld);

/* --- scalar and array entry points with the reference's argument lists ----- */
/* compute_kernel(dim,distance,position_difference,average_smooth_length,w,temp_dwdx)
 * -- Taylor_Green_Vortex/compute_kernel.f90:21-27. dim must be 2. */
int sph2d_compute_kernel(const int32_t* dim, const double* distance, const double* position_difference,
                         const double* average_smooth_length, double* w, double* dwdx);
/* Array form over n pairs: dx is (n,2) column-major, dwdx (n,2). */
int sph2d_compute_kernel_n(int32_t n, const double* distance, const double* dx, const double* h,
                           double* w, double* dwdx);
/* status_equation(rho,energy,type,press,c) -- Taylor_Green_Vortex/status_equation.f90:19-24. */
int sph2d_status_equation(const double* rho, const double* energy, const int32_t* type,
                          double* press, double* c);
/* Array form over n particles (EOS per single_step_compute.f90:62-73 skips type 120). */
int sph2d_status_equation_n(const sph2d_params* p, int32_t n, const double* rho, const double* energy,
                            const int32_t* type, double* press, double* c);

/* Reorder_Particle_Index -- Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/
 * Reorder_Particle_Index.f90:54-287 (argument-less there: module arrays), on DEVICE
 * arrays of n_total entries. Fluid particles (type 2) among the first n_fluid that lie
 * outside [0,box_x]x[0,box_y] (lower bound the REAL(4) -1.0E-10) become ill (120); then
 * the particles are stably partitioned into fluid, wavemaker (130), fixed ghost (-3),
 * inlet (-201), outlet (-202), every other type dropped. x, y, type and every listed
 * column (distinct device arrays; positions and velocities as two columns each) are
 * permuted in place; entries past the kept total are left as they were.
 * counts[6] = (fluid, wavemaker, fixed ghost, inlet, outlet, total kept); perm (device,
 * optional) receives the source index of every kept slot. Runs on `stream` (a
 * hipStream_t; NULL = the null stream) and synchronises it before returning. */
int sph2d_reorder_particle_index(int32_t n_total, int32_t n_fluid, double box_x, double box_y,
                                 double* x, double* y, int32_t* type, int32_t n_f64, double* const* f64_cols,
                                 int32_t n_i32, int32_t* const* i32_cols, int32_t* perm, int32_t* counts,
                                 void* stream);

/* No counterpart in the reference: sph2d_reorder_particle_index for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors, but every
 * scratch buffer (category tables, scan tiles, counters, staging column, and perm
 * when NULL) lives in a file-scope device array of the library, one copy per GPU,
 * held until the process ends (~24.5 MiB): no hipMalloc, hipFree or other
 * allocation. At most 2,097,152 particles (n_total larger: SPH2D_ERR_CAPACITY) and
 * 64 columns of each kind. Calls that overlap on one GPU share that array and race;
 * the caller serialises them. */
int sph2d_reorder_particle_index_noalloc(int32_t n_total, int32_t n_fluid, double box_x, double box_y,
                                         double* x, double* y, int32_t* type, int32_t n_f64,
                                         double* const* f64_cols, int32_t n_i32, int32_t* const* i32_cols,
                                         int32_t* perm, int32_t* counts, void* stream);

/* --- one entry per replaced routine (context form) ------------------------- */
/* Distribute_Period_Boundary_Particle.f90:56-206: wrap into the box and append ghosts. */
int sph2d_distribute_period_boundary(sph2d_ctx* ctx);
/* Devide_block_and_send_tasks.f90 (one rank): binning into 4h cells, local order
 * (cell column, cell row, gid) and the global -> subdomain copy; the analytic TGV
 * fields at step start (:88-116) when i_time_step is a sampling step. */
int sph2d_devide_block(sph2d_ctx* ctx, int32_t i_time_step);
/* single_step_compute.f90:62-171: EOS, pair search, free surface, solver, rate assembly. */
int sph2d_single_step_compute(sph2d_ctx* ctx, int32_t i_time_step);
/* Reverse of the last sph2d_single_step_compute (no reference counterpart: the
 * adjoint of the evaluation above, for gradients through the step). Given weights
 * per fluid particle (gid order) of its outputs -- lam_drhodt (n), lam_acc (ld,2)
 * for dv/dt, lam_avg (ld,2) for the XSPH velocity or NULL -- writes the gradient of
 * sum(weights * outputs) with respect to every fluid particle's position (ld,2),
 * velocity (ld,2) and density (n), its periodic ghosts' shares added (any output
 * may be NULL). Held fixed: the neighbour set, the free-surface flags, the limiter
 * branches and the masses. FP64 form, modified or plain linear Riemann solver,
 * periodic box (no walls, no damping zone); the evaluation must be the context's
 * last one. The first call allocates the reverse's device buffers (16 doubles per
 * local and 10 per global particle slot), kept until sph2d_ctx_destroy. */
int sph2d_single_step_compute_vjp(sph2d_ctx* ctx, int32_t n, const double* lam_drhodt, const double* lam_acc,
                                  const double* lam_avg, int32_t ld, double* grad_pos, double* grad_vel,
                                  double* grad_rho);
/* The same reverse with device arrays (no reference counterpart: for callers whose
 * weights and gradients live in device memory, e.g. a training loop): every pointer
 * is device memory of the context's GPU, in the same (ld,2) layouts; the weights are
 * packed and the gradients unpacked on the device. Enqueued on the context's stream
 * (sph2d_get_stream); the gradients are complete when that stream reaches them. */
int sph2d_single_step_compute_vjp_device(sph2d_ctx* ctx, int32_t n, const double* lam_drhodt,
                                         const double* lam_acc, const double* lam_avg, int32_t ld,
                                         double* grad_pos, double* grad_vel, double* grad_rho);
/* --- the evaluation routine by routine (single_step_compute.f90:62-171) --------
 * sph2d_find_particle_pair, sph2d_define_freesurface and sph2d_solver, called in
 * that order after sph2d_devide_block (or a stage update), give the same results,
 * bit for bit, as sph2d_single_step_compute. */
/* find_particle_pair.f90:21-163, with the EOS and volumes of single_step_compute.f90:
 * 62-80 before it: the pair set of the current local state (1e-7 <= r <= 3 hbar, one
 * of the two fluid), recorded per particle. The same neighbour walk forms the sums
 * the later routines read: fra_v and sum dW of Define_freesurface_and_Normailized_
 * derivates.f90:79-159, the SPS strain -> tau (Solver...SPS.f90:144-296) and grad rho
 * (:508-526). */
int sph2d_find_particle_pair(sph2d_ctx* ctx, int32_t i_time_step);
/* The pair list of the last sph2d_find_particle_pair as the reference holds it
 * (find_particle_pair.f90:64-123): pair_i ascending, then the candidates j > i in the
 * stencil order self, left, right, upper left, up, upper right, lower left, down,
 * lower right (Distribute_Backgroud_Grid.f90:102-112), each cell's particles in
 * local order; 1-based LOCAL indices (sph2d_download_local "name" maps them to
 * gids). capacity 0 (or NULL arrays) only counts. */
int sph2d_download_pair_list(sph2d_ctx* ctx, int64_t capacity, int32_t* pair_i, int32_t* pair_j,
                             int64_t* pair_number);
/* Define_freesurface_and_Normailized_derivates.f90:21-444 after sph2d_find_particle_pair:
 * surface flags (fra_v <= 0.65), unit normals with the geometric un-flag tests,
 * near-surface flags; the renormalisation inverse is the identity, as in the reference. */
int sph2d_define_freesurface(sph2d_ctx* ctx);
/* Solver_Normalized_<variant>_Riemann_Delta[_SPS].f90 and the rate assembly of
 * single_step_compute.f90:152-260 (gravity, end damping) after sph2d_define_freesurface.
 * variant: SPH2D_SOLVER_*; MUSCL needs a context created with the MUSCL solver (its
 * gradients are formed in the pair walk). sps_on 0 gives the Delta forms without SPS
 * (see sph2d_params.sps_on; it clears this evaluation's tau); 1 needs a context
 * created with sps_on 1. */
int sph2d_solver(sph2d_ctx* ctx, int32_t variant, int32_t sps_on);
/* Stage updates of time_iteration.f90 on the fluid particles, with the rates of the
 * last evaluation by local index (the reference's re-use). scheme SPH2D_INTEGRATOR_PC:
 * stage 1 = predictor (:553-576: before_* = state, Euler step, temp_* = the rates),
 * stage 2 = corrector (:590-599). scheme SPH2D_INTEGRATOR_RK: stages 1, 2, 3 = the
 * K_1..K_3 updates (:825-836, :859-870, :889-900; stage 1 also sets before_*),
 * stage 4 = the final combination (:918-928). stage 0 (either scheme) = before_* =
 * state only. */
int sph2d_stage_update(sph2d_ctx* ctx, int32_t scheme, int32_t stage);
/* drhodt(ld) and particle_acceleration(ld,2) of the last evaluation for the n fluid
 * particles in global (reference) order (gid = reference index - 1). Either may be NULL. */
int sph2d_download_rates(sph2d_ctx* ctx, int32_t n, double* drhodt, double* acc, int32_t ld);

/* Reinitialize_particle_density.f90:66-182 (MLS order 6 or 1). */
int sph2d_reinitialize_particle_density(sph2d_ctx* ctx, int32_t i_time_step);
/* Shift_particle_position.f90:56-136 with the last evaluation's pair data. */
int sph2d_shift_particle_position(sph2d_ctx* ctx, int32_t i_time_step);
/* Refresh_domain.f90:136-300 (one rank): subdomain -> global arrays. */
int sph2d_refresh_domain(sph2d_ctx* ctx, int32_t i_time_step);
/* One full step of time_iteration.f90 (PC :509-657 or RK-variant :779-1049),
 * including ghosts, ordering, reinitialisation, shift and refresh. */
int sph2d_time_step(sph2d_ctx* ctx, int32_t i_time_step);
/* Steps first..first+nsteps-1 with no host synchronisation. */
int sph2d_run(sph2d_ctx* ctx, int32_t first_step, int32_t nsteps);
/* Sampling_subroutine.f90:340-461: out[0]=KE/KE0, [1]=max|v|/U, then (Linf,L1,L2)
 * for u, v, p/(rho0 U^2), |v| at [2..4], [5..7], [8..10], [11..13]. */
int sph2d_sample(sph2d_ctx* ctx, double* out14);

/* --- output and restart (the reference's on-disk formats) ------------------ */
/* Normalisation of the output routines (information_module.f90:56,150-152):
 * domain_size_y, velocity_0 = U_Module, pressure_0 = rho0*U^2, time_0 = Lx/U. */
typedef struct {
  double domain_size_y, velocity_0, pressure_0, time_0;
} sph2d_scales;
int sph2d_get_scales(sph2d_ctx* ctx, sph2d_scales* out);

/* Vorticity (Refresh_domain.f90:320-351: sum over the last evaluation's pairs of
 * V_j (v_i - v_j) x Normalized_dwdx) and the step-start analytic fields
 * (Devide_block_and_send_tasks.f90:88-116: u, v, p, |v| as columns of the (ld,4)
 * array ana), in global order, of the last sampling step (i_time_step a multiple of
 * sample_every: both are computed on those steps only); *step receives that step.
 * Any output may be NULL. */
int sph2d_download_output_fields(sph2d_ctx* ctx, int32_t n, double* vort_z, double* ana, int32_t ld,
                                 int32_t* step);

/* Output_for_tecplot.f90:64-291 for the current state: a ZONE appended to the
 * video file (video_result.dat, time_iteration.f90:147-151), one appended to the
 * free-surface file (Free_Surface.dat, :137-140) and the #!TDV101 binary file; a
 * NULL path skips that file. i_time_step must be the last sampling step. */
int sph2d_output_for_tecplot(sph2d_ctx* ctx, int32_t i_time_step, const char* video_path,
                             const char* surface_path, const char* binary_path);
/* Output_for_paraview.f90:188-263 for the current state (VTK XML, ASCII). */
int sph2d_output_for_paraview(sph2d_ctx* ctx, int32_t i_time_step, const char* path);

/* Post-processing grid (Distribute_Grid_For_Postproceeding.f90, information_module.f90:
 * 176-188): element_number = int(box / dx) per direction, element_dx = box_x /
 * element_number_x; node k = j + (i-1)*(element_number_x+1) (1-based) sits at
 * ((j-1) element_dx, (i-1) element_dx). */
int sph2d_grid_size(sph2d_ctx* ctx, int32_t* element_number_x, int32_t* element_number_y, double* element_dx);
/* Calculation_For_PostProceeding.f90:96-300 after the last step (which must be a
 * sampling step): order-6 MLS interpolation of velocity (ld,2), pressure and
 * vorticity onto the nodes, the level-set value Fai, and the analytic u, v, p, |v|
 * at the nodes as columns of ana (ld,4). Any output may be NULL. */
int sph2d_calculation_for_postproceeding(sph2d_ctx* ctx, int32_t i_time_step, int32_t n_nodes, double* vel,
                                         double* press, double* vort_z, double* fai, double* ana, int32_t ld);
/* Output_for_Grid_PostProceding.f90: a ZONE appended to grid_path
 * (Post_Proceeding_Grid.dat) and the #!TDV101 grid file binary_path (NULL skips). */
int sph2d_output_for_grid_postproceding(sph2d_ctx* ctx, int32_t i_time_step, const char* grid_path,
                                        const char* binary_path);

/* Sampling_subroutine.f90:21-177 (Sampling) at n points pos (ld,2) after the last
 * step (a multiple of sample_every): over the fluid particles within 3h of each
 * point, the average pressure, the order-6 MLS pressure (less g rho_j (y - y_j) Fai_j
 * when gravity != 0, Gravity_Switch 'on') and the MLS velocity vel (ld,2); zeros
 * for a point with 3 or fewer particles in its support. Any output may be NULL. */
int sph2d_sampling(sph2d_ctx* ctx, int32_t i_time_step, int32_t n, const double* pos, int32_t ld, double gravity,
                   double* average_press, double* mls_press, double* vel);
/* Sampling_subroutine.f90:246-279 after any step: the wave front (largest x of the
 * fluid particles, from 0) and the elevation at n_probes wave probes at x = probe_x
 * (largest y of the fluid particles within 0.55 dx, from 0), unscaled. */
int sph2d_wave_probes(sph2d_ctx* ctx, int32_t n_probes, const double* probe_x, double* wave_front,
                      double* elevation);
/* Sampling_subroutine.f90:466-512: for each of n_lines lines from first(2,k) to
 * last(2,k) (n_points points), sph2d_sampling at the points and the analytic field,
 * written to dir/<k>th_<t>_SamplingLine_Results.dat. */
int sph2d_output_sampling_lines(sph2d_ctx* ctx, int32_t i_time_step, const char* dir, int32_t n_lines,
                                const double* first, const double* last, int32_t n_points);

/* --- batched form (no reference counterpart) ------------------------------------- */
/* Independent contexts (uneven sizes, any Re) advanced together. When the contexts
 * share the predictor-corrector, precision form, solver and shifting switch (no
 * XSPH, no damping), every step after step 1 that is not a sampling step replays a
 * hipGraph of one launch per kernel covering every instance (one graph for the
 * repeating steps, one for MLS steps); otherwise the repeating steps replay a graph
 * with every instance's step as a parallel branch. Other steps are enqueued per context.
 * Each instance's results are bit-identical to running it alone. The contexts must
 * share integrator, mls_on, mls_every and sample_every, outlive the batch and not
 * be used by other calls while sph2d_batch_run runs; on return, later calls on a
 * context are ordered after the batch's work. */
typedef struct sph2d_batch sph2d_batch;
int sph2d_batch_create(int32_t n, sph2d_ctx* const* ctxs, sph2d_batch** out);
int sph2d_batch_run(sph2d_batch* batch, int32_t first_step, int32_t nsteps);
/* Nodes of the captured step graph (0 before the first replayable step) and replays so far. */
int sph2d_batch_info(sph2d_batch* batch, int32_t* graph_nodes, int32_t* graph_replays);
/* Configurations of the fused launches: config = order + 2 * block_order + 4 * shape,
 * with order 0 (instances as given) or 1 (by local capacity, largest first) in the
 * block tables, block_order 0 (XCD-aware) or 1 (linear) over the neighbour passes,
 * and shape 0..3 the tile of the FP32-mixed neighbour passes: particles per block
 * and LDS candidate-stage slots of passes A and C, {256, 1024, 1088},
 * {128, 640, 704}, {64, 384, 416}, {512, 1792, 1856}. FP64 batches use shape 0. */
#define SPH2D_BATCH_CONFIGS 16
/* Autotuner of the fused launches: times trial_steps real steps (after one untimed
 * step that captures the graphs) of each configuration (16 for FP32-mixed
 * instances, the 4 of shape 0 for FP64) and keeps the fastest. Advances every
 * instance by that many times (trial_steps + 1) steps from first_step; trial_ms
 * (SPH2D_BATCH_CONFIGS, may be NULL) receives ms per step (0 for configurations not
 * tried), chosen the configuration. Every configuration computes the same numbers. */
int sph2d_batch_autotune(sph2d_batch* batch, int32_t first_step, int32_t trial_steps, double* trial_ms,
                         int32_t* chosen);
/* Sets the configuration of the fused launches (0 at creation); the next replayable
 * step captures new graphs. */
int sph2d_batch_configure(sph2d_batch* batch, int32_t config);
int sph2d_batch_config(sph2d_batch* batch, int32_t* config);
int sph2d_batch_destroy(sph2d_batch* batch);

/* --- piston wavemaker (Wavemaker_Subroutine.f90, PistonType) ------------------- */
enum { SPH2D_WAVE_SOLITARY = 0, SPH2D_WAVE_STOKES2 = 1, SPH2D_WAVE_AIRY = 2 };
typedef struct {
  int32_t wave_type;         /* SPH2D_WAVE_* (wave_type, information_module.f90:276-280) */
  int32_t soft_steps;        /* wave_maker_soft_time (:306) */
  int32_t standing_steps;    /* standing_time_step */
  double dt, g;              /* dt (:309), g */
  double water_depth;        /* water_depth (:194) */
  double wave_height;        /* wave_height (:211) */
  double wave_number;        /* Wave_number (:202), Stokes_2nd and Airy */
  double wave_frequency;     /* Wave_Frequency (:201), rad/s */
  double phase;              /* Wave_initial_phase_angle (:208) */
} sph2d_wavemaker;
/* Wavemaker_Subroutine.f90:66-138 and wavemaker_plate_motion.f90 (host only): the
 * paddle position, velocity and acceleration at step i_time_step. Solitary: Goring's
 * paddle (fixed-point iteration to 1e-4 from x = 0); Stokes_2nd and Airy: the piston
 * laws with stroke H/m1, m1 = 2 sinh^2(kd)/(sinh(kd)cosh(kd) + kd), scaled by the
 * soft-start coefficient, which is 0 before step soft_steps + standing_steps and 1
 * from then on (:69-73 divide integers). */
int sph2d_wavemaker_motion(const sph2d_wavemaker* w, int32_t i_time_step, double* position, double* velocity,
                           double* acceleration);
/* Wavemaker_Subroutine.f90:196-215 on the device state: fluid particles behind the
 * paddle are mirrored in front of it; those within dx of it take its velocity. */
int sph2d_wavemaker_pull_back(sph2d_ctx* ctx, double position, double velocity);
/* Wavemaker_Subroutine.f90:140-215 on the device state, before a step: the water
 * height next to the paddle, the wavemaker particles placed behind the paddle at
 * its velocity with the hydrostatic density below that height, then the pull-back. */
int sph2d_wavemaker_update(sph2d_ctx* ctx, double position, double velocity);

/* --- wave-theory wavemaker and its damping (relaxation) zone ---------------------------
 * Parallel_Taylor_Green_Vortices_2D_Release.f90:113-143: the zone's analytic fields start
 * as the state, the initial positions of fluid and wavemaker particles are kept, fluid
 * particles with x <= zone_length (WavemakerDampingLengh) are in the zone with the soft
 * coefficient 1 - (exp((x/L)^3.5) - 1)/(exp(1.0 REAL(4)) - 1). zone_length <= 0 keeps
 * only the initial positions (the wave-theory wavemaker needs them) and no zone. */
int sph2d_relaxation_init(sph2d_ctx* ctx, double zone_length);
/* Wavemaker_Subroutine.f90:262-396 and :617-700 (WaveTheory, Stokes_2nd or Airy) before a
 * step: fluid below y = 0 typed 120; the wavemaker particles moved, at the wave's velocity
 * and pressure (rho from the inverse Tait law, mass rho0 V0 below the elevation, else 0);
 * the zone's analytic position, velocity, pressure and density. Uses w's wave_type
 * (SPH2D_WAVE_STOKES2 or SPH2D_WAVE_AIRY), soft_steps, standing_steps, dt, g, water_depth,
 * wave_height, wave_number, wave_frequency (period 2 pi / it) and phase. */
int sph2d_wave_theory_update(sph2d_ctx* ctx, const sph2d_wavemaker* w, int32_t i_time_step);
/* wave_maker_damping_zone.f90:61-95 on the subdomain state: velocity, position and pressure
 * of zone particles blended toward the analytic fields, rho = rho0 (1 + 7 p/(c0^2 rho0))^
 * 0.142857142. sph2d_time_step runs it after the corrector once a zone is set. */
int sph2d_wave_maker_damping_zone(sph2d_ctx* ctx, int32_t i_time_step);

/* --- differentiable wave tank ------------------------------------------------------
 * A discrete adjoint of the predictor-corrector step (FP64 form) of a walled tank with
 * piston wavemaker particles, as config 4's gradient of a probe-pressure misfit with
 * respect to paddle coefficients needs it. No counterpart in the reference: it
 * differentiates the steps of time_iteration.f90:541-657 (PC), Wavemaker_Subroutine.f90
 * :140-215 (paddle placement and pull-back), Assign_boundary_particles.f90:157-299
 * (walls), single_step_compute.f90:62-171 with its end damping (:174-260),
 * Shift_particle_position.f90 and the probe MLS of Sampling_subroutine.f90:21-177. */
typedef struct {
  int32_t struct_size;       /* sizeof(sph2d_adjoint_cfg), checked */
  int32_t n_harmonics;       /* 1..16: stroke x_p(t) = sum_k a_k sin(2 pi k f0 t), u_p = dx_p/dt */
  double f0;                 /* base frequency, Hz */
  int32_t standing_steps;    /* >= 1: the paddle rests for i_time_step <= standing_steps;
                                t = dt (i_time_step - standing_steps) (standing_time_step) */
  int32_t sample_every;      /* the probes are sampled after every sample_every-th step of the window */
  int32_t n_probes;          /* 1..8 */
  double probe_x[8], probe_y[8];
  int32_t checkpoint_every;  /* K: one snapshot every K steps; the reverse recomputes each segment */
} sph2d_adjoint_cfg;
/* Runs steps first..first+n_steps-1 (first >= 2: step 1's extra evaluation stays outside
 * the window), each after sph2d_wavemaker_update with the stroke of coefficients a, keeping a
 * snapshot every K steps; probe_p (host, may be NULL) receives the MLS pressure (with the
 * hydrostatic correction when gravity is on) of each probe after every sample_every-th
 * step: n_steps / sample_every rows of n_probes. Needs the FP64 form, the predictor-
 * corrector, walls, wavemaker particles, no periodic box, MLS reinitialisation off,
 * epsilon_xsph = 0, an _SPS linear Riemann solver and the reference's rate re-use. */
int sph2d_forward(sph2d_ctx* ctx, const sph2d_adjoint_cfg* cfg, const double* a, int32_t first_step,
                  int32_t n_steps, double* probe_p);
/* The reverse of the last sph2d_forward: dl_dprobe (host, the shape of probe_p) are the
 * weights of the probe pressures; dl_da (host, n_harmonics) receives sum of
 * dl_dprobe * d(probe_p)/d(a). The context's state is the forward's final state again
 * afterwards. What the reverse holds fixed: neighbour sets, free-surface and
 * near-surface flags, limiter, clamp, damping-zone and pull-back branches, the water
 * height's arg-max particle, which walls and wavemaker particles carry mass, and the SPS
 * term's particle_mass read by local index. */
int sph2d_backward(sph2d_ctx* ctx, const double* dl_dprobe, double* dl_da);

/* --- device memory through this library's HIP runtime ------------------------------
 * For callers that pass device arrays (sph2d_reorder_particle_index and its no-allocation
 * form) but hold no HIP runtime binding of their own (Fortran, ctypes): hipMalloc,
 * hipFree and a synchronous hipMemcpy (to_device 1: host to device, 0: device to host)
 * of the runtime this library is linked against. No reference counterpart. */
int sph2d_device_alloc(size_t bytes, void** out);
int sph2d_device_free(void* p);
int sph2d_device_copy(void* dst, const void* src, size_t bytes, int32_t to_device);

/* Host-array writers (no device; Fortran layout, ld >= n). Text follows the
 * reference's edit descriptors exactly; files named "appended" get the
 * reference's header lines when they are new or empty. */
/* Fw.d ('F') or Ew.dE3 ('E') of one value into out (cap > w). */
int sph2d_format_fortran(double value, char kind, int32_t w, int32_t d, char* out, int32_t cap);
/* One REAL(8) list-directed item as ifort writes it (23 characters; cap > 23). */
int sph2d_format_list_directed(double value, char* out, int32_t cap);
/* write(unit,*) of n REAL(8) values appended to path (Wave_front_results.dat and
 * the wave gauge files, Sampling_subroutine.f90:283,287). */
int sph2d_write_list_directed(const char* path, int32_t n, const double* values);
/* write(unit,'(20F20.8)') of n values appended to path (the pressure probe files,
 * Sampling_subroutine.f90:318-322). */
int sph2d_write_f20_8(const char* path, int32_t n, const double* values);
/* Sampling_subroutine.f90:300-322: Kalman filter of n measurements; last_x, last_p
 * carry the state between samplings (set on current_time_number == 1). */
int sph2d_kalman_filter(int32_t n, int32_t current_time_number, const double* measured, double* last_x,
                        double* last_p, double* filtered);
/* Initialize_SPH_Calculation.f90:385-398: n_points points from first(2) to last(2),
 * columns x, y of out (ld,2). */
int sph2d_sampling_line_points(int32_t n_points, const double* first, const double* last, double* out, int32_t ld);
/* Initialize_SPH_Calculation.f90:405-436: dir/<line>th_SamplingLine_Position.dat. */
int sph2d_write_sampling_line_positions(const char* dir, int32_t line, int32_t n_points, const double* pos,
                                        int32_t ld);
/* Sampling_subroutine.f90:490-512: dir/<line>th_<t_nondim, F7.4>_SamplingLine_Results.dat
 * (20f12.6: x y, analytic u v p, MLS u v p, average p, normalised by sc). */
int sph2d_write_sampling_line_results(const char* dir, int32_t line, double t_nondim, int32_t n_points,
                                      const double* pos, const double* ana_vel, const double* ana_press,
                                      const double* vel, const double* mls_press, const double* average_press,
                                      int32_t ld, const sph2d_scales* sc);
/* Parallel_Taylor_Green_Vortices_2D_Release.f90:185-199 (Final_Doamin_Results.dat). */
int sph2d_write_final_results(const char* path, int32_t n, const double* pos, const double* vel,
                              const double* rho, const double* press, int32_t ld);
/* Output_for_tecplot.f90:70-137 (appended ZONE; ana as in sph2d_download_output_fields). */
int sph2d_write_tecplot_zone(const char* path, double current_time, int32_t n, const double* pos,
                             const double* vel, const double* press, const double* vort_z, const double* ana,
                             int32_t ld, const sph2d_scales* scales);
/* Output_for_tecplot.f90:141-158 (appended ZONE of the surface particles given). */
int sph2d_write_surface_zone(const char* path, double current_time, int32_t n, const double* pos,
                             const double* normal, int32_t ld, const sph2d_scales* scales);
/* Output_for_tecplot.f90:162-291 (#!TDV101, point-packed doubles X Y VX VY P U). */
int sph2d_write_tecplot_binary(const char* path, double time, int32_t n, const double* pos, const double* vel,
                               const double* press, int32_t ld, const sph2d_scales* scales);
/* Output_for_paraview.f90:188-263. */
int sph2d_write_paraview(const char* path, int32_t n, const double* pos, const double* vel, const double* press,
                         int32_t ld, const sph2d_scales* scales);
/* Restart record: time_iteration.f90:143-144,495-498 (appended ZONE, 9F20.8,3I12);
 * wavemaker_layer may be NULL (zeros). */
int sph2d_write_latest_results_zone(const char* path, int32_t i_time_step, int32_t n, const double* pos,
                                    const double* vel, const double* press, const double* rho,
                                    const double* mass, const double* energy, const double* c,
                                    const int32_t* initial_type, const int32_t* wavemaker_layer, int32_t ld);
/* Distribute_Latest_Particle.f90:60-215: reads zone `zone` (1-based; 0 = the last),
 * stepping back while the zone holds a NaN. Outputs may be NULL; *n_out receives the
 * particle count and *step_out the step the zone was saved at. */
int sph2d_read_latest_results(const char* path, int32_t zone, int32_t capacity, int32_t* n_out, double* pos,
                              double* vel, double* press, double* rho, double* mass, double* energy, double* c,
                              int32_t* initial_type, int32_t* wavemaker_layer, int32_t* step_out, int32_t ld);
/* Distribute_Grid_For_Postproceeding.f90:60-95 (either path may be NULL). */
int sph2d_write_initial_grid(const char* grid_path, const char* node_path, int32_t element_number_x,
                             int32_t element_number_y, double element_dx);
/* Output_for_Grid_PostProceding.f90:110-138 (appended ZONE + connectivity). */
int sph2d_write_grid_zone(const char* path, double current_time, int32_t element_number_x,
                          int32_t element_number_y, double element_dx, const double* vel, const double* press,
                          const double* vort_z, const double* fai, const double* ana, int32_t ld,
                          const sph2d_scales* scales);
/* Output_for_Grid_PostProceding.f90:140-291 (#!TDV101, X Y VX VY P U Fai). */
int sph2d_write_grid_binary(const char* path, double time, int32_t element_number_x, int32_t element_number_y,
                            double element_dx, const double* vel, const double* press, const double* fai,
                            int32_t ld, const sph2d_scales* scales);
/* Sampling_subroutine.f90:360-461 (format 20F20.8): appends one line to each of the
 * six history files in dir from sph2d_sample's 14 values. */
int sph2d_write_sampling(const char* dir, double nondim_time, const double* out14);

/* --- parity / inspection -------------------------------------------------- */
/* Download a named subdomain (local-order) array of the last evaluation (in the
 * FP32-mixed form after sph2d_time_step/sph2d_run, u v rho mass p c volume are
 * current on sampling and MLS steps only: the other steps' evaluations read the
 * FP32 records instead):
 * f64: x y u v rho mass p c volume fra_v sum_dwdx_x sum_dwdx_y tau_xx tau_xy tau_yy
 *      grad_rho_x grad_rho_y drhodt ax ay shift_x shift_y avg_x avg_y (XSPH velocity,
 *      when epsilon_xsph is not 0) normal_x normal_y
 *      internal_x internal_y visc_x visc_y art_internal_x art_internal_y sps_x sps_y
 *      lap_x lap_y (the last 10 only after sph2d_set_debug(ctx,1))
 *      grad_p_x grad_p_y grad_u_x grad_u_y grad_v_x grad_v_y (SPH2D_SOLVER_MUSCL only)
 * i32: type name(gid, 0-based) surface near_surface
 *      neighbour_count (particles within 3h, from the last evaluation's recorded
 *      acceptance masks; lists of at most 256 candidates) */
int sph2d_download_local(sph2d_ctx* ctx, const char* name, void* out, int32_t n);
/* Pair set of the current state (find_particle_pair.f90:64-123): unordered pairs
 * as 0-based gids (gid_lo < gid_hi), sorted ascending by (lo,hi). */
int sph2d_download_pairs(sph2d_ctx* ctx, int64_t capacity, int32_t* gid_lo, int32_t* gid_hi,
                         int64_t* n_pairs);
/* 1: also write the acceleration components of the solver (for parity tests). */
int sph2d_set_debug(sph2d_ctx* ctx, int32_t on);
int sph2d_get_diagnostics(sph2d_ctx* ctx, sph2d_diag* out);
/* In-step pass timing (measurement hook, no reference counterpart): while on,
 * every force evaluation records HIP events around pass A and the force pass;
 * sph2d_get_pass_timing synchronises and returns the average durations (ms) over
 * the evaluations since the last call: out3 = (force pass, pass A, pass A start to
 * force pass end), and their number. */
int sph2d_set_pass_timing(sph2d_ctx* ctx, int32_t on);
int sph2d_get_pass_timing(sph2d_ctx* ctx, double* out3, int32_t* n_evaluations);
/* Average duration (ms, HIP events on the context's stream) of `reps` back-to-back
 * launches of one part of the last evaluation: pass 0 = force pass (Riemann,
 * delta-SPH, SPS, shift sums), 1 = neighbour-sum pass A, 2 = the whole evaluation.
 * No reference counterpart: measurement hook for the roofline report. */
int sph2d_time_pass(sph2d_ctx* ctx, int32_t pass, int32_t reps, double* avg_ms);
/* Device stream of the context (hipStream_t) for callers that time or order work. */
int sph2d_get_stream(sph2d_ctx* ctx, void** stream);

#ifdef __cplusplus
}
#endif
#endif
