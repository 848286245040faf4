/*
 * sph2d_oracle.h -- CPU restatement of the reference WCSPH step (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it. It restates, in
 * plain C and in the reference's own order of operations, the Fortran path
 *   Taylor_Green_Vortex/time_iteration.f90 -> single_step_compute.f90 -> ...
 * with every reference quirk kept (identity renormalisation, SPS mass index,
 * stale-index rate re-use, frozen ghosts, REAL(4) literals).
 * Arrays are 0-based here; "gid" is the reference's 1-based global index - 1.
 */
#ifndef SPH2D_ORACLE_H
#define SPH2D_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime form of the compile-time parameters of
 * Taylor_Green_Vortex/information_module.f90 (defaults in orc_default_params). */
typedef struct {
  int32_t nx, ny;            /* interior_particle_number_x/_y              (:75-76)  */
  double box_x, box_y;       /* domain_size_x/_y == boundary_size_x/_y     (:54-57)  */
  double rho0;               /* water_rho_0                                (:31)     */
  double c0sq;               /* square_c_0                                 (:145)    */
  double nu;                 /* water_kinematic_viscosity (REAL(4) 1.0E-2) (:37)     */
  double U;                  /* U_Module                                   (:52)     */
  double h_factor;           /* smooth_length_factor                       (:134)    */
  double dt;                 /* dt                                         (:309)    */
  int32_t integrator;        /* leapforg_or_eular: 2 = PC, 3 = RK-variant  (:303)    */
  int32_t shift_on;          /* ShiftParticle_Switch                       (:349)    */
  int32_t mls_on;            /* RefreshDensityByMLS_Switch                 (:324)    */
  int32_t mls_every;         /* MLS_time_step                              (:332)    */
  int32_t mls_order;         /* RefreshDensity_MLS_Order (1 or 6)          (:330)    */
  int32_t solver;            /* 0 = Modified linear Riemann (default), see ORC_SOLVER_* */
  double alpha, beta, fai;   /* Monaghan artificial viscosity              (:210-212) */
  double delta, delta_alpha; /* delta-SPH                                  (:217-218) */
  double epsilon_xsph;       /* XSPH epsilon                               (:207)    */
  double shift_beta, shift_R;/* shift_belta, shift_R                       (:350-351) */
  double mls_alpha;          /* slighted_modified_alpha (REAL(4) 0.0001)   (:359)    */
  int32_t rate_reuse;        /* 0 = reference (stale local index), 1 = carried by gid  */
  /* end damping zone (single_step_compute.f90:174-260; EndDampingZone_Switch 'on' and */
  /* make_wave_or_not 1): 0 off, 1 Cosin_Type, 2 OpenFoam_Type, 3 DualSPH_Type          */
  int32_t damping;
  double damping_length;     /* EndZoneDampingLengh                        (:291)    */
  double damping_amplitude;  /* EndZoneDampingAmplitude (Cosin_Type)       (:294)    */
  double water_depth;        /* water_depth                                (:194)    */
  /* fixed ghost walls around [0,box_x]x[0,box_y] instead of the periodic ghosts:   */
  /* fix_ghost_layer layers (:86-93 DistributeFixedBoundaryParticleOrNot = 1 and     */
  /* DistributePeriodBoundaryParticleOrNot = 0); 0 = periodic box                   */
  int32_t walls;
  double gravity;            /* g (9.81) when Gravity_Switch is 'on', else 0  (:32-33) */
  /* piston wavemaker particles (fix_wave_maker_or_not, wave_maker_layer :213-214,  */
  /* PistonType): columns of type-130 particles left of x = 0, before the walls    */
  int32_t wavemaker_layers;
  int32_t sps_on;            /* 1: _SPS solvers; 0: Delta solvers without SPS (XSPH factor 1) */
  int32_t renorm;            /* 0: identity for the dyadic inverse (the reference); 1: the inverse */
  /* ---- the moving-square program (Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/, */
  /* "MS"): ms = 1 selects its semantics wherever they differ from the TGV program --    */
  /* linear EOS (Status_Equation.f90:56-90, EquationOfState = 1), pair bound 1.0E-10     */
  /* (Find_particle_pair.f90:73-75), 4 dx cells (Information_Module.f90:189), no free    */
  /* surface (Has_FreeSurface_Or_Not = 0), the fixed-ghost MLS with mass and the         */
  /* characteristic correction (Assign_boundary_particles.f90:611-760), XSPH velocity    */
  /* in the continuity equation of every MS solver, the MS predictor-corrector           */
  /* (SPH_Time_Iteration.f90:313-387), the MS density reinitialisation and the moving    */
  /* body with its force. Layout by orc_init_moving_square. */
  int32_t ms;
  double background_p;       /* Background_Pressure (MS Information_Module.f90:487) */
  double release_time;       /* Shock_wave_release_time (:715): characteristic wall correction for 0 < t <= it */
  double body_cx, body_cy, body_len;   /* Block_Center_X/_Y, Block_Length (:152-154) */
  double body_speed, body_ramp;        /* prescribed motion u(t) = body_speed min(t / body_ramp, 1) in x */
  int32_t sample_every;      /* Sampling_time_step (:748): the body force is computed on these steps */
  double max_shift_dx;       /* max_shift / interior_dx: 0.25 (TGV information_module.f90:353), 0.15 (MS :679) */
} orc_params;

enum { ORC_SOLVER_MODIFIED_LINEAR_RIEMANN = 0,   /* Solver_Normalized_Modified_Linear_Riemann_Delta_SPS.f90 */
       ORC_SOLVER_LINEAR_RIEMANN = 1,            /* Solver_Normalized_Linear_Riemann_Delta_SPS.f90 */
       ORC_SOLVER_HLLC = 2,                      /* Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90 */
       ORC_SOLVER_MUSCL = 3,                     /* Solver_Normalized_MUSCL_Linear_Riemann_Delta_SPS.f90 */
       /* MS solvers (ms = 1; MS ORC_SOLVER_MODIFIED_LINEAR_RIEMANN with sps_on = 0 is
        * MS/Solver_Normalized_Modified_Linear_Riemann_Delta.f90, the MS default) */
       ORC_SOLVER_MS_MUSCL_HLLC = 4,             /* MS/Solver_Normalized_MUSCL_HLLC_Riemann_Delta.f90 (+ the SPS
                                                    block of MS/Solver_Normalized_HLLC_Riemann_Delta_SPS.f90) */
       ORC_SOLVER_MS_HLLC = 5 };                 /* MS/Solver_Normalized_HLLC_Riemann_Delta_SPS.f90 (sps_on = 0:
                                                    MS/Solver_Normalized_HLLC_Riemann_Delta.f90) */

typedef struct orc_ctx orc_ctx;

void orc_default_params(orc_params* p);
orc_ctx* orc_create(const orc_params* p);
void orc_destroy(orc_ctx* c);

/* Scalar restatements (reference argument order). */
void orc_compute_kernel(int32_t dim, double r, const double* dx, double h, double* w, double* dwdx);
void orc_status_equation(const orc_ctx* c, double rho, double e, int32_t type, double* p, double* cs);
int32_t orc_brinv(double* a_colmajor, int32_t n);
int32_t orc_matrix_inversion_less_three(int32_t n, const double* a_colmajor, double* res_colmajor);
/* function_module.f90:34-80 (MUSCL_Limiter). */
void orc_muscl_limiter(double delphi_i, double delphi_j, double d_phi_ij,
                       double* del_phi_lim_i, double* del_phi_lim_j);
/* HLLC star state of one pair in the reference's pair order (i = left, j = right),
 * e = unit vector from i to j, x_over_t = interior_dx/dt
 * (Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90:362-490). */
void orc_hllc_star(const double* v_left, double p_left, double rho_left, double c_left,
                   const double* v_right, double p_right, double rho_right, double c_right,
                   const double* e, double x_over_t, double* v_star, double* p_star);

/* Reorder_Particle_Index, Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/
 * Reorder_Particle_Index.f90:54-287, on host arrays of n_total entries: ill marking of
 * fluid outside the box among the first n_fluid, the per-type index lists in index
 * order, the copy back as fluid, wavemaker, fixed ghost, inlet, outlet. Columns are
 * permuted in place (entries past the kept total unchanged); perm[k] = source index;
 * counts[6] = (fluid, wavemaker, fixed ghost, inlet, outlet, total kept). */
void orc_reorder_particle_index(int32_t n_total, int32_t n_fluid, double box_x, double box_y,
                                double* x, double* y, int32_t* type, int32_t n_f64, double** f64_cols,
                                int32_t n_i32, int32_t** i32_cols, int32_t* perm, int32_t* counts);

/* Initial lattice, Distribute_Initial_Particle.f90:204-249. */
void orc_init_tgv(orc_ctx* c);
/* Overwrite the fluid state (global order, n fluid particles). */
int32_t orc_set_fluid_state(orc_ctx* c, int32_t n, const double* x, const double* y,
                            const double* u, const double* v, const double* rho,
                            const double* mass, const int32_t* type);

/* Pieces of one time step. */
int32_t orc_wall_positions(const orc_ctx* c, double* x, double* y);   /* fixed ghost positions (NULL: count) */
void orc_period_boundary(orc_ctx* c);                       /* Distribute_Period_Boundary_Particle */
void orc_devide_block(orc_ctx* c, int32_t i_time_step);     /* Devide_block_and_send_tasks (1 rank) */
void orc_single_step_compute(orc_ctx* c, int32_t i_time_step);
void orc_time_step(orc_ctx* c, int32_t i_time_step);        /* one full step of time_iteration */
void orc_reinitialize_density(orc_ctx* c);                  /* Reinitialize_particle_density */
void orc_shift(orc_ctx* c);                                 /* Shift_particle_position */
void orc_shift_sums(orc_ctx* c);                            /* its pair sums only (:78-103) */
void orc_refresh(orc_ctx* c);                               /* Refresh_domain (1 rank) */
/* Calculation_For_PostProceeding on the (nex+1)x(ney+1) element nodes, spacing edx;
 * out = 9 blocks of n_nodes: u, v, p, vorticity_z, Fai, analytic u, v, p, |v|. */
void orc_grid_postprocess(orc_ctx* c, int32_t i_time_step, int32_t nex, int32_t ney, double edx, double* out);
void orc_sampling_points(orc_ctx* c, int32_t n, const double* px, const double* py, double gravity, double* out);
void orc_wave_probes(orc_ctx* c, int32_t n_probes, const double* probe_x, double* out);
void orc_wavemaker_pull_back(orc_ctx* c, double xp, double up);
void orc_wavemaker_update(orc_ctx* c, double xp, double up);
/* Wave-theory wavemaker (wave_type 1 Stokes_2nd, 2 Airy) and its damping zone */
void orc_relaxation_init(orc_ctx* c, double zone_length);
void orc_wave_theory_update(orc_ctx* c, int32_t wave_type, int32_t soft_steps, int32_t standing_steps, double dt,
                            double g, double d, double H, double k, double om, double phase, int32_t it);
void orc_relaxation_blend(orc_ctx* c);

/* The moving-square program (ms = 1). Layout of MS/Initial_Particle_From_First_Step.f90
 * :60-620 with MS/SPH_Subroutine_Module.f90:871-925: fluid on the nx x ny lattice of
 * [0,box_x]x[0,box_y] (rows in y) without the points of the square [body_cx -/+ L/2] x
 * [body_cy -/+ L/2], rho = rho0, p = Background_Pressure, v = 0, m = rho0 dx dy; then
 * `walls` layers of fixed ghosts (type -3) around the tank (bottom, right, top, left) and
 * `walls` layers inside the square's edges (type -3, initial type -4 = body). */
void orc_init_moving_square(orc_ctx* c);
/* one MS time step (SPH_Time_Iteration.f90:58-540 with Iteration_Method = 2) */
void orc_ms_time_step(orc_ctx* c, int32_t i_time_step);
/* Body_Motion_Solver.f90:62-273 with the prescribed ramp in place of the spline file:
 * t = present time; moves the body particles (global arrays) and sets the body velocity
 * and acceleration used by the wall assignment. */
void orc_ms_body_motion(orc_ctx* c, double t);
/* MS single_step_compute (Single_step_compute.f90:35-283) at step i_time_step, time t. */
void orc_ms_single_step_compute(orc_ctx* c, int32_t i_time_step, double t);
/* Body force of the last MS evaluation that computed it (Calculate_Body_Force.f90:51-142,
 * Single_step_compute.f90:219-249): out[6] = pressure x, y, viscous x, y, total x, y
 * (force on the water; Body_Force.dat divides by force_0 = 0.5 rho0 U^2 L). */
void orc_ms_body_force(const orc_ctx* c, double* out);
/* initial types of the global particles (particle_initial_type: 2, -3, -4 body) */
const int32_t* orc_ms_initial_type(const orc_ctx* c);

/* Sampling_subroutine.f90:340-461. out[0]=KE, [1]=max|v|, [2..4]=u Linf,L1,L2,
 * [5..7]=v, [8..10]=p/(rho0 U^2), [11..13]=|v|. */
void orc_sample(const orc_ctx* c, double* out);
double orc_initial_ke(const orc_ctx* c);

/* Sizes. */
int32_t orc_n_fluid(const orc_ctx* c);
int32_t orc_n_global(const orc_ctx* c);      /* fluid + periodic ghosts */
int32_t orc_n_local(const orc_ctx* c);       /* subdomain particles */
int32_t orc_n_singular(const orc_ctx* c);    /* singular dyadic tensors of fluid particles, last evaluation */
int64_t orc_pair_number(const orc_ctx* c);

/* Named array access: returns a pointer into the context (valid until the next call
 * that reallocates), and the element count in *n. Names are the reference's. */
const double* orc_get_f64(const orc_ctx* c, const char* name, int64_t* n);
const int32_t* orc_get_i32(const orc_ctx* c, const char* name, int64_t* n);
double* orc_mut_f64(orc_ctx* c, const char* name, int64_t* n);
int32_t* orc_mut_i32(orc_ctx* c, const char* name, int64_t* n);

#ifdef __cplusplus
}
#endif
#endif
