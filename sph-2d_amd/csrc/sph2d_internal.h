// sph2d_internal.h -- device-side layout of one context and the kernel launchers.
//
// Layout in HBM (one context = one simulation):
//   global order (gid = reference global index - 1): fluid 0..N-1, then the
//   periodic ghosts regenerated every step (Distribute_Period_Boundary_Particle)
//   local order (the reference's subdomain order on one rank): 4h cells taken
//   column by column (x), row by row (y), particles of a cell by ascending gid.
// Positions and velocities are double2 (x,y) so a neighbour costs one 16-byte
// load; (rho, volume) and (p, c) are packed the same way.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include "../../include/sph2d.h"

namespace sph2d {

constexpr int kBlock = 256;
constexpr int kNbListCap = 96;                  // recorded neighbours per particle (24 words)
constexpr int kNbListWords = kNbListCap / 4;
constexpr int kNbOverflow = 1 << 30;            // nb_count flag: the list is not complete

// Constants of one run, passed by value to every kernel.
struct DevParams {
  double box_x, box_y;
  double rho0, c0sq, sqrt_c0sq, eos_b;   // eos_b = c0sq*rho0/7.0 (status_equation.f90:61)
  double nu, U;
  double dx, h, dt;
  double three_h;                        // kernel_scale*smooth_length (3*h)
  double kfactor;                        // compute_kernel.f90:123 factor for h
  double e9;                             // REAL(4) exp(-9.0) widened
  double r_min;                          // REAL(4) 1.0E-7 widened (find_particle_pair.f90:87)
  double support_volume;                 // PI*(3h)**2 (Define_freesurface...f90:65)
  double fra_limit;                      // REAL(4) 0.65 widened
  double chain_d, chain_ox, chain_oy;    // 4h cells, origin -5 cells
  int chain_nx, chain_ny;
  double alpha, beta, fai, delta, delta_alpha, epsilon_xsph;
  int walls;                             // fixed ghost layers (0: periodic box)
  int wm_layers;                         // wavemaker particle columns (0: none)
  double g;                              // gravity (0: off)
  double wall_h, wall_kfactor, vol0;     // hbar = 1.2 (h + h) / 2 and its kernel factor; dx dy
  int damping;                           // SPH2D_DAMPING_*
  double damp_len, damp_amp, water_depth, e32;   // e32 = exp(1.0) in REAL(4)
  double shift_beta, shift_R, max_shift, w_dx, h2;
  double sps_dl, sps_c012, sps_c23, sps_c23b;  // REAL(4) constants of the SPS closure
  double mls_alpha;
  double surf_lim;                       // REAL(4) 1.414 * h
  double inv_h, inv_c0, inv_sv, inv_wdx, cut2;  // reciprocals and the squared-distance pre-test
  float cut2_32;                         // FP32 screen: above this r^2 a pair is certainly outside 3h
  double r2_max, r2_min;                 // fl(sqrt(r2)) <= 3h  <=>  r2 <= r2_max; >= 1e-7 <=> r2 >= r2_min
  double x_over_t;                       // interior_dx/dt, the HLLC fan sample point
  int solver, write_debug;
  int mls_order;
  int mixed;                             // 1: FP32-mixed evaluation (mirrors are written)
  int sps;                               // 1: SPS stress (the _SPS solvers); 0: tau = 0
  int renorm;                            // 1: the dyadic tensor's 2x2 inverse normalises grad W
  double xsph_f;                         // XSPH factor: 2 in the _SPS solvers, 1 without SPS
  double rebin_m;                        // how far a particle may leave its binned cell (3x3 stencil covers 3h)
  // the moving-square program (Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/): linear
  // EOS, no free surface, the fixed-ghost MLS with mass and the characteristic correction,
  // XSPH in the continuity equation, the moving body and its force
  int ms;
  double pb;                             // Background_Pressure
  double release_time;                   // Shock_wave_release_time
  double t_eval;                         // physical time of the evaluation in progress
  double cust_u, cust_a;                 // the body's velocity and acceleration (x)
  int body_force;                        // 1: this evaluation forms the body force
};

// Device counters (one int array, indices below).
enum Counter {
  C_NFLUID = 0, C_NGHOST, C_NLOCAL, C_NXG, C_NYF, C_NYX, C_NSURF, C_NILL, C_ERR,
  C_MAXCELL, C_REBIN, C_MLSFB, C_NSING, C_NPAIR_LO, C_NPAIR_HI, C_COUNT = 16
};
enum ErrFlag { E_GHOST_CAP = 1, E_LOCAL_CAP = 2, E_CELL = 4 };

struct Buffers {
  // global order
  double2 *g_pos, *g_vel;
  double *g_rho, *g_mass, *g_p, *g_c, *g_e;
  int *g_type, *g_key, *g_slot;
  double *ana;                            // 4 per fluid particle: u, v, p, |v|
  double *g_vort;                         // vorticity (z) per fluid gid, on sampling steps
  int *g_btype;                           // boundary type of each fixed ghost (-6, -7, -8), by gid
  int *gblk, *cnt_yx, *off_yx, *src_x;   // gblk: per block of gids, x-ghost then y-ghost counts
  // cells (column-major key over the global grid)
  int *cell_count, *cell_start, *cell_members;
  // local order
  int *name, *l_type, *l_cell;
  int *l_of_g;                            // local index of each gid (-1: not in the subdomain)
  double2 *l_pos, *l_vel, *l_rv, *l_pc;
  float2 *l_pos32;                        // FP32 copy of l_pos for the neighbour pre-test
  double *l_mass;
  // one force evaluation
  double *fra_v;
  double2 *sdw, *grad, *nrm, *shift, *avg;
  double *tau;                            // 3 per particle: xx, xy, yy
  double *tinv;                           // 4 per particle (renormalisation on): T^-1 column-major
  double2 *mg;                            // MUSCL only, 3 per particle: grad p, grad u, grad v
  int *surf, *near_;
  // accepted neighbours of the last pass A: the candidate slots (0..255) in ascending
  // order, one byte each, 4 per word, word w of particle s at nb_list[w * kstride + s];
  // nb_count[s] their number, with kNbOverflow set when they exceeded kNbListCap
  uint32_t* nb_list;
  int* nb_count;
  double *drhodt;
  double2 *acc;
  double *dbg;                            // 10 per particle (debug components)
  // integrator (local index, persist across steps)
  double *b_rho, *t_drho, *k_drho;        // k_drho: 3 stages
  double2 *b_pos, *b_vel, *t_acc, *k_vel, *k_acc;
  double *mls_rho;
  // scratch
  int *scan_blocks;
  int *counters;
  double *red;                            // reduction partials
  int *pair_count, *pair_off;
  int *pair_lo, *pair_hi;
  // FP32 neighbour records of the mixed form (local order), 48 bytes per particle:
  //   m_r0 = (u, v, rho - rho0, m), m negated for particles that are not fluid
  //   m_r1 = (tau_xx, tau_xy, tau_yy, particle_mass at the local index)
  //   m_r2 = (grad rho x, grad rho y, x - x_cell, y - y_cell): offsets from the
  //          corner of the particle's binned cell, so neighbour offsets keep
  //          FP32 precision relative to h
  float4 *m_r0, *m_r1, *m_r2;
  float2 *m_pc;                           // (p, c) from the FP64 Tait law of the particle's rho
  float2 *m_mg;                           // MUSCL only: FP32 copy of mg

  int kstride;                            // stride of the 3 RK stage slabs (= local capacity)
  // the moving-square program: initial type per gid (-4 body), the body-force parts per
  // local particle (pressure, viscous) and its near-body flag, per-block partial sums
  int *g_init;
  double2 *bp, *bv;
  int *near_body;
  double *bpart;
};

// Records the message sph2d_last_error returns and passes the code through.
int set_error(int code, const std::string& msg);

// Launchers (sph2d_fp64.hip). All run on `s` and take sizes from device counters.
void launch_init_tgv(hipStream_t s, const DevParams& P, Buffers& B, int nx, int ny);
// Reverse of the last FP64 evaluation (sph2d_adjoint.hip); lam: 5 per fluid gid,
// w: 6 and gs, ga: 5 per local particle, G: 5 per gid (result in the first n_fluid).
void launch_vjp_pack(hipStream_t st, int n, int ld, const double* ldr, const double* lacc, const double* lavg,
                     double* lam);
void launch_vjp_unpack(hipStream_t st, int n, int ld, const double* G, double* gpos, double* gvel, double* grho);
void launch_eval_vjp(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, int cap_fluid, int cap_ghost,
                     int cap_global, const double* lam, double* w, double* gs, double* ga, double* G);
// the slot form of the reverse evaluation (weights w: 10 per local particle) and the
// reverse of whole steps (sph2d_adjstep.hip)
void launch_eval_vjp_slots(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, int do_shift, double* w,
                           double* gs, double* ga, double* G, double* wsum);
void launch_adj_zero(hipStream_t s, double* p, size_t n);
void launch_adj_stage1(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, const double* lamG,
                       const double* lamR, double* W, double* scr, int shift_on);
void launch_adj_stage2(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, const double* G,
                       const double* scr, const double* lamR, double* lamG, double* lamRp);
void launch_adj_wm_init(hipStream_t st, unsigned long long* hbits, int* argmax);
void launch_adj_wm_flags(hipStream_t st, const DevParams& P, Buffers& B, int cap_fluid, double xp, int* flags,
                         const unsigned long long* hbits, int* argmax);
void launch_adj_wm_reverse(hipStream_t st, const DevParams& P, Buffers& B, int n_total, int n_wm, const double* G,
                           const int* flags, const unsigned long long* hbits, double* lamG, double* c3,
                           double* part, const int* argmax, double* out2);
constexpr int kAdjCopies = 64;                   // copies of one k_adj_copy launch
struct AdjCopy {
  const void* src[kAdjCopies];
  void* dst[kAdjCopies];
  size_t words[kAdjCopies];
  int n;
};
void launch_adj_copy(hipStream_t st, const AdjCopy& C);
void launch_adj_probe(hipStream_t st, const DevParams& P, Buffers& B, double ex, double ey, double gravity,
                      const double* lam, double* lamG);
void launch_wavemaker_height(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, double xp,
                             unsigned long long* height_bits);
void launch_period_boundary(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, int cap_ghost);
// predict 1: the local copy also applies the predictor-corrector's predictor (k_gather_predict);
// 2: and keeps the step-start state in before_* for k_stage's corrector; 3: as 1 without
// the FP64 local v, p, c, rho, volume, mass (FP32-mixed steps whose later kernels do not read them)
void launch_devide_block(hipStream_t s, const DevParams& P, Buffers& B, int cap_global, int n_cells,
                         bool analytic, double t_nondim, int predict = 0);
// ev (optional, 4 events): recorded before and after pass A and before and after the force pass.
void launch_eval(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, hipEvent_t* ev = nullptr);
// FP32-mixed passes A and C (sph2d_eval32.hip)
void launch_eval32_a(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_eval32_c(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
// One part of launch_eval: 0 = pass A (pair masks and neighbour sums), 1 = the free
// surface (flags, normals, near-surface flags), 2 = the force pass.
void launch_eval_part(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int part);
// The pair list in the reference's order (1-based local indices): per-particle counts
// into pair_count, scan into pair_off and the total into counters[C_NPAIR_LO], then
// the list (at most capacity entries) into pair_lo (i) and pair_hi (j).
void launch_count_pair_list(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_pair_list(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int64_t capacity);
// tau = 0 for every local particle (the Delta solvers without SPS)
void launch_clear_tau(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
// pass 0: force pass only, 1: pass A only, 2: the whole evaluation
void launch_eval_pass(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int pass);
void launch_stage(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int kind);
void launch_mls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
// Calculation_For_PostProceeding.f90 on the element nodes (sph2d_fp64.hip)
struct GridArgs {
  int nnx, nny, n_nodes, n_surf;
  double edx, t_nondim, Re;
  const int* surf_gid;                    // surface particles (gid order) and their normals
  const double2* surf_nrm;
  double* out;                            // 9 blocks of n_nodes
};
void launch_grid_mls(hipStream_t s, const DevParams& P, Buffers& B, const GridArgs& G);
void launch_end_damping(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_assign_walls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
// The moving-square program (sph2d_ms.hip). The fixed ghosts of the tank and the body
// (Assign_boundary_particles.f90:611-760); one MS evaluation after the ghosts are
// assigned (pass A: gradients, strain and tau, XSPH; force pass with the body-force
// parts); the body force totals into out6 (device, 6 doubles); the MS predictor and
// corrector (SPH_Time_Iteration.f90:313-347); the order-3 density MLS into mls_rho;
// the body's move by (ddx, ddy); fluid outside the tank typed 120.
void launch_ms_assign_walls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_ms_eval(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_ms_body_sum(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, double* out6);
void launch_ms_stage(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int kind);
void launch_ms_mls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_ms_body_move(hipStream_t s, const DevParams& P, Buffers& B, int n_total, double ddx, double ddy);
void launch_ms_ill(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid);
// host layout of the moving-square tank (x, y, btype, init may be null: count only)
int ms_ghost_layout(double box_x, double box_y, double gdx, double gdy, int layers, double cx, double cy,
                    double len, double* x, double* y, int* btype, int* init);
enum MsStage { MS_PREDICT = 0, MS_CORRECT = 1 };
void launch_wall_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
// Distribute_Initial_Particle.f90:317-495 and Distribute_Fixed_Boundary_Particle.f90:52-104:
// positions and boundary types of the fixed ghost layers (x, y, type may be null: count only)
int wall_layout(double box_x, double box_y, double dx, double dy, int layers, double* x, double* y, int* btype);

// ---- batched launches: one launch covers the blocks of many independent instances ----
struct Inst {
  DevParams P;
  Buffers B;
  int cap_fluid, cap_ghost, cap_global, cap_local, n_cells;
};
// Block-count classes of the step's launches (blocks of one instance in each).
enum BatchClass {
  BC_FLUID,      // grid_for(cap_fluid): wrap/count, ghost emit
  BC_TWO,        // 2 (of 1024 threads): the two ghost-count scans
  BC_ONE,        // 1: scan tops
  BC_FILL,       // grid_for(max(cap_ghost, n_cells)): the step's zeroing
  BC_SCAN_G,     // scan tiles over cap_ghost
  BC_GHOST,      // grid_for(cap_ghost): corner ghosts
  BC_GLOBAL,     // grid_for(cap_global): keys, scatter, local copy
  BC_SCAN_C,     // scan tiles over n_cells
  BC_RANK,       // cells / (kBlock / 64): rank in cell
  BC_LOCAL,      // grid_for(cap_local): evaluation passes, corrector
  BC_NBR,        // cap_local / kNbShape[shape].block: the FP32-mixed neighbour passes
  BC_COUNT
};
// Tile shapes of the batched FP32-mixed neighbour passes, chosen by the batch
// autotuner: particles per block (consecutive in local order, so a block covers a
// run of cells of one cell column) and the LDS stage capacities of pass A (24 B per
// candidate slot) and pass C (32 B). Shape 0 is the single-instance launch's.
struct NbShape {
  int block, cap_a, cap_c;
};
constexpr int kNbShapes = 4;
constexpr NbShape kNbShape[kNbShapes] = {{256, 1024, 1088}, {128, 640, 704}, {64, 384, 416}, {512, 1792, 1856}};
struct BatchArgs {
  const Inst* inst;
  const int* start;                       // n + 1 first blocks of the instances (this class)
  int n;
  int xcd;                                // neighbour passes: XCD-aware block order over the launch
};
struct BatchDev {
  const Inst* inst = nullptr;
  int n = 0;
  int xcd = 1;
  int shape = 0;                          // kNbShape entry of the mixed neighbour passes
  const int* start[BC_COUNT] = {};
  int total[BC_COUNT] = {};
  const void* q32 = nullptr;              // n FP32 parameter sets (mixed form)
};
int batch_blocks(int cls, const Inst& I, int shape);
size_t q32_bytes();
void q32_fill(const DevParams& P, void* out);
// One predictor-corrector step after step 1 that is not a sampling step, for every
// instance (same form, solver and shifting switch; no damping, no XSPH); mls: the
// step reinitialises the density (every instance has the same MLS schedule).
void launch_batch_step(hipStream_t s, const BatchDev& D, int mixed, int mls, int shift_on, int solver);
void launch_batch_eval(hipStream_t s, const BatchDev& D, int mixed, int solver);
void launch_batch_eval32(hipStream_t s, const BatchDev& D, int pass, int solver);
void launch_wavemaker_pull_back(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, double xp, double up);
// Wavemaker_Subroutine.f90:140-193: water height next to the paddle (bit pattern at
// height_bits, zeroed by the caller), then the n_wm wavemaker particles after the fluid
void launch_wavemaker_place(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, int n_wm, double xp,
                            double up, unsigned long long* height_bits);
// Sampling_subroutine.f90 at sampling points (sph2d_fp64.hip)
// The wavemaker damping (relaxation) zone and the wave-theory wavemaker: per-gid arrays
// (In_WaveMakerDampingZone_OrNot, position_soft_coefficient, WaveMaker_Analytical_*,
// Initial_particle_Position) and one step's wave quantities.
struct RelaxArgs {
  int* in_zone;
  double* coef;
  double2 *an_pos, *an_vel, *init_pos;
  double *an_p, *an_rho;
};
struct WaveArgs {
  int wave_type;                          // 1 Stokes_2nd, 2 Airy
  double soft, t, g, d, H, k, om, phase, T, L, A, skd, ckd;
};
void launch_wave_theory(hipStream_t s, const DevParams& P, Buffers& B, const RelaxArgs& R, const WaveArgs& W,
                        int cap_fluid, int n_wm, int relax_on);
void launch_relax_blend(hipStream_t s, const DevParams& P, Buffers& B, const RelaxArgs& R, int cap_local);
void launch_relax_init(hipStream_t s, const DevParams& P, Buffers& B, const RelaxArgs& R, int n_total,
                       double zone_length);

struct ProbeArgs {
  int n;
  double gravity;                         // g when Gravity_Switch is on, else 0
  const double2* pos;                     // n sampling positions
  double* out;                            // 4 blocks of n: average p, MLS p, u, v
};
void launch_probe_mls(hipStream_t s, const DevParams& P, Buffers& B, const ProbeArgs& Q);
void launch_wave_probes(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, const double* probe_x,
                        int n_probes, unsigned long long* out_bits);
// Refresh_domain.f90:320-351 from the last evaluation (sph2d_eval64.hip)
void launch_vorticity(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_shift_apply(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
void launch_shift_apply_mode(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int mode);
void launch_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);
// MLS commit (mode >= 1), shifting (mode <= 1, when shift_on) and refresh in one pass
void launch_shift_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int mode, int shift_on);
// predictor-corrector corrector, shifting (when shift_on) and refresh in one pass (no MLS this step)
void launch_correct_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int shift_on);
void launch_sample(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, double* host_out14);
void launch_pairs(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int64_t capacity,
                  int64_t* n_pairs);
void launch_kernel_n(hipStream_t s, int n, const double* r, const double* dx, const double* h,
                     double* w, double* dwdx, double e9);
void launch_eos_n(hipStream_t s, int n, double rho0, double c0sq, const double* rho,
                  const int* type, double* p, double* c);
void launch_fill_i32(hipStream_t s, int* p, int n, int v);
// out[i] = sum_{k<i} in[k]; the total goes to *total. `blocks` holds scan_block_count(n) ints.
void launch_exclusive_scan(hipStream_t s, const int* in, int n, int* out, int* blocks, int* total);
int scan_block_count(int n);
void launch_max_cell(hipStream_t s, Buffers& B, int n_cells);
// Reorder_Particle_Index (sph2d_reorder.hip): blk and off hold 5*G ints (G = ceil(n/256)),
// scan_blocks scan_block_count(5*G), total 1, counts 6, perm n_total.
void launch_reorder_partition(hipStream_t s, int n_total, int n_fluid, int* type, const double* x,
                              const double* y, double box_x, double box_y, int* blk, int* off,
                              int* scan_blocks, int* total, int* counts, int* perm);
// Scratch of the no-allocation Reorder_Particle_Index (sph2d_reorder.hip): the
// device address of this device's copy of the file-scope array, looked up once per
// device; bytes = its size, max_n = the largest n_total it serves.
int reorder_scratch(void** ptr, size_t* bytes, int* max_n);
// Copy of the first *total entries of src to dst (device counts, no host size).
void launch_reorder_copy_f64(hipStream_t s, int n_total, const int* total, const double* src, double* dst);
void launch_reorder_copy_i32(hipStream_t s, int n_total, const int* total, const int* src, int* dst);
void launch_reorder_gather_f64(hipStream_t s, int n_total, const int* perm, const int* total, const double* src,
                               double* dst);
void launch_reorder_gather_i32(hipStream_t s, int n_total, const int* perm, const int* total, const int* src,
                               int* dst);
void launch_count_pairs(hipStream_t s, const DevParams& P, Buffers& B, int cap_local);

// Stage kinds for launch_stage.
enum StageKind {
  ST_SAVE_BEFORE = 0,   // before_* = state
  ST_PC_PREDICT,        // time_iteration.f90:553-566 (+ save temp_*)
  ST_PC_CORRECT,        // :590-599
  ST_RK_HALF0,          // :825-836 (+ save K_1)
  ST_RK_HALF1,          // :859-870 (+ save K_2)
  ST_RK_FULL2,          // :889-900 (+ save K_3)
  ST_RK_FINAL,          // :918-928
  ST_EOS                // status_equation + volume for all local particles
};

}  // namespace sph2d
