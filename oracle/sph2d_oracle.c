/*
 * sph2d_oracle.c -- CPU restatement of the reference WCSPH step (TEST INFRASTRUCTURE).
 *
 * The checker for the HIP path, never the product. It follows the Fortran
 * reference (Taylor_Green_Vortex/ unless stated) line by line: the same
 * global/subdomain arrays, the same local ordering, the same half pair list
 * and scatter order, the same REAL(4) literals and integer-power chains
 * (flang 22 expands x**n with a constant n into x*x*...*x left to right;
 * that chain is checked against the compiled reference in tests/).
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, like the reference build).
 */
#include "sph2d_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* information_module.f90:29  PI (REAL(8) parameter) */
#define PI_MODULE 3.14159265358979324
/* compute_kernel.f90:32  PI=3.14159265358 is a REAL(4) literal stored in REAL(8) */
#define PI_KERNEL ((double)3.14159265358f)

typedef struct { double x, y; } v2;

struct orc_ctx {
  orc_params P;
  /* derived constants (information_module.f90) */
  double dx, dy, h, chain_d, chain_ox, chain_oy, e0, max_shift;
  int chain_nx, chain_ny;
  int n_fluid;                 /* particle_ture_number */
  int n_ghost;                 /* Period_Boundary_Particle_Number, or the fixed ghosts */
  int n_wavemaker;             /* wave_maker_particle_number (first of the ghosts)     */
  int *gbtype;                 /* Boundary_particle_type of fixed ghosts (-6, -7, -8)  */
  int cap_global, cap_local;
  /* global arrays: particle_* (Public_variable_module.f90:33-49) */
  double *gx, *gy, *gu, *gv, *gmass, *grho, *gh, *gc, *gp, *ge;
  int *gtype;
  double *ana_u, *ana_v, *ana_p, *ana_vm;        /* Analytical_particle_* */
  double ke0;
  /* subdomain arrays (local order) */
  int n_local;
  int *name;                    /* Full_subdomain_particle_name (gid) */
  double *x, *y, *u, *v, *mass, *rho, *hh, *c, *p, *e, *vol;
  int *type;
  /* background grid of the subdomain */
  int sub_nx, sub_ny, sub_ncell, sub_start;
  int *cell_count, *cell_list;  /* subdomain_number, subdomain_particle_in_chain (cap 300) */
  int *near_n, *near_name;      /* subdomain_chain_near_mesh_number/_name (9 per cell) */
  int *chain_no;                /* subdomain_particle_chain_number */
  int *cand_n, *cand;           /* subdomain_effect_particle_number / _particle (cap 300) */
  /* pairs */
  int64_t npair, cap_pair;
  int *pi, *pj;
  double *pw, *pdwx, *pdwy, *pdx, *pdy, *pvx, *pvy, *prd, *pc_, *prho_, *ph_;
  double *ndix, *ndiy, *ndjx, *ndjy;            /* Normalized_dwdx_i/_j */
  /* free surface */
  double *fra_v, *sum_w, *sdwx, *sdwy, *T, *Tinv, *nrmx, *nrmy;
  int *surf, *near;
  int n_singular;
  /* solver */
  double *Sxx, *Sxy, *Syy, *txx, *txy, *tyy;
  double *drhodt, *dedt, *ax, *ay;
  double *ia_x, *ia_y, *av_x, *av_y, *ai_x, *ai_y, *sps_x, *sps_y, *lap_x, *lap_y, *ext_x, *ext_y;
  double *avg_x, *avg_y, *corr, *grx, *gry;
  double *gpx, *gpy, *gux, *guy, *gvx, *gvy;    /* MUSCL: grad_pressure, grad_velocity_u/_v */
  /* first-order rounding scale of drhodt, ax, ay per particle (test metric only, no
   * reference counterpart): the sum over the pair terms of each term evaluated with
   * the magnitudes of the operands that cancel inside it (v_i - v*, the pressure
   * and its dissipation, psi = (rho_j - rho_i) - grad rho . x_ji with the rounding
   * scale of grad rho); mag_grho is that scale of grad rho */
  double *mag_drho, *mag_ax, *mag_ay, *mag_grho;
  /* shift */
  double *shx, *shy;
  /* Refresh_domain's vorticity (z), local order; global vorticity, surface flags and
   * normals after the refresh */
  double *vort, *gvort, *gnx, *gny;
  int *gsurf;
  /* particle_in_chain of the step start: row-major cell k holds gchain_members[off[k] ..] */
  int *gchain_off, *gchain_members, gchain_ncell;
  /* time integration (local index arrays, persist across steps) */
  double *b_rho, *b_e, *b_x, *b_y, *b_u, *b_v;
  double *t_u, *t_v, *t_drho, *t_de, *t_ax, *t_ay;              /* PC temp_* */
  double *k_u[3], *k_v[3], *k_drho[3], *k_de[3], *k_ax[3], *k_ay[3]; /* RK K_1..K_3 */
  double *mls_rho;
  /* wavemaker damping (relaxation) zone and wave-theory wavemaker, global arrays:
   * In_WaveMakerDampingZone_OrNot, position_soft_coefficient, WaveMaker_Analytical_*,
   * Initial_particle_Position (Public_variable_module.f90:61-66) */
  int relax_on;
  int *rz_in;
  double *rz_c, *an_x, *an_y, *an_u, *an_v, *an_p, *an_rho, *x0, *y0;
  /* the moving-square program: particle_initial_type (global), Block_Center_X/_Y, the
   * body's Custom_velocity(1) and Custom_accleration(1), Real_time, the time of the
   * evaluation in progress, the body force (pressure, viscous, total; x, y) and its
   * per-particle parts (subdomain_Near_Body_Or_Not, Body_Force_*_Acceleration) */
  int *ginit;
  double body_x, body_y, cust_u, cust_a, real_time, t_eval;
  double bforce[6];
  int *near_body;
  double *bp_x, *bp_y, *bv_x, *bv_y;
};

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */
static void ms_reinitialize_density(orc_ctx* c);
static void* xcalloc(size_t n, size_t s) {
  void* p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "orc: out of memory\n"); abort(); }
  return p;
}
static double ipow7(double x) { return x * x * x * x * x * x * x; }
static double ipow3(double x) { return x * x * x; }
static double ipow4(double x) { return x * x * x * x; }

void orc_default_params(orc_params* p) {
  memset(p, 0, sizeof(*p));
  p->nx = 100; p->ny = 100;
  p->box_x = 1.0; p->box_y = 1.0;
  p->rho0 = 1000.0;
  p->c0sq = 100.0;
  p->nu = (double)1.0E-2f;
  p->U = 1.0;
  p->h_factor = 1.20;
  p->dt = 0.0001;
  p->integrator = 3;
  p->shift_on = 1;
  p->mls_on = 1;
  p->mls_every = 20;
  p->mls_order = 6;
  p->solver = ORC_SOLVER_MODIFIED_LINEAR_RIEMANN;
  p->alpha = 0.0; p->beta = 2.0; p->fai = 0.1;
  p->delta = 0.1; p->delta_alpha = 0.02;
  p->epsilon_xsph = 0.0;
  p->shift_beta = 0.5; p->shift_R = 0.2;
  p->mls_alpha = (double)0.0001f;
  p->damping = 0;                                  /* make_wave_or_not = 0 in the TGV case */
  p->damping_length = 3.0;
  p->damping_amplitude = (double)0.3f;
  p->water_depth = 1.0;                            /* water_depth = domain_size_y */
  p->walls = 0;                                    /* periodic box (TGV default)     */
  p->gravity = 0.0;                                /* Gravity_Switch = 'off'         */
  p->wavemaker_layers = 0;                         /* make_wave_or_not = 0           */
  p->sps_on = 1;                                   /* Solver_..._SPS (single_step_compute.f90:137) */
  p->ms = 0;
  p->background_p = 0.0;                           /* MS Information_Module.f90:486-487 */
  p->release_time = 10.0;                          /* MS :715 */
  p->body_cx = 2.5; p->body_cy = 2.5; p->body_len = 1.0;
  p->body_speed = 1.0; p->body_ramp = 0.5;
  p->sample_every = 100;
  p->max_shift_dx = 0.25;                          /* information_module.f90:353 */
}

/* ------------------------------------------------------------------ */
/* compute_kernel.f90:21-133 (model: improved Gaussian, support 3h)   */
/* ------------------------------------------------------------------ */
void orc_compute_kernel(int32_t dim, double distance, const double* pd, double hbar,
                        double* w, double* dwdx) {
  double q = distance / hbar;
  /* factor=1.0/(PI**(dim/2.0)*(h**dim)*(1-10*exp(-9.0)))  -- last factor is REAL(4) */
  float c4 = 1.0f - 10.0f * expf(-9.0f);
  double hd = (dim == 2) ? hbar * hbar : (dim == 1 ? hbar : hbar * hbar * hbar);
  double pid = (dim == 2) ? PI_KERNEL : pow(PI_KERNEL, (double)((float)dim / 2.0f));
  double factor = 1.0 / (pid * hd * (double)c4);
  for (int i = 0; i < dim; ++i) dwdx[i] = 0.0;
  *w = 0.0;
  if (distance <= 3.0 * hbar) {
    double eq = exp(-(q * q));
    double e9 = (double)expf(-9.0f);
    *w = (eq - e9) * factor;
    for (int i = 0; i < dim; ++i)
      dwdx[i] = (eq * (-2.0 * q)) * pd[i] / (distance * hbar) * factor;
  }
}

/* status_equation.f90:19-82 (water branch, Tait gamma=7) */
void orc_status_equation(const orc_ctx* c, double rho, double e, int32_t type,
                         double* p, double* cs) {
  (void)e;
  if (abs(type) == 1) {            /* air branch: c only, p untouched (:46-50) */
    float gamma = 1.4f;
    *cs = sqrt(((double)gamma - 1.0) * rho);
    return;
  }
  if (c->P.ms) {                   /* MS Status_Equation.f90:62-64, EquationOfState = 1 */
    *p = c->P.c0sq * (rho - c->P.rho0) + c->P.background_p;
    *cs = sqrt(c->P.c0sq) * ipow3(rho / c->P.rho0);
    return;
  }
  double b = c->P.c0sq * c->P.rho0 / 7.0;
  double r = rho / c->P.rho0;
  *p = b * (ipow7(r) - 1.0);
  *cs = sqrt(c->P.c0sq) * ipow3(r);
}

/* function_module.f90:157-249  matrix_inversion_less_three (column-major n x n).
 * Returns check_error (1 = ok). On det==0 only the diagonal of res is set to 1. */
int32_t orc_matrix_inversion_less_three(int32_t n, const double* A, double* res) {
#define M(a, i, j) a[(i) + (j) * n]
  double inv[9] = {0};
  double det = 0.0;
  for (int k = 0; k < n; ++k) M(res, k, k) = 1.0;
  if (n == 2) {
    double a1 = M(A, 0, 0), b1 = M(A, 0, 1), a2 = M(A, 1, 0), b2 = M(A, 1, 1);
    det = (a1 * b2 - b1 * a2);
    if (det != 0) {
      inv[0] = b2 / det; inv[2] = -b1 / det; inv[1] = -a2 / det; inv[3] = a1 / det;
    }
  } else if (n == 3) {
    double a1 = M(A,0,0), b1 = M(A,0,1), c1 = M(A,0,2);
    double a2 = M(A,1,0), b2 = M(A,1,1), c2 = M(A,1,2);
    double a3 = M(A,2,0), b3 = M(A,2,1), c3 = M(A,2,2);
    det = (a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3) + a3 * (b1 * c2 - c1 * b2));
    if (det != 0) {
      inv[0 + 0*3] = (b2*c3 - c2*b3) / det; inv[0 + 1*3] = (b3*c1 - c3*b1) / det;
      inv[0 + 2*3] = (b1*c2 - c1*b2) / det; inv[1 + 0*3] = (c2*a3 - a2*c3) / det;
      inv[1 + 1*3] = (c3*a1 - a3*c1) / det; inv[1 + 2*3] = (c1*a2 - a1*c2) / det;
      inv[2 + 0*3] = (a2*b3 - b2*a3) / det; inv[2 + 1*3] = (a3*b1 - b3*a1) / det;
      inv[2 + 2*3] = (a1*b2 - b1*a2) / det;
    }
  } else {
    return 0;
  }
  if (det != 0) { memcpy(res, inv, sizeof(double) * n * n); return 1; }
  return 0;
#undef M
}

/* function_module.f90:568-639  BRINV: in-place Gauss-Jordan inverse with full
 * pivoting (column-major). Returns L (1 = ok, 0 = singular: D+1.0==1.0). */
int32_t orc_brinv(double* A, int32_t N) {
#define M(i, j) A[(i) + (j) * N]
  int IS[16], JS[16];
  double T, D;
  for (int K = 0; K < N; ++K) {
    D = 0.0;
    for (int I = K; I < N; ++I)
      for (int J = K; J < N; ++J)
        if (fabs(M(I, J)) > D) { D = fabs(M(I, J)); IS[K] = I; JS[K] = J; }
    if (D + 1.0 == 1.0) return 0;
    for (int J = 0; J < N; ++J) { T = M(K, J); M(K, J) = M(IS[K], J); M(IS[K], J) = T; }
    for (int I = 0; I < N; ++I) { T = M(I, K); M(I, K) = M(I, JS[K]); M(I, JS[K]) = T; }
    M(K, K) = 1.0 / M(K, K);
    for (int J = 0; J < N; ++J) if (J != K) M(K, J) = M(K, J) * M(K, K);
    for (int I = 0; I < N; ++I)
      if (I != K)
        for (int J = 0; J < N; ++J)
          if (J != K) M(I, J) = M(I, J) - M(I, K) * M(K, J);
    for (int I = 0; I < N; ++I) if (I != K) M(I, K) = -M(I, K) * M(K, K);
  }
  for (int K = N - 1; K >= 0; --K) {
    for (int J = 0; J < N; ++J) { T = M(K, J); M(K, J) = M(JS[K], J); M(JS[K], J) = T; }
    for (int I = 0; I < N; ++I) { T = M(I, K); M(I, K) = M(I, IS[K]); M(I, IS[K]) = T; }
  }
  return 1;
#undef M
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
static void alloc_local(orc_ctx* c, int cap) {
  c->cap_local = cap;
#define A(f) c->f = xcalloc(cap, sizeof(double))
#define I(f) c->f = xcalloc(cap, sizeof(int))
  I(name); A(x); A(y); A(u); A(v); A(mass); A(rho); A(hh); A(c); A(p); A(e); A(vol); I(type);
  I(chain_no); I(cand_n); c->cand = xcalloc((size_t)cap * 300, sizeof(int));
  A(fra_v); A(sum_w); A(sdwx); A(sdwy); c->T = xcalloc((size_t)cap * 4, sizeof(double));
  c->Tinv = xcalloc((size_t)cap * 4, sizeof(double)); A(nrmx); A(nrmy); I(surf); I(near);
  A(Sxx); A(Sxy); A(Syy); A(txx); A(txy); A(tyy);
  A(drhodt); A(dedt); A(ax); A(ay);
  A(ia_x); A(ia_y); A(av_x); A(av_y); A(ai_x); A(ai_y); A(sps_x); A(sps_y);
  A(lap_x); A(lap_y); A(ext_x); A(ext_y); A(avg_x); A(avg_y); A(corr); A(grx); A(gry);
  A(gpx); A(gpy); A(gux); A(guy); A(gvx); A(gvy);
  A(mag_drho); A(mag_ax); A(mag_ay); A(mag_grho);
  A(shx); A(shy); A(vort);
  A(b_rho); A(b_e); A(b_x); A(b_y); A(b_u); A(b_v);
  A(t_u); A(t_v); A(t_drho); A(t_de); A(t_ax); A(t_ay);
  for (int k = 0; k < 3; ++k) { A(k_u[k]); A(k_v[k]); A(k_drho[k]); A(k_de[k]); A(k_ax[k]); A(k_ay[k]); }
  A(mls_rho);
  I(near_body); A(bp_x); A(bp_y); A(bv_x); A(bv_y);
#undef A
#undef I
}

orc_ctx* orc_create(const orc_params* p) {
  orc_ctx* c = xcalloc(1, sizeof(orc_ctx));
  c->P = *p;
  /* information_module.f90:77-81,133-168 */
  c->dx = p->box_x / p->nx;
  c->dy = p->box_y / p->ny;
  c->h = p->h_factor * c->dx;
  c->chain_d = (3 + 1) * c->h;
  if (p->ms) c->chain_d = 4 * c->dx;                             /* MS Information_Module.f90:189 */
  c->chain_ox = -(5 * c->chain_d);
  c->chain_oy = -(5 * c->chain_d);
  c->chain_nx = (int)(p->box_x / c->chain_d + 2 * 5);
  c->chain_ny = (int)(p->box_y / c->chain_d + 2 * 5);
  c->e0 = p->c0sq / (7.0 - 1.0);
  c->max_shift = p->max_shift_dx * c->dx;
  int n = p->nx * p->ny;
  c->cap_global = (int)(n * 2.5) + 64;      /* n_total = 2.5 * interior number (:124-125) */
  c->n_fluid = 0;
#define A(f) c->f = xcalloc(c->cap_global, sizeof(double))
  A(gx); A(gy); A(gu); A(gv); A(gmass); A(grho); A(gh); A(gc); A(gp); A(ge);
  A(ana_u); A(ana_v); A(ana_p); A(ana_vm); A(gvort); A(gnx); A(gny);
#undef A
  c->gtype = xcalloc(c->cap_global, sizeof(int));
  c->gbtype = xcalloc(c->cap_global, sizeof(int));
  c->gsurf = xcalloc(c->cap_global, sizeof(int));
  c->ginit = xcalloc(c->cap_global, sizeof(int));
  alloc_local(c, c->cap_global);
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->gx, c->gy, c->gu, c->gv, c->gmass, c->grho, c->gh, c->gc, c->gp, c->ge,
                  c->ana_u, c->ana_v, c->ana_p, c->ana_vm, c->gtype, c->gbtype, c->gvort, c->gnx, c->gny, c->gsurf,
                  c->gchain_off, c->gchain_members,
                  c->name, c->x, c->y, c->u, c->v, c->mass, c->rho, c->hh, c->c, c->p, c->e,
                  c->vol, c->type, c->chain_no, c->cand_n, c->cand, c->fra_v, c->sum_w, c->sdwx,
                  c->sdwy, c->T, c->Tinv, c->nrmx, c->nrmy, c->surf, c->near, c->Sxx, c->Sxy,
                  c->Syy, c->txx, c->txy, c->tyy, c->drhodt, c->dedt, c->ax, c->ay, c->ia_x,
                  c->ia_y, c->av_x, c->av_y, c->ai_x, c->ai_y, c->sps_x, c->sps_y, c->lap_x,
                  c->lap_y, c->ext_x, c->ext_y, c->avg_x, c->avg_y, c->corr, c->grx, c->gry,
                  c->gpx, c->gpy, c->gux, c->guy, c->gvx, c->gvy, c->mag_drho, c->mag_ax, c->mag_ay, c->mag_grho, c->shx, c->shy, c->vort, c->b_rho, c->b_e, c->b_x, c->b_y, c->b_u, c->b_v, c->t_u,
                  c->t_v, c->t_drho, c->t_de, c->t_ax, c->t_ay, c->mls_rho,
                  c->rz_in, c->rz_c, c->an_x, c->an_y, c->an_u, c->an_v, c->an_p, c->an_rho, c->x0, c->y0,
                  c->cell_count, c->cell_list, c->near_n, c->near_name,
                  c->pi, c->pj, c->pw, c->pdwx, c->pdwy, c->pdx, c->pdy, c->pvx, c->pvy, c->prd,
                  c->pc_, c->prho_, c->ph_, c->ndix, c->ndiy, c->ndjx, c->ndjy,
                  c->ginit, c->near_body, c->bp_x, c->bp_y, c->bv_x, c->bv_y};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  for (int k = 0; k < 3; ++k) {
    free(c->k_u[k]); free(c->k_v[k]); free(c->k_drho[k]); free(c->k_de[k]);
    free(c->k_ax[k]); free(c->k_ay[k]);
  }
  free(c);
}

/* ------------------------------------------------------------------ */
/* Distribute_Initial_Particle.f90:317-495: fix_ghost_layer layers of fixed ghost */
/* particles around the box (spacing fix_ghost_dx = dx, fix_ghost_dy = dy): per    */
/* layer a top run (right to left), a bottom run (left to right), a right run      */
/* (upwards) and a left run (downwards), each without its end point; the first     */
/* layer's point count k comes from walking until within dx/2 of the run's end,    */
/* and grows by 2 per layer. Distribute_Fixed_Boundary_Particle.f90:52-104 types   */
/* them: -6 top and bottom rows, -7 corners, -8 left and right columns.            */
/* Appended after the fluid as type -3: v = 0, rho = rho0, p = 0, c = c0, e = e0,  */
/* mass = water_particle_mass_0, which the TGV branch never sets (0).              */
/* ------------------------------------------------------------------ */
static int wall_run(double x0, double y0, double x1, double y1, double sx, double sy, double step, double gdx,
                    int* k, int first, double* x, double* y, int at) {
  if (first) {
    int j = 0;
    for (;;) {
      ++j;
      double px = x0 + sx * ((j - 1) * step), py = y0 + sy * ((j - 1) * step);
      if (sqrt((x1 - px) * (x1 - px) + (y1 - py) * (y1 - py)) <= gdx / 2.0) break;
    }
    *k = j;
  }
  double dist = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0)) / (*k - 1);
  int n = 0;
  for (int j = 1; j <= *k - 1; ++j) {
    if (x) { x[at + n] = x0 + sx * ((j - 1) * dist); y[at + n] = y0 + sy * ((j - 1) * dist); }
    ++n;
  }
  *k = *k + 2;
  return n;
}

static int wall_positions(double bx, double by, double gdx, double gdy, int layers, double* x, double* y) {
  int n = 0, k = 0;
  for (int i = 1; i <= layers; ++i)                                  /* Top: :323-362 */
    n += wall_run(bx + (i - 1) * gdy + gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0,
                  -(i - 1) * gdy - gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0, -1, 0, gdx, gdx, &k, i == 1, x, y, n);
  for (int i = 1; i <= layers; ++i)                                  /* Bottom: :366-403 */
    n += wall_run(-(i - 1) * gdy - gdy / 2.0, -(i - 1) * gdy - gdy / 2.0,
                  bx + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0, 1, 0, gdx, gdx, &k, i == 1, x, y, n);
  for (int i = 1; i <= layers; ++i)                                  /* Right: :406-442 */
    n += wall_run(bx + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0,
                  bx + (i - 1) * gdy + gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0, 0, 1, gdx, gdx, &k, i == 1, x, y, n);
  for (int i = 1; i <= layers; ++i)                                  /* Left: :445-480 */
    n += wall_run(-(i - 1) * gdy - gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0,
                  -(i - 1) * gdy - gdy / 2.0, -(i - 1) * gdy - gdy / 2.0, 0, -1, gdy, gdx, &k, i == 1, x, y, n);
  return n;
}

/* Distribute_Initial_Particle.f90:254-311 (PistonType): for rows i (y = (i-1/2) dy  */
/* <= boundary_size_y) and columns j = 1..wave_maker_layer, a particle at           */
/* x = -(j-1) dx - dx/2 with rho = rho0 (-6.0 g (y - Ly)/c0^2 + 1.0)**(1.0/6.0)     */
/* (REAL(4) literals; g = 9.81 whatever Gravity_Switch says), at rest, type 130,   */
/* mass V0 rho(i) with i the wavemaker's own index into the fluid's densities      */
/* (:300-310). Appended right after the fluid; the layer goes to gbtype.           */
#define WAVEMAKER_G 9.81
static void append_wavemaker(orc_ctx* c) {
  const orc_params* P = &c->P;
  const int N = c->n_fluid;
  const double vol0 = c->dx * c->dy;
  int w = 0;
  for (int i = 1; i <= P->ny + 200; ++i)
    for (int j = 1; j <= P->wavemaker_layers; ++j) {
      double px = -(j - 1) * c->dx - c->dx / 2.0;
      double py = (i - 1) * c->dy + c->dy / 2.0;
      if (py <= P->box_y) {
        int k = N + w;
        c->gx[k] = px; c->gy[k] = py;
        c->grho[k] = P->rho0 * pow(-6.0 * WAVEMAKER_G * (py - P->box_y) / P->c0sq + 1.0, (double)(1.0f / 6.0f));
        c->gp[k] = P->c0sq * P->rho0 * (pow(c->grho[k] / P->rho0, 7.0) - 1.0) / 7.0;
        c->gbtype[k] = j;
        ++w;
      }
    }
  for (int q = 0; q < w; ++q) {
    int k = N + q;
    c->gu[k] = 0.0; c->gv[k] = 0.0;
    c->gmass[k] = vol0 * c->grho[q];
    c->gtype[k] = 130; c->gh[k] = c->h; c->gc[k] = sqrt(P->c0sq); c->ge[k] = c->e0;
  }
  c->n_wavemaker = w;
  c->n_ghost = w;
}

/* Wavemaker_Subroutine.f90:140-193 on the global arrays: the water height next to */
/* the paddle (largest fluid y within kernel_scale dx of x_p, from 0, plus dx/2),  */
/* then each wavemaker particle at x_p - (layer-1) dx - dx/2 with the paddle's     */
/* velocity and, below that height, rho0 (-6.0 g (y - height)/c0^2 + 1.0)**(1/6),  */
/* p from the Tait law with exponent 7.0, mass rho V0 (else rho0, p 0, mass 0);    */
/* then :196-215, the fluid pull-back.                                              */
void orc_wavemaker_update(orc_ctx* c, double xp, double up) {
  const orc_params* P = &c->P;
  const int N = c->n_fluid;
  const double vol0 = c->dx * c->dy;
  double left = 0.0;
  for (int k = 0; k < N; ++k) {
    double d = fabs(xp - c->gx[k]);
    if (d <= 3 * c->dx && c->gy[k] > left) left = c->gy[k];
  }
  left = left + 0.5 * c->dx;
  for (int q = 0; q < c->n_wavemaker; ++q) {
    int i = N + q;
    c->gx[i] = xp - (c->gbtype[i] - 1) * c->dx - c->dx / 2.0;
    c->gu[i] = up;
    if (c->gy[i] > left) {
      c->grho[i] = P->rho0; c->gp[i] = 0.0; c->gmass[i] = 0.0;
    } else {
      c->grho[i] = P->rho0 * pow(-6.0 * WAVEMAKER_G * (c->gy[i] - left) / P->c0sq + 1.0, (double)(1.0f / 6.0f));
      c->gp[i] = P->c0sq * P->rho0 * (pow(c->grho[i] / P->rho0, 7.0) - 1.0) / 7.0;
      c->gmass[i] = c->grho[i] * vol0;
    }
  }
  orc_wavemaker_pull_back(c, xp, up);
}

static void append_walls(orc_ctx* c) {
  const orc_params* P = &c->P;
  int N = c->n_fluid + c->n_wavemaker;
  int nw = wall_positions(P->box_x, P->box_y, c->dx, c->dy, P->walls, NULL, NULL);
  if (N + nw > c->cap_global) { fprintf(stderr, "orc: wall capacity\n"); abort(); }
  wall_positions(P->box_x, P->box_y, c->dx, c->dy, P->walls, c->gx + N, c->gy + N);
  for (int q = 0; q < nw; ++q) {
    int m = N + q;
    double px = c->gx[m], py = c->gy[m];
    int bt;
    if (py <= 0.0) bt = (px < 0.0 || px > P->box_x) ? -7 : -6;
    else if (py >= P->box_y) bt = (px < 0.0 || px > P->box_x) ? -7 : -6;
    else bt = -8;
    c->gbtype[m] = bt;
    c->gu[m] = 0.0; c->gv[m] = 0.0; c->grho[m] = P->rho0; c->gmass[m] = 0.0;
    c->gtype[m] = -3; c->gh[m] = c->h; c->gp[m] = 0.0; c->gc[m] = sqrt(P->c0sq); c->ge[m] = c->e0;
  }
  c->n_ghost = c->n_wavemaker + nw;
}

int32_t orc_wall_positions(const orc_ctx* c, double* x, double* y) {
  return wall_positions(c->P.box_x, c->P.box_y, c->dx, c->dy, c->P.walls, x, y);
}

/* ------------------------------------------------------------------ */
/* Distribute_Initial_Particle.f90:204-249  (TGV lattice)             */
/* ------------------------------------------------------------------ */
void orc_init_tgv(orc_ctx* c) {
  const orc_params* P = &c->P;
  int k = 0;
  for (int i = 1; i <= P->ny; ++i) {
    for (int j = 1; j <= P->nx; ++j) {
      double x = (j - 1) * c->dx + c->dx / 2.0;
      double y = (i - 1) * c->dy + c->dy / 2.0;
      c->gx[k] = x; c->gy[k] = y;
      c->gu[k] = P->U * sin(2 * PI_MODULE * x) * cos(2 * PI_MODULE * y);
      c->gv[k] = -P->U * cos(2 * PI_MODULE * x) * sin(2 * PI_MODULE * y);
      double pr = 0.25 * P->rho0 * (P->U * P->U) * (cos(4 * PI_MODULE * x) + cos(4 * PI_MODULE * y));
      c->gp[k] = pr;
      /* exponent 0.142857142 is a REAL(4) literal */
      c->grho[k] = P->rho0 * pow(1 + 7 * pr / (P->c0sq * P->rho0), (double)0.142857142f);
      ++k;
    }
  }
  c->n_fluid = k;
  double vol0 = c->dx * c->dy;
  for (int i = 0; i < c->n_fluid; ++i) {
    c->gmass[i] = vol0 * c->grho[i];
    c->gtype[i] = 2;
    c->gh[i] = c->h;
    c->gc[i] = sqrt(P->c0sq);
    c->ge[i] = c->e0;
  }
  /* Initialize_SPH_Calculation.f90:490-496 */
  double ke = 0.0;
  for (int i = 0; i < c->n_fluid; ++i)
    ke = ke + 0.5 * c->gmass[i] * (c->gu[i] * c->gu[i] + c->gv[i] * c->gv[i]);
  c->ke0 = ke;
  c->n_ghost = 0;
  c->n_wavemaker = 0;
  if (P->wavemaker_layers > 0) append_wavemaker(c);
  if (P->walls > 0) append_walls(c);
}

int32_t orc_set_fluid_state(orc_ctx* c, int32_t n, const double* x, const double* y,
                            const double* u, const double* v, const double* rho,
                            const double* mass, const int32_t* type) {
  if (n + 8 > c->cap_global / 2) return -1;
  for (int i = 0; i < n; ++i) {
    c->gx[i] = x[i]; c->gy[i] = y[i]; c->gu[i] = u[i]; c->gv[i] = v[i];
    c->grho[i] = rho[i]; c->gmass[i] = mass[i]; c->gtype[i] = type ? type[i] : 2;
    c->gh[i] = c->h; c->gc[i] = sqrt(c->P.c0sq); c->ge[i] = c->e0; c->gp[i] = 0.0;
  }
  c->n_fluid = n;
  c->n_ghost = 0;
  double ke = 0.0;
  for (int i = 0; i < n; ++i) ke = ke + 0.5 * c->gmass[i] * (c->gu[i] * c->gu[i] + c->gv[i] * c->gv[i]);
  c->ke0 = ke;
  c->n_wavemaker = 0;
  if (c->P.wavemaker_layers > 0) append_wavemaker(c);
  if (c->P.walls > 0) append_walls(c);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Distribute_Period_Boundary_Particle.f90:56-206                     */
/* ------------------------------------------------------------------ */
static void copy_ghost(orc_ctx* c, int j, int i, double x, double y) {
  c->gx[j] = x; c->gy[j] = y;
  c->gu[j] = c->gu[i]; c->gv[j] = c->gv[i];
  c->grho[j] = c->grho[i]; c->gmass[j] = c->gmass[i];
  c->gtype[j] = -101;
  c->gh[j] = c->gh[i]; c->gp[j] = c->gp[i]; c->gc[j] = c->gc[i]; c->ge[j] = c->ge[i];
}

void orc_period_boundary(orc_ctx* c) {
  if (c->P.walls > 0) return;                         /* fixed walls: no periodic ghosts */
  const double xmin = 0.0, xmax = c->P.box_x, lx = c->P.box_x;
  const double ymin = 0.0, ymax = c->P.box_y, ly = c->P.box_y;
  const double r3h = 3 * c->h;   /* kernel_scale*smooth_length */
  int N = c->n_fluid;
  for (int i = 0; i < N; ++i) {                       /* :56-90 wrap */
    if (c->gtype[i] == 120) continue;
    if (c->gx[i] > xmax) c->gx[i] = c->gx[i] - lx;
    else if (c->gx[i] < xmin) c->gx[i] = c->gx[i] + lx;
    if (c->gy[i] > ymax) c->gy[i] = c->gy[i] - ly;
    else if (c->gy[i] < ymin) c->gy[i] = c->gy[i] + ly;
  }
  int ng = 0;
  for (int i = 0; i < N; ++i) {                       /* :98-148 x-direction ghosts */
    if (c->gtype[i] == 120) continue;
    if (c->gx[i] <= (xmin + r3h) && c->gx[i] >= xmin) {
      ++ng; copy_ghost(c, N + ng - 1, i, c->gx[i] + lx, c->gy[i]);
    }
    if (c->gx[i] >= (xmax - r3h) && c->gx[i] <= xmax) {
      ++ng; copy_ghost(c, N + ng - 1, i, c->gx[i] - lx, c->gy[i]);
    }
  }
  int nx_g = ng;
  for (int i = 0; i < N + nx_g; ++i) {                /* :153-206 y-direction (and corners) */
    if (c->gtype[i] == 120) continue;
    if (c->gy[i] <= (ymin + r3h) && c->gy[i] >= ymin) {
      ++ng; copy_ghost(c, N + ng - 1, i, c->gx[i], c->gy[i] + ly);
    }
    if (c->gy[i] >= (ymax - r3h) && c->gy[i] <= ymax) {
      ++ng; copy_ghost(c, N + ng - 1, i, c->gx[i], c->gy[i] - ly);
    }
  }
  c->n_ghost = ng;
}

/* ------------------------------------------------------------------ */
/* Devide_block_and_send_tasks.f90 on one rank:                       */
/*  analytic fields (:88-116), global binning (:123-162), local order */
/*  column-major by cell then ascending gid (:488-633), copy (:666-)  */
/* Distribute_Backgroud_Grid.f90:102-200: the near cells of row-major cell i (1-based)
 * of a gx_n x gy_n grid, in the order self, L, R, UL, U, UR, DL, D, DR (edges 6,
 * corners 4); returns their number. */
static int stencil_names(int i, int gx_n, int gy_n, int* nm) {
  int mx = (i - 1) % gx_n + 1, my = (i - 1) / gx_n + 1;
  if (mx > 1 && mx < gx_n && my > 1 && my < gy_n) {
    nm[0] = i; nm[1] = i - 1; nm[2] = i + 1; nm[3] = i - 1 + gx_n; nm[4] = i + gx_n;
    nm[5] = i + 1 + gx_n; nm[6] = i - 1 - gx_n; nm[7] = i - gx_n; nm[8] = i + 1 - gx_n;
    return 9;
  } else if (mx == 1 && my > 1 && my < gy_n) {
    nm[0] = i; nm[1] = i + 1; nm[2] = i + gx_n; nm[3] = i + 1 + gx_n; nm[4] = i - gx_n; nm[5] = i + 1 - gx_n;
    return 6;
  } else if (mx == gx_n && my > 1 && my < gy_n) {
    nm[0] = i; nm[1] = i - 1; nm[2] = i + gx_n; nm[3] = i - 1 + gx_n; nm[4] = i - gx_n; nm[5] = i - 1 - gx_n;
    return 6;
  } else if (my == gy_n && mx > 1 && mx < gx_n) {
    nm[0] = i; nm[1] = i - 1; nm[2] = i + 1; nm[3] = i - 1 - gx_n; nm[4] = i - gx_n; nm[5] = i + 1 - gx_n;
    return 6;
  } else if (my == 1 && mx > 1 && mx < gx_n) {
    nm[0] = i; nm[1] = i - 1; nm[2] = i + 1; nm[3] = i - 1 + gx_n; nm[4] = i + gx_n; nm[5] = i + 1 + gx_n;
    return 6;
  } else if (mx == 1 && my == gy_n) {
    nm[0] = i; nm[1] = i + 1; nm[2] = i - gx_n; nm[3] = i + 1 - gx_n;
    return 4;
  } else if (mx == gx_n && my == gy_n) {
    nm[0] = i; nm[1] = i - 1; nm[2] = i - gx_n; nm[3] = i - 1 - gx_n;
    return 4;
  } else if (mx == 1 && my == 1) {
    nm[0] = i; nm[1] = i + 1; nm[2] = i + gx_n; nm[3] = i + 1 + gx_n;
    return 4;
  } else if (mx == gx_n && my == 1) {
    nm[0] = i; nm[1] = i - 1; nm[2] = i + gx_n; nm[3] = i - 1 + gx_n;
    return 4;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
void orc_devide_block(orc_ctx* c, int32_t i_time_step) {
  const orc_params* P = &c->P;
  int N = c->n_fluid;
  int NG = N + c->n_ghost;
  for (int i = 0; i < N; ++i) {                       /* :75-83 */
    if (c->gtype[i] == 120) continue;
    if (c->gy[i] < 0.0) c->gtype[i] = 120;
  }
  double t = i_time_step * P->dt / (P->box_x / P->U);  /* Nondimensionalized_real_time */
  double Re = P->U * P->box_x / P->nu;
  for (int i = 0; i < N; ++i) {                       /* :88-116 */
    double x = c->gx[i], y = c->gy[i];
    c->ana_u[i] = P->U * sin(2 * PI_MODULE * x) * cos(2 * PI_MODULE * y) *
                  exp(-8 * (PI_MODULE * PI_MODULE) / Re * t);
    c->ana_v[i] = -P->U * cos(2 * PI_MODULE * x) * sin(2 * PI_MODULE * y) *
                  exp(-8 * (PI_MODULE * PI_MODULE) / Re * t);
    c->ana_p[i] = 0.25 * P->rho0 * (P->U * P->U) *
                  (cos(4 * PI_MODULE * x) + cos(4 * PI_MODULE * y)) *
                  exp(-16 * (PI_MODULE * PI_MODULE) / Re * t);
    c->ana_vm[i] = sqrt(c->ana_u[i] * c->ana_u[i] + c->ana_v[i] * c->ana_v[i]);
  }
  /* :123-162 binning of all global particles into 4h cells, row-major id */
  int nx = c->chain_nx, ny = c->chain_ny, ncell = nx * ny;
  int* cnt = xcalloc(ncell + 1, sizeof(int));
  int* cellof = xcalloc(NG, sizeof(int));
  for (int i = 0; i < NG; ++i) {
    cellof[i] = -1;
    if (c->gtype[i] == 120 || c->gtype[i] == 0) continue;
    int xn = (int)floor((c->gx[i] - c->chain_ox) / c->chain_d + 1);
    int yn = (int)floor((c->gy[i] - c->chain_oy) / c->chain_d + 1);
    int k = (yn - 1) * nx + xn;
    if (k <= ncell && k >= 1 && xn >= 1 && xn <= nx) {
      cellof[i] = k; cnt[k]++;
    } else {
      c->gtype[i] = 120;
    }
  }
  /* single rank: start/end column = first/last non-empty column (:281-300) */
  int* colcnt = xcalloc(nx + 2, sizeof(int));
  for (int i = 0; i < NG; ++i)
    if (cellof[i] > 0) colcnt[(cellof[i] - 1) % nx + 1]++;
  int start = 1, end = nx;
  for (start = 1; start <= nx; ++start) if (colcnt[start] != 0) break;
  for (end = nx; end >= 1; --end) if (colcnt[end] != 0) break;
  /* stable bucket: row-major cell k holds gids in ascending order */
  int* off = xcalloc(ncell + 2, sizeof(int));
  for (int k = 1; k <= ncell; ++k) off[k + 1] = off[k] + cnt[k];
  int* members = xcalloc(NG + 1, sizeof(int));
  int* fill = xcalloc(ncell + 2, sizeof(int));
  for (int i = 0; i < NG; ++i)
    if (cellof[i] > 0) { int k = cellof[i]; members[off[k] + fill[k]++] = i; }
  /* local order: column k (x) ascending, row j (y) ascending, members ascending gid */
  int n = 0;
  for (int col = start; col <= end; ++col)
    for (int row = 1; row <= ny; ++row) {
      int k = col + (row - 1) * nx;
      for (int m = 0; m < cnt[k]; ++m) c->name[n++] = members[off[k] + m];
    }
  c->n_local = n;
  for (int k = 0; k < n; ++k) {                       /* :666-765 copy global -> subdomain */
    int i = c->name[k];
    c->x[k] = c->gx[i]; c->y[k] = c->gy[i];
    c->u[k] = c->gu[i]; c->v[k] = c->gv[i];
    c->mass[k] = c->gmass[i]; c->type[k] = c->gtype[i];
    c->rho[k] = c->grho[i]; c->hh[k] = c->gh[i]; c->c[k] = c->gc[i];
    c->p[k] = c->gp[i]; c->e[k] = c->ge[i];
  }
  /* subdomain background grid (:304-360 + Distribute_Backgroud_Grid.f90:72-200) */
  c->sub_nx = (end - start + 1) + 2;
  c->sub_ny = ny;
  c->sub_ncell = c->sub_nx * c->sub_ny;
  free(c->cell_count); free(c->cell_list); free(c->near_n); free(c->near_name);
  c->cell_count = xcalloc(c->sub_ncell + 1, sizeof(int));
  c->cell_list = xcalloc((size_t)(c->sub_ncell + 1) * 300, sizeof(int));
  c->near_n = xcalloc(c->sub_ncell + 1, sizeof(int));
  c->near_name = xcalloc((size_t)(c->sub_ncell + 1) * 9, sizeof(int));
  for (int i = 1; i <= c->sub_ncell; ++i) c->near_n[i] = stencil_names(i, c->sub_nx, c->sub_ny, &c->near_name[i * 9]);
  /* the subdomain grid origin (:330-333) */
  c->sub_start = start;
  /* the global binning (particle_in_chain, :144-148) stays for the post-processing */
  free(c->gchain_off); free(c->gchain_members);
  c->gchain_off = off; c->gchain_members = members; c->gchain_ncell = ncell;
  free(cnt); free(cellof); free(colcnt); free(fill);
}

/* ------------------------------------------------------------------ */
/* Assign_boundary_particles.f90:75-152,303-321                       */
/* ------------------------------------------------------------------ */
/* Assign_boundary_particles.f90:157-299: each fixed ghost takes the pressure and   */
/* velocity of its non-ghost candidates within 3 hbar, hbar = 1.2 (h_i + h_j)/2    */
/* (with gravity, each p_j carried to the wall's height: p_j - g rho_j (y_i - y_j)) */
/* (1.2 a REAL(4) literal), by MLS of order 1 (MLS_Interpolation_In_Subdomain_     */
/* Domain.f90 with MLS_order_Boundary = 1: Fai_k = (1/sum W) W_k, with 3 or more   */
/* particles, else zero); the velocity is mirrored by the boundary type (-6: v,    */
/* -7: u and v, -8: u); then rho from the inverse Tait law (REAL(4) exponent),     */
/* mass = rho V0 and c = c0 (rho/rho0)^3, or the rest state when p is 0.           */
static void assign_walls(orc_ctx* c) {
  const orc_params* P = &c->P;
  const double vol0 = c->dx * c->dy;
  int* tn = xcalloc(300, sizeof(int));
  double* wk = xcalloc(300, sizeof(double));
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != -3) continue;
    double u = 0.0, v = 0.0, pr = 0.0;
    int nt = 0;
    for (int L = 0; L < c->cand_n[i]; ++L) {
      int j = c->cand[(size_t)i * 300 + L];
      if (c->type[j] == -3) continue;
      double pd[2] = {c->x[i] - c->x[j], c->y[i] - c->y[j]};
      double d = sqrt(pd[0] * pd[0] + pd[1] * pd[1]);
      double hb = (double)1.2f * (c->hh[i] + c->hh[j]) / 2.0;
      if (d <= 3 * hb && d >= (double)1.0E-7f) {
        double w, dw[2];
        orc_compute_kernel(2, d, pd, hb, &w, dw);
        tn[nt] = j; wk[nt] = w; ++nt;
      }
    }
    if (nt >= 3) {
      double a = 0.0;
      for (int k = 0; k < nt; ++k) a = a + wk[k] * 1.0;
      double ainv = 1.0 / a;
      for (int k = 0; k < nt; ++k) {
        double fai = 1.0 * (ainv * wk[k]);
        int j = tn[k];
        if (P->gravity != 0.0) pr = pr + c->p[j] * fai - P->gravity * c->rho[j] * (c->y[i] - c->y[j]) * fai;
        else pr = pr + c->p[j] * fai;
        u = u + c->u[j] * fai;
        v = v + c->v[j] * fai;
      }
    }
    int bt = c->gbtype[c->name[i]];
    if (bt == -6) v = -v;
    else if (bt == -7) { u = -u; v = -v; }
    else if (bt == -8) u = -u;
    if (pr != 0.0) {
      c->e[i] = c->e0;
      c->rho[i] = P->rho0 * pow(1 + 7 * pr / (P->c0sq * P->rho0), (double)0.142857142f);
      c->mass[i] = c->rho[i] * vol0;
      c->c[i] = sqrt(P->c0sq) * ipow3(c->rho[i] / P->rho0);
    } else {
      u = 0.0; v = 0.0; pr = 0.0;
      c->e[i] = c->e0; c->rho[i] = P->rho0; c->mass[i] = 0.0; c->c[i] = sqrt(P->c0sq);
    }
    c->u[i] = u; c->v[i] = v; c->p[i] = pr;
  }
  free(tn); free(wk);
}

/* MS Assign_boundary_particles.f90:611-760 (DistributeFixedBoundaryParticleOrNot = 1,    */
/* ExternalBoundaryCondition = 0, MLS_order_Boundary = 1): every fixed ghost i (walls     */
/* and body) starts from v = 0, p = 0, m = 0; its non-ghost candidates within 3 hbar,    */
/* hbar = 1.2 (REAL(4)) smooth_length, no lower bound, weigh W m_j / rho_j; with 3 or    */
/* more of them the order-1 MLS (Fai_k = (1/sum W) W_k) gives                            */
/*   P = sum p_j Fai + rho_j Fai g_ext . (x_i - x_j), v = sum v_j Fai, m = sum m_j Fai    */
/* (g_ext = gravity, plus the body's acceleration for body particles). Walls, while     */
/* 0 < t <= Shock_wave_release_time, get the characteristic correction of their normal  */
/* velocity and pressure with reference state (u, p) = (0, 0) (boundary type -6: y,     */
/* -7: x; -8 keeps p = 0: the reference sets p only in those two branches), otherwise    */
/* v, p = the MLS values; the mirror then negates both components (no-slip). Body       */
/* particles take the MLS pressure and the body's velocity. Finally rho = (p - p_b) /    */
/* c0^2 + rho0 and c = c0 (rho/rho0)^3, or the rest state when m is 0.                   */
static void ms_assign_walls(orc_ctx* c) {
  const orc_params* P = &c->P;
  const double t = c->t_eval;
  const double hb = (double)1.2f * c->h;
  const double sqc0 = sqrt(P->c0sq);
  int* tn = xcalloc(300, sizeof(int));
  double* wk = xcalloc(300, sizeof(double));
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != -3) continue;
    const int g = c->name[i];
    const int body = c->ginit[g] == -4;
    double iu = 0.0, iv = 0.0, ip = 0.0, m = 0.0, pr = 0.0;
    int nt = 0;
    for (int L = 0; L < c->cand_n[i]; ++L) {
      int j = c->cand[(size_t)i * 300 + L];
      if (c->type[j] == -3) continue;
      double pd[2] = {c->x[i] - c->x[j], c->y[i] - c->y[j]};
      double d = sqrt(pd[0] * pd[0] + pd[1] * pd[1]);
      if (d <= 3.0 * hb) {
        double w, dw[2];
        orc_compute_kernel(2, d, pd, hb, &w, dw);
        tn[nt] = j; wk[nt] = w * c->mass[j] / c->rho[j]; ++nt;
      }
    }
    const double gx = body ? 0.0 + c->cust_a : 0.0;                /* Gravity_grad + Custom_accleration */
    const double gy = body ? -P->gravity + 0.0 : -P->gravity;
    if (nt >= 3) {
      double a = 0.0;
      for (int k = 0; k < nt; ++k) a = a + wk[k] * 1.0;
      double ainv = 1.0 / a;
      for (int k = 0; k < nt; ++k) {
        double fai = 1.0 * (ainv * wk[k]);
        int j = tn[k];
        double dot = gx * (c->x[i] - c->x[j]) + gy * (c->y[i] - c->y[j]);
        ip = ip + c->p[j] * fai + c->rho[j] * fai * dot;
        iu = iu + c->u[j] * fai;
        iv = iv + c->v[j] * fai;
        m = m + c->mass[j] * fai;
      }
    }
    double u, v;
    if (!body) {
      if (0.0 < t && t <= P->release_time) {
        const double irho = (ip - P->background_p) / P->c0sq + P->rho0;
        const double tc = sqc0 * ipow3(irho / P->rho0);
        u = iu; v = iv;
        const int bt = c->gbtype[g];
        if (bt == -6) {
          double J2 = irho * tc * (iv - 0.0) + (ip - 0.0);
          double J3 = -(irho * tc * (iv - 0.0)) + (ip - 0.0);
          v = 0.0 + (J2 - J3) / (2 * irho * tc);
          pr = 0.0 + 0.5 * (J2 + J3);
        } else if (bt == -7) {
          double J2 = irho * tc * (iu - 0.0) + (ip - 0.0);
          double J3 = -(irho * tc * (iu - 0.0)) + (ip - 0.0);
          u = 0.0 + (J2 - J3) / (2 * irho * tc);
          pr = 0.0 + 0.5 * (J2 + J3);
        }
      } else {
        u = iu; v = iv; pr = ip;
      }
      u = -u; v = -v;                                           /* -6, -7, -8 no-slip mirrors */
    } else {
      pr = ip;
      u = c->cust_u; v = 0.0;                                   /* Custom_velocity */
    }
    if (m != 0.0) {
      c->e[i] = c->e0;
      c->rho[i] = (pr - P->background_p) / P->c0sq + P->rho0;
      c->c[i] = sqc0 * ipow3(c->rho[i] / P->rho0);
    } else {
      u = 0.0; v = 0.0; pr = 0.0;
      c->e[i] = c->e0; c->rho[i] = P->rho0; c->c[i] = sqc0;
    }
    c->u[i] = u; c->v[i] = v; c->p[i] = pr; c->mass[i] = m;
  }
  free(tn); free(wk);
}

static void assign_boundary_particles(orc_ctx* c) {
  int n = c->n_local;
  double o_sub = c->chain_ox + (c->sub_start - 1) * c->chain_d;
  memset(c->cell_count, 0, sizeof(int) * (c->sub_ncell + 1));
  memset(c->chain_no, 0, sizeof(int) * n);
  for (int i = 0; i < n; ++i) {
    if (c->type[i] == 120 || c->type[i] == 0) continue;
    int xn = (int)floor((c->x[i] - o_sub) / c->chain_d + 1) + 1;
    int yn = (int)floor((c->y[i] - c->chain_oy) / c->chain_d + 1);
    int k = (yn - 1) * c->sub_nx + xn;
    if (k <= c->sub_ncell && k >= 1) {
      c->chain_no[i] = k;
      c->cell_count[k]++;
      if (c->cell_count[k] <= 300) c->cell_list[(size_t)k * 300 + c->cell_count[k] - 1] = i;
    } else {
      c->type[i] = 120;
    }
  }
  for (int i = 0; i < n; ++i) {
    int mesh = c->chain_no[i];
    int nn = c->near_n[mesh];
    int cnt = 0;
    for (int k = 0; k < nn; ++k) {
      int nm = c->near_name[mesh * 9 + k];
      int pm = c->cell_count[nm] < 300 ? c->cell_count[nm] : 300;
      for (int v = 0; v < pm; ++v) {
        int j = c->cell_list[(size_t)nm * 300 + v];
        ++cnt;
        if (cnt <= 300) c->cand[(size_t)i * 300 + cnt - 1] = j;
      }
    }
    c->cand_n[i] = cnt < 300 ? cnt : 300;
  }
  if (c->P.walls > 0 && c->P.ms) ms_assign_walls(c);             /* MS :611-760 */
  else if (c->P.walls > 0) assign_walls(c);                     /* :157-299 */
  for (int k = 0; k < n; ++k) c->vol[k] = c->mass[k] / c->rho[k];
}

/* ------------------------------------------------------------------ */
/* find_particle_pair.f90:61-125                                       */
/* ------------------------------------------------------------------ */
static void grow_pairs(orc_ctx* c, int64_t need) {
  if (need <= c->cap_pair) return;
  int64_t cap = need * 3 / 2 + 1024;
#define R(f, t) c->f = realloc(c->f, cap * sizeof(t)); if (!c->f) abort();
  R(pi, int); R(pj, int); R(pw, double); R(pdwx, double); R(pdwy, double);
  R(pdx, double); R(pdy, double); R(pvx, double); R(pvy, double); R(prd, double);
  R(pc_, double); R(prho_, double); R(ph_, double);
  R(ndix, double); R(ndiy, double); R(ndjx, double); R(ndjy, double);
#undef R
  c->cap_pair = cap;
}

static void find_particle_pair(orc_ctx* c) {
  int n = c->n_local;
  c->npair = 0;
  grow_pairs(c, (int64_t)n * 40 + 16);
  /* MS Find_particle_pair.f90:73-75: 1.0E-10 (REAL(4)) */
  const double rmin = c->P.ms ? (double)1.0E-10f : (double)1.0E-7f;
  for (int i = 0; i < n - 1; ++i) {
    for (int L = 0; L < c->cand_n[i]; ++L) {
      int j = c->cand[(size_t)i * 300 + L];
      if (j <= i || j >= n) continue;
      if (c->type[i] != 2 && c->type[j] != 2) continue;
      double pd[2] = {c->x[i] - c->x[j], c->y[i] - c->y[j]};
      double r = sqrt(pd[0] * pd[0] + pd[1] * pd[1]);
      double hb = (c->hh[i] + c->hh[j]) / 2.0;
      if (r <= (3 * hb) && r >= rmin) {
        grow_pairs(c, c->npair + 1);
        int64_t k = c->npair++;
        double w, dw[2];
        orc_compute_kernel(2, r, pd, hb, &w, dw);
        c->pi[k] = i; c->pj[k] = j; c->pw[k] = w;
        c->pdwx[k] = dw[0]; c->pdwy[k] = dw[1];
        c->pdx[k] = pd[0]; c->pdy[k] = pd[1];
        c->pvx[k] = c->u[i] - c->u[j]; c->pvy[k] = c->v[i] - c->v[j];
        c->prd[k] = c->rho[i] - c->rho[j];
        c->pc_[k] = (c->c[i] + c->c[j]) / 2.0;
        c->prho_[k] = (c->rho[i] + c->rho[j]) / 2.0;
        c->ph_[k] = hb;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Define_freesurface_and_Normailized_derivates.f90:65-440            */
/* ------------------------------------------------------------------ */
static void define_freesurface(orc_ctx* c) {
  int n = c->n_local;
  const double sv = PI_MODULE * ((3 * c->h) * (3 * c->h));     /* :65 support_volume */
  memset(c->fra_v, 0, sizeof(double) * n);
  memset(c->sum_w, 0, sizeof(double) * n);
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :79-93 */
    int i = c->pi[k], j = c->pj[k];
    c->fra_v[i] = c->fra_v[i] + c->vol[j] / sv;
    c->fra_v[j] = c->fra_v[j] + c->vol[i] / sv;
    if (c->type[i] == c->type[j]) {
      c->sum_w[i] = c->sum_w[i] + c->pw[k] * c->vol[j];
      c->sum_w[j] = c->sum_w[j] + c->pw[k] * c->vol[i];
    }
  }
  int* slist = xcalloc(n + 1, sizeof(int));
  int ns = 0;
  memset(c->surf, 0, sizeof(int) * n);
  for (int i = 0; i < n; ++i) {                                   /* :119-131 */
    if (c->P.ms) break;                  /* MS :42-51: Has_FreeSurface_Or_Not = 0 */
    if (c->type[i] != 2) continue;
    if (c->fra_v[i] <= (double)0.65f) { slist[ns++] = i; c->surf[i] = 1; }
  }
  memset(c->T, 0, sizeof(double) * 4 * n);
  memset(c->sdwx, 0, sizeof(double) * n);
  memset(c->sdwy, 0, sizeof(double) * n);
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :141-159 */
    int i = c->pi[k], j = c->pj[k];
    double pd[2] = {-c->pdx[k], -c->pdy[k]};
    double dw[2] = {c->pdwx[k], c->pdwy[k]};
    double* sd_i[2] = {&c->sdwx[i], &c->sdwy[i]};
    double* sd_j[2] = {&c->sdwx[j], &c->sdwy[j]};
    for (int L = 0; L < 2; ++L) {
      for (int o = 0; o < 2; ++o) {
        c->T[i * 4 + L + 2 * o] = c->T[i * 4 + L + 2 * o] + pd[L] * dw[o] * c->vol[j];
        c->T[j * 4 + L + 2 * o] = c->T[j * 4 + L + 2 * o] + pd[L] * dw[o] * c->vol[i];
      }
      *sd_i[L] = *sd_i[L] + dw[L] * c->vol[j];
      *sd_j[L] = *sd_j[L] - dw[L] * c->vol[i];
    }
  }
  c->n_singular = 0;
  for (int i = 0; i < n; ++i) {                                   /* :168-188 */
    double* Ti = &c->Tinv[i * 4];
    Ti[0] = 1.0; Ti[1] = 0.0; Ti[2] = 0.0; Ti[3] = 1.0;
    if (c->type[i] != 2) continue;
    double A[4] = {c->T[i * 4 + 0], c->T[i * 4 + 1], c->T[i * 4 + 2], c->T[i * 4 + 3]};
    double Ainv[4] = {1.0, 0.0, 0.0, 1.0};
    int ok = orc_matrix_inversion_less_three(2, A, Ainv);
    if (!ok) c->n_singular++;
    /* check_error==1 -> the caller replaces the inverse by the identity (:180-186);
     * check_error==0 keeps diag=1 and the previous (identity) off-diagonals. With
     * renorm (a mode of this library, no reference counterpart) the inverse is kept. */
    Ti[0] = 1.0; Ti[1] = 0.0; Ti[2] = 0.0; Ti[3] = 1.0;
    if (c->P.renorm && ok) { Ti[0] = Ainv[0]; Ti[1] = Ainv[1]; Ti[2] = Ainv[2]; Ti[3] = Ainv[3]; }
  }
  for (int L = 0; L < ns; ++L) {                                  /* :295-376 */
    int i = slist[L];
    double* Ti = &c->Tinv[i * 4];
    double r1 = c->sdwx[i], r2 = c->sdwy[i];
    double n1 = Ti[0] * r1 + Ti[2] * r2;
    double n2 = Ti[1] * r1 + Ti[3] * r2;
    double len = 0.0;
    len = len + n1 * n1;
    len = len + n2 * n2;
    len = sqrt(len);
    if (isnan(len)) { c->surf[i] = 0; continue; }
    if (len != 0.0) { n1 = -n1 / len; n2 = -n2 / len; }
    else { c->surf[i] = 0; continue; }
    c->nrmx[i] = n1; c->nrmy[i] = n2;
    double vx = c->x[i] + c->h * n1, vy = c->y[i] + c->h * n2;
    for (int k = 0; k < c->cand_n[i]; ++k) {
      int j = c->cand[(size_t)i * 300 + k];
      if (c->type[j] != 2) continue;
      double pdx = c->x[j] - c->x[i], pdy = c->y[j] - c->y[i];
      double JI = sqrt(pdx * pdx + pdy * pdy);
      double Nd = n1 * pdx + n2 * pdy;
      double qx = c->x[j] - vx, qy = c->y[j] - vy;
      double JT = sqrt(qx * qx + qy * qy);
      double lim = (double)1.414f * c->h;
      if (JI >= lim && JT < c->h) {
        c->surf[i] = 0;
      } else if (JI < lim && acos(Nd / JI) <= (PI_MODULE / 4.0)) {
        c->surf[i] = 0;
      }
    }
  }
  free(slist);
  memset(c->near, 0, sizeof(int) * n);
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :418-440 */
    int i = c->pi[k], j = c->pj[k];
    if (c->surf[i] == 1) c->near[j] = 1;
    if (c->surf[j] == 1) c->near[i] = 1;
  }
}

/* ------------------------------------------------------------------ */
/* function_module.f90:34-80: limited MUSCL slopes (Toro, beta_lim=1.3) */
/* ------------------------------------------------------------------ */
void orc_muscl_limiter(double delphi_i, double delphi_j, double d_phi_ij,
                       double* del_phi_lim_i, double* del_phi_lim_j) {
  const double beta_lim = 1.3;                                    /* real(8) 1.3d0 */
  const double diff_factor = 0.5;
  double lim, fi, fj, a, b;
  if (delphi_j > 0.0) {                                           /* :47-61 */
    a = beta_lim * delphi_i < delphi_j ? beta_lim * delphi_i : delphi_j;
    b = delphi_i < beta_lim * delphi_j ? delphi_i : beta_lim * delphi_j;
    lim = 0.0;
    if (a > lim) lim = a;
    if (b > lim) lim = b;
    if (lim > 0.5 * d_phi_ij && fabs(lim) > 0.0) {
      if (delphi_i > delphi_j) { fi = diff_factor; fj = 0.0; }
      else { fi = 0.0; fj = diff_factor; }
      *del_phi_lim_i = d_phi_ij * fi;
      *del_phi_lim_j = d_phi_ij * fj;
    } else {
      *del_phi_lim_i = lim;
      *del_phi_lim_j = lim;
    }
  } else {                                                        /* :62-76 */
    a = beta_lim * delphi_i > delphi_j ? beta_lim * delphi_i : delphi_j;
    b = delphi_i > beta_lim * delphi_j ? delphi_i : beta_lim * delphi_j;
    lim = 0.0;
    if (a < lim) lim = a;
    if (b < lim) lim = b;
    if (lim < 0.5 * d_phi_ij && fabs(lim) > 0.0) {
      if (delphi_i < delphi_j) { fi = diff_factor; fj = 0.0; }
      else { fi = 0.0; fj = diff_factor; }
      *del_phi_lim_i = d_phi_ij * fi;
      *del_phi_lim_j = d_phi_ij * fj;
    } else {
      *del_phi_lim_i = lim;
      *del_phi_lim_j = lim;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90:362-490:        */
/* rotate to the pair frame, acoustic U*, P*, wave speeds S_L, S_R     */
/* (Q = 1 unless P* exceeds that side's pressure), sample the fan at   */
/* x/t, rotate back. Rho_asterisk (:387) is computed there but unused. */
/* ------------------------------------------------------------------ */
void orc_hllc_star(const double* vl, double PL, double rhoL, double cL,
                   const double* vr, double PR, double rhoR, double cR,
                   const double* e, double x_over_t, double* vstar, double* pstar) {
  const double sn = e[1], cs = e[0];
  const double L1 = vl[0] * cs + vl[1] * sn, L2 = -vl[0] * sn + vl[1] * cs;
  const double R1 = vr[0] * cs + vr[1] * sn, R2 = -vr[0] * sn + vr[1] * cs;
  const double Us = (L1 * cL * rhoL + R1 * cR * rhoR + PL - PR) / (rhoL * cL + rhoR * cR);
  const double Ps = (PL * cL * rhoL + PR * cR * rhoR - cL * rhoL * cR * rhoR * (R1 - L1)) /
                    (rhoL * cL + rhoR * cR);
  const double ex7 = (double)0.142857142f;
  const double RsL = rhoL * pow(Ps / PL, ex7);
  const double RsR = rhoR * pow(Ps / PR, ex7);
  double QL = 1.0, QR = 1.0, t;
  if (!(Ps <= PL)) {
    t = ((rhoL / RsL) * (PL - Ps)) / ((rhoL - RsL) * (cL * cL));
    QL = t >= 0.0 ? sqrt(t) : 1.0;
  }
  const double SL = L1 - cL * QL;
  if (!(Ps <= PR)) {
    t = ((rhoR / RsR) * (PR - Ps)) / ((rhoR - RsR) * (cR * cR));
    QR = t >= 0.0 ? sqrt(t) : 1.0;
  }
  const double SR = R1 + cR * QR;
  const double S = Us;
  double H1, H2;
  if (SL > x_over_t) {
    H1 = L1; H2 = L2;
  } else if (S >= x_over_t && x_over_t >= SL) {
    const double q = rhoL * ((SL - L1) / (SL - S));
    H1 = (S * q) / q; H2 = (L2 * q) / q;
  } else if (SR >= x_over_t && x_over_t >= S) {
    const double q = rhoR * ((SR - R1) / (SR - S));
    H1 = (S * q) / q; H2 = (R2 * q) / q;
  } else {
    H1 = R1; H2 = R2;
  }
  vstar[0] = H1 * cs - H2 * sn;
  vstar[1] = H1 * sn + H2 * cs;
  *pstar = Ps;
}

/* ------------------------------------------------------------------ */
/* Solver_Normalized_Modified_Linear_Riemann_Delta_SPS.f90:126-635    */
/* (and Solver_Normalized_Linear_Riemann_Delta_SPS.f90:362-371)       */
/* ------------------------------------------------------------------ */
static void solver_riemann_delta_sps(orc_ctx* c) {
  const orc_params* P = &c->P;
  int n = c->n_local;
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :126-139 */
    int i = c->pi[k], j = c->pj[k];
    double* Ti = &c->Tinv[i * 4];
    double* Tj = &c->Tinv[j * 4];
    double d1 = c->pdwx[k], d2 = c->pdwy[k];
    c->ndix[k] = Ti[0] * d1 + Ti[2] * d2;
    c->ndiy[k] = Ti[1] * d1 + Ti[3] * d2;
    c->ndjx[k] = Tj[0] * d1 + Tj[2] * d2;
    c->ndjy[k] = Tj[1] * d1 + Tj[3] * d2;
  }
  memset(c->Sxx, 0, sizeof(double) * n); memset(c->Sxy, 0, sizeof(double) * n);
  memset(c->Syy, 0, sizeof(double) * n);
  memset(c->txx, 0, sizeof(double) * n); memset(c->txy, 0, sizeof(double) * n);
  memset(c->tyy, 0, sizeof(double) * n);
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :144-182 */
    int i = c->pi[k], j = c->pj[k];
    double vd1 = -c->pvx[k], vd2 = -c->pvy[k];
    c->Sxx[i] = c->Sxx[i] + vd1 * c->ndix[k] * c->vol[j];
    c->Sxy[i] = c->Sxy[i] + (vd1 * c->ndiy[k] + vd2 * c->ndix[k]) * c->vol[j];
    c->Syy[i] = c->Syy[i] + vd2 * c->ndiy[k] * c->vol[j];
    c->Sxx[j] = c->Sxx[j] + vd1 * c->ndjx[k] * c->vol[i];
    c->Sxy[j] = c->Sxy[j] + (vd1 * c->ndjy[k] + vd2 * c->ndjx[k]) * c->vol[i];
    c->Syy[j] = c->Syy[j] + vd2 * c->ndjy[k] * c->vol[i];
  }
  const double dL = (double)(sqrtf(2.0f) / 2.0f) * c->dx;        /* :266 */
  const double c012 = (double)0.12f;
  const double c23 = (double)(2.0f / 3.0f);
  const double c23b = (double)((2.0f / 3.0f) * 0.0066f);
  for (int i = 0; i < n; ++i) {                                   /* :256-296 */
    if (c->type[i] != 2) continue;
    if (!P->sps_on) continue;          /* Solver_..._Delta.f90 has no SPS stress: tau stays 0 */
    double master = c->Sxx[i] * c->Sxx[i] + c->Syy[i] * c->Syy[i] + 0.0;
    double last = c->Sxy[i] * c->Sxy[i] + 0.0 + 0.0;
    double abs2 = 2 * master + last;
    double t0 = c012 * dL;
    double nut = (t0 * t0) * sqrt(abs2);
    double kk = c23 * nut * (c->Sxx[i] + c->Syy[i]);
    double blin = c23b * (dL * dL) * abs2;
    c->txx[i] = (2 * nut * c->Sxx[i] - kk - blin) * c->rho[i];
    c->txy[i] = (nut * c->Sxy[i]) * c->rho[i];
    c->tyy[i] = (2 * nut * c->Syy[i] - kk - blin) * c->rho[i];
  }
  if (P->solver == ORC_SOLVER_MUSCL) {
    /* MUSCL_..._SPS.f90:145-185: grad p, grad u, grad v with the renormalised
     * dwdx (identity inverse), summed like grad rho */
    memset(c->gpx, 0, sizeof(double) * n); memset(c->gpy, 0, sizeof(double) * n);
    memset(c->gux, 0, sizeof(double) * n); memset(c->guy, 0, sizeof(double) * n);
    memset(c->gvx, 0, sizeof(double) * n); memset(c->gvy, 0, sizeof(double) * n);
    for (int64_t k = 0; k < c->npair; ++k) {
      int i = c->pi[k], j = c->pj[k];
      double* Ti = &c->Tinv[i * 4];
      double* Tj = &c->Tinv[j * 4];
      double d1 = c->pdwx[k], d2 = c->pdwy[k];
      double mi1 = Ti[0] * d1 + Ti[2] * d2, mi2 = Ti[1] * d1 + Ti[3] * d2;
      double mj1 = Tj[0] * d1 + Tj[2] * d2, mj2 = Tj[1] * d1 + Tj[3] * d2;
      double pd = c->p[j] - c->p[i];
      double ud = -c->pvx[k], vd = -c->pvy[k];
      c->gpx[i] = c->gpx[i] + pd * c->vol[j] * mi1; c->gpy[i] = c->gpy[i] + pd * c->vol[j] * mi2;
      c->gpx[j] = c->gpx[j] + pd * c->vol[i] * mj1; c->gpy[j] = c->gpy[j] + pd * c->vol[i] * mj2;
      c->gux[i] = c->gux[i] + ud * c->vol[j] * mi1; c->guy[i] = c->guy[i] + ud * c->vol[j] * mi2;
      c->gux[j] = c->gux[j] + ud * c->vol[i] * mj1; c->guy[j] = c->guy[j] + ud * c->vol[i] * mj2;
      c->gvx[i] = c->gvx[i] + vd * c->vol[j] * mi1; c->gvy[i] = c->gvy[i] + vd * c->vol[j] * mi2;
      c->gvx[j] = c->gvx[j] + vd * c->vol[i] * mj1; c->gvy[j] = c->gvy[j] + vd * c->vol[i] * mj2;
    }
  }
  const double sqc0 = sqrt(P->c0sq);
  const double x_over_t = c->dx / P->dt;                         /* HLLC :444 interior_dx/dt */
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :300-502 */
    int i = c->pi[k], j = c->pj[k];
    double PL = c->p[i], PR = c->p[j];
    double Pb = (PL + PR) / 2.0;
    double Iv[2] = {c->u[i], c->v[i]}, Jv[2] = {c->u[j], c->v[j]};
    double vb[2] = {(Iv[0] + Jv[0]) / 2.0, (Iv[1] + Jv[1]) / 2.0};
    double xij[2] = {c->pdx[k], c->pdy[k]};
    double dist = sqrt(xij[0] * xij[0] + xij[1] * xij[1]);
    double ev[2] = {-(xij[0] / dist), -(xij[1] / dist)};
    if (P->solver == ORC_SOLVER_MUSCL) {                         /* MUSCL :464-504 */
      double xr[2] = {-xij[0], -xij[1]};                         /* x_j - x_i */
      double li, lj, dpi, dpj;
      dpi = 0.5 * (c->gpx[i] * xr[0] + c->gpy[i] * xr[1]);
      dpj = -0.5 * (c->gpx[j] * xr[0] + c->gpy[j] * xr[1]);
      orc_muscl_limiter(dpi, dpj, c->p[j] - c->p[i], &li, &lj);
      PL = c->p[i] + li;
      PR = c->p[j] - lj;
      dpi = 0.5 * (c->gux[i] * xr[0] + c->guy[i] * xr[1]);
      dpj = -0.5 * (c->gux[j] * xr[0] + c->guy[j] * xr[1]);
      orc_muscl_limiter(dpi, dpj, -c->pvx[k], &li, &lj);
      Iv[0] = c->u[i] + li;
      Jv[0] = c->u[j] - lj;
      dpi = 0.5 * (c->gvx[i] * xr[0] + c->gvy[i] * xr[1]);
      dpj = -0.5 * (c->gvx[j] * xr[0] + c->gvy[j] * xr[1]);
      orc_muscl_limiter(dpi, dpj, -c->pvy[k], &li, &lj);
      Iv[1] = c->v[i] + li;
      Jv[1] = c->v[j] - lj;
    }
    double UL = Iv[0] * ev[0] + Iv[1] * ev[1];
    double UR = Jv[0] * ev[0] + Jv[1] * ev[1];
    double Ub = (UL + UR) / 2.0;
    double cb = c->pc_[k], rb = c->prho_[k];
    double Us, vs[2], Ps;
    if (P->solver == ORC_SOLVER_HLLC) {
      orc_hllc_star(Iv, PL, c->rho[i], c->c[i], Jv, PR, c->rho[j], c->c[j], ev, x_over_t, vs, &Ps);
    } else if (P->solver == ORC_SOLVER_LINEAR_RIEMANN) {
      double zi = c->rho[i] * c->c[i], zj = c->rho[j] * c->c[j];
      Us = zi * UR + zj * UL + PL - PR;
      Us = Us / (zi + zj);
      vs[0] = Us * ev[0]; vs[1] = Us * ev[1];
      Ps = zi * PR + zj * PL + c->rho[i] * c->c[i] * c->rho[j] * c->c[j] * (UL - UR);
      Ps = Ps / (zi + zj);
    } else {
      Us = Ub + 0.5 * (PL - PR) / (rb * sqc0);
      vs[0] = (Us - Ub) * ev[0] + vb[0];
      vs[1] = (Us - Ub) * ev[1] + vb[1];
      double d = UL - UR;
      double bR = 3.0 * (d > 0.0 ? d : 0.0);
      bR = bR < cb ? bR : cb;
      Ps = Pb + 0.5 * bR * rb * (UL - UR);
    }
    double di[2] = {c->ndix[k], c->ndiy[k]}, dj[2] = {c->ndjx[k], c->ndjy[k]};
    double Ivs[2], Jvs[2], IPs, JPs;
    if (c->near[i] == 1) { Ivs[0] = vb[0]; Ivs[1] = vb[1]; IPs = Pb; }
    else { Ivs[0] = vs[0]; Ivs[1] = vs[1]; IPs = Ps; }
    if (c->near[j] == 1) { Jvs[0] = vb[0]; Jvs[1] = vb[1]; JPs = Pb; }
    else { Jvs[0] = vs[0]; Jvs[1] = vs[1]; JPs = Ps; }
    double ti = (Iv[0] - Ivs[0]) * di[0] + (Iv[1] - Ivs[1]) * di[1];
    double tj = -((Jv[0] - Jvs[0]) * dj[0] + (Jv[1] - Jvs[1]) * dj[1]);
    c->drhodt[i] = c->drhodt[i] + 2 * c->rho[i] * c->vol[j] * ti;
    c->drhodt[j] = c->drhodt[j] + 2 * c->rho[j] * c->vol[i] * tj;
    {
      /* v* is built from v_i, v_j and (P_L - P_R)/(rho c): its rounding scale */
      const double vsc = fabs(Iv[0]) + fabs(Iv[1]) + fabs(Jv[0]) + fabs(Jv[1]) +
                         (fabs(PL) + fabs(PR)) / (rb * sqc0);
      const double ai = (fabs(di[0]) + fabs(di[1])) * vsc, aj = (fabs(dj[0]) + fabs(dj[1])) * vsc;
      c->mag_drho[i] += 2 * c->rho[i] * c->vol[j] * ai;
      c->mag_drho[j] += 2 * c->rho[j] * c->vol[i] * aj;
    }
    double rij = 1.0 / (c->rho[i] * c->rho[j]);
    double* iax[2] = {&c->ia_x[i], &c->ia_y[i]};
    double* jax[2] = {&c->ia_x[j], &c->ia_y[j]};
    double* imag[2] = {&c->mag_ax[i], &c->mag_ay[i]};
    double* jmag[2] = {&c->mag_ax[j], &c->mag_ay[j]};
    /* P* = mean pressure plus a dissipation ~ rho c (U_L - U_R): its rounding scale */
    const double psc = fabs(PL) + fabs(PR) + rb * cb * (fabs(UL) + fabs(UR));
    for (int L = 0; L < 2; ++L) {
      double tpi = -2 * IPs * di[L] * rij;
      double tpj = -2 * JPs * dj[L] * rij;
      *iax[L] = *iax[L] + c->mass[j] * tpi;
      *jax[L] = *jax[L] - c->mass[i] * tpj;
      *imag[L] += c->mass[j] * 2 * psc * fabs(di[L]) * rij;
      *jmag[L] += c->mass[i] * 2 * psc * fabs(dj[L]) * rij;
    }
    double hb = c->ph_[k];
    double vij[2] = {c->pvx[k], c->pvy[k]};
    double r2 = xij[0] * xij[0] + xij[1] * xij[1];
    double vx = vij[0] * xij[0] + vij[1] * xij[1];
    if (vx < 0.0) {                                              /* :392-417 */
      double fhb = P->fai * hb;
      double phi = hb * vx / (r2 + fhb * fhb);
      double Pi = (-(P->alpha * cb * phi) + P->beta * (phi * phi)) / rb;
      double* iav[2] = {&c->av_x[i], &c->av_y[i]};
      double* jav[2] = {&c->av_x[j], &c->av_y[j]};
      for (int L = 0; L < 2; ++L) {
        double tvi = -Pi * di[L];
        double tvj = -Pi * dj[L];
        *iav[L] = *iav[L] + c->mass[j] * tvi;
        *jav[L] = *jav[L] - c->mass[i] * tvj;
        *imag[L] += fabs(c->mass[j] * tvi);
        *jmag[L] += fabs(c->mass[i] * tvj);
      }
    }
    double Ri = xij[0] * di[0] + xij[1] * di[1];                  /* :432-436 */
    c->lap_x[i] = c->lap_x[i] + c->vol[j] * 2 * P->nu * Ri / r2 * vij[0];
    c->lap_y[i] = c->lap_y[i] + c->vol[j] * 2 * P->nu * Ri / r2 * vij[1];
    double Rj = xij[0] * dj[0] + xij[1] * dj[1];
    c->lap_x[j] = c->lap_x[j] - c->vol[i] * 2 * P->nu * Rj / r2 * vij[0];
    c->lap_y[j] = c->lap_y[j] - c->vol[i] * 2 * P->nu * Rj / r2 * vij[1];
    c->mag_ax[i] += fabs(c->vol[j] * 2 * P->nu * Ri / r2 * vij[0]);
    c->mag_ay[i] += fabs(c->vol[j] * 2 * P->nu * Ri / r2 * vij[1]);
    c->mag_ax[j] += fabs(c->vol[i] * 2 * P->nu * Rj / r2 * vij[0]);
    c->mag_ay[j] += fabs(c->vol[i] * 2 * P->nu * Rj / r2 * vij[1]);
    double tmp = P->delta_alpha * hb * cb * vx / r2;              /* :442-446 */
    double* iai[2] = {&c->ai_x[i], &c->ai_y[i]};
    double* jai[2] = {&c->ai_x[j], &c->ai_y[j]};
    for (int L = 0; L < 2; ++L) {
      *iai[L] = *iai[L] + tmp * di[L] * P->rho0 / c->rho[i] * c->vol[j];
      *jai[L] = *jai[L] - tmp * dj[L] * P->rho0 / c->rho[j] * c->vol[i];
      *imag[L] += fabs(tmp * di[L] * P->rho0 / c->rho[i] * c->vol[j]);
      *jmag[L] += fabs(tmp * dj[L] * P->rho0 / c->rho[j] * c->vol[i]);
    }
    /* :450-492 SPS with raw dwdx and the GLOBAL particle_mass at LOCAL index */
    double d1 = c->pdwx[k], d2 = c->pdwy[k];
    double tp = 0.0;
    tp = tp + (c->txx[i] + c->txx[j]) * d1;
    tp = tp + (c->txy[i] + c->txy[j]) * d2;
    tp = tp * rij;
    c->sps_x[i] = c->sps_x[i] + c->gmass[j] * tp;
    c->sps_x[j] = c->sps_x[j] - c->gmass[i] * tp;
    c->mag_ax[i] += fabs(c->gmass[j] * tp);
    c->mag_ax[j] += fabs(c->gmass[i] * tp);
    tp = tp + (c->tyy[i] + c->tyy[j]) * d2 + (c->txy[i] + c->txy[j]) * d1;
    tp = tp * rij;
    c->sps_y[i] = c->sps_y[i] + c->gmass[j] * tp;
    c->sps_y[j] = c->sps_y[j] - c->gmass[i] * tp;
    c->mag_ay[i] += fabs(c->gmass[j] * tp);
    c->mag_ay[j] += fabs(c->gmass[i] * tp);
    if (P->solver == ORC_SOLVER_MODIFIED_LINEAR_RIEMANN) {       /* :495-500 XSPH */
      /* 2 epsilon in the _SPS solver; epsilon in Solver_..._Delta.f90:139-140 */
      const double xf = P->sps_on ? 2.0 : 1.0;
      double rs = c->rho[i] + c->rho[j];
      c->avg_x[i] = c->avg_x[i] - P->epsilon_xsph * xf * c->mass[j] * vij[0] / rs * c->pw[k];
      c->avg_y[i] = c->avg_y[i] - P->epsilon_xsph * xf * c->mass[j] * vij[1] / rs * c->pw[k];
      c->avg_x[j] = c->avg_x[j] + P->epsilon_xsph * xf * c->mass[i] * vij[0] / rs * c->pw[k];
      c->avg_y[j] = c->avg_y[j] + P->epsilon_xsph * xf * c->mass[i] * vij[1] / rs * c->pw[k];
    }
  }
  memset(c->grx, 0, sizeof(double) * n); memset(c->gry, 0, sizeof(double) * n);
  memset(c->mag_grho, 0, sizeof(double) * n);
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :508-526 */
    int i = c->pi[k], j = c->pj[k];
    double rd = -c->prd[k];
    double* Ti = &c->Tinv[i * 4];
    double* Tj = &c->Tinv[j * 4];
    double d1 = c->pdwx[k], d2 = c->pdwy[k];
    double mi1 = Ti[0] * d1 + Ti[2] * d2, mi2 = Ti[1] * d1 + Ti[3] * d2;
    double mj1 = Tj[0] * d1 + Tj[2] * d2, mj2 = Tj[1] * d1 + Tj[3] * d2;
    c->grx[i] = c->grx[i] + rd * c->vol[j] * mi1;
    c->gry[i] = c->gry[i] + rd * c->vol[j] * mi2;
    c->grx[j] = c->grx[j] + rd * c->vol[i] * mj1;
    c->gry[j] = c->gry[j] + rd * c->vol[i] * mj2;
    c->mag_grho[i] += fabs(rd) * c->vol[j] * (fabs(mi1) + fabs(mi2));
    c->mag_grho[j] += fabs(rd) * c->vol[i] * (fabs(mj1) + fabs(mj2));
  }
  for (int64_t k = 0; k < c->npair; ++k) {                       /* :552-583 */
    int i = c->pi[k], j = c->pj[k];
    double rd = -c->prd[k];
    double hb = c->ph_[k], cb = c->pc_[k];
    double pd[2] = {-c->pdx[k], -c->pdy[k]};
    double r2 = pd[0] * pd[0] + pd[1] * pd[1];
    double sg[2] = {c->grx[i] + c->grx[j], c->gry[i] + c->gry[j]};
    double psi = rd - 0.5 * (sg[0] * pd[0] + sg[1] * pd[1]);
    double ti = (pd[0] * c->ndix[k] + pd[1] * c->ndiy[k]) / r2;
    double tj = (pd[0] * c->ndjx[k] + pd[1] * c->ndjy[k]) / r2;
    c->corr[i] = c->corr[i] + 2 * P->delta * hb * cb * psi * ti * c->vol[j];
    c->corr[j] = c->corr[j] - 2 * P->delta * hb * cb * psi * tj * c->vol[i];
    {
      /* psi cancels to second order: its scale is |rho_j - rho_i| plus the grad rho
       * products and their own rounding scale */
      const double psc = fabs(rd) + 0.5 * (fabs(sg[0] * pd[0]) + fabs(sg[1] * pd[1])) +
                         0.5 * (c->mag_grho[i] + c->mag_grho[j]) * (fabs(pd[0]) + fabs(pd[1]));
      c->mag_drho[i] += fabs(2 * P->delta * hb * cb * psc * ti * c->vol[j]);
      c->mag_drho[j] += fabs(2 * P->delta * hb * cb * psc * tj * c->vol[i]);
    }
  }
  for (int i = 0; i < n; ++i) {                                   /* :627-635 */
    if (c->type[i] != 2) continue;
    c->drhodt[i] = c->drhodt[i] + c->corr[i];
  }
}

/* ------------------------------------------------------------------ */
/* single_step_compute.f90:62-171                                      */
/* ------------------------------------------------------------------ */
/* single_step_compute.f90:174-260: fluid particles with Lx - x <= L and x <= Lx  */
/* get their acceleration, XSPH average velocity and subdomain velocity scaled:   */
/* by 1 - 0.5 A (1 - cos(PI (x + L - Lx) / L)) (Cosin_Type, :181-201), by          */
/* (exp(((Lx - x) / L)**3.5) - 1) / (exp(1.0) - 1) (y / depth) (OpenFoam_Type,     */
/* exp(1.0) a REAL(4) intrinsic, :203-228) or by                                  */
/* 1 - 6 ((x + L - Lx) / L)**2 ((depth - y) / depth) (DualSPH_Type, :230-255);     */
/* damping 4 is the linear ramp 1 - A (x + L - Lx) / L, applied like Cosin_Type     */
/* (a damping form of this library with no reference counterpart).                */
static void end_damping_zone(orc_ctx* c) {
  const orc_params* P = &c->P;
  const double Lx = P->box_x, L = P->damping_length, d = P->water_depth;
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != 2) continue;
    const double x = c->x[i], y = c->y[i];
    if (!((Lx - x) <= L && x <= Lx)) continue;
    double f;
    if (P->damping == 1) {
      const double coef = 0.5 * P->damping_amplitude * (1 - cos(PI_MODULE * (x + L - Lx) / L));
      f = 1 - coef;
    } else if (P->damping == 2) {
      f = (exp(pow((Lx - x) / L, 3.5)) - 1.0) / ((double)expf(1.0f) - 1.0) * (y / d);
    } else if (P->damping == 4) {
      f = 1 - P->damping_amplitude * ((x + L - Lx) / L);
    } else {
      const double q = (x + L - Lx) / L;
      f = 1 - 6 * (q * q) * ((d - y) / d);
    }
    c->ax[i] = c->ax[i] * f; c->ay[i] = c->ay[i] * f;
    c->mag_ax[i] = c->mag_ax[i] * fabs(f); c->mag_ay[i] = c->mag_ay[i] * fabs(f);
    c->avg_x[i] = c->avg_x[i] * f; c->avg_y[i] = c->avg_y[i] * f;
    c->u[i] = c->u[i] * f; c->v[i] = c->v[i] * f;
  }
}

void orc_single_step_compute(orc_ctx* c, int32_t i_time_step) {
  (void)i_time_step;
  int n = c->n_local;
  for (int i = 0; i < n; ++i) {                                   /* :62-73 EOS */
    if (c->type[i] == 120) continue;
    orc_status_equation(c, c->rho[i], c->e[i], c->type[i], &c->p[i], &c->c[i]);
  }
  memset(c->vol, 0, sizeof(double) * n);
  assign_boundary_particles(c);                                  /* :80 */
  find_particle_pair(c);                                         /* :93 */
  define_freesurface(c);                                         /* :108 */
  size_t nb = sizeof(double) * n;                                /* :113-126 */
  memset(c->drhodt, 0, nb); memset(c->dedt, 0, nb);
  memset(c->ia_x, 0, nb); memset(c->ia_y, 0, nb);
  memset(c->ai_x, 0, nb); memset(c->ai_y, 0, nb);
  memset(c->av_x, 0, nb); memset(c->av_y, 0, nb);
  memset(c->ext_x, 0, nb); memset(c->ext_y, 0, nb);
  memset(c->avg_x, 0, nb); memset(c->avg_y, 0, nb);
  memset(c->corr, 0, nb);
  memset(c->mag_drho, 0, nb); memset(c->mag_ax, 0, nb); memset(c->mag_ay, 0, nb);
  memset(c->lap_x, 0, nb); memset(c->lap_y, 0, nb);
  memset(c->sps_x, 0, nb); memset(c->sps_y, 0, nb);
  solver_riemann_delta_sps(c);                                   /* :137 */
  memset(c->ax, 0, nb); memset(c->ay, 0, nb);                    /* :152-171 */
  for (int i = 0; i < n; ++i) {
    if (c->type[i] != 2) continue;
    if (c->P.gravity != 0.0) c->ext_y[i] = c->ext_y[i] - c->P.gravity;   /* :161-163 */
    c->mag_ay[i] += fabs(c->ext_y[i]);
    c->ax[i] = c->ia_x[i] + c->ext_x[i] + c->av_x[i] + c->ai_x[i] + c->sps_x[i] + c->lap_x[i];
    c->ay[i] = c->ia_y[i] + c->ext_y[i] + c->av_y[i] + c->ai_y[i] + c->sps_y[i] + c->lap_y[i];
    c->dedt[i] = c->dedt[i] + 0.0 + 0.0;
  }
  if (c->P.damping != 0) end_damping_zone(c);                     /* :174-260 */
}

/* ------------------------------------------------------------------ */
/* Reinitialize_particle_density.f90:66-182 +                          */
/* MLS_Interpolation_In_Subdomain_Domain.f90:68-297 (orders 1 and 6)   */
/* ------------------------------------------------------------------ */
static void mls6_fai(const orc_ctx* c, double xi, double yi, int ne, const int* nm,
                     const double* wk, double* fai) {
  const double al = c->P.mls_alpha;
  const double h = c->h;
  /* P (6 x (ne+6)), B = P*W, A = B*P^T with k ascending */
  int nc = ne + 6;
  double* Pm = xcalloc((size_t)6 * nc, sizeof(double));
  double* Bm = xcalloc((size_t)6 * nc, sizeof(double));
  for (int k = 0; k < ne; ++k) {
    int j = nm[k];
    double a = (c->x[j] - xi) / h, b = (c->y[j] - yi) / h;
    Pm[0 + 6 * k] = 1;
    Pm[1 + 6 * k] = a;
    Pm[2 + 6 * k] = b;
    Pm[3 + 6 * k] = a * a;
    Pm[4 + 6 * k] = (c->x[j] - xi) / h * (c->y[j] - yi) / h;
    Pm[5 + 6 * k] = b * b;
  }
  for (int L = 0; L < 6; ++L) Pm[L + 6 * (ne + L)] = al;
  for (int k = 0; k < nc; ++k) {
    double wkk = k < ne ? wk[k] : al;
    for (int m = 0; m < 6; ++m) Bm[m + 6 * k] = Pm[m + 6 * k] * wkk;
  }
  double A[36] = {0}, Ai[36];
  for (int l = 0; l < 6; ++l)
    for (int m = 0; m < 6; ++m) {
      double s = 0.0;
      for (int k = 0; k < nc; ++k) s = s + Bm[m + 6 * k] * Pm[l + 6 * k];
      A[m + 6 * l] = s;
    }
  memcpy(Ai, A, sizeof(A));
  int ok = orc_brinv(Ai, 6);
  if (ok) {
    for (int k = 0; k < ne; ++k) {
      double s = 0.0;
      for (int l = 0; l < 6; ++l) s = s + Ai[0 + 6 * l] * Bm[l + 6 * k];
      fai[k] = s;
    }
  } else {
    double sw = 0.0;
    for (int k = 0; k < ne; ++k) sw = sw + wk[k];
    for (int k = 0; k < ne; ++k) fai[k] = wk[k] / sw;
  }
  if (ne <= 8) {
    double sw = 0.0;
    for (int k = 0; k < ne; ++k) sw = sw + wk[k];
    for (int k = 0; k < ne; ++k) fai[k] = wk[k] / sw;
  }
  free(Pm); free(Bm);
}

void orc_reinitialize_density(orc_ctx* c) {
  if (c->P.ms) { ms_reinitialize_density(c); return; }
  int n = c->n_local;
  memcpy(c->mls_rho, c->rho, sizeof(double) * n);
  memset(c->rho, 0, sizeof(double) * n);
  if (c->P.mls_order == 1) {
    for (int64_t k = 0; k < c->npair; ++k) {
      int i = c->pi[k], j = c->pj[k];
      if (c->type[i] == c->type[j]) {
        if (c->type[i] == 2) c->rho[i] = c->rho[i] + c->mass[j] * c->pw[k] / c->sum_w[i];
        if (c->type[j] == 2) c->rho[j] = c->rho[j] + c->mass[i] * c->pw[k] / c->sum_w[j];
      }
    }
    return;
  }
  int nm[300]; double wk[300], fai[300];
  for (int i = 0; i < n; ++i) {
    if (c->type[i] != 2) continue;
    c->rho[i] = 0.0;
    int ne = 0;
    for (int L = 0; L < c->cand_n[i]; ++L) {
      int j = c->cand[(size_t)i * 300 + L];
      double pd[2] = {c->x[i] - c->x[j], c->y[i] - c->y[j]};
      double r = sqrt(pd[0] * pd[0] + pd[1] * pd[1]);
      double hb = (c->hh[i] + c->hh[j]) / 2.0;
      if (r <= (3 * hb) && r >= (double)1.0e-7f) {
        double w, dw[2];
        orc_compute_kernel(2, r, pd, hb, &w, dw);
        nm[ne] = j; wk[ne] = w; ++ne;
      }
    }
    mls6_fai(c, c->x[i], c->y[i], ne, nm, wk, fai);
    for (int L = 0; L < ne; ++L) c->rho[i] = c->rho[i] + c->mls_rho[nm[L]] * fai[L];
  }
}

/* ------------------------------------------------------------------ */
/* Shift_particle_position.f90:56-136 (pairs of the last evaluation)   */
/* ------------------------------------------------------------------ */
void orc_shift_sums(orc_ctx* c) {
  const orc_params* P = &c->P;
  int n = c->n_local;
  double pd0[2] = {0.0, 0.0}, wdx, dwd[2];
  orc_compute_kernel(2, c->dx, pd0, c->h, &wdx, dwd);
  memset(c->shx, 0, sizeof(double) * n); memset(c->shy, 0, sizeof(double) * n);
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    double f = P->shift_beta * (c->h * c->h) * (1 + P->shift_R * ipow4(c->pw[k] / wdx));
    c->shx[i] = c->shx[i] - f * c->ndix[k] * c->vol[j];
    c->shx[j] = c->shx[j] + f * c->ndjx[k] * c->vol[i];
    c->shy[i] = c->shy[i] - f * c->ndiy[k] * c->vol[j];
    c->shy[j] = c->shy[j] + f * c->ndjy[k] * c->vol[i];
  }
}

void orc_shift(orc_ctx* c) {
  const orc_params* P = &c->P;
  int n = c->n_local;
  orc_shift_sums(c);
  for (int i = 0; i < n; ++i) {
    if (c->type[i] != 2) continue;
    if (c->near[i] == 1) continue;
    double a = sqrt(c->shx[i] * c->shx[i] + c->shy[i] * c->shy[i]);
    if (a > c->max_shift) {
      c->shx[i] = c->shx[i] / a * c->max_shift;
      c->shy[i] = c->shy[i] / a * c->max_shift;
    }
    c->x[i] = c->x[i] + c->shx[i];
    c->y[i] = c->y[i] + c->shy[i];
    if (P->ms) {                          /* MS Shift_particle_position.f90:117: DOT_PRODUCT */
      c->rho[i] = c->rho[i] + (c->grx[i] * c->shx[i] + c->gry[i] * c->shy[i]);
    } else {
      c->rho[i] = c->rho[i] + c->grx[i] * c->shx[i];
      c->rho[i] = c->rho[i] + c->gry[i] * c->shy[i];
    }
    c->p[i] = (P->c0sq * P->rho0 / 7.0) * (ipow7(c->rho[i] / P->rho0) - 1.0);
  }
}

/* Refresh_domain.f90:320-335: vorticity from the last evaluation's pair cache,
 * vort_i += V_j (v_i - v_j) x Normalized_dwdx_i, vort_j += V_i (v_i - v_j) x
 * Normalized_dwdx_j; Cross_Product(...)(3) = X1*Y2 - X2*Y1 (function_module.f90:150). */
static void refresh_vorticity(orc_ctx* c) {
  memset(c->vort, 0, sizeof(double) * c->n_local);
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    double vx = c->pvx[k], vy = c->pvy[k];
    c->vort[i] = c->vort[i] + c->vol[j] * (vx * c->ndiy[k] - vy * c->ndix[k]);
    c->vort[j] = c->vort[j] + c->vol[i] * (vx * c->ndjy[k] - vy * c->ndjx[k]);
  }
}

/* Refresh_domain.f90:136-300 on one rank: global arrays zeroed, then filled. */
void orc_refresh(orc_ctx* c) {
  refresh_vorticity(c);
  {
    int NG0 = c->n_fluid + c->n_ghost;
    memset(c->gvort, 0, sizeof(double) * NG0); memset(c->gnx, 0, sizeof(double) * NG0);
    memset(c->gny, 0, sizeof(double) * NG0); memset(c->gsurf, 0, sizeof(int) * NG0);
    for (int k = 0; k < c->n_local; ++k) {          /* :300-351 particle_vorticity, Normal_vector */
      int g = c->name[k];
      c->gvort[g] = c->vort[k]; c->gnx[g] = c->nrmx[k]; c->gny[g] = c->nrmy[k]; c->gsurf[g] = c->surf[k];
    }
  }
  int NG = c->n_fluid + c->n_ghost;
  size_t nb = sizeof(double) * NG;
  memset(c->gx, 0, nb); memset(c->gy, 0, nb); memset(c->gu, 0, nb); memset(c->gv, 0, nb);
  memset(c->gp, 0, nb); memset(c->gmass, 0, nb); memset(c->grho, 0, nb); memset(c->gh, 0, nb);
  memset(c->gc, 0, nb); memset(c->ge, 0, nb); memset(c->gtype, 0, sizeof(int) * NG);
  for (int k = 0; k < c->n_local; ++k) {
    int g = c->name[k];
    c->gx[g] = c->x[k]; c->gy[g] = c->y[k]; c->gu[g] = c->u[k]; c->gv[g] = c->v[k];
    c->gp[g] = c->p[k]; c->gmass[g] = c->mass[k]; c->grho[g] = c->rho[k];
    c->gh[g] = c->hh[k]; c->gc[g] = c->c[k]; c->ge[g] = c->e[k]; c->gtype[g] = c->type[k];
  }
}

/* ------------------------------------------------------------------ */
/* time_iteration.f90: PC (:509-657) and RK-variant (:779-1049) steps  */
/* ------------------------------------------------------------------ */
static void save_before(orc_ctx* c) {
  size_t nb = sizeof(double) * c->n_local;
  memcpy(c->b_rho, c->rho, nb); memcpy(c->b_u, c->u, nb); memcpy(c->b_v, c->v, nb);
  memcpy(c->b_x, c->x, nb); memcpy(c->b_y, c->y, nb); memcpy(c->b_e, c->e, nb);
}

/* stage with rates r at fraction f of dt: the RK-variant writes drhodt*dt/2 etc. */
static void stage_half(orc_ctx* c) {
  double dt = c->P.dt;
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != 2) continue;
    c->rho[i] = c->b_rho[i] + c->drhodt[i] * dt / 2;
    c->e[i] = c->b_e[i] + c->dedt[i] * dt / 2;
    c->u[i] = c->b_u[i] + 0.5 * dt * c->ax[i] + c->avg_x[i];
    c->x[i] = c->b_x[i] + 0.5 * dt * c->u[i];
    c->v[i] = c->b_v[i] + 0.5 * dt * c->ay[i] + c->avg_y[i];
    c->y[i] = c->b_y[i] + 0.5 * dt * c->v[i];
  }
}
static void stage_full(orc_ctx* c) {
  double dt = c->P.dt;
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != 2) continue;
    c->rho[i] = c->b_rho[i] + c->drhodt[i] * dt;
    c->e[i] = c->b_e[i] + c->dedt[i] * dt;
    c->u[i] = c->b_u[i] + dt * c->ax[i] + c->avg_x[i];
    c->x[i] = c->b_x[i] + dt * c->u[i];
    c->v[i] = c->b_v[i] + dt * c->ay[i] + c->avg_y[i];
    c->y[i] = c->b_y[i] + dt * c->v[i];
  }
}
static void save_k(orc_ctx* c, int s) {
  size_t nb = sizeof(double) * c->n_local;
  memcpy(c->k_u[s], c->u, nb); memcpy(c->k_v[s], c->v, nb);
  memcpy(c->k_drho[s], c->drhodt, nb); memcpy(c->k_de[s], c->dedt, nb);
  memcpy(c->k_ax[s], c->ax, nb); memcpy(c->k_ay[s], c->ay, nb);
}
static void rk_final(orc_ctx* c) {
  double dt = c->P.dt;
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != 2) continue;
    c->rho[i] = c->b_rho[i] + (c->k_drho[0][i] + 2 * c->k_drho[1][i] + 2 * c->k_drho[2][i] + c->drhodt[i]) * dt / 6.0;
    c->e[i] = c->b_e[i] + (c->k_de[0][i] + 2 * c->k_de[1][i] + 2 * c->k_de[2][i] + c->dedt[i]) * dt / 6.0;
    c->u[i] = c->b_u[i] + (c->k_ax[0][i] + 2 * c->k_ax[1][i] + 2 * c->k_ax[2][i] + c->ax[i]) * dt / 6.0;
    c->x[i] = c->b_x[i] + (c->k_u[0][i] + 2 * c->k_u[1][i] + 2 * c->k_u[2][i] + c->u[i]) * dt / 6.0 + 0.5 * dt * dt * c->ax[i];
    c->v[i] = c->b_v[i] + (c->k_ay[0][i] + 2 * c->k_ay[1][i] + 2 * c->k_ay[2][i] + c->ay[i]) * dt / 6.0;
    c->y[i] = c->b_y[i] + (c->k_v[0][i] + 2 * c->k_v[1][i] + 2 * c->k_v[2][i] + c->v[i]) * dt / 6.0 + 0.5 * dt * dt * c->ay[i];
  }
}
static void pc_predict(orc_ctx* c) { stage_full(c); }
static void pc_save(orc_ctx* c) {
  size_t nb = sizeof(double) * c->n_local;
  memcpy(c->t_u, c->u, nb); memcpy(c->t_v, c->v, nb); memcpy(c->t_drho, c->drhodt, nb);
  memcpy(c->t_de, c->dedt, nb); memcpy(c->t_ax, c->ax, nb); memcpy(c->t_ay, c->ay, nb);
}
static void pc_correct(orc_ctx* c) {
  double dt = c->P.dt;
  for (int i = 0; i < c->n_local; ++i) {
    if (c->type[i] != 2) continue;
    c->rho[i] = c->b_rho[i] + 0.5 * (c->drhodt[i] + c->t_drho[i]) * dt;
    c->e[i] = c->b_e[i] + 0.5 * (c->dedt[i] + c->t_de[i]) * dt;
    c->u[i] = c->b_u[i] + 0.5 * dt * (c->ax[i] + c->t_ax[i]) + c->avg_x[i];
    c->x[i] = c->b_x[i] + 0.5 * dt * (c->u[i] + c->t_u[i]) + 0.5 * dt * dt * c->ax[i];
    c->v[i] = c->b_v[i] + 0.5 * dt * (c->ay[i] + c->t_ay[i]) + c->avg_y[i];
    c->y[i] = c->b_y[i] + 0.5 * dt * (c->v[i] + c->t_v[i]) + 0.5 * dt * dt * c->ay[i];
  }
}

/* Carry the previous step's rates with the particle (by gid) instead of by local
 * index -- a deliberate deviation from time_iteration.f90:542-555,942-957. */
static void remap_rates(orc_ctx* c, const int* old_name, int old_n) {
  int NG = c->n_fluid + c->n_ghost + 1;
  int cap = c->cap_global;
  int* inv = xcalloc(cap, sizeof(int));
  for (int i = 0; i < cap; ++i) inv[i] = -1;
  for (int k = 0; k < old_n; ++k) inv[old_name[k]] = k;
  double *d = xcalloc(c->n_local, 8), *ax = xcalloc(c->n_local, 8), *ay = xcalloc(c->n_local, 8);
  for (int k = 0; k < c->n_local; ++k) {
    int g = c->name[k];
    int o = (g < cap) ? inv[g] : -1;
    if (o >= 0 && g < c->n_fluid) { d[k] = c->drhodt[o]; ax[k] = c->ax[o]; ay[k] = c->ay[o]; }
  }
  memcpy(c->drhodt, d, 8 * c->n_local); memcpy(c->ax, ax, 8 * c->n_local); memcpy(c->ay, ay, 8 * c->n_local);
  free(d); free(ax); free(ay); free(inv); (void)NG;
}

/* ------------------------------------------------------------------ */
/* Wave-theory wavemaker and its damping (relaxation) zone             */
/* ------------------------------------------------------------------ */
/* Parallel_Taylor_Green_Vortices_2D_Release.f90:113-143: the analytic fields start as */
/* the particle state; fluid particles with x <= L are in the zone with the soft        */
/* coefficient 1 - (exp((x/L)**3.5) - 1.0)/(exp(1.0) - 1.0) (exp(1.0) a REAL(4)          */
/* intrinsic); Initial_particle_Position (here for fluid and wavemaker particles) is    */
/* the current position. zone_length <= 0 keeps only the initial positions (the        */
/* wave-theory wavemaker particles need them) and leaves the zone off.                  */
void orc_relaxation_init(orc_ctx* c, double zone_length) {
  const int n = c->n_fluid + c->n_wavemaker;
  if (!c->rz_in) {
    c->rz_in = xcalloc(c->cap_global, sizeof(int));
#define A(f) c->f = xcalloc(c->cap_global, sizeof(double));
    A(rz_c) A(an_x) A(an_y) A(an_u) A(an_v) A(an_p) A(an_rho) A(x0) A(y0)
#undef A
  }
  for (int i = 0; i < n; ++i) {
    c->x0[i] = c->gx[i]; c->y0[i] = c->gy[i];
    c->an_x[i] = c->gx[i]; c->an_y[i] = c->gy[i];
    c->an_u[i] = c->gu[i]; c->an_v[i] = c->gv[i];
    c->an_rho[i] = c->grho[i]; c->an_p[i] = c->gp[i];
    c->rz_in[i] = 0; c->rz_c[i] = 0.0;
    if (i < c->n_fluid && zone_length > 0.0 && c->gx[i] <= zone_length) {
      c->rz_in[i] = 1;
      c->rz_c[i] = (1 - (exp(pow(c->gx[i] / zone_length, 3.5)) - 1.0) / ((double)expf(1.0f) - 1.0));
    }
  }
  c->relax_on = zone_length > 0.0;
}

/* rho0 (1 + 7 p / (c0^2 rho0))**0.142857142 (REAL(4) exponent) */
static double inv_tait(const orc_params* P, double p) {
  return P->rho0 * pow(1 + 7 * p / (P->c0sq * P->rho0), (double)0.142857142f);
}

/* Wavemaker_Subroutine.f90:262-346 (Stokes_2nd) and :348-396 (Airy) for the wavemaker   */
/* particles, :268-276 (fluid below y = 0 typed 120), and :617-700 (the zone's analytic */
/* fields, Stokes_2nd and Airy) at step it, with time_soft_coefficient of :283-287     */
/* (integer division) and current_time = dt (it - standing). wave_type 1 Stokes_2nd,    */
/* 2 Airy; T = 2 pi / omega, wave length 2 pi / k, amplitude H / 2.                      */
void orc_wave_theory_update(orc_ctx* c, int32_t wave_type, int32_t soft_steps, int32_t standing_steps, double dt,
                            double g, double d, double H, double k, double om, double phase, int32_t it) {
  const orc_params* P = &c->P;
  const double pi = PI_MODULE;
  const int N = c->n_fluid;
  for (int i = 0; i < N; ++i) {
    if (c->gtype[i] == 120) continue;
    if (c->gy[i] < 0.0) c->gtype[i] = 120;
  }
  const double t = dt * (it - standing_steps);
  double soft;
  if (it <= soft_steps + standing_steps) soft = 0.5 * (sin((-0.5 + (it - standing_steps) / soft_steps) * pi) + 1);
  else soft = 1.0;
  if (!(t >= 0.0)) return;
  const double T = 2 * pi / om, L = 2 * pi / k, A = H / 2.0;
  const double skd = sinh(k * d), ckd = cosh(k * d);
  const double V0 = c->dx * c->dy;
  for (int q = 0; q < c->n_wavemaker; ++q) {                     /* :300-346, :354-396 */
    const int i = N + q;
    const double ang = k * c->gx[i] - om * t + phase;
    double el, u, v;
    const double yo = c->gy[i];
    if (wave_type == 1) {
      el = soft * (0.5 * H * cos(ang) + pi * H * H / (8 * L) * ckd / (skd * skd * skd) * (2 + cosh(2 * k * d)) * cos(2 * ang));
      u = pi * H / T * cosh(k * yo) * cos(ang) / skd + 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * cosh(2 * k * yo) * cos(2 * ang) / (skd * skd * skd * skd);
      v = pi * H / T * sinh(k * yo) * sin(ang) / skd + 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * sinh(2 * k * yo) * sin(2 * ang) / (skd * skd * skd * skd);
      c->gu[i] = soft * u;
      c->gv[i] = soft * v;
      c->gx[i] = c->x0[i] - soft * (A * cosh(k * yo) * sin(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * cosh(2 * k * yo) * sin(2 * ang) / (skd * skd * skd * skd));
      c->gy[i] = c->y0[i] + soft * (A * sinh(k * yo) * cos(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * sinh(2 * k * yo) * cos(2 * ang) / (skd * skd * skd * skd));
    } else {
      el = soft * 0.5 * H * cos(ang);
      u = pi * H / T * cosh(k * yo) * cos(ang) / skd;
      v = pi * H / T * sinh(k * yo) * sin(ang) / skd;
      c->gu[i] = soft * u;
      c->gv[i] = soft * v;
      c->gx[i] = c->x0[i] - soft * A * cosh(k * yo) * sin(ang) / skd;
      c->gy[i] = c->y0[i] + soft * A * sinh(k * yo) * cos(ang) / skd;
    }
    const double y = c->gy[i];
    if (y <= el + d) {
      double pr;
      if (wave_type == 1)
        pr = P->rho0 * g * (el + d - y) + soft * (P->rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd + 3.0 / 4.0 * P->rho0 * g * H * (pi * H / L) * (1.0 / sinh(2 * k * d)) * (cosh(2 * k * y) / (skd * skd) - (double)(1.0f / 3.0f)) * cos(2 * ang) - 1.0 / 4.0 * P->rho0 * g * H * (pi * H / L) * cosh(2 * k * y) / sinh(2 * k * d));
      else
        pr = P->rho0 * g * (d + el - y) + soft * (P->rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd);
      c->gp[i] = pr;
      c->grho[i] = inv_tait(P, pr);
      c->gmass[i] = P->rho0 * V0;
      c->gc[i] = sqrt(P->c0sq);
    } else {
      c->gp[i] = 0.0; c->grho[i] = P->rho0; c->gmass[i] = 0.0; c->gc[i] = sqrt(P->c0sq);
    }
  }
  if (!c->relax_on) return;
  for (int i = 0; i < N; ++i) {                                  /* :617-700 */
    if (c->gtype[i] == 120 || c->rz_in[i] != 1) continue;
    const double px = c->an_x[i], py = c->an_y[i];
    const double ang = k * px - om * t + phase;
    double el;
    if (wave_type == 1) {
      el = soft * (0.5 * H * cos(ang) + pi * H * H * (1.0 / tanh(k * d)) * (1 + 3.0 / (2 * (skd * skd * skd)) * cos(2 * ang)));
      c->an_u[i] = pi * H / T * cosh(k * py) * cos(ang) / skd - 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * cosh(2 * k * py) * cos(2 * ang) / (skd * skd * skd * skd);
      c->an_v[i] = pi * H / T * sinh(k * py) * sin(ang) / skd - 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * sinh(2 * k * py) * sin(2 * ang) / (skd * skd * skd * skd);
      c->an_u[i] = soft * c->an_u[i];
      c->an_v[i] = soft * c->an_v[i];
      c->an_x[i] = c->x0[i] - soft * (A * cosh(k * py) * sin(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * cosh(2 * k * py) * sin(2 * ang) / (skd * skd * skd * skd));
      c->an_y[i] = c->y0[i] + soft * (A * sinh(k * py) * cos(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * sinh(2 * k * py) * cos(2 * ang) / (skd * skd * skd * skd));
    } else {
      el = soft * 0.5 * H * cos(ang);
      c->an_u[i] = pi * H / T * cosh(k * py) * cos(ang) / skd;
      c->an_v[i] = pi * H / T * sinh(k * py) * sin(ang) / skd;
      c->an_u[i] = soft * c->an_u[i];
      c->an_v[i] = soft * c->an_v[i];
      c->an_x[i] = c->x0[i] - soft * A * cosh(k * py) * sin(ang) / skd;
      c->an_y[i] = c->y0[i] + soft * A * sinh(k * py) * cos(ang) / skd;
    }
    const double y = c->an_y[i];
    if (y <= el + d) {
      if (wave_type == 1)
        c->an_p[i] = P->rho0 * g * (el + d - y) + soft * (P->rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd + 3.0 / 4.0 * P->rho0 * g * H * (pi * H / L) * (1.0 / sinh(2 * k * d)) * (cosh(2 * k * y) / (skd * skd) - (double)(1.0f / 3.0f)) * cos(2 * ang) - 1.0 / 4.0 * P->rho0 * g * H * (pi * H / L) * cosh(2 * k * y) / sinh(2 * k * d));
      else
        c->an_p[i] = P->rho0 * g * (el + d - y) + soft * (P->rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd);
      c->an_rho[i] = inv_tait(P, c->an_p[i]);
    } else {
      c->an_p[i] = 0.0;
      c->an_rho[i] = P->rho0;
    }
  }
}

/* wave_maker_damping_zone.f90:61-95 on the subdomain (after the corrector): velocity,  */
/* position and pressure blended toward the analytic fields with the particle's soft     */
/* coefficient, then rho from the inverse Tait law (REAL(4) exponent).                   */
void orc_relaxation_blend(orc_ctx* c) {
  for (int s = 0; s < c->n_local; ++s) {
    const int g = c->name[s];
    if (g < 0 || g >= c->n_fluid || c->rz_in[g] != 1) continue;
    const double f = c->rz_c[g];
    c->u[s] = f * c->an_u[g] + (1 - f) * c->u[s];
    c->v[s] = f * c->an_v[g] + (1 - f) * c->v[s];
    c->x[s] = f * c->an_x[g] + (1 - f) * c->x[s];
    c->y[s] = f * c->an_y[g] + (1 - f) * c->y[s];
    c->p[s] = f * c->an_p[g] + (1 - f) * c->p[s];
    c->rho[s] = inv_tait(&c->P, c->p[s]);
  }
}

void orc_time_step(orc_ctx* c, int32_t it) {
  int* old_name = NULL; int old_n = c->n_local;
  if (c->P.rate_reuse == 1 && it > 1) {
    old_name = xcalloc(old_n + 1, sizeof(int));
    memcpy(old_name, c->name, sizeof(int) * old_n);
  }
  orc_period_boundary(c);
  orc_devide_block(c, it);
  if (old_name) { remap_rates(c, old_name, old_n); free(old_name); }
  if (c->P.integrator == 2) {                                    /* :509-657 */
    save_before(c);
    if (it == 1) orc_single_step_compute(c, it);
    pc_predict(c);
    pc_save(c);
    orc_single_step_compute(c, it);
    pc_correct(c);
    if (c->relax_on) orc_relaxation_blend(c);                    /* :660-667 */
  } else {                                                       /* :779-1049 */
    save_before(c);
    if (it == 1) orc_single_step_compute(c, it);
    stage_half(c); save_k(c, 0);
    orc_single_step_compute(c, it);
    stage_half(c); save_k(c, 1);
    orc_single_step_compute(c, it);
    stage_full(c); save_k(c, 2);
    orc_single_step_compute(c, it);
    rk_final(c);
    if (c->relax_on) orc_relaxation_blend(c);                    /* :1052-1059 */
  }
  if (c->P.mls_on && (it % c->P.mls_every) == 0) orc_reinitialize_density(c);
  if (c->P.shift_on) orc_shift(c);
  orc_refresh(c);
}

/* ------------------------------------------------------------------ */
/* Sampling_subroutine.f90:340-461                                     */
/* ------------------------------------------------------------------ */
void orc_sample(const orc_ctx* c, double* out) {
  const orc_params* P = &c->P;
  int N = c->n_fluid;
  double ke = 0.0, vmax = 0.0;
  double *eu = xcalloc(N, 8), *ev = xcalloc(N, 8), *ep = xcalloc(N, 8), *em = xcalloc(N, 8);
  double p0 = P->rho0 * P->U * P->U;
  for (int i = 0; i < N; ++i) {
    double vv = c->gu[i] * c->gu[i] + c->gv[i] * c->gv[i];
    ke = ke + 0.5 * c->gmass[i] * vv;
    double vm = sqrt(vv);
    if (vm >= vmax) vmax = vm;
    eu[i] = (c->gu[i] - c->ana_u[i]) / P->U;
    ev[i] = (c->gv[i] - c->ana_v[i]) / P->U;
    ep[i] = (c->gp[i] - c->ana_p[i]) / p0;
    em[i] = (vm - c->ana_vm[i]) / P->U;
  }
  out[0] = ke / c->ke0;
  out[1] = vmax / P->U;
  double* arrs[4] = {eu, ev, ep, em};
  for (int a = 0; a < 4; ++a) {
    double n1 = 0.0, n2 = 0.0, ninf = 0.0;
    for (int i = 0; i < N; ++i) {
      double e = fabs(arrs[a][i]);
      if (e >= ninf) ninf = e;
      n1 = n1 + e;
      n2 = n2 + e * e;
    }
    out[2 + 3 * a] = ninf;
    out[3 + 3 * a] = n1 / N;
    out[4 + 3 * a] = sqrt(n2 / N);
  }
  free(eu); free(ev); free(ep); free(em);
}

double orc_initial_ke(const orc_ctx* c) { return c->ke0; }
int32_t orc_n_fluid(const orc_ctx* c) { return c->n_fluid; }
int32_t orc_n_global(const orc_ctx* c) { return c->n_fluid + c->n_ghost; }
int32_t orc_n_local(const orc_ctx* c) { return c->n_local; }
int32_t orc_n_singular(const orc_ctx* c) { return c->n_singular; }
int64_t orc_pair_number(const orc_ctx* c) { return c->npair; }

/* ------------------------------------------------------------------ */
/* named array access                                                  */
/* ------------------------------------------------------------------ */
typedef struct { const char* name; size_t off; int kind; } fld;  /* kind 0 local, 1 global, 2 pair, 3 local*4 */
#define FL(nm, f, k) {nm, offsetof(struct orc_ctx, f), k}
#include <stddef.h>
static const fld F64[] = {
  FL("x", x, 0), FL("y", y, 0), FL("u", u, 0), FL("v", v, 0), FL("mass", mass, 0),
  FL("rho", rho, 0), FL("h", hh, 0), FL("c", c, 0), FL("p", p, 0), FL("e", e, 0),
  FL("volume", vol, 0), FL("fra_v", fra_v, 0), FL("sum_w", sum_w, 0),
  FL("sum_dwdx_x", sdwx, 0), FL("sum_dwdx_y", sdwy, 0), FL("normal_x", nrmx, 0),
  FL("normal_y", nrmy, 0), FL("S_xx", Sxx, 0), FL("S_xy", Sxy, 0), FL("S_yy", Syy, 0),
  FL("tau_xx", txx, 0), FL("tau_xy", txy, 0), FL("tau_yy", tyy, 0),
  FL("drhodt", drhodt, 0), FL("dedt", dedt, 0), FL("ax", ax, 0), FL("ay", ay, 0),
  FL("internal_x", ia_x, 0), FL("internal_y", ia_y, 0), FL("visc_x", av_x, 0),
  FL("visc_y", av_y, 0), FL("art_internal_x", ai_x, 0), FL("art_internal_y", ai_y, 0),
  FL("sps_x", sps_x, 0), FL("sps_y", sps_y, 0), FL("lap_x", lap_x, 0), FL("lap_y", lap_y, 0),
  FL("avg_vel_x", avg_x, 0), FL("avg_vel_y", avg_y, 0), FL("rho_correct", corr, 0),
  FL("grad_rho_x", grx, 0), FL("grad_rho_y", gry, 0), FL("shift_x", shx, 0),
  FL("shift_y", shy, 0), FL("vorticity", vort, 0), FL("g_vorticity", gvort, 1), FL("g_normal_x", gnx, 1), FL("g_normal_y", gny, 1),
  FL("grad_p_x", gpx, 0), FL("grad_p_y", gpy, 0), FL("grad_u_x", gux, 0), FL("grad_u_y", guy, 0),
  FL("grad_v_x", gvx, 0), FL("grad_v_y", gvy, 0),
  FL("mag_drhodt", mag_drho, 0), FL("mag_ax", mag_ax, 0), FL("mag_ay", mag_ay, 0),
  FL("T", T, 3), FL("T_inv", Tinv, 3),
  FL("g_x", gx, 1), FL("g_y", gy, 1), FL("g_u", gu, 1), FL("g_v", gv, 1),
  FL("g_mass", gmass, 1), FL("g_rho", grho, 1), FL("g_h", gh, 1), FL("g_c", gc, 1),
  FL("g_p", gp, 1), FL("g_e", ge, 1),
  FL("ana_u", ana_u, 4), FL("ana_v", ana_v, 4), FL("ana_p", ana_p, 4), FL("ana_vm", ana_vm, 4),
  FL("pair_w", pw, 2), FL("pair_dwdx_x", pdwx, 2), FL("pair_dwdx_y", pdwy, 2),
  FL("position_soft_coefficient", rz_c, 1), FL("an_x", an_x, 1), FL("an_y", an_y, 1), FL("an_u", an_u, 1),
  FL("an_v", an_v, 1), FL("an_p", an_p, 1), FL("an_rho", an_rho, 1), FL("initial_x", x0, 1), FL("initial_y", y0, 1),
  FL("body_pressure_x", bp_x, 0), FL("body_pressure_y", bp_y, 0), FL("body_viscous_x", bv_x, 0),
  FL("body_viscous_y", bv_y, 0),
};
static const fld I32[] = {
  FL("type", type, 0), FL("name", name, 0), FL("surface", surf, 0), FL("near_surface", near, 0),
  FL("candidate_n", cand_n, 0), FL("g_type", gtype, 1), FL("g_btype", gbtype, 1), FL("g_surface", gsurf, 1), FL("pair_i", pi, 2), FL("pair_j", pj, 2),
  FL("in_damping_zone", rz_in, 1), FL("g_initial_type", ginit, 1), FL("near_body", near_body, 0),
};
static int64_t fld_count(const orc_ctx* c, int kind) {
  switch (kind) {
    case 0: return c->n_local;
    case 1: return c->n_fluid + c->n_ghost;
    case 2: return c->npair;
    case 3: return (int64_t)c->n_local * 4;
    case 4: return c->n_fluid;
  }
  return 0;
}
double* orc_mut_f64(orc_ctx* c, const char* name, int64_t* n) {
  for (size_t k = 0; k < sizeof(F64) / sizeof(F64[0]); ++k)
    if (!strcmp(F64[k].name, name)) {
      if (n) *n = fld_count(c, F64[k].kind);
      return *(double**)((char*)c + F64[k].off);
    }
  if (n) *n = -1;
  return NULL;
}
int32_t* orc_mut_i32(orc_ctx* c, const char* name, int64_t* n) {
  for (size_t k = 0; k < sizeof(I32) / sizeof(I32[0]); ++k)
    if (!strcmp(I32[k].name, name)) {
      if (n) *n = fld_count(c, I32[k].kind);
      return *(int32_t**)((char*)c + I32[k].off);
    }
  if (n) *n = -1;
  return NULL;
}
const double* orc_get_f64(const orc_ctx* c, const char* name, int64_t* n) {
  return orc_mut_f64((orc_ctx*)c, name, n);
}
const int32_t* orc_get_i32(const orc_ctx* c, const char* name, int64_t* n) {
  return orc_mut_i32((orc_ctx*)c, name, n);
}

/* ------------------------------------------------------------------ */
/* Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/                 */
/* Reorder_Particle_Index.f90:54-287                                   */
/* ------------------------------------------------------------------ */
void orc_reorder_particle_index(int32_t n_total, int32_t n_fluid, double box_x, double box_y,
                                double* x, double* y, int32_t* type, int32_t n_f64, double** f64_cols,
                                int32_t n_i32, int32_t** i32_cols, int32_t* perm, int32_t* counts) {
  const double lo = (double)-1.0E-10f;
  for (int i = 0; i < n_fluid; ++i) {                             /* :56-64 */
    if (type[i] == 2) {
      if (x[i] <= lo || x[i] > box_x || y[i] <= lo || y[i] > box_y) type[i] = 120;
    }
  }
  int* lists[5];                                                  /* :84-115 */
  int nl[5] = {0, 0, 0, 0, 0};
  for (int c = 0; c < 5; ++c) lists[c] = xcalloc(n_total > 0 ? n_total : 1, sizeof(int));
  for (int i = 0; i < n_total; ++i) {
    int c = -1;
    if (type[i] == 2) c = 0;                 /* Water_Particle_Label */
    else if (type[i] == -3) c = 2;           /* Fix_Ghost_Particle_Label */
    else if (type[i] == -201) c = 3;         /* Inlet_Particle_Label */
    else if (type[i] == -202) c = 4;         /* Outlet_Particle_Label */
    else if (type[i] == 130) c = 1;          /* Wave_Maker_Particle_Label */
    if (c >= 0) lists[c][nl[c]++] = i;
  }
  int k = 0;                                                      /* :157-287 */
  for (int c = 0; c < 5; ++c)
    for (int L = 0; L < nl[c]; ++L) perm[k++] = lists[c][L];
  double* tf = xcalloc(k > 0 ? k : 1, sizeof(double));
  int* ti = xcalloc(k > 0 ? k : 1, sizeof(int));
  double* fc[2] = {x, y};
  for (int col = 0; col < 2 + n_f64; ++col) {
    double* a = col < 2 ? fc[col] : f64_cols[col - 2];
    for (int m = 0; m < k; ++m) tf[m] = a[perm[m]];
    memcpy(a, tf, sizeof(double) * (size_t)k);
  }
  for (int col = 0; col < 1 + n_i32; ++col) {
    int32_t* a = col == 0 ? type : i32_cols[col - 1];
    for (int m = 0; m < k; ++m) ti[m] = a[perm[m]];
    memcpy(a, ti, sizeof(int) * (size_t)k);
  }
  for (int c = 0; c < 5; ++c) { counts[c] = nl[c]; free(lists[c]); }
  counts[5] = k;
  free(tf);
  free(ti);
}


/* ------------------------------------------------------------------ */
/* The support of a point in the global chain (Calculation_For_PostProceeding.f90 */
/* :184-250, Sampling_subroutine.f90:55-111): fluid particles (gid < N) of the     */
/* 9-cell stencil of the point's cell, in stencil then chain order, at distance    */
/* within [1e-7 (REAL(4)), 3h]; tn = their gids, wk = W. Returns the count.        */
/* ------------------------------------------------------------------ */
static int point_support(orc_ctx* c, double ex, double ey, int N, int* eff, int* tn, double* wk) {
  const double h = c->h;
  int xn = (int)floor((ex - c->chain_ox) / c->chain_d + 1);
  int yn = (int)floor((ey - c->chain_oy) / c->chain_d + 1);
  if (xn < 1 || xn > c->chain_nx || yn < 1 || yn > c->chain_ny) return 0;
  int mesh = (yn - 1) * c->chain_nx + xn;
  int nm[9];
  int nnear = stencil_names(mesh, c->chain_nx, c->chain_ny, nm);
  int ne = 0;
  for (int q = 0; q < nnear; ++q) {
    int cell = nm[q];
    for (int m = c->gchain_off[cell]; m < c->gchain_off[cell + 1]; ++m) {
      int jj = c->gchain_members[m];
      if (jj < N && ne < 300) eff[ne++] = jj;
    }
  }
  int nt = 0;
  for (int q = 0; q < ne; ++q) {
    int jj = eff[q];
    double pd[2] = {ex - c->gx[jj], ey - c->gy[jj]};
    double d = sqrt(pd[0] * pd[0] + pd[1] * pd[1]);
    if (d <= 3.0 * h && d >= (double)1.0e-7f) {
      double w, dw[2];
      orc_compute_kernel(2, d, pd, h, &w, dw);
      tn[nt] = jj; wk[nt] = w; ++nt;
    }
  }
  return nt;
}

/* MLS_Interpolation_In_Gobal_Domain.f90:110-250, order 6, at (ex, ey): Fai of */
/* the nt support particles (Shepard W / sum W with 10 or fewer, or when the   */
/* moment matrix has no inverse). pm: 6 x nt scratch.                          */
static void point_mls_fai(orc_ctx* c, double ex, double ey, int nt, const int* tn, const double* wk, double* pm,
                          double* fai) {
  const double h = c->h, al = c->P.mls_alpha;
  for (int q = 0; q < nt; ++q) {
    int jj = tn[q];
    double a = (c->gx[jj] - ex) / h, b = (c->gy[jj] - ey) / h;
    double* col = &pm[6 * q];
    col[0] = 1; col[1] = a; col[2] = b; col[3] = a * a; col[4] = a * b; col[5] = b * b;
  }
  double A[36];                                       /* column-major 6x6 */
  for (int m = 0; m < 6; ++m)
    for (int l = 0; l < 6; ++l) {
      double v = 0.0;
      for (int q = 0; q < nt; ++q) v = v + (pm[6 * q + m] * wk[q]) * pm[6 * q + l];
      for (int L = 0; L < 6; ++L) v = v + ((m == L ? al : 0.0) * al) * (l == L ? al : 0.0);
      A[m + 6 * l] = v;
    }
  int ok = orc_brinv(A, 6);
  double sw = 0.0;
  for (int q = 0; q < nt; ++q) sw = sw + wk[q];
  for (int q = 0; q < nt; ++q) {
    if (ok) {
      double v = 0.0;
      for (int m = 0; m < 6; ++m) v = v + A[0 + 6 * m] * (pm[6 * q + m] * wk[q]);
      fai[q] = v;
    } else {
      fai[q] = wk[q] / sw;
    }
    if (nt <= 10) fai[q] = wk[q] / sw;
  }
}

/* ------------------------------------------------------------------ */
/* Calculation_For_PostProceeding.f90:96-300 (one rank) on the element nodes of  */
/* Distribute_Grid_For_Postproceeding.f90:30-39 (node k = j + (i-1) nnx at       */
/* ((j-1) edx, (i-1) edx)), with MLS_Interpolation_In_Gobal_Domain.f90 order 6.  */
/* Uses the step-start binning (particle_in_chain) and the refreshed global       */
/* arrays, as the reference does after Refresh_domain. out: 9 blocks of n_nodes:  */
/* u, v, p, vorticity_z, Fai, analytic u, v, p, |v| (Devide_block...:104-115).    */
/* ------------------------------------------------------------------ */
void orc_grid_postprocess(orc_ctx* c, int32_t i_time_step, int32_t nex, int32_t ney, double edx, double* out) {
  const orc_params* P = &c->P;
  const int nnx = nex + 1, nny = ney + 1, nn = nnx * nny, N = c->n_fluid;
  const double h = c->h, pi = PI_MODULE;
  const double t = i_time_step * P->dt / (P->box_x / P->U);
  const double Re = P->U * P->box_x / P->nu;
  memset(out, 0, sizeof(double) * 9 * (size_t)nn);
  /* :96-102 surface particles in gid order */
  int ns = 0;
  int* sname = xcalloc(N + 1, sizeof(int));
  for (int i = 0; i < N; ++i) if (c->gsurf[i] == 1) sname[ns++] = i;
  int* eff = xcalloc(300, sizeof(int));
  int* tn = xcalloc(300, sizeof(int));
  double* wk = xcalloc(300, sizeof(double));
  double* pm = xcalloc(6 * 306, sizeof(double));       /* P(m,k), column k */
  double* fai = xcalloc(300, sizeof(double));
  for (int k = 0; k < nn; ++k) {
    const int j0 = k % nnx, i0 = k / nnx;
    const double ex = j0 * edx, ey = i0 * edx;          /* both coordinates use element_dx */
    out[5 * (size_t)nn + k] = P->U * sin(2 * pi * ex) * cos(2 * pi * ey) * exp(-8 * (pi * pi) / Re * t);
    out[6 * (size_t)nn + k] = -P->U * cos(2 * pi * ex) * sin(2 * pi * ey) * exp(-8 * (pi * pi) / Re * t);
    out[7 * (size_t)nn + k] = 0.25 * P->rho0 * (P->U * P->U) * (cos(4 * pi * ex) + cos(4 * pi * ey)) *
                              exp(-16 * (pi * pi) / Re * t);
    out[8 * (size_t)nn + k] = sqrt(out[5 * (size_t)nn + k] * out[5 * (size_t)nn + k] +
                                   out[6 * (size_t)nn + k] * out[6 * (size_t)nn + k]);
    /* :125-170 level set from the nearest surface particle */
    double f = 1.0;
    if (ns > 0) {
      double best = 1000000000.0, dfx = 0.0, dfy = 0.0;
      int idx = -1;
      for (int q = 0; q < ns; ++q) {
        int jj = sname[q];
        double pdx = ex - c->gx[jj], pdy = ey - c->gy[jj];
        double d = sqrt(pdx * pdx + pdy * pdy);
        if (d <= best) { best = d; idx = jj; dfx = -pdx; dfy = -pdy; }
      }
      double det = dfx * c->gnx[idx] + dfy * c->gny[idx];
      if (det <= -2 * h) f = -1.0;
      else if (det >= 2 * h) f = 1.0;
      else f = det / (2 * h);
    }
    out[4 * (size_t)nn + k] = f;
    if (f <= -0.9) continue;
    int nt = point_support(c, ex, ey, N, eff, tn, wk);
    if (nt < 3) continue;
    point_mls_fai(c, ex, ey, nt, tn, wk, pm, fai);
    double su = 0.0, sv = 0.0, sp = 0.0, so = 0.0;
    for (int q = 0; q < nt; ++q) {
      int jj = tn[q];
      sp = sp + c->gp[jj] * fai[q];
      su = su + c->gu[jj] * fai[q];
      sv = sv + c->gv[jj] * fai[q];
      so = so + c->gvort[jj] * fai[q];
    }
    out[0 * (size_t)nn + k] = su; out[1 * (size_t)nn + k] = sv;
    out[2 * (size_t)nn + k] = sp; out[3 * (size_t)nn + k] = so;
  }
  free(sname); free(eff); free(tn); free(wk); free(pm); free(fai);
}

/* ------------------------------------------------------------------ */
/* Sampling_subroutine.f90:21-177 (Sampling) at n points (px, py) of the       */
/* state after Refresh_domain: with more than 3 particles in the support, the   */
/* average pressure over the support, the order-6 MLS pressure (:160-164, minus */
/* g rho_j (y_s - y_j) Fai_j when gravity != 0) and the MLS velocity; zeros     */
/* otherwise. out: 4 blocks of n (average p, MLS p, u, v).                      */
/* ------------------------------------------------------------------ */
void orc_sampling_points(orc_ctx* c, int32_t n, const double* px, const double* py, double gravity, double* out) {
  const int N = c->n_fluid;
  int* eff = xcalloc(300, sizeof(int));
  int* tn = xcalloc(300, sizeof(int));
  double* wk = xcalloc(300, sizeof(double));
  double* pm = xcalloc(6 * 306, sizeof(double));
  double* fai = xcalloc(300, sizeof(double));
  memset(out, 0, sizeof(double) * 4 * (size_t)n);
  for (int k = 0; k < n; ++k) {
    const double sx = px[k], sy = py[k];
    int nt = point_support(c, sx, sy, N, eff, tn, wk);
    if (nt <= 3) continue;
    double avg = 0.0, mp = 0.0, su = 0.0, sv = 0.0;
    for (int q = 0; q < nt; ++q) avg = avg + c->gp[tn[q]];
    point_mls_fai(c, sx, sy, nt, tn, wk, pm, fai);
    for (int q = 0; q < nt; ++q) {
      int jj = tn[q];
      if (gravity != 0.0) mp = mp + c->gp[jj] * fai[q] - gravity * c->grho[jj] * (sy - c->gy[jj]) * fai[q];
      else mp = mp + c->gp[jj] * fai[q];
      su = su + c->gu[jj] * fai[q];
      sv = sv + c->gv[jj] * fai[q];
    }
    out[k] = avg / nt;
    out[(size_t)n + k] = mp;
    out[2 * (size_t)n + k] = su;
    out[3 * (size_t)n + k] = sv;
  }
  free(eff); free(tn); free(wk); free(pm); free(fai);
}

/* Sampling_subroutine.f90:246-279: the wave front (largest x of the fluid        */
/* particles, from 0) and each probe's elevation (largest y, from 0, of the fluid */
/* particles within 0.55 (REAL(4)) dx of the probe's x). out: 1 + n_probes.        */
void orc_wave_probes(orc_ctx* c, int32_t n_probes, const double* probe_x, double* out) {
  const int N = c->n_fluid;
  double front = 0.0;
  for (int j = 0; j < n_probes; ++j) out[1 + j] = 0.0;
  for (int i = 0; i < N; ++i) {
    for (int j = 0; j < n_probes; ++j) {
      double d = fabs(probe_x[j] - c->gx[i]);
      if (d <= (double)0.55f * c->dx && c->gy[i] > out[1 + j]) out[1 + j] = c->gy[i];
    }
    if (c->gx[i] >= front) front = c->gx[i];
  }
  out[0] = front;
}

/* Wavemaker_Subroutine.f90:196-215 on the global arrays: fluid particles (not typed */
/* 120) with x < xp are mirrored to 2 xp - x; then those with x - xp < dx take u = up. */
void orc_wavemaker_pull_back(orc_ctx* c, double xp, double up) {
  for (int i = 0; i < c->n_fluid; ++i) {
    if (c->gtype[i] == 120) continue;
    if (c->gx[i] < xp) c->gx[i] = 2 * xp - c->gx[i];
    if ((c->gx[i] - xp) < c->dx) c->gu[i] = up;
  }
}

/* ================================================================== */
/* The moving-square program (Moving_Square_Or_Cylinder_Inside_a_     */
/* Rectangular_Box/, "MS"), ms = 1. The local order stays the rule of  */
/* orc_devide_block (4 dx cells column by column, gid order within a   */
/* cell): MS's two-dimensional rank decomposition has no one-rank form */
/* to restate, and only the HLLC pair orientation and the stale rates  */
/* of the predictor depend on it.                                      */
/* ================================================================== */

/* One straight run of layer points (MS/Initial_Particle_From_First_Step.f90:178-620):
 * the first layer's count k comes from walking from (x0, y0) in steps of gdx along
 * (sx, sy) until within gdx/2 of (x1, y1); the run places k - 1 points at the spacing
 * |(x1, y1) - (x0, y0)| / (k - 1); k then changes by dk (+2 outside walls, -2 body). */
static int ms_run(double x0, double y0, double x1, double y1, double sx, double sy, double gdx, int* k,
                  int first, int dk, double* x, double* y, int at) {
  if (first) {
    int j = 0;
    for (;;) {
      ++j;
      double px = x0 + sx * ((j - 1) * gdx), py = y0 + sy * ((j - 1) * gdx);
      if (sqrt((x1 - px) * (x1 - px) + (y1 - py) * (y1 - py)) <= gdx / 2.0) break;
    }
    *k = j;
  }
  double dist = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0)) / (*k - 1);
  int n = 0;
  for (int j = 1; j <= *k - 1; ++j) {
    if (x) { x[at + n] = x0 + sx * ((j - 1) * dist); y[at + n] = y0 + sy * ((j - 1) * dist); }
    ++n;
  }
  *k = *k + dk;
  return n;
}

/* Fixed ghosts of the tank (:178-405) then of the square (:419-620); bt: boundary
 * type, it: initial type (-3 wall, -4 body). Returns the count (x == NULL: count only). */
static int ms_ghost_layout(const orc_params* P, double gdx, double gdy, double* x, double* y, int* bt, int* it) {
  const double bx = P->box_x, by = P->box_y;
  const int L = P->walls;
  int n = 0, k = 0, n0;
  n0 = n;
  for (int i = 1; i <= L; ++i)                                       /* bottom :178-219 */
    n += ms_run(-(i - 1) * gdy - gdy / 2.0, -(i - 1) * gdy - gdy / 2.0, bx + (i - 1) * gdy + gdy / 2.0,
                -(i - 1) * gdy - gdy / 2.0, 1, 0, gdx, &k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= x[q] && x[q] <= bx) ? -6 : -7; it[q] = -3; }
  n0 = n;
  for (int i = 1; i <= L; ++i)                                       /* right :243-279 */
    n += ms_run(bx + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0, bx + (i - 1) * gdy + gdy / 2.0,
                by + (i - 1) * gdy + gdy / 2.0, 0, 1, gdx, &k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= y[q] && y[q] <= by) ? -8 : -7; it[q] = -3; }
  n0 = n;
  for (int i = 1; i <= L; ++i)                                       /* top :300-351 */
    n += ms_run(bx + (i - 1) * gdy + gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0,
                by + (i - 1) * gdy + gdy / 2.0, -1, 0, gdx, &k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= x[q] && x[q] <= bx) ? -6 : -7; it[q] = -3; }
  n0 = n;
  for (int i = 1; i <= L; ++i)                                       /* left :355-405 */
    n += ms_run(-(i - 1) * gdy - gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0,
                -(i - 1) * gdy - gdy / 2.0, 0, -1, gdx, &k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= y[q] && y[q] <= by) ? -8 : -7; it[q] = -3; }
  const double cx = P->body_cx, cy = P->body_cy, hl = 0.5 * P->body_len;
  /* the square: per side a run whose first point is a corner (-7), the rest -6 or -8 */
  const double sgn[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
  for (int side = 0; side < 4; ++side) {
    n0 = n;
    int first_of_layer[16];
    int nl = 0;
    for (int i = 1; i <= L; ++i) {
      /* (Block_Center -/+ 0.5 Block_Length) +/- (i-1) dy +/- dy/2, left to right */
      const double lo_x = (cx - hl) + (i - 1) * gdy + gdy / 2.0, hi_x = (cx + hl) - (i - 1) * gdy - gdy / 2.0;
      const double lo_y = (cy - hl) + (i - 1) * gdy + gdy / 2.0, hi_y = (cy + hl) - (i - 1) * gdy - gdy / 2.0;
      double x0, y0, x1, y1;
      if (side == 0) { x0 = lo_x; y0 = lo_y; x1 = hi_x; y1 = lo_y; }          /* :419-467 */
      else if (side == 1) { x0 = hi_x; y0 = lo_y; x1 = hi_x; y1 = hi_y; }     /* :468-516 */
      else if (side == 2) { x0 = hi_x; y0 = hi_y; x1 = lo_x; y1 = hi_y; }     /* :517-565 */
      else { x0 = lo_x; y0 = hi_y; x1 = lo_x; y1 = lo_y; }                    /* :566-614 */
      if (nl < 16) first_of_layer[nl++] = n;
      n += ms_run(x0, y0, x1, y1, sgn[side][0], sgn[side][1], gdx, &k, i == 1, -2, x, y, n);
    }
    if (x) {
      for (int q = n0; q < n; ++q) { bt[q] = (side == 0 || side == 2) ? -6 : -8; it[q] = -4; }
      for (int l = 0; l < nl; ++l) if (first_of_layer[l] < n) bt[first_of_layer[l]] = -7;
    }
  }
  return n;
}

void orc_init_moving_square(orc_ctx* c) {
  const orc_params* P = &c->P;
  const double cx = P->body_cx, cy = P->body_cy, hl = 0.5 * P->body_len;
  int k = 0;
  for (int i = 1; i <= P->ny; ++i)                    /* SPH_Subroutine_Module.f90:889-918 */
    for (int j = 1; j <= P->nx; ++j) {
      double x = (j - 1) * c->dx + c->dx / 2.0;
      double y = (i - 1) * c->dy + c->dy / 2.0;
      if (x >= (cx - hl) && x <= (cx + hl) && y >= (cy - hl) && y <= (cy + hl)) continue;
      if (k >= c->cap_global) { fprintf(stderr, "orc: capacity\n"); abort(); }
      c->gx[k] = x; c->gy[k] = y;
      c->gu[k] = 0.0; c->gv[k] = 0.0;
      c->gp[k] = 0.0 + P->background_p;
      c->grho[k] = (c->gp[k] - P->background_p) / P->c0sq + P->rho0;
      ++k;
    }
  c->n_fluid = k;
  const double vol0 = c->dx * c->dy;                  /* Initial_Particle_From_First_Step.f90:81-88 */
  const double m0 = P->rho0 * vol0;
  for (int i = 0; i < k; ++i) {
    c->gmass[i] = m0; c->gtype[i] = 2; c->ginit[i] = 2; c->gh[i] = c->h;
    c->gc[i] = sqrt(P->c0sq); c->ge[i] = c->e0;
  }
  c->ke0 = 0.0;
  c->n_wavemaker = 0;
  int ng = ms_ghost_layout(P, c->dx, c->dy, NULL, NULL, NULL, NULL);
  if (k + ng > c->cap_global) { fprintf(stderr, "orc: ghost capacity\n"); abort(); }
  ms_ghost_layout(P, c->dx, c->dy, c->gx + k, c->gy + k, c->gbtype + k, c->ginit + k);
  for (int q = 0; q < ng; ++q) {                      /* :621-632 */
    int m = k + q;
    c->gu[m] = 0.0; c->gv[m] = 0.0; c->grho[m] = P->rho0; c->gmass[m] = m0;
    c->gtype[m] = -3; c->gh[m] = c->h; c->gp[m] = 0.0 + P->background_p; c->gc[m] = sqrt(P->c0sq);
    c->ge[m] = c->e0;
  }
  c->n_ghost = ng;
  c->body_x = cx; c->body_y = cy;
  c->cust_u = 0.0; c->cust_a = 0.0;
  c->real_time = 0.0;
  memset(c->bforce, 0, sizeof(c->bforce));
}

const int32_t* orc_ms_initial_type(const orc_ctx* c) { return c->ginit; }

/* Body_Motion_Solver.f90:62-273 (Free_motion_or_Custom_motion = 1) with the spline of
 * Input_Data/Motion_Body.dat replaced by the prescribed ramp u(t) = U min(t/T, 1):
 * Custom_position(1) = x0 + U t^2 / (2T) (t < T), x0 + U T / 2 + U (t - T) after;
 * Custom_velocity(1) = u(t), Custom_accleration(1) = U / T (t < T) or 0; the body
 * particles move by the change of the centre (:245-271). */
void orc_ms_body_motion(orc_ctx* c, double t) {
  const orc_params* P = &c->P;
  const double U = P->body_speed, T = P->body_ramp;
  double xc, u, a;
  if (t < T) { xc = P->body_cx + 0.5 * U * t * t / T; u = U * t / T; a = U / T; }
  else { xc = P->body_cx + 0.5 * U * T + U * (t - T); u = U; a = 0.0; }
  c->cust_u = u; c->cust_a = a;
  const double ddx = xc - c->body_x, ddy = P->body_cy - c->body_y;
  c->body_x = xc; c->body_y = P->body_cy;
  const int NG = c->n_fluid + c->n_ghost;
  for (int i = 0; i < NG; ++i)
    if (c->ginit[i] == -4) { c->gx[i] = c->gx[i] + ddx; c->gy[i] = c->gy[i] + ddy; }
}

/* The MS solvers. First grad rho and grad v (and grad p for MUSCL-HLLC) with the       */
/* normalised kernel gradients, the SPS stress of the fluid (sps_on, the block of MS/   */
/* Solver_Normalized_HLLC_Riemann_Delta_SPS.f90:95-235), the XSPH velocity epsilon      */
/* m_j (v_j - v_i) / (rho_i + rho_j) W of every particle; then per pair the star state: */
/*  - MUSCL-HLLC (MS/Solver_Normalized_MUSCL_HLLC_Riemann_Delta.f90:148-412): p, u, v    */
/*    reconstructed with the limited slopes, velocities rotated into the pair frame,     */
/*    U* = (U_L z_L + U_R z_R + P_L - P_R) / (z_L + z_R), P* acoustic, rho* from the      */
/*    linear EOS, HLLC fan sampled at the rotated mean velocity, P* = c0^2 (rho_HLLC -    */
/*    rho0) + p_b;                                                                        */
/*  - HLLC (MS/Solver_Normalized_HLLC_Riemann_Delta_SPS.f90:240-330): no reconstruction,  */
/*    U* = (U_R z_L + U_L z_R + P_L - P_R) / (z_L + z_R) (the reference's own form);      */
/*  - modified linear (MS/Solver_Normalized_Modified_Linear_Riemann_Delta.f90:118-170);  */
/* the continuity with (v_i - v* + XSPH_i - XSPH_j), momentum, Monaghan viscosity,      */
/* Morris laplacian, delta-alpha term, SPS (raw dwdx, particle_mass at the local index, */
/* temp_p carried from x into y), and the delta-SPH density correction.                  */
static void ms_solver(orc_ctx* c) {
  const orc_params* P = &c->P;
  const int n = c->n_local;
  const int sv = P->solver;
  size_t nb = sizeof(double) * n;
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    double* Ti = &c->Tinv[i * 4];
    double* Tj = &c->Tinv[j * 4];
    double d1 = c->pdwx[k], d2 = c->pdwy[k];
    c->ndix[k] = Ti[0] * d1 + Ti[2] * d2; c->ndiy[k] = Ti[1] * d1 + Ti[3] * d2;
    c->ndjx[k] = Tj[0] * d1 + Tj[2] * d2; c->ndjy[k] = Tj[1] * d1 + Tj[3] * d2;
  }
  memset(c->grx, 0, nb); memset(c->gry, 0, nb);
  memset(c->gux, 0, nb); memset(c->guy, 0, nb); memset(c->gvx, 0, nb); memset(c->gvy, 0, nb);
  memset(c->gpx, 0, nb); memset(c->gpy, 0, nb); memset(c->mag_grho, 0, nb);
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    double rd = -c->prd[k], ud = -c->pvx[k], vd = -c->pvy[k];
    double Vi = c->vol[i], Vj = c->vol[j];
    double a1 = c->ndix[k], a2 = c->ndiy[k], b1 = c->ndjx[k], b2 = c->ndjy[k];
    c->grx[i] = c->grx[i] + rd * Vj * a1; c->gry[i] = c->gry[i] + rd * Vj * a2;
    c->grx[j] = c->grx[j] + rd * Vi * b1; c->gry[j] = c->gry[j] + rd * Vi * b2;
    c->mag_grho[i] += fabs(rd) * Vj * (fabs(a1) + fabs(a2));
    c->mag_grho[j] += fabs(rd) * Vi * (fabs(b1) + fabs(b2));
    c->gux[i] = c->gux[i] + ud * Vj * a1; c->guy[i] = c->guy[i] + ud * Vj * a2;
    c->gux[j] = c->gux[j] + ud * Vi * b1; c->guy[j] = c->guy[j] + ud * Vi * b2;
    c->gvx[i] = c->gvx[i] + vd * Vj * a1; c->gvy[i] = c->gvy[i] + vd * Vj * a2;
    c->gvx[j] = c->gvx[j] + vd * Vi * b1; c->gvy[j] = c->gvy[j] + vd * Vi * b2;
    if (sv == ORC_SOLVER_MS_MUSCL_HLLC) {
      double pd = c->p[j] - c->p[i];
      c->gpx[i] = c->gpx[i] + pd * Vj * a1; c->gpy[i] = c->gpy[i] + pd * Vj * a2;
      c->gpx[j] = c->gpx[j] + pd * Vi * b1; c->gpy[j] = c->gpy[j] + pd * Vi * b2;
    }
  }
  memset(c->Sxx, 0, nb); memset(c->Sxy, 0, nb); memset(c->Syy, 0, nb);
  memset(c->txx, 0, nb); memset(c->txy, 0, nb); memset(c->tyy, 0, nb);
  if (P->sps_on) {
    for (int64_t k = 0; k < c->npair; ++k) {
      int i = c->pi[k], j = c->pj[k];
      double vd1 = -c->pvx[k], vd2 = -c->pvy[k];
      c->Sxx[i] = c->Sxx[i] + vd1 * c->ndix[k] * c->vol[j];
      c->Sxy[i] = c->Sxy[i] + (vd1 * c->ndiy[k] + vd2 * c->ndix[k]) * c->vol[j];
      c->Syy[i] = c->Syy[i] + vd2 * c->ndiy[k] * c->vol[j];
      c->Sxx[j] = c->Sxx[j] + vd1 * c->ndjx[k] * c->vol[i];
      c->Sxy[j] = c->Sxy[j] + (vd1 * c->ndjy[k] + vd2 * c->ndjx[k]) * c->vol[i];
      c->Syy[j] = c->Syy[j] + vd2 * c->ndjy[k] * c->vol[i];
    }
    const double dL = (double)(sqrtf(2.0f) / 2.0f) * c->dx;
    const double c012 = (double)0.12f, c23 = (double)(2.0f / 3.0f), c23b = (double)((2.0f / 3.0f) * 0.0066f);
    for (int i = 0; i < n; ++i) {
      if (c->type[i] != 2) continue;
      double master = c->Sxx[i] * c->Sxx[i] + c->Syy[i] * c->Syy[i] + 0.0;
      double last = c->Sxy[i] * c->Sxy[i] + 0.0 + 0.0;
      double abs2 = 2 * master + last;
      double t0 = c012 * dL;
      double nut = (t0 * t0) * sqrt(abs2);
      double kk = c23 * nut * (c->Sxx[i] + c->Syy[i]);
      double blin = c23b * (dL * dL) * abs2;
      c->txx[i] = (2 * nut * c->Sxx[i] - kk - blin) * c->rho[i];
      c->txy[i] = (nut * c->Sxy[i]) * c->rho[i];
      c->tyy[i] = (2 * nut * c->Syy[i] - kk - blin) * c->rho[i];
    }
  }
  memset(c->avg_x, 0, nb); memset(c->avg_y, 0, nb);
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    double rs = c->rho[i] + c->rho[j], w = c->pw[k];
    c->avg_x[i] = c->avg_x[i] - P->epsilon_xsph * c->mass[j] * c->pvx[k] / rs * w;
    c->avg_y[i] = c->avg_y[i] - P->epsilon_xsph * c->mass[j] * c->pvy[k] / rs * w;
    c->avg_x[j] = c->avg_x[j] + P->epsilon_xsph * c->mass[i] * c->pvx[k] / rs * w;
    c->avg_y[j] = c->avg_y[j] + P->epsilon_xsph * c->mass[i] * c->pvy[k] / rs * w;
  }
  const double sqc0 = sqrt(P->c0sq);
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    double PL = c->p[i], PR = c->p[j];
    double Iv[2] = {c->u[i], c->v[i]}, Jv[2] = {c->u[j], c->v[j]};
    double xij[2] = {c->pdx[k], c->pdy[k]};                      /* x_i - x_j */
    double cb = c->pc_[k], rb = c->prho_[k];
    double vs[2], Ps, Pbar, vb[2];
    if (sv == ORC_SOLVER_MS_MUSCL_HLLC || sv == ORC_SOLVER_MS_HLLC) {
      double e[2];
      if (sv == ORC_SOLVER_MS_MUSCL_HLLC) {
        double xr[2] = {-xij[0], -xij[1]};                        /* x_j - x_i */
        double dist = sqrt(xr[0] * xr[0] + xr[1] * xr[1]);
        e[0] = xr[0] / dist; e[1] = xr[1] / dist;
        double li, lj;
        orc_muscl_limiter(0.5 * (c->gpx[i] * xr[0] + c->gpy[i] * xr[1]),
                          -(0.5 * (c->gpx[j] * xr[0] + c->gpy[j] * xr[1])), c->p[j] - c->p[i], &li, &lj);
        PL = c->p[i] + li; PR = c->p[j] - lj;
        orc_muscl_limiter(0.5 * (c->gux[i] * xr[0] + c->guy[i] * xr[1]),
                          -(0.5 * (c->gux[j] * xr[0] + c->guy[j] * xr[1])), -c->pvx[k], &li, &lj);
        Iv[0] = c->u[i] + li; Jv[0] = c->u[j] - lj;
        orc_muscl_limiter(0.5 * (c->gvx[i] * xr[0] + c->gvy[i] * xr[1]),
                          -(0.5 * (c->gvx[j] * xr[0] + c->gvy[j] * xr[1])), -c->pvy[k], &li, &lj);
        Iv[1] = c->v[i] + li; Jv[1] = c->v[j] - lj;
      } else {
        double dist = sqrt(xij[0] * xij[0] + xij[1] * xij[1]);
        e[0] = -xij[0] / dist; e[1] = -xij[1] / dist;
      }
      Pbar = (PL + PR) / 2.0;
      vb[0] = (Iv[0] + Jv[0]) / 2.0; vb[1] = (Iv[1] + Jv[1]) / 2.0;
      /* Transfer_Matrix = [e1 e2; -e2 e1] */
      double L1 = e[0] * Iv[0] + e[1] * Iv[1], L2 = -e[1] * Iv[0] + e[0] * Iv[1];
      double R1 = e[0] * Jv[0] + e[1] * Jv[1], R2 = -e[1] * Jv[0] + e[0] * Jv[1];
      double B1 = e[0] * vb[0] + e[1] * vb[1];
      double rL = c->rho[i], rR = c->rho[j], cL = c->c[i], cR = c->c[j];
      double Us = (sv == ORC_SOLVER_MS_MUSCL_HLLC)
                      ? (L1 * cL * rL + R1 * cR * rR + PL - PR) / (rL * cL + rR * cR)
                      : (R1 * cL * rL + L1 * cR * rR + PL - PR) / (rL * cL + rR * cR);
      double Pst = (PR * cL * rL + PL * cR * rR - cL * rL * cR * rR * (R1 - L1)) / (rL * cL + rR * cR);
      double rst = (Pst - P->background_p) / P->c0sq + P->rho0;
      double QL = 1.0, QR = 1.0;
      if (!(Pst <= PL)) QL = (rL > rst) ? sqrt((rst / rL) * (Pst - PL) / (rL - rst)) / cL : 1.0;
      if (!(Pst <= PR)) QR = (rR > rst) ? sqrt((rst / rR) * (Pst - PR) / (rR - rst)) / cR : 1.0;
      double SL = L1 - cL * QL, SR = R1 + cR * QR, xot = B1;
      double rH, H1, H2;
      if (SL > xot) { rH = rL; H1 = L1; H2 = L2; }
      else if (Us >= xot && xot >= SL) {
        double q = rL * ((SL - L1) / (SL - Us));
        rH = q; H1 = (Us * q) / rH; H2 = (L2 * q) / rH;
      } else if (SR >= xot && xot >= Us) {
        double q = rR * ((SR - R1) / (SR - Us));
        rH = q; H1 = (Us * q) / rH; H2 = (R2 * q) / rH;
      } else { rH = rR; H1 = R1; H2 = R2; }
      vs[0] = e[0] * H1 + (-e[1]) * H2;                          /* TRANSPOSE(Transfer_Matrix) */
      vs[1] = e[1] * H1 + e[0] * H2;
      Ps = P->c0sq * (rH - P->rho0) + P->background_p;
      if (sv == ORC_SOLVER_MS_MUSCL_HLLC) { Iv[0] = c->u[i]; Iv[1] = c->v[i]; Jv[0] = c->u[j]; Jv[1] = c->v[j]; }
    } else {                                                      /* modified linear */
      Pbar = (PL + PR) / 2.0;
      vb[0] = (Iv[0] + Jv[0]) / 2.0; vb[1] = (Iv[1] + Jv[1]) / 2.0;
      double dist = sqrt(xij[0] * xij[0] + xij[1] * xij[1]);
      double e[2] = {-xij[0] / dist, -xij[1] / dist};
      double UL = Iv[0] * e[0] + Iv[1] * e[1], UR = Jv[0] * e[0] + Jv[1] * e[1];
      double Ub = (UL + UR) / 2.0;
      double Us = Ub + 0.5 * (PL - PR) / (rb * sqc0);
      vs[0] = (Us - Ub) * e[0] + vb[0];
      vs[1] = (Us - Ub) * e[1] + vb[1];
      double d = UL - UR;
      double bR = 3.0 * (d > 0.0 ? d : 0.0);
      bR = bR < cb ? bR : cb;
      Ps = Pbar + 0.5 * bR * rb * (UL - UR);
    }
    double di[2] = {c->ndix[k], c->ndiy[k]}, dj[2] = {c->ndjx[k], c->ndjy[k]};
    double Ivs[2], Jvs[2], IPs, JPs;
    if (c->near[i] == 1) { Ivs[0] = vb[0]; Ivs[1] = vb[1]; IPs = Pbar; }
    else { Ivs[0] = vs[0]; Ivs[1] = vs[1]; IPs = Ps; }
    if (c->near[j] == 1) { Jvs[0] = vb[0]; Jvs[1] = vb[1]; JPs = Pbar; }
    else { Jvs[0] = vs[0]; Jvs[1] = vs[1]; JPs = Ps; }
    double qi[2] = {Iv[0] - Ivs[0] + c->avg_x[i] - c->avg_x[j], Iv[1] - Ivs[1] + c->avg_y[i] - c->avg_y[j]};
    double qj[2] = {Jv[0] - Jvs[0] + c->avg_x[j] - c->avg_x[i], Jv[1] - Jvs[1] + c->avg_y[j] - c->avg_y[i]};
    double ti = qi[0] * di[0] + qi[1] * di[1];
    double tj = -(qj[0] * dj[0] + qj[1] * dj[1]);
    c->drhodt[i] = c->drhodt[i] + 2 * c->rho[i] * c->vol[j] * ti;
    c->drhodt[j] = c->drhodt[j] + 2 * c->rho[j] * c->vol[i] * tj;
    {
      const double vsc = fabs(Iv[0]) + fabs(Iv[1]) + fabs(Jv[0]) + fabs(Jv[1]) + fabs(vs[0]) + fabs(vs[1]) +
                         fabs(c->avg_x[i]) + fabs(c->avg_y[i]) + fabs(c->avg_x[j]) + fabs(c->avg_y[j]);
      c->mag_drho[i] += 2 * c->rho[i] * c->vol[j] * (fabs(di[0]) + fabs(di[1])) * vsc;
      c->mag_drho[j] += 2 * c->rho[j] * c->vol[i] * (fabs(dj[0]) + fabs(dj[1])) * vsc;
    }
    double rij = 1.0 / (c->rho[i] * c->rho[j]);
    double* iax[2] = {&c->ia_x[i], &c->ia_y[i]};
    double* jax[2] = {&c->ia_x[j], &c->ia_y[j]};
    double* imag[2] = {&c->mag_ax[i], &c->mag_ay[i]};
    double* jmag[2] = {&c->mag_ax[j], &c->mag_ay[j]};
    const double psc = fabs(PL) + fabs(PR) + fabs(Ps) +
                       rb * cb * (fabs(Iv[0]) + fabs(Iv[1]) + fabs(Jv[0]) + fabs(Jv[1]));
    for (int L = 0; L < 2; ++L) {
      double tpi = -2 * IPs * di[L] * rij;
      double tpj = -2 * JPs * dj[L] * rij;
      *iax[L] = *iax[L] + c->mass[j] * tpi;
      *jax[L] = *jax[L] - c->mass[i] * tpj;
      *imag[L] += c->mass[j] * 2 * psc * fabs(di[L]) * rij;
      *jmag[L] += c->mass[i] * 2 * psc * fabs(dj[L]) * rij;
    }
    double hb = c->ph_[k];
    double vij[2] = {c->pvx[k], c->pvy[k]};
    double r2 = xij[0] * xij[0] + xij[1] * xij[1];
    double vx = vij[0] * xij[0] + vij[1] * xij[1];
    if (vx < 0.0) {
      double fhb = P->fai * hb;
      double phi = hb * vx / (r2 + fhb * fhb);
      double Pi = (-(P->alpha * cb * phi) + P->beta * (phi * phi)) / rb;
      double* iav[2] = {&c->av_x[i], &c->av_y[i]};
      double* jav[2] = {&c->av_x[j], &c->av_y[j]};
      for (int L = 0; L < 2; ++L) {
        double tvi = -Pi * di[L];
        double tvj = -Pi * dj[L];
        *iav[L] = *iav[L] + c->mass[j] * tvi;
        *jav[L] = *jav[L] - c->mass[i] * tvj;
        *imag[L] += fabs(c->mass[j] * tvi);
        *jmag[L] += fabs(c->mass[i] * tvj);
      }
    }
    double Ri = xij[0] * di[0] + xij[1] * di[1];
    double Rj = xij[0] * dj[0] + xij[1] * dj[1];
    c->lap_x[i] = c->lap_x[i] + c->vol[j] * 2 * P->nu * Ri / r2 * vij[0];
    c->lap_y[i] = c->lap_y[i] + c->vol[j] * 2 * P->nu * Ri / r2 * vij[1];
    c->lap_x[j] = c->lap_x[j] - c->vol[i] * 2 * P->nu * Rj / r2 * vij[0];
    c->lap_y[j] = c->lap_y[j] - c->vol[i] * 2 * P->nu * Rj / r2 * vij[1];
    c->mag_ax[i] += fabs(c->vol[j] * 2 * P->nu * Ri / r2 * vij[0]);
    c->mag_ay[i] += fabs(c->vol[j] * 2 * P->nu * Ri / r2 * vij[1]);
    c->mag_ax[j] += fabs(c->vol[i] * 2 * P->nu * Rj / r2 * vij[0]);
    c->mag_ay[j] += fabs(c->vol[i] * 2 * P->nu * Rj / r2 * vij[1]);
    double tmp = P->delta_alpha * hb * cb * vx / r2;
    double* iai[2] = {&c->ai_x[i], &c->ai_y[i]};
    double* jai[2] = {&c->ai_x[j], &c->ai_y[j]};
    for (int L = 0; L < 2; ++L) {
      *iai[L] = *iai[L] + tmp * di[L] * P->rho0 / c->rho[i] * c->vol[j];
      *jai[L] = *jai[L] - tmp * dj[L] * P->rho0 / c->rho[j] * c->vol[i];
      *imag[L] += fabs(tmp * di[L] * P->rho0 / c->rho[i] * c->vol[j]);
      *jmag[L] += fabs(tmp * dj[L] * P->rho0 / c->rho[j] * c->vol[i]);
    }
    if (P->sps_on) {                                  /* HLLC_Riemann_Delta_SPS.f90:260-282 */
      double d1 = c->pdwx[k], d2 = c->pdwy[k];
      double tp = 0.0;
      tp = tp + (c->txx[i] + c->txx[j]) * d1;
      tp = tp + (c->txy[i] + c->txy[j]) * d2;
      c->sps_x[i] = c->sps_x[i] + c->gmass[j] * tp * rij;
      c->sps_x[j] = c->sps_x[j] - c->gmass[i] * tp * rij;
      c->mag_ax[i] += fabs(c->gmass[j] * tp * rij);
      c->mag_ax[j] += fabs(c->gmass[i] * tp * rij);
      tp = tp + (c->tyy[i] + c->tyy[j]) * d2 + (c->txy[i] + c->txy[j]) * d1;
      c->sps_y[i] = c->sps_y[i] + c->gmass[j] * tp * rij;
      c->sps_y[j] = c->sps_y[j] - c->gmass[i] * tp * rij;
      c->mag_ay[i] += fabs(c->gmass[j] * tp * rij);
      c->mag_ay[j] += fabs(c->gmass[i] * tp * rij);
    }
  }
  for (int64_t k = 0; k < c->npair; ++k) {                       /* delta-SPH density correction */
    int i = c->pi[k], j = c->pj[k];
    double rd = -c->prd[k];
    double hb = c->ph_[k], cb = c->pc_[k];
    double pd[2] = {-c->pdx[k], -c->pdy[k]};
    double r2 = pd[0] * pd[0] + pd[1] * pd[1];
    double sg[2] = {c->grx[i] + c->grx[j], c->gry[i] + c->gry[j]};
    double psi = rd - 0.5 * (sg[0] * pd[0] + sg[1] * pd[1]);
    double ti = (pd[0] * c->ndix[k] + pd[1] * c->ndiy[k]) / r2;
    double tj = (pd[0] * c->ndjx[k] + pd[1] * c->ndjy[k]) / r2;
    c->corr[i] = c->corr[i] + 2 * P->delta * hb * cb * psi * ti * c->vol[j];
    c->corr[j] = c->corr[j] - 2 * P->delta * hb * cb * psi * tj * c->vol[i];
    const double psc = fabs(rd) + 0.5 * (fabs(sg[0] * pd[0]) + fabs(sg[1] * pd[1])) +
                       0.5 * (c->mag_grho[i] + c->mag_grho[j]) * (fabs(pd[0]) + fabs(pd[1]));
    c->mag_drho[i] += fabs(2 * P->delta * hb * cb * psc * ti * c->vol[j]);
    c->mag_drho[j] += fabs(2 * P->delta * hb * cb * psc * tj * c->vol[i]);
  }
}

/* Calculate_Body_Force.f90:51-142 (laplacianTerm = 1) and its sums in                  */
/* Single_step_compute.f90:219-249: fluid particles paired with a body particle j get  */
/* -V_j (p_i + p_j) grad W_i and rho_i V_j 2 nu (x_ij . grad W_i) / r^2 v_ij; the force */
/* is the sum of V_i times those (and V_i rho_i g for the total) over those particles.  */
static void ms_body_force(orc_ctx* c) {
  const orc_params* P = &c->P;
  const int n = c->n_local;
  memset(c->near_body, 0, sizeof(int) * n);
  size_t nb = sizeof(double) * n;
  memset(c->bp_x, 0, nb); memset(c->bp_y, 0, nb); memset(c->bv_x, 0, nb); memset(c->bv_y, 0, nb);
  for (int64_t k = 0; k < c->npair; ++k) {
    int i = c->pi[k], j = c->pj[k];
    const int bi = c->ginit[c->name[i]] == -4, bj = c->ginit[c->name[j]] == -4;
    if (!bi && !bj) continue;
    double vij[2] = {c->pvx[k], c->pvy[k]}, xij[2] = {c->pdx[k], c->pdy[k]};
    double r2 = xij[0] * xij[0] + xij[1] * xij[1];
    if (bj) {
      c->near_body[i] = 1;
      c->bp_x[i] = c->bp_x[i] - c->vol[j] * (c->p[i] + c->p[j]) * c->ndix[k];
      c->bp_y[i] = c->bp_y[i] - c->vol[j] * (c->p[i] + c->p[j]) * c->ndiy[k];
      double R = xij[0] * c->ndix[k] + xij[1] * c->ndiy[k];
      c->bv_x[i] = c->bv_x[i] + c->rho[i] * c->vol[j] * 2 * P->nu * R / r2 * vij[0];
      c->bv_y[i] = c->bv_y[i] + c->rho[i] * c->vol[j] * 2 * P->nu * R / r2 * vij[1];
    }
    if (bi) {
      c->near_body[j] = 1;
      c->bp_x[j] = c->bp_x[j] + c->vol[i] * (c->p[i] + c->p[j]) * c->ndjx[k];
      c->bp_y[j] = c->bp_y[j] + c->vol[i] * (c->p[i] + c->p[j]) * c->ndjy[k];
      double R = xij[0] * c->ndjx[k] + xij[1] * c->ndjy[k];
      c->bv_x[j] = c->bv_x[j] - c->rho[j] * c->vol[i] * 2 * P->nu * R / r2 * vij[0];
      c->bv_y[j] = c->bv_y[j] - c->rho[j] * c->vol[i] * 2 * P->nu * R / r2 * vij[1];
    }
  }
  double f[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    if (c->near_body[i] != 1) continue;
    double ex = 0.0 + c->rho[i] * 0.0, ey = 0.0 + c->rho[i] * (-P->gravity);
    f[0] = f[0] + c->vol[i] * c->bp_x[i];
    f[1] = f[1] + c->vol[i] * c->bp_y[i];
    f[2] = f[2] + c->vol[i] * c->bv_x[i];
    f[3] = f[3] + c->vol[i] * c->bv_y[i];
    f[4] = f[4] + c->vol[i] * (c->bp_x[i] + c->bv_x[i] + ex);
    f[5] = f[5] + c->vol[i] * (c->bp_y[i] + c->bv_y[i] + ey);
  }
  memcpy(c->bforce, f, sizeof(f));
}

void orc_ms_body_force(const orc_ctx* c, double* out) { memcpy(out, c->bforce, sizeof(c->bforce)); }

void orc_ms_single_step_compute(orc_ctx* c, int32_t i_time_step, double t) {
  const int n = c->n_local;
  c->t_eval = t;
  for (int i = 0; i < n; ++i) {                                   /* :35-45 EOS */
    if (c->type[i] == 120) continue;
    orc_status_equation(c, c->rho[i], c->e[i], c->type[i], &c->p[i], &c->c[i]);
  }
  memset(c->vol, 0, sizeof(double) * n);
  assign_boundary_particles(c);                                  /* :49 */
  find_particle_pair(c);                                         /* :62 */
  define_freesurface(c);                                         /* :74 */
  size_t nb = sizeof(double) * n;                                /* :80-93 */
  memset(c->drhodt, 0, nb); memset(c->dedt, 0, nb);
  memset(c->ia_x, 0, nb); memset(c->ia_y, 0, nb);
  memset(c->ai_x, 0, nb); memset(c->ai_y, 0, nb);
  memset(c->av_x, 0, nb); memset(c->av_y, 0, nb);
  memset(c->ext_x, 0, nb); memset(c->ext_y, 0, nb);
  memset(c->corr, 0, nb);
  memset(c->mag_drho, 0, nb); memset(c->mag_ax, 0, nb); memset(c->mag_ay, 0, nb);
  memset(c->lap_x, 0, nb); memset(c->lap_y, 0, nb);
  memset(c->sps_x, 0, nb); memset(c->sps_y, 0, nb);
  ms_solver(c);                                                  /* :111-140 */
  if (c->P.sample_every > 0 && (i_time_step % c->P.sample_every) == 0) ms_body_force(c);   /* :219-249 */
  memset(c->ax, 0, nb); memset(c->ay, 0, nb);                    /* :253-283 */
  for (int i = 0; i < n; ++i) {
    if (c->type[i] != 2) continue;
    c->ext_y[i] = c->ext_y[i] + (-c->P.gravity);
    c->mag_ay[i] += fabs(c->ext_y[i]);
    c->ax[i] = c->ia_x[i] + c->ext_x[i] + c->av_x[i] + c->ai_x[i] + c->sps_x[i] + c->lap_x[i];
    c->ay[i] = c->ia_y[i] + c->ext_y[i] + c->av_y[i] + c->ai_y[i] + c->sps_y[i] + c->lap_y[i];
    c->dedt[i] = c->dedt[i] + 0.0 + 0.0;
    c->drhodt[i] = c->drhodt[i] + c->corr[i];
  }
}

/* MS Reinitialize_particle_density.f90:29-73: order 1 from the pairs (same-type pairs, */
/* sum_w of the free-surface routine; every other density 0), or order 3 per fluid      */
/* particle over its candidates within 3h (no lower bound: the particle itself counts), */
/* W weighted by V_j, MLS_Interpolation_In_Subdomain_Domain (SPH_Subroutine_Module.f90  */
/* :2565-2792: basis (1, dx/h, dy/h), alpha rows, the 3x3 inverse of                    */
/* matrix_inversion_less_three, Shepard when singular); NaN keeps the old density;      */
/* other particles keep theirs.                                                          */
static void ms_reinitialize_density(orc_ctx* c) {
  const orc_params* P = &c->P;
  const int n = c->n_local;
  memcpy(c->mls_rho, c->rho, sizeof(double) * n);
  memset(c->rho, 0, sizeof(double) * n);
  if (P->mls_order == 1) {
    for (int64_t k = 0; k < c->npair; ++k) {
      int i = c->pi[k], j = c->pj[k];
      if (c->type[i] == c->type[j]) {
        if (c->type[i] == 2) c->rho[i] = c->rho[i] + c->mass[j] * c->pw[k] / c->sum_w[i];
        if (c->type[j] == 2) c->rho[j] = c->rho[j] + c->mass[i] * c->pw[k] / c->sum_w[j];
      }
    }
    return;
  }
  int nm[300]; double wk[300], p2[300], p3[300];
  const double al = P->mls_alpha;
  for (int i = 0; i < n; ++i) {
    if (c->type[i] != 2) { c->rho[i] = c->mls_rho[i]; continue; }
    int ne = 0;
    for (int L = 0; L < c->cand_n[i]; ++L) {
      int j = c->cand[(size_t)i * 300 + L];
      double pd[2] = {c->x[i] - c->x[j], c->y[i] - c->y[j]};
      double r = sqrt(pd[0] * pd[0] + pd[1] * pd[1]);
      double hb = (c->hh[i] + c->hh[j]) / 2.0;
      if (r <= 3.0 * hb) {
        double w, dw[2];
        orc_compute_kernel(2, r, pd, hb, &w, dw);
        nm[ne] = j; wk[ne] = w * c->vol[j];
        p2[ne] = (c->x[j] - c->x[i]) / c->h; p3[ne] = (c->y[j] - c->y[i]) / c->h;
        ++ne;
      }
    }
    /* A = P W P^T + alpha^3 I (the alpha columns come last), column-major */
    double A[9] = {0};
    for (int k = 0; k < ne; ++k) {
      double pk[3] = {1.0, p2[k], p3[k]};
      for (int m = 0; m < 3; ++m)
        for (int l = 0; l < 3; ++l) A[m + 3 * l] = A[m + 3 * l] + pk[m] * wk[k] * pk[l];
    }
    for (int l = 0; l < 3; ++l) A[l + 3 * l] = A[l + 3 * l] + (al * al) * al;
    double Ai[9];
    int ok = orc_matrix_inversion_less_three(3, A, Ai);
    double rho = 0.0;
    if (ok) {
      for (int k = 0; k < ne; ++k) {
        double b[3] = {1.0 * wk[k], p2[k] * wk[k], p3[k] * wk[k]};
        double fai = Ai[0] * b[0] + Ai[3] * b[1] + Ai[6] * b[2];
        rho = rho + c->mls_rho[nm[k]] * fai;
      }
    } else {
      double sw = 0.0;
      for (int k = 0; k < ne; ++k) sw = sw + wk[k];
      for (int k = 0; k < ne; ++k) rho = rho + c->mls_rho[nm[k]] * (wk[k] / sw);
    }
    c->rho[i] = isnan(rho) ? c->mls_rho[i] : rho;
  }
}

void orc_ms_time_step(orc_ctx* c, int32_t it) {
  const orc_params* P = &c->P;
  const double dt = P->dt;
  c->real_time = c->real_time + dt;                              /* :58-62 */
  orc_ms_body_motion(c, c->real_time);                           /* :113-115 */
  for (int i = 0; i < c->n_fluid; ++i) {                         /* Reorder_Particle_Index.f90:80-95 */
    if (c->gtype[i] != 2) continue;
    double x = c->gx[i], y = c->gy[i];
    if (x < (double)-1.0E-10f || x > P->box_x || y < (double)-1.0E-10f || y > P->box_y) c->gtype[i] = 120;
  }
  orc_devide_block(c, it);
  const int n = c->n_local;
  save_before(c);                                                /* :313-317 */
  if (it == 1) orc_ms_single_step_compute(c, it, c->real_time);
  size_t nb = sizeof(double) * n;                                /* :321-324 */
  memcpy(c->t_drho, c->drhodt, nb); memcpy(c->t_de, c->dedt, nb);
  memcpy(c->t_ax, c->ax, nb); memcpy(c->t_ay, c->ay, nb);
  memcpy(c->t_u, c->u, nb); memcpy(c->t_v, c->v, nb);
  for (int i = 0; i < n; ++i) {                                  /* :326-333 */
    if (c->type[i] != 2) continue;
    c->rho[i] = c->b_rho[i] + c->drhodt[i] * dt;
    c->e[i] = c->b_e[i] + c->dedt[i] * dt;
    c->x[i] = c->b_x[i] + dt * (c->u[i] + c->avg_x[i]) + 0.5 * c->ax[i] * (dt * dt);
    c->y[i] = c->b_y[i] + dt * (c->v[i] + c->avg_y[i]) + 0.5 * c->ay[i] * (dt * dt);
    c->u[i] = c->b_u[i] + dt * c->ax[i];
    c->v[i] = c->b_v[i] + dt * c->ay[i];
  }
  orc_ms_single_step_compute(c, it, c->real_time);               /* :336-337 */
  for (int i = 0; i < n; ++i) {                                  /* :339-346 */
    if (c->type[i] != 2) continue;
    c->rho[i] = c->b_rho[i] + 0.5 * (c->drhodt[i] + c->t_drho[i]) * dt;
    c->e[i] = c->b_e[i] + 0.5 * (c->dedt[i] + c->t_de[i]) * dt;
    c->x[i] = c->b_x[i] + 0.5 * dt * (c->u[i] + c->t_u[i]) + c->avg_x[i] * dt + 0.5 * c->ax[i] * (dt * dt);
    c->y[i] = c->b_y[i] + 0.5 * dt * (c->v[i] + c->t_v[i]) + c->avg_y[i] * dt + 0.5 * c->ay[i] * (dt * dt);
    c->u[i] = c->b_u[i] + 0.5 * dt * (c->ax[i] + c->t_ax[i]);
    c->v[i] = c->b_v[i] + 0.5 * dt * (c->ay[i] + c->t_ay[i]);
  }
  if (P->mls_on && (it % P->mls_every) == 0) ms_reinitialize_density(c);   /* :508-510 */
  if (P->shift_on) orc_shift(c);                                 /* :513-515 */
  orc_refresh(c);                                                /* :518 */
}
