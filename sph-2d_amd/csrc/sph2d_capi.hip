// sph2d_capi.hip -- C-ABI (include/sph2d.h): context, host driver of the step
// (the restated time_iteration.f90), uploads/downloads, diagnostics.
#include "sph2d_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace sph2d;

namespace {

// The message of the calling thread's last failure, in a fixed buffer: recording
// an error allocates nothing (the no-allocation entry point relies on that).
thread_local char g_err[512];

int fail(int code, const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

int fail(int code, const std::string& msg) { return fail(code, msg.c_str()); }

}  // namespace

// the message of sph2d_last_error for the other translation units
int sph2d::set_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HIP_OR_FAIL(x)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess)                                                                   \
      return fail(SPH2D_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));           \
  } while (0)

constexpr double kPiModule = 3.14159265358979324;  // information_module.f90:29

}  // namespace

struct AdjTape;
void tape_free(AdjTape* t);

struct sph2d_ctx {
  sph2d_params prm{};
  DevParams P{};
  Buffers B{};
  hipStream_t stream = nullptr;
  int cap_fluid = 0, cap_ghost = 0, cap_global = 0, cap_local = 0, n_cells = 0;
  int n_fluid = 0;
  int n_wm = 0;                     // wavemaker particles (after the fluid)
  double ke0 = 0.0;
  bool have_state = false;
  bool debug = false;
  bool evaluated = false;
  // The reverse evaluation needs the local order, state, masks and flags of the last
  // sph2d_single_step_compute unchanged; every entry that rewrites any of them clears this.
  bool vjp_ready = false;
  // the evaluation routine by routine: 0 none since the last state change, 1 pairs
  // found, 2 free surface defined, 3 solved
  int eval_part = 0;
  bool mls_pending = false;
  double* d_out = nullptr;          // 16 doubles for reductions
  std::vector<void*> allocs;
  std::vector<size_t> alloc_bytes;  // bytes of each entry of allocs
  // in-step pass timing: 4 events per evaluation (pass A start/end, force pass start/end)
  int ana_step = 0, vort_step = 0;   // last steps whose analytic fields / vorticity are held
  // post-processing grid (allocated on first use)
  double* grid_out = nullptr;
  size_t grid_cap = 0;
  int* surf_gid = nullptr;
  double2* surf_nrm = nullptr;
  size_t surf_cap = 0;
  // sampling points and wave probes (allocated on first use, in doubles)
  double* probe_buf = nullptr;
  size_t probe_cap = 0;
  // reverse evaluation: weights, per-local adjoints and the by-gid result (first use)
  double* adj_buf = nullptr;
  // sph2d_forward's checkpoints and sph2d_backward's work space (first use)
  AdjTape* tape = nullptr;
  // wave-theory wavemaker and its damping (relaxation) zone (sph2d_relaxation_init)
  RelaxArgs relax{};
  bool relax_ready = false, relax_on = false;
  bool pass_timing = false;
  // the moving-square program: Real_time, the body centre, velocity and acceleration,
  // the step of the last body force and its device totals
  double ms_time = 0.0, body_x = 0.0, body_y = 0.0, cust_u = 0.0, cust_a = 0.0;
  int bforce_step = 0;
  double* d_bforce = nullptr;
  std::vector<hipEvent_t> events;
  size_t events_used = 0;

  template <class T>
  int alloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fail(SPH2D_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    // Zeroed on the context's stream: a null-stream hipMemset is not ordered with the
    // non-blocking stream's kernels and could land after them.
    e = hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), stream);
    if (e != hipSuccess) return fail(SPH2D_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    allocs.push_back(q);
    alloc_bytes.push_back(std::max<size_t>(n, 1) * sizeof(T));
    *p = static_cast<T*>(q);
    return SPH2D_OK;
  }
};

namespace {

// host-side kernel value (compute_kernel.f90:123-129) for W(dx, h) of the shift (:56-68)
double host_w(double r, double h) {
  const float c4 = 1.0f - 10.0f * expf(-9.0f);
  const double factor = 1.0 / ((double)3.14159265358f * (h * h) * (double)c4);
  const double q = r / h;
  if (r <= 3.0 * h) return (std::exp(-(q * q)) - (double)expf(-9.0f)) * factor;
  return 0.0;
}

void make_devparams(const sph2d_params& p, DevParams& P) {
  P = DevParams{};
  P.box_x = p.box_x; P.box_y = p.box_y;
  P.rho0 = p.rho0; P.c0sq = p.c0sq; P.sqrt_c0sq = std::sqrt(p.c0sq);
  P.eos_b = p.c0sq * p.rho0 / 7.0;
  P.nu = p.nu; P.U = p.U;
  P.dx = p.dx; P.h = p.h; P.dt = p.dt;
  P.three_h = 3 * p.h;
  const float c4 = 1.0f - 10.0f * expf(-9.0f);
  P.kfactor = 1.0 / ((double)3.14159265358f * (p.h * p.h) * (double)c4);
  P.e9 = (double)expf(-9.0f);
  P.r_min = p.ms ? (double)1.0E-10f : (double)1.0E-7f;   // MS Find_particle_pair.f90:73-75
  P.support_volume = kPiModule * ((3 * p.h) * (3 * p.h));
  P.fra_limit = (double)0.65f;
  P.chain_d = p.ms ? 4 * p.dx : (3 + 1) * p.h;      // MS Information_Module.f90:189: 4 interior_dx
  P.chain_ox = -(5 * P.chain_d);
  P.chain_oy = -(5 * P.chain_d);
  P.chain_nx = (int)(p.box_x / P.chain_d + 2 * 5);
  P.chain_ny = (int)(p.box_y / P.chain_d + 2 * 5);
  P.alpha = p.alpha; P.beta = p.beta; P.fai = p.fai;
  P.delta = p.delta; P.delta_alpha = p.delta_alpha; P.epsilon_xsph = p.epsilon_xsph;
  P.shift_beta = p.shift_beta; P.shift_R = p.shift_R; P.max_shift = p.max_shift;
  P.w_dx = host_w(p.dx, p.h);
  P.h2 = p.h * p.h;
  P.sps_dl = (double)(sqrtf(2.0f) / 2.0f) * p.dx;
  P.sps_c012 = (double)0.12f;
  P.sps_c23 = (double)(2.0f / 3.0f);
  {
    volatile float a = 2.0f / 3.0f, b = 0.0066f;
    P.sps_c23b = (double)(float)(a * b);
  }
  P.mls_alpha = p.mls_alpha;
  P.walls = p.walls;
  P.wm_layers = p.wavemaker_layers;
  P.g = p.gravity;
  P.wall_h = (double)1.2f * (p.h + p.h) / 2.0;     // Assign_boundary_particles.f90:182
  P.wall_kfactor = 1.0 / ((double)3.14159265358f * (P.wall_h * P.wall_h) * (double)c4);
  P.vol0 = p.dx * (p.box_y / p.ny);                // intial_volume_0 = interior_dx*interior_dy
  P.damping = p.damping;
  P.damp_len = p.damping_length;
  P.damp_amp = p.damping_amplitude;
  P.water_depth = p.water_depth;
  P.e32 = (double)expf(1.0f);
  P.surf_lim = (double)1.414f * p.h;
  P.inv_h = 1.0 / p.h;
  P.inv_c0 = 1.0 / std::sqrt(p.c0sq);
  P.inv_sv = 1.0 / P.support_volume;
  P.inv_wdx = 1.0 / P.w_dx;
  // any pair with r <= 3h has r^2 below this bound, so the exact test r <= 3h stays decisive
  P.cut2 = P.three_h * P.three_h * (1.0 + 1e-10);
  {
    // FP32 coordinates are within 2^-24*|x| of the FP64 ones; positions stay within
    // the box widened by the ghost band, so each FP32 offset is within 4 ulp of that span
    const double span = std::max(p.box_x, p.box_y) + 8 * p.h;
    const double eps = 4.0 * span * std::ldexp(1.0, -24);
    const double bound = (P.three_h + 2 * eps) * (P.three_h + 2 * eps) * (1.0 + 1e-6);
    P.cut2_32 = std::nextafter((float)bound, 1e30f);
  }
  {
    // The reference keeps a pair when r = sqrt(dx*dx+dy*dy) satisfies 1e-7 <= r <= 3h.
    // sqrt is correctly rounded and monotone, so these are bounds on r2 itself.
    double x = P.three_h * P.three_h;
    while (std::sqrt(x) > P.three_h) x = std::nextafter(x, 0.0);
    while (std::sqrt(std::nextafter(x, 1e300)) <= P.three_h) x = std::nextafter(x, 1e300);
    P.r2_max = x;
    double y = P.r_min * P.r_min;
    while (std::sqrt(y) < P.r_min) y = std::nextafter(y, 1e300);
    while (std::sqrt(std::nextafter(y, 0.0)) >= P.r_min) y = std::nextafter(y, 0.0);
    P.r2_min = y;
  }
  P.x_over_t = p.dx / p.dt;            // HLLC: interior_dx/dt
  P.solver = p.solver;
  P.mls_order = p.mls_order;
  P.write_debug = 0;
  P.mixed = p.precision == SPH2D_FP32_MIXED ? 1 : 0;
  P.sps = p.sps_on;
  P.renorm = p.renormalization;
  P.xsph_f = p.sps_on ? 2.0 : 1.0;
  P.rebin_m = 0.5 * (P.chain_d - P.three_h);
  P.ms = p.ms;
  P.pb = p.background_p;
  P.release_time = p.release_time;
  P.t_eval = 0.0;
  P.cust_u = 0.0;
  P.cust_a = 0.0;
  P.body_force = 0;
}

int check_params(const sph2d_params* p) {
  if (!p) return fail(SPH2D_ERR_ARGUMENT, "params is NULL");
  if (p->struct_size != (int32_t)sizeof(sph2d_params))
    return fail(SPH2D_ERR_ARGUMENT, "params.struct_size mismatch (header/library version)");
  if (!(p->h > 0) || !(p->dx > 0) || !(p->box_x > 0) || !(p->box_y > 0) || !(p->dt > 0))
    return fail(SPH2D_ERR_ARGUMENT, "h, dx, box and dt must be positive");
  if (p->integrator != SPH2D_INTEGRATOR_PC && p->integrator != SPH2D_INTEGRATOR_RK)
    return fail(SPH2D_ERR_UNSUPPORTED, "integrator must be 2 (predictor-corrector) or 3 (RK-variant)");
  if (p->solver < SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN || p->solver > SPH2D_SOLVER_MS_HLLC)
    return fail(SPH2D_ERR_UNSUPPORTED, "solver must be one of SPH2D_SOLVER_*");
  if (p->ms != 0 && p->ms != 1) return fail(SPH2D_ERR_ARGUMENT, "ms must be 0 or 1");
  if (!p->ms && p->solver > SPH2D_SOLVER_MUSCL)
    return fail(SPH2D_ERR_UNSUPPORTED, "the MS solvers belong to the moving-square program (ms = 1)");
  if (p->ms) {
    if (p->solver != SPH2D_SOLVER_MS_MUSCL_HLLC && p->solver != SPH2D_SOLVER_MS_HLLC &&
        p->solver != SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN)
      return fail(SPH2D_ERR_UNSUPPORTED, "ms = 1 takes SPH2D_SOLVER_MS_MUSCL_HLLC, _MS_HLLC or _MODIFIED_LINEAR_RIEMANN");
    if (p->integrator != SPH2D_INTEGRATOR_PC || !p->walls || p->periodic || p->wavemaker_layers || p->renormalization ||
        p->damping != SPH2D_DAMPING_OFF)
      return fail(SPH2D_ERR_UNSUPPORTED, "ms = 1 runs the predictor-corrector in a walled tank without wavemaker, "
                                         "damping or renormalisation");
    if (p->mls_on && p->mls_order != 3) return fail(SPH2D_ERR_UNSUPPORTED, "ms = 1 reinitialises with MLS order 3");
    if (!(p->body_len > 0) || !(p->body_ramp > 0) || !std::isfinite(p->body_speed))
      return fail(SPH2D_ERR_ARGUMENT, "body_len and body_ramp must be positive");
    if (4 * p->dx < 3 * p->h) return fail(SPH2D_ERR_ARGUMENT, "ms = 1 needs 3 h <= 4 dx (the 4 dx cells)");
  }
  if (p->precision != SPH2D_FP64 && p->precision != SPH2D_FP32_MIXED)
    return fail(SPH2D_ERR_UNSUPPORTED, "precision must be SPH2D_FP64 or SPH2D_FP32_MIXED");
  if (p->rate_reuse != SPH2D_RATE_REUSE_REFERENCE)
    return fail(SPH2D_ERR_UNSUPPORTED, "only the reference's rate re-use is built");
  if (!p->ms && p->mls_on && p->mls_order != 6)
    return fail(SPH2D_ERR_UNSUPPORTED, "only MLS order 6 (the reference default) is built");
  if (p->walls < 0 || p->walls > 64) return fail(SPH2D_ERR_ARGUMENT, "walls must be in [0, 64]");
  if (!std::isfinite(p->gravity) || p->gravity < 0) return fail(SPH2D_ERR_ARGUMENT, "gravity must be >= 0");
  if (p->wavemaker_layers < 0 || p->wavemaker_layers > 64)
    return fail(SPH2D_ERR_ARGUMENT, "wavemaker_layers must be in [0, 64]");
  if (p->wavemaker_layers && !p->walls) return fail(SPH2D_ERR_ARGUMENT, "wavemaker particles need fixed walls");
  if (p->periodic && p->walls) return fail(SPH2D_ERR_ARGUMENT, "fixed walls need periodic = 0");
  if (!p->periodic && !p->walls)
    return fail(SPH2D_ERR_UNSUPPORTED, "the box is periodic or has fixed ghost walls (walls > 0)");
  if (p->mls_every <= 0 || p->sample_every <= 0)
    return fail(SPH2D_ERR_ARGUMENT, "mls_every and sample_every must be positive");
  if (p->damping < SPH2D_DAMPING_OFF || p->damping > SPH2D_DAMPING_LINEAR)
    return fail(SPH2D_ERR_ARGUMENT, "damping must be one of SPH2D_DAMPING_*");
  if (p->sps_on != 0 && p->sps_on != 1) return fail(SPH2D_ERR_ARGUMENT, "sps_on must be 0 or 1");
  if (p->renormalization != 0 && p->renormalization != 1)
    return fail(SPH2D_ERR_ARGUMENT, "renormalization must be 0 or 1");
  if (p->renormalization && (p->precision != SPH2D_FP64 || p->solver == SPH2D_SOLVER_MUSCL))
    return fail(SPH2D_ERR_UNSUPPORTED, "the 2x2 renormalisation is built for the FP64 form without MUSCL");
  if (p->damping != SPH2D_DAMPING_OFF && (!(p->damping_length > 0) || !(p->water_depth > 0)))
    return fail(SPH2D_ERR_ARGUMENT, "damping_length and water_depth must be positive");
  return SPH2D_OK;
}

int grid_n(int n) { return (n + kBlock - 1) / kBlock; }

int read_counters(sph2d_ctx* c, int* host) {
  HIP_OR_FAIL(hipMemcpyAsync(host, c->B.counters, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  if (host[C_ERR] & E_GHOST_CAP) return fail(SPH2D_ERR_CAPACITY, "periodic ghost capacity exceeded");
  return SPH2D_OK;
}

int compute_ke0(sph2d_ctx* c) {
  launch_sample(c->stream, c->P, c->B, c->cap_fluid, c->d_out);
  double out[16];
  HIP_OR_FAIL(hipMemcpyAsync(out, c->d_out, sizeof(double) * 14, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  c->ke0 = out[0];
  return SPH2D_OK;
}

// True when [p, p + bytes) lies inside one device allocation of this process.
bool device_range(const void* p, size_t bytes) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
  if (a.type != hipMemoryTypeDevice) return false;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) { (void)hipGetLastError(); return false; }
  const char* b = static_cast<const char*>(base);
  const char* q = static_cast<const char*>(p);
  return q >= b && q + bytes <= b + size;
}

struct DeviceScratch {
  std::vector<void*> ptrs;
  ~DeviceScratch() { for (void* p : ptrs) (void)hipFree(p); }
  template <typename T>
  hipError_t get(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(q);
    *p = static_cast<T*>(q);
    return e;
  }
};

}  // namespace

extern "C" {

const char* sph2d_last_error(void) { return g_err; }

int sph2d_default_params(sph2d_params* p) {
  if (!p) return fail(SPH2D_ERR_ARGUMENT, "params is NULL");
  std::memset(p, 0, sizeof(*p));
  p->struct_size = sizeof(sph2d_params);
  p->nx = 100; p->ny = 100;
  p->box_x = 1.0; p->box_y = 1.0;
  p->rho0 = 1000.0;
  p->c0sq = 100.0;
  p->nu = (double)1.0E-2f;
  p->U = 1.0;
  p->dx = 1.0 / 100;
  p->h = 1.20 * p->dx;
  p->dt = 0.0001;
  p->integrator = SPH2D_INTEGRATOR_RK;
  p->solver = SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN;
  p->precision = SPH2D_FP64;
  p->rate_reuse = SPH2D_RATE_REUSE_REFERENCE;
  p->shift_on = 1;
  p->mls_on = 1;
  p->mls_every = 20;
  p->mls_order = 6;
  p->sample_every = 100;
  p->periodic = 1;
  p->alpha = 0.0; p->beta = 2.0; p->fai = 0.1;
  p->delta = 0.1; p->delta_alpha = 0.02;
  p->epsilon_xsph = 0.0;
  p->shift_beta = 0.5; p->shift_R = 0.2;
  p->max_shift = 0.25 * p->dx;
  p->mls_alpha = (double)0.0001f;
  p->damping = SPH2D_DAMPING_OFF;                  // make_wave_or_not = 0 in the TGV case
  p->damping_length = 3.0;
  p->damping_amplitude = (double)0.3f;
  p->water_depth = p->box_y;                       // water_depth = domain_size_y
  p->walls = 0;                                    // DistributeFixedBoundaryParticleOrNot = 0
  p->gravity = 0.0;                                // Gravity_Switch = 'off'
  p->wavemaker_layers = 0;                         // make_wave_or_not = 0
  p->sps_on = 1;                                   // Solver_..._SPS (single_step_compute.f90:137)
  p->renormalization = 0;                          // the identity (:168-188)
  p->ms = 0;                                       // the Taylor-Green program
  p->background_p = 0.0;                           // MS Information_Module.f90:486-487
  p->release_time = 10.0;                          // :715
  p->body_cx = 2.5; p->body_cy = 2.5; p->body_len = 1.0;
  p->body_speed = 1.0; p->body_ramp = 0.5;
  return SPH2D_OK;
}

int wavemaker_count(const sph2d_params& p);

int sph2d_ctx_create(const sph2d_params* p, int32_t n_fluid_capacity, sph2d_ctx** out) {
  if (!out) return fail(SPH2D_ERR_ARGUMENT, "out is NULL");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  if (n_fluid_capacity <= 0) return fail(SPH2D_ERR_ARGUMENT, "capacity must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SPH2D_ERR_HIP, "no HIP device: the sph2d library runs on the GPU only");
  sph2d_ctx* c = new sph2d_ctx();
  c->prm = *p;
  make_devparams(*p, c->P);
  // periodic ghosts: 2*(m+1) layers along every edge, m = 3h/dx (plus corners), x3 margin
  const int m = (int)std::ceil(3 * p->h / p->dx) + 1;
  const long ghost_bound = 2L * m * (long)(p->nx + p->ny + 4 * m);
  c->cap_fluid = n_fluid_capacity;
  c->cap_ghost = (int)std::max(16384L, 3 * ghost_bound);
  if (p->walls) {
    const int nw = p->ms ? ms_ghost_layout(p->box_x, p->box_y, p->dx, p->box_y / p->ny, p->walls, p->body_cx,
                                           p->body_cy, p->body_len, nullptr, nullptr, nullptr, nullptr)
                         : wall_layout(p->box_x, p->box_y, p->dx, p->box_y / p->ny, p->walls, nullptr, nullptr, nullptr);
    c->cap_ghost = std::max(c->cap_ghost, wavemaker_count(*p) + nw + 64);
  }
  c->cap_global = c->cap_fluid + c->cap_ghost;
  c->cap_local = c->cap_global;
  c->n_cells = c->P.chain_nx * c->P.chain_ny;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(SPH2D_ERR_HIP, "hipStreamCreate failed"); }
  Buffers& B = c->B;
  const size_t G = c->cap_global, L = c->cap_local, F = c->cap_fluid, H = c->cap_ghost;
  rc = 0;
  rc |= c->alloc(&B.g_pos, G); rc |= c->alloc(&B.g_vel, G);
  rc |= c->alloc(&B.g_rho, G); rc |= c->alloc(&B.g_mass, G); rc |= c->alloc(&B.g_p, G);
  rc |= c->alloc(&B.g_c, G); rc |= c->alloc(&B.g_e, G);
  rc |= c->alloc(&B.g_type, G); rc |= c->alloc(&B.g_key, G); rc |= c->alloc(&B.g_slot, G);
  rc |= c->alloc(&B.ana, 4 * F); rc |= c->alloc(&B.g_vort, F);
  rc |= c->alloc(&B.g_btype, G);
  rc |= c->alloc(&B.gblk, 2 * ((F + 255) / 256) + 2);
  rc |= c->alloc(&B.cnt_yx, H); rc |= c->alloc(&B.off_yx, H);
  rc |= c->alloc(&B.src_x, H);
  rc |= c->alloc(&B.cell_count, c->n_cells); rc |= c->alloc(&B.cell_start, c->n_cells + 1);
  rc |= c->alloc(&B.cell_members, G); rc |= c->alloc(&B.l_of_g, G);
  rc |= c->alloc(&B.name, L); rc |= c->alloc(&B.l_type, L); rc |= c->alloc(&B.l_cell, L);
  rc |= c->alloc(&B.l_pos, L); rc |= c->alloc(&B.l_vel, L); rc |= c->alloc(&B.l_rv, L);
  rc |= c->alloc(&B.l_pc, L); rc |= c->alloc(&B.l_pos32, L);
  rc |= c->alloc(&B.l_mass, L);
  rc |= c->alloc(&B.fra_v, L); rc |= c->alloc(&B.sdw, L); rc |= c->alloc(&B.grad, L);
  rc |= c->alloc(&B.nrm, L); rc |= c->alloc(&B.shift, L); rc |= c->alloc(&B.avg, L);
  rc |= c->alloc(&B.tau, 3 * L); rc |= c->alloc(&B.surf, L); rc |= c->alloc(&B.near_, L);
  if (p->renormalization) rc |= c->alloc(&B.tinv, 4 * L);
  rc |= c->alloc(&B.nb_list, (size_t)kNbListWords * L);
  rc |= c->alloc(&B.nb_count, L);
  rc |= c->alloc(&B.drhodt, L); rc |= c->alloc(&B.acc, L);
  rc |= c->alloc(&B.b_rho, L); rc |= c->alloc(&B.t_drho, L); rc |= c->alloc(&B.k_drho, 3 * L);
  rc |= c->alloc(&B.b_pos, L); rc |= c->alloc(&B.b_vel, L);
  rc |= c->alloc(&B.t_acc, L); rc |= c->alloc(&B.k_vel, 3 * L); rc |= c->alloc(&B.k_acc, 3 * L);
  rc |= c->alloc(&B.mls_rho, L);
  rc |= c->alloc(&B.scan_blocks, (std::max<size_t>(G, (size_t)c->n_cells + 1) + 1023) / 1024 + 16);
  rc |= c->alloc(&B.counters, C_COUNT);
  rc |= c->alloc(&B.red, (size_t)grid_n(c->cap_fluid) * 14 + 16);
  rc |= c->alloc(&B.pair_count, L); rc |= c->alloc(&B.pair_off, L);
  rc |= c->alloc(&c->d_out, 16);
  if (c->P.mixed) {
    rc |= c->alloc(&B.m_r0, L); rc |= c->alloc(&B.m_r1, L); rc |= c->alloc(&B.m_r2, L);
    rc |= c->alloc(&B.m_pc, L);
  }
  B.mg = nullptr; B.m_mg = nullptr;
  if (p->solver == SPH2D_SOLVER_MUSCL) {
    rc |= c->alloc(&B.mg, 3 * L);
    if (c->P.mixed) rc |= c->alloc(&B.m_mg, 3 * L);
  }
  B.dbg = nullptr; B.pair_lo = nullptr; B.pair_hi = nullptr;
  B.kstride = (int)L;
  B.g_init = nullptr; B.bp = nullptr; B.bv = nullptr; B.near_body = nullptr; B.bpart = nullptr;
  if (p->ms) {
    if (!B.mg) rc |= c->alloc(&B.mg, 3 * L);
    rc |= c->alloc(&B.g_init, G);
    rc |= c->alloc(&B.bp, L); rc |= c->alloc(&B.bv, L); rc |= c->alloc(&B.near_body, L);
    rc |= c->alloc(&B.bpart, 6 * (size_t)grid_n((int)L) + 6);
    rc |= c->alloc(&c->d_bforce, 6);
  }
  if (rc) {
    std::string msg = g_err;
    sph2d_ctx_destroy(c);
    return fail(SPH2D_ERR_HIP, "allocation failed: " + msg);
  }
  *out = c;
  return SPH2D_OK;
}

int sph2d_ctx_destroy(sph2d_ctx* c) {
  if (!c) return SPH2D_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  tape_free(c->tape);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SPH2D_OK;
}

// The counters are set by kernels, not by a host copy: nothing host-side must be
// kept alive, and every later kernel on the stream sees them in order.
static int set_nfluid(sph2d_ctx* c, int n) {
  launch_fill_i32(c->stream, c->B.counters, C_COUNT, 0);
  launch_fill_i32(c->stream, c->B.counters + C_NFLUID, 1, n);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  c->n_fluid = n;
  return SPH2D_OK;
}

// Distribute_Initial_Particle.f90:254-311 (PistonType): wave_maker_layer columns at
// x = -(j-1) dx - dx/2 for the rows y = (i-1/2) dy <= boundary_size_y, rows outer,
// with rho = rho0 (-6.0 g (y - Ly)/c0^2 + 1.0)**(1.0/6.0) (REAL(4) literals, g = 9.81),
// p from the Tait law with exponent 7.0, at rest, type 130, mass V0 rho(q) with q the
// particle's own index into the fluid's densities (:300-310); the layer goes to btype.
static void wavemaker_layout(const sph2d_params& p, const double* fluid_rho, std::vector<double2>& pos,
                             std::vector<double>& rho, std::vector<double>& pr, std::vector<double>& mass,
                             std::vector<int>& layer) {
  const double dy = p.box_y / p.ny;
  for (int i = 1; i <= p.ny + 200; ++i)
    for (int j = 1; j <= p.wavemaker_layers; ++j) {
      const double x = -(j - 1) * p.dx - p.dx / 2.0, y = (i - 1) * dy + dy / 2.0;
      if (!(y <= p.box_y)) continue;
      const double r = p.rho0 * std::pow(-6.0 * 9.81 * (y - p.box_y) / p.c0sq + 1.0, (double)(1.0f / 6.0f));
      pos.push_back(make_double2(x, y));
      rho.push_back(r);
      pr.push_back(p.c0sq * p.rho0 * (std::pow(r / p.rho0, 7.0) - 1.0) / 7.0);
      layer.push_back(j);
    }
  for (size_t q = 0; q < pos.size(); ++q) mass.push_back(p.dx * dy * fluid_rho[q]);
}

// wave_maker_particle_number of the layout below
int wavemaker_count(const sph2d_params& p) {
  if (p.wavemaker_layers <= 0) return 0;
  const double dy = p.box_y / p.ny;
  int rows = 0;
  for (int i = 1; i <= p.ny + 200; ++i) rows += ((i - 1) * dy + dy / 2.0) <= p.box_y;
  return rows * p.wavemaker_layers;
}

// Distribute_Initial_Particle.f90:483-495: the fixed ghost layers after the n fluid
// particles and the wavemaker particles (type -3, at rest, rho0, p = 0, c0, e0; mass
// water_particle_mass_0, which the TGV branch leaves 0); both counted as the ghosts
// of every step. fluid_rho: the host densities of the fluid (NULL: read back).
static int upload_walls(sph2d_ctx* c, int n, const double* fluid_rho) {
  const sph2d_params& p = c->prm;
  const double dy = p.box_y / p.ny;
  Buffers& B = c->B;
  const int nwm = wavemaker_count(p);
  if (nwm > 0) {
    std::vector<double> frho(nwm);
    if (nwm > n) return fail(SPH2D_ERR_ARGUMENT, "fewer fluid particles than wavemaker particles");
    if (fluid_rho) std::copy(fluid_rho, fluid_rho + nwm, frho.begin());
    else {
      HIP_OR_FAIL(hipMemcpyAsync(frho.data(), B.g_rho, nwm * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    }
    std::vector<double2> pos, vel(nwm, make_double2(0.0, 0.0));
    std::vector<double> rho, pr, mass, cs(nwm, std::sqrt(p.c0sq)), e(nwm, p.c0sq / (7.0 - 1.0));
    std::vector<int> layer, t(nwm, 130);
    wavemaker_layout(p, frho.data(), pos, rho, pr, mass, layer);
    HIP_OR_FAIL(hipMemcpyAsync(B.g_pos + n, pos.data(), nwm * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_vel + n, vel.data(), nwm * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_rho + n, rho.data(), nwm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_mass + n, mass.data(), nwm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_p + n, pr.data(), nwm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_c + n, cs.data(), nwm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_e + n, e.data(), nwm * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_type + n, t.data(), nwm * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_btype + n, layer.data(), nwm * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  }
  c->n_wm = nwm;
  n += nwm;
  const int nw = wall_layout(p.box_x, p.box_y, p.dx, dy, p.walls, nullptr, nullptr, nullptr);
  if (nwm + nw > c->cap_ghost) return fail(SPH2D_ERR_CAPACITY, "wall layers exceed the ghost capacity");
  std::vector<double> x(nw), y(nw), zero(nw, 0.0), rho(nw, p.rho0), cs(nw, std::sqrt(p.c0sq)),
      e(nw, p.c0sq / (7.0 - 1.0));
  std::vector<int> bt(nw), t(nw, -3);
  wall_layout(p.box_x, p.box_y, p.dx, dy, p.walls, x.data(), y.data(), bt.data());
  std::vector<double2> pos(nw), vel(nw, make_double2(0.0, 0.0));
  for (int q = 0; q < nw; ++q) pos[q] = make_double2(x[q], y[q]);
  HIP_OR_FAIL(hipMemcpyAsync(B.g_pos + n, pos.data(), nw * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_vel + n, vel.data(), nw * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_rho + n, rho.data(), nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_mass + n, zero.data(), nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_p + n, zero.data(), nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_c + n, cs.data(), nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_e + n, e.data(), nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_type + n, t.data(), nw * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_btype + n, bt.data(), nw * sizeof(int), hipMemcpyHostToDevice, c->stream));
  launch_fill_i32(c->stream, B.counters + C_NGHOST, 1, nwm + nw);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SPH2D_OK;
}

int sph2d_init_tgv(sph2d_ctx* c) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  const int n = c->prm.nx * c->prm.ny;
  if (n > c->cap_fluid) return fail(SPH2D_ERR_CAPACITY, "nx*ny exceeds the context capacity");
  int rc = set_nfluid(c, n);
  if (rc) return rc;
  launch_init_tgv(c->stream, c->P, c->B, c->prm.nx, c->prm.ny);
  HIP_OR_FAIL(hipGetLastError());
  if (c->prm.walls) {
    rc = upload_walls(c, n, nullptr);
    if (rc) return rc;
  }
  rc = compute_ke0(c);
  if (rc) return rc;
  c->have_state = true;
  c->evaluated = false;
  c->vjp_ready = false;
  c->eval_part = 0;
  return SPH2D_OK;
}

// ---- the moving-square program --------------------------------------------------------
int sph2d_moving_square_layout(const sph2d_params* p, int32_t cap, double* x, double* y, int32_t* btype,
                               int32_t* init, int32_t* n_fluid, int32_t* n_total) {
  if (!p) return fail(SPH2D_ERR_ARGUMENT, "params is NULL");
  if (!(p->dx > 0) || p->nx <= 0 || p->ny <= 0 || p->walls < 0 || p->walls > 64)
    return fail(SPH2D_ERR_ARGUMENT, "bad lattice or wall layers");
  const double dx = p->box_x / p->nx, dy = p->box_y / p->ny;
  const double hl = 0.5 * p->body_len, cx = p->body_cx, cy = p->body_cy;
  int k = 0;
  for (int i = 1; i <= p->ny; ++i)                       // SPH_Subroutine_Module.f90:889-918
    for (int j = 1; j <= p->nx; ++j) {
      const double px = (j - 1) * dx + dx / 2.0, py = (i - 1) * dy + dy / 2.0;
      if (px >= (cx - hl) && px <= (cx + hl) && py >= (cy - hl) && py <= (cy + hl)) continue;
      if (k < cap) {
        if (x) x[k] = px;
        if (y) y[k] = py;
        if (btype) btype[k] = 0;
        if (init) init[k] = 2;
      }
      ++k;
    }
  const int ng = ms_ghost_layout(p->box_x, p->box_y, dx, dy, p->walls, cx, cy, p->body_len, nullptr, nullptr,
                                 nullptr, nullptr);
  if (n_fluid) *n_fluid = k;
  if (n_total) *n_total = k + ng;
  if (x || y || btype || init) {
    if (k + ng > cap) return fail(SPH2D_ERR_CAPACITY, "cap is smaller than the layout");
    std::vector<double> gx(ng), gy(ng);
    std::vector<int> gb(ng), gi(ng);
    ms_ghost_layout(p->box_x, p->box_y, dx, dy, p->walls, cx, cy, p->body_len, gx.data(), gy.data(), gb.data(),
                    gi.data());
    for (int q = 0; q < ng; ++q) {
      if (x) x[k + q] = gx[q];
      if (y) y[k + q] = gy[q];
      if (btype) btype[k + q] = gb[q];
      if (init) init[k + q] = gi[q];
    }
  }
  return SPH2D_OK;
}

int sph2d_init_moving_square(sph2d_ctx* c) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  const sph2d_params& p = c->prm;
  if (!p.ms) return fail(SPH2D_ERR_STATE, "the context was not created with ms = 1");
  int nf = 0, nt = 0;
  int rc = sph2d_moving_square_layout(&p, 0, nullptr, nullptr, nullptr, nullptr, &nf, &nt);
  if (rc) return rc;
  if (nf > c->cap_fluid) return fail(SPH2D_ERR_CAPACITY, "the tank's fluid exceeds the context capacity");
  if (nt - nf > c->cap_ghost) return fail(SPH2D_ERR_CAPACITY, "the fixed ghosts exceed the ghost capacity");
  std::vector<double> x(nt), y(nt);
  std::vector<int> bt(nt), init(nt);
  rc = sph2d_moving_square_layout(&p, nt, x.data(), y.data(), bt.data(), init.data(), &nf, &nt);
  if (rc) return rc;
  rc = set_nfluid(c, nf);
  if (rc) return rc;
  const double dy = p.box_y / p.ny, m0 = p.rho0 * (p.dx * dy);
  const double rho = (0.0 + p.background_p - p.background_p) / p.c0sq + p.rho0;   // :900-903
  std::vector<double2> pos(nt), vel(nt, make_double2(0.0, 0.0));
  for (int q = 0; q < nt; ++q) pos[q] = make_double2(x[q], y[q]);
  std::vector<double> vr(nt, rho), vm(nt, m0), vp(nt, 0.0 + p.background_p), vc(nt, std::sqrt(p.c0sq)),
      ve(nt, p.c0sq / (7.0 - 1.0));
  for (int q = nf; q < nt; ++q) vr[q] = p.rho0;
  std::vector<int> t(nt, 2);
  for (int q = nf; q < nt; ++q) t[q] = -3;
  Buffers& B = c->B;
  hipStream_t st = c->stream;
  HIP_OR_FAIL(hipMemcpyAsync(B.g_pos, pos.data(), nt * sizeof(double2), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_vel, vel.data(), nt * sizeof(double2), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_rho, vr.data(), nt * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_mass, vm.data(), nt * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_p, vp.data(), nt * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_c, vc.data(), nt * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_e, ve.data(), nt * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_type, t.data(), nt * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_btype, bt.data(), nt * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_init, init.data(), nt * sizeof(int), hipMemcpyHostToDevice, st));
  launch_fill_i32(st, B.counters + C_NGHOST, 1, nt - nf);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipStreamSynchronize(st));
  c->n_wm = 0;
  c->ke0 = 0.0;
  c->ms_time = 0.0;
  c->body_x = p.body_cx; c->body_y = p.body_cy;
  c->cust_u = 0.0; c->cust_a = 0.0;
  c->bforce_step = 0;
  c->have_state = true;
  c->evaluated = false;
  c->vjp_ready = false;
  c->eval_part = 0;
  return SPH2D_OK;
}

int sph2d_body_motion(sph2d_ctx* c, double t) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  if (!c->prm.ms) return fail(SPH2D_ERR_STATE, "the context was not created with ms = 1");
  if (!c->have_state) return fail(SPH2D_ERR_STATE, "no particle state: call sph2d_init_moving_square");
  const double U = c->prm.body_speed, T = c->prm.body_ramp;
  double xc, u, a;
  if (t < T) { xc = c->prm.body_cx + 0.5 * U * t * t / T; u = U * t / T; a = U / T; }
  else { xc = c->prm.body_cx + 0.5 * U * T + U * (t - T); u = U; a = 0.0; }
  c->ms_time = t;                                 // Real_time of the evaluations that follow
  c->cust_u = u; c->cust_a = a;
  const double ddx = xc - c->body_x, ddy = c->prm.body_cy - c->body_y;
  c->body_x = xc; c->body_y = c->prm.body_cy;
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_ms_body_move(c->stream, c->P, c->B, c->cap_global, ddx, ddy);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_body_force(sph2d_ctx* c, double* out6, int32_t* step) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  if (!c->prm.ms) return fail(SPH2D_ERR_STATE, "the context was not created with ms = 1");
  if (step) *step = c->bforce_step;
  if (!out6) return SPH2D_OK;
  if (c->bforce_step == 0) {
    for (int k = 0; k < 6; ++k) out6[k] = 0.0;
    return SPH2D_OK;
  }
  HIP_OR_FAIL(hipMemcpyAsync(out6, c->d_bforce, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SPH2D_OK;
}

int sph2d_upload_state(sph2d_ctx* c, int32_t n, const double* pos, const double* vel,
                       const double* rho, const double* mass, const double* energy,
                       const double* h, const int32_t* type, int32_t ld) {
  if (!c || !pos || !vel || !rho || !mass) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  if (n <= 0 || n > c->cap_fluid) return fail(SPH2D_ERR_CAPACITY, "n outside (0, capacity]");
  if (ld < n) return fail(SPH2D_ERR_ARGUMENT, "ld < n");
  std::vector<double2> p2(n), v2(n);
  std::vector<double> e(n), pr(n, 0.0), cs(n, std::sqrt(c->prm.c0sq));
  std::vector<int> t(n, 2);
  for (int i = 0; i < n; ++i) {
    p2[i] = make_double2(pos[i], pos[ld + i]);
    v2[i] = make_double2(vel[i], vel[ld + i]);
    e[i] = energy ? energy[i] : c->prm.c0sq / (7.0 - 1.0);
    if (h && h[i] != c->prm.h)
      return fail(SPH2D_ERR_UNSUPPORTED, "per-particle smoothing lengths must all equal params.h");
    if (type) {
      t[i] = type[i];
      if (t[i] != 2 && t[i] != 120)
        return fail(SPH2D_ERR_UNSUPPORTED, "only fluid (2) and ill (120) particle types are built");
    }
    if (!std::isfinite(pos[i]) || !std::isfinite(pos[ld + i]) || !std::isfinite(vel[i]) ||
        !std::isfinite(vel[ld + i]) || !std::isfinite(mass[i]) || !std::isfinite(rho[i]) || !(rho[i] > 0))
      return fail(SPH2D_ERR_NONFINITE, "non-finite or non-positive state at particle " + std::to_string(i));
  }
  Buffers& B = c->B;
  if (c->prm.ms && (!c->have_state || n != c->n_fluid))
    return fail(SPH2D_ERR_STATE, "the moving-square tank takes the fluid state of its layout: call "
                                 "sph2d_init_moving_square first and pass n = its fluid count");
  HIP_OR_FAIL(hipMemcpyAsync(B.g_pos, p2.data(), n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_vel, v2.data(), n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_rho, rho, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_mass, mass, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->prm.ms) {                 // the fluid only: the tank's and body's ghosts stay
    HIP_OR_FAIL(hipMemcpyAsync(B.g_e, e.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(B.g_type, t.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    c->evaluated = false;
    c->vjp_ready = false;
    c->eval_part = 0;
    return SPH2D_OK;
  }
  HIP_OR_FAIL(hipMemcpyAsync(B.g_e, e.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_p, pr.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_c, cs.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(B.g_type, t.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  int rc = set_nfluid(c, n);
  if (rc) return rc;
  if (c->prm.walls) {
    if (n + c->cap_ghost > c->cap_global) return fail(SPH2D_ERR_CAPACITY, "no room for the wall layers");
    rc = upload_walls(c, n, rho);
    if (rc) return rc;
  }
  rc = compute_ke0(c);
  if (rc) return rc;
  c->have_state = true;
  c->evaluated = false;
  c->vjp_ready = false;
  c->eval_part = 0;
  return SPH2D_OK;
}

int sph2d_download_state(sph2d_ctx* c, int32_t n, double* pos, double* vel, double* rho,
                         double* press, double* cs, double* mass, int32_t* type, int32_t ld) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  if (n != c->n_fluid) return fail(SPH2D_ERR_ARGUMENT, "n must equal the number of fluid particles");
  if (ld < n) return fail(SPH2D_ERR_ARGUMENT, "ld < n");
  Buffers& B = c->B;
  std::vector<double2> tmp(n);
  if (pos) {
    HIP_OR_FAIL(hipMemcpyAsync(tmp.data(), B.g_pos, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) { pos[i] = tmp[i].x; pos[ld + i] = tmp[i].y; }
  }
  if (vel) {
    HIP_OR_FAIL(hipMemcpyAsync(tmp.data(), B.g_vel, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) { vel[i] = tmp[i].x; vel[ld + i] = tmp[i].y; }
  }
  if (rho) HIP_OR_FAIL(hipMemcpyAsync(rho, B.g_rho, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (press) HIP_OR_FAIL(hipMemcpyAsync(press, B.g_p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (cs) HIP_OR_FAIL(hipMemcpyAsync(cs, B.g_c, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (mass) HIP_OR_FAIL(hipMemcpyAsync(mass, B.g_mass, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (type) HIP_OR_FAIL(hipMemcpyAsync(type, B.g_type, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SPH2D_OK;
}

// ---- scalar and array forms ---------------------------------------------------
static hipStream_t scalar_stream() {
  static hipStream_t s = nullptr;
  if (!s) (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  return s;
}

int sph2d_compute_kernel_n(int32_t n, const double* r, const double* dx, const double* h,
                           double* w, double* dwdx) {
  if (n <= 0 || !r || !dx || !h || !w || !dwdx) return fail(SPH2D_ERR_ARGUMENT, "bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SPH2D_ERR_HIP, "no HIP device: the sph2d library runs on the GPU only");
  hipStream_t s = scalar_stream();
  double *dr, *ddx, *dh, *dw, *ddw;
  HIP_OR_FAIL(hipMalloc(&dr, n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&ddx, 2 * n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&dh, n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&dw, n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&ddw, 2 * n * sizeof(double)));
  HIP_OR_FAIL(hipMemcpyAsync(dr, r, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(hipMemcpyAsync(ddx, dx, 2 * n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(hipMemcpyAsync(dh, h, n * sizeof(double), hipMemcpyHostToDevice, s));
  launch_kernel_n(s, n, dr, ddx, dh, dw, ddw, (double)expf(-9.0f));
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(w, dw, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(hipMemcpyAsync(dwdx, ddw, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(hipStreamSynchronize(s));
  (void)hipFree(dr); (void)hipFree(ddx); (void)hipFree(dh); (void)hipFree(dw); (void)hipFree(ddw);
  return SPH2D_OK;
}

int sph2d_compute_kernel(const int32_t* dim, const double* distance, const double* position_difference,
                         const double* average_smooth_length, double* w, double* dwdx) {
  if (!dim || *dim != 2) return fail(SPH2D_ERR_UNSUPPORTED, "only dim = 2 is built");
  return sph2d_compute_kernel_n(1, distance, position_difference, average_smooth_length, w, dwdx);
}

int sph2d_status_equation_n(const sph2d_params* p, int32_t n, const double* rho, const double* energy,
                            const int32_t* type, double* press, double* cs) {
  (void)energy;
  if (n <= 0 || !rho || !type || !press || !cs) return fail(SPH2D_ERR_ARGUMENT, "bad argument");
  sph2d_params d;
  if (!p) { sph2d_default_params(&d); p = &d; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SPH2D_ERR_HIP, "no HIP device: the sph2d library runs on the GPU only");
  hipStream_t s = scalar_stream();
  double *dr, *dp, *dc;
  int* dt;
  HIP_OR_FAIL(hipMalloc(&dr, n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&dp, n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&dc, n * sizeof(double)));
  HIP_OR_FAIL(hipMalloc(&dt, n * sizeof(int)));
  HIP_OR_FAIL(hipMemcpyAsync(dr, rho, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(hipMemcpyAsync(dt, type, n * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(hipMemcpyAsync(dp, press, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(hipMemcpyAsync(dc, cs, n * sizeof(double), hipMemcpyHostToDevice, s));
  launch_eos_n(s, n, p->rho0, p->c0sq, dr, dt, dp, dc);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(press, dp, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(hipMemcpyAsync(cs, dc, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(hipStreamSynchronize(s));
  (void)hipFree(dr); (void)hipFree(dp); (void)hipFree(dc); (void)hipFree(dt);
  return SPH2D_OK;
}

int sph2d_status_equation(const double* rho, const double* energy, const int32_t* type,
                          double* press, double* cs) {
  return sph2d_status_equation_n(nullptr, 1, rho, energy, type, press, cs);
}

// ---- Reorder_Particle_Index (device arrays) -------------------------------------
int sph2d_reorder_particle_index(int32_t n_total, int32_t n_fluid, double box_x, double box_y, double* x,
                                 double* y, int32_t* type, int32_t n_f64, double* const* f64_cols, int32_t n_i32,
                                 int32_t* const* i32_cols, int32_t* perm, int32_t* counts, void* stream) {
  if (n_total <= 0 || n_fluid < 0 || n_fluid > n_total || n_f64 < 0 || n_i32 < 0 || !counts)
    return fail(SPH2D_ERR_ARGUMENT, "bad sizes or NULL counts");
  if ((n_f64 > 0 && !f64_cols) || (n_i32 > 0 && !i32_cols)) return fail(SPH2D_ERR_ARGUMENT, "NULL column list");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SPH2D_ERR_HIP, "no HIP device: the sph2d library runs on the GPU only");
  const size_t nd = (size_t)n_total * sizeof(double), ni = (size_t)n_total * sizeof(int32_t);
  if (!device_range(x, nd) || !device_range(y, nd) || !device_range(type, ni))
    return fail(SPH2D_ERR_ARGUMENT, "x, y and type must be device arrays of n_total entries");
  for (int k = 0; k < n_f64; ++k)
    if (!device_range(f64_cols[k], nd))
      return fail(SPH2D_ERR_ARGUMENT, "f64 column " + std::to_string(k) + " is not a device array of n_total entries");
  for (int k = 0; k < n_i32; ++k)
    if (!device_range(i32_cols[k], ni))
      return fail(SPH2D_ERR_ARGUMENT, "i32 column " + std::to_string(k) + " is not a device array of n_total entries");
  if (perm && !device_range(perm, ni)) return fail(SPH2D_ERR_ARGUMENT, "perm must be a device array of n_total entries");
  {
    std::vector<const void*> all = {x, y, type};
    for (int k = 0; k < n_f64; ++k) all.push_back(f64_cols[k]);
    for (int k = 0; k < n_i32; ++k) all.push_back(i32_cols[k]);
    if (perm) all.push_back(perm);
    std::sort(all.begin(), all.end());
    if (std::adjacent_find(all.begin(), all.end()) != all.end())
      return fail(SPH2D_ERR_ARGUMENT, "x, y, type, perm and the columns must be distinct arrays");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int G = (n_total + kBlock - 1) / kBlock;
  DeviceScratch sc;
  int *blk, *off, *sblk, *misc, *dperm;
  double* tmp;
  HIP_OR_FAIL(sc.get(&blk, 5 * (size_t)G));
  HIP_OR_FAIL(sc.get(&off, 5 * (size_t)G));
  HIP_OR_FAIL(sc.get(&sblk, (size_t)scan_block_count(5 * G) + 1));
  HIP_OR_FAIL(sc.get(&misc, 8));
  HIP_OR_FAIL(sc.get(&tmp, (size_t)n_total));
  if (perm) dperm = perm;
  else HIP_OR_FAIL(sc.get(&dperm, (size_t)n_total));
  int* total = misc;
  int* dcounts = misc + 1;
  launch_reorder_partition(s, n_total, n_fluid, type, x, y, box_x, box_y, blk, off, sblk, total, dcounts, dperm);
  HIP_OR_FAIL(hipGetLastError());
  int hc[6];
  HIP_OR_FAIL(hipMemcpyAsync(hc, dcounts, sizeof(hc), hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(hipStreamSynchronize(s));
  const size_t kept = (size_t)hc[5];
  auto permute_f64 = [&](double* col) -> int {
    launch_reorder_gather_f64(s, n_total, dperm, total, col, tmp);
    HIP_OR_FAIL(hipMemcpyAsync(col, tmp, kept * sizeof(double), hipMemcpyDeviceToDevice, s));
    return SPH2D_OK;
  };
  auto permute_i32 = [&](int32_t* col) -> int {
    int* t = reinterpret_cast<int*>(tmp);
    launch_reorder_gather_i32(s, n_total, dperm, total, col, t);
    HIP_OR_FAIL(hipMemcpyAsync(col, t, kept * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    return SPH2D_OK;
  };
  int rc = permute_f64(x);
  if (!rc) rc = permute_f64(y);
  if (!rc) rc = permute_i32(type);
  for (int k = 0; k < n_f64 && !rc; ++k) rc = permute_f64(f64_cols[k]);
  for (int k = 0; k < n_i32 && !rc; ++k) rc = permute_i32(i32_cols[k]);
  if (rc) return rc;
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipStreamSynchronize(s));
  for (int k = 0; k < 6; ++k) counts[k] = hc[k];
  return SPH2D_OK;
}

// The same routine with its scratch in the library's per-device static array: no
// hipMalloc, hipFree or other allocation, at most 2^21 particles; calls that overlap
// on one GPU share that array and race.
int sph2d_reorder_particle_index_noalloc(int32_t n_total, int32_t n_fluid, double box_x, double box_y, double* x,
                                         double* y, int32_t* type, int32_t n_f64, double* const* f64_cols,
                                         int32_t n_i32, int32_t* const* i32_cols, int32_t* perm, int32_t* counts,
                                         void* stream) {
  if (n_total <= 0 || n_fluid < 0 || n_fluid > n_total || n_f64 < 0 || n_i32 < 0 || !counts)
    return fail(SPH2D_ERR_ARGUMENT, "bad sizes or NULL counts");
  if ((n_f64 > 0 && !f64_cols) || (n_i32 > 0 && !i32_cols)) return fail(SPH2D_ERR_ARGUMENT, "NULL column list");
  constexpr int kMaxCols = 64;
  if (n_f64 > kMaxCols || n_i32 > kMaxCols) return fail(SPH2D_ERR_ARGUMENT, "at most 64 columns of each kind");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(SPH2D_ERR_HIP, "no HIP device: the sph2d library runs on the GPU only");
  void* base = nullptr;
  size_t bytes = 0;
  int max_n = 0;
  int rc = reorder_scratch(&base, &bytes, &max_n);
  if (rc) return rc;
  if (n_total > max_n)
    return fail(SPH2D_ERR_CAPACITY, "n_total exceeds the 2097152 particles the no-allocation scratch holds");
  const size_t nd = (size_t)n_total * sizeof(double), ni = (size_t)n_total * sizeof(int32_t);
  if (!device_range(x, nd) || !device_range(y, nd) || !device_range(type, ni))
    return fail(SPH2D_ERR_ARGUMENT, "x, y and type must be device arrays of n_total entries");
  for (int k = 0; k < n_f64; ++k)
    if (!device_range(f64_cols[k], nd)) return fail(SPH2D_ERR_ARGUMENT, "an f64 column is not a device array of n_total entries");
  for (int k = 0; k < n_i32; ++k)
    if (!device_range(i32_cols[k], ni)) return fail(SPH2D_ERR_ARGUMENT, "an i32 column is not a device array of n_total entries");
  if (perm && !device_range(perm, ni)) return fail(SPH2D_ERR_ARGUMENT, "perm must be a device array of n_total entries");
  {
    const void* all[3 + 2 * kMaxCols + 1];
    int na = 0;
    all[na++] = x; all[na++] = y; all[na++] = type;
    for (int k = 0; k < n_f64; ++k) all[na++] = f64_cols[k];
    for (int k = 0; k < n_i32; ++k) all[na++] = i32_cols[k];
    if (perm) all[na++] = perm;
    for (int a = 0; a < na; ++a)
      for (int b = a + 1; b < na; ++b)
        if (all[a] == all[b]) return fail(SPH2D_ERR_ARGUMENT, "x, y, type, perm and the columns must be distinct arrays");
  }
  // carve the scratch as sph2d_reorder.hip sized it (each piece 256-byte aligned)
  const size_t Gmax = ((size_t)max_n + kBlock - 1) / kBlock;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  char* p = static_cast<char*>(base);
  int* blk = reinterpret_cast<int*>(p); p += al(sizeof(int) * 5 * Gmax);
  int* off = reinterpret_cast<int*>(p); p += al(sizeof(int) * 5 * Gmax);
  int* sblk = reinterpret_cast<int*>(p); p += al(sizeof(int) * 1024);
  int* misc = reinterpret_cast<int*>(p); p += al(sizeof(int) * 8);
  double* tmp = reinterpret_cast<double*>(p); p += al(sizeof(double) * (size_t)max_n);
  int* sperm = reinterpret_cast<int*>(p); p += al(sizeof(int) * (size_t)max_n);
  const int G = (n_total + kBlock - 1) / kBlock;
  if ((size_t)(p - static_cast<char*>(base)) > bytes || scan_block_count(5 * G) + 1 > 1024)
    return fail(SPH2D_ERR_CAPACITY, "no-allocation scratch layout exceeds its array");
  int* dperm = perm ? perm : sperm;
  int* total = misc;
  int* dcounts = misc + 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_fill_i32(s, misc, 8, 0);                    // the previous call's counters are stale
  launch_reorder_partition(s, n_total, n_fluid, type, x, y, box_x, box_y, blk, off, sblk, total, dcounts, dperm);
  HIP_OR_FAIL(hipGetLastError());
  for (int k = -3; k < n_f64 + n_i32; ++k) {
    if (k == -1 || k >= n_f64) {
      int32_t* col = k == -1 ? type : i32_cols[k - n_f64];
      int* t = reinterpret_cast<int*>(tmp);
      launch_reorder_gather_i32(s, n_total, dperm, total, col, t);
      launch_reorder_copy_i32(s, n_total, total, t, col);
    } else {
      double* col = k == -3 ? x : (k == -2 ? y : f64_cols[k]);
      launch_reorder_gather_f64(s, n_total, dperm, total, col, tmp);
      launch_reorder_copy_f64(s, n_total, total, tmp, col);
    }
  }
  HIP_OR_FAIL(hipGetLastError());
  int hc[6];
  HIP_OR_FAIL(hipMemcpyAsync(hc, dcounts, sizeof(hc), hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(hipStreamSynchronize(s));
  for (int k = 0; k < 6; ++k) counts[k] = hc[k];
  return SPH2D_OK;
}

// ---- step pieces --------------------------------------------------------------
static int need_state(sph2d_ctx* c) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  if (!c->have_state) return fail(SPH2D_ERR_STATE, "no particle state: call sph2d_init_tgv or sph2d_upload_state");
  return SPH2D_OK;
}

int sph2d_distribute_period_boundary(sph2d_ctx* c) {
  int rc = need_state(c);
  if (rc) return rc;
  c->vjp_ready = false;
  c->eval_part = 0;
  if (c->prm.walls) return SPH2D_OK;                 // fixed walls: no periodic ghosts
  launch_period_boundary(c->stream, c->P, c->B, c->cap_fluid, c->cap_ghost);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

static int devide_block(sph2d_ctx* c, int32_t it, int predict) {
  int rc = need_state(c);
  if (rc) return rc;
  c->vjp_ready = false;
  c->eval_part = 0;
  const bool ana = !c->prm.ms && (it % c->prm.sample_every) == 0;   // the TGV analytic fields
  const double t = it * c->prm.dt / (c->prm.box_x / c->prm.U);   // Nondimensionalized_real_time
  launch_fill_i32(c->stream, c->B.counters + C_NILL, 1, 0);
  launch_devide_block(c->stream, c->P, c->B, c->cap_global, c->n_cells, ana, t, predict);
  if (ana) c->ana_step = it;
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_devide_block(sph2d_ctx* c, int32_t it) { return devide_block(c, it, 0); }

// One evaluation of the moving-square program (MS Single_step_compute.f90:35-283): EOS
// for every particle, the fixed ghosts, pass A and the force pass, and on sampling
// steps the body force.
static void do_eval_ms(sph2d_ctx* c, int it) {
  hipStream_t s = c->stream;
  launch_stage(s, c->P, c->B, c->cap_local, ST_EOS);
  c->P.t_eval = c->ms_time;
  c->P.cust_u = c->cust_u;
  c->P.cust_a = c->cust_a;
  c->P.body_force = (it > 0 && it % c->prm.sample_every == 0) ? 1 : 0;
  launch_ms_assign_walls(s, c->P, c->B, c->cap_local);
  launch_ms_eval(s, c->P, c->B, c->cap_local);
  if (c->P.body_force) {
    launch_ms_body_sum(s, c->P, c->B, c->cap_local, c->d_bforce);
    c->bforce_step = it;
  }
  c->evaluated = true;
}

static void do_eval(sph2d_ctx* c, bool with_eos, int it) {
  if (c->prm.ms) { do_eval_ms(c, it); return; }
  if (with_eos) launch_stage(c->stream, c->P, c->B, c->cap_local, ST_EOS);
  c->P.write_debug = c->debug ? 1 : 0;
  hipEvent_t* ev = nullptr;
  if (c->pass_timing) {
    while (c->events.size() < c->events_used + 4) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) break;
      c->events.push_back(e);
    }
    if (c->events.size() >= c->events_used + 4) {
      ev = &c->events[c->events_used];
      c->events_used += 4;
    }
  }
  if (c->P.walls) launch_assign_walls(c->stream, c->P, c->B, c->cap_local);   // single_step_compute.f90:80
  launch_eval(c->stream, c->P, c->B, c->cap_local, ev);
  if (c->P.damping) launch_end_damping(c->stream, c->P, c->B, c->cap_local);
  c->evaluated = true;
  // Refresh_domain's vorticity uses the step's last evaluation; it is kept on
  // sampling steps (each evaluation of such a step overwrites it).
  if (it > 0 && it % c->prm.sample_every == 0) {
    launch_vorticity(c->stream, c->P, c->B, c->cap_local);
    c->vort_step = it;
  }
}

int sph2d_single_step_compute(sph2d_ctx* c, int32_t it) {
  int rc = need_state(c);
  if (rc) return rc;
  do_eval(c, true, it);
  HIP_OR_FAIL(hipGetLastError());
  c->vjp_ready = true;
  c->eval_part = 3;
  return SPH2D_OK;
}

// ---- the evaluation routine by routine ----------------------------------------------
int sph2d_find_particle_pair(sph2d_ctx* c, int32_t it) {
  (void)it;
  int rc = need_state(c);
  if (rc) return rc;
  c->vjp_ready = false;
  launch_stage(c->stream, c->P, c->B, c->cap_local, ST_EOS);     // single_step_compute.f90:62-73
  c->P.write_debug = c->debug ? 1 : 0;
  if (c->P.walls) launch_assign_walls(c->stream, c->P, c->B, c->cap_local);   // :80
  launch_eval_part(c->stream, c->P, c->B, c->cap_local, 0);
  HIP_OR_FAIL(hipGetLastError());
  c->eval_part = 1;
  return SPH2D_OK;
}

int sph2d_download_pair_list(sph2d_ctx* c, int64_t capacity, int32_t* pi, int32_t* pj, int64_t* np_out) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!np_out) return fail(SPH2D_ERR_ARGUMENT, "pair_number is NULL");
  if (c->eval_part < 1) return fail(SPH2D_ERR_STATE, "no pair search on the current state: call sph2d_find_particle_pair");
  launch_count_pair_list(c->stream, c->P, c->B, c->cap_local);
  int cnt[C_COUNT];
  rc = read_counters(c, cnt);
  if (rc) return rc;
  const int64_t np = (int64_t)(unsigned)cnt[C_NPAIR_LO];
  *np_out = np;
  if (!pi || !pj || capacity <= 0) return SPH2D_OK;
  if (capacity < np) return fail(SPH2D_ERR_CAPACITY, "pair capacity too small");
  DeviceScratch tmp;
  int *di = nullptr, *dj = nullptr;
  HIP_OR_FAIL(tmp.get(&di, (size_t)np));
  HIP_OR_FAIL(tmp.get(&dj, (size_t)np));
  Buffers B = c->B;
  B.pair_lo = di; B.pair_hi = dj;
  launch_pair_list(c->stream, c->P, B, c->cap_local, np);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(pi, di, np * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(pj, dj, np * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SPH2D_OK;
}

int sph2d_define_freesurface(sph2d_ctx* c) {
  int rc = need_state(c);
  if (rc) return rc;
  if (c->eval_part != 1) return fail(SPH2D_ERR_STATE, "call sph2d_find_particle_pair first");
  launch_eval_part(c->stream, c->P, c->B, c->cap_local, 1);
  HIP_OR_FAIL(hipGetLastError());
  c->eval_part = 2;
  return SPH2D_OK;
}

int sph2d_solver(sph2d_ctx* c, int32_t variant, int32_t sps_on) {
  int rc = need_state(c);
  if (rc) return rc;
  if (c->eval_part != 2) return fail(SPH2D_ERR_STATE, "call sph2d_define_freesurface first");
  if (variant < SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN || variant > SPH2D_SOLVER_MUSCL)
    return fail(SPH2D_ERR_UNSUPPORTED, "variant must be one of SPH2D_SOLVER_*");
  if (sps_on != 0 && sps_on != 1) return fail(SPH2D_ERR_ARGUMENT, "sps_on must be 0 or 1");
  if (variant == SPH2D_SOLVER_MUSCL && c->P.solver != SPH2D_SOLVER_MUSCL)
    return fail(SPH2D_ERR_UNSUPPORTED, "the MUSCL gradients are formed by the pair walk of a context created "
                                       "with the MUSCL solver");
  if (sps_on && !c->P.sps)
    return fail(SPH2D_ERR_UNSUPPORTED, "tau is formed by the pair walk of a context created with sps_on = 1");
  DevParams Q = c->P;
  Q.solver = variant;
  Q.sps = sps_on;
  Q.xsph_f = sps_on ? 2.0 : 1.0;
  Q.write_debug = c->debug ? 1 : 0;
  if (!sps_on && c->P.sps) launch_clear_tau(c->stream, Q, c->B, c->cap_local);
  launch_eval_part(c->stream, Q, c->B, c->cap_local, 2);
  if (Q.damping) launch_end_damping(c->stream, Q, c->B, c->cap_local);   // single_step_compute.f90:174-260
  HIP_OR_FAIL(hipGetLastError());
  c->evaluated = true;
  c->eval_part = 3;
  c->vjp_ready = variant == c->P.solver && sps_on == c->P.sps;
  return SPH2D_OK;
}

int sph2d_stage_update(sph2d_ctx* c, int32_t scheme, int32_t stage) {
  int rc = need_state(c);
  if (rc) return rc;
  int kind = -1;
  if (stage == 0) kind = ST_SAVE_BEFORE;
  else if (scheme == SPH2D_INTEGRATOR_PC && stage == 1) kind = ST_PC_PREDICT;
  else if (scheme == SPH2D_INTEGRATOR_PC && stage == 2) kind = ST_PC_CORRECT;
  else if (scheme == SPH2D_INTEGRATOR_RK && stage >= 1 && stage <= 4)
    kind = stage == 1 ? ST_RK_HALF0 : (stage == 2 ? ST_RK_HALF1 : (stage == 3 ? ST_RK_FULL2 : ST_RK_FINAL));
  if (kind < 0 || (scheme != SPH2D_INTEGRATOR_PC && scheme != SPH2D_INTEGRATOR_RK))
    return fail(SPH2D_ERR_ARGUMENT, "scheme must be SPH2D_INTEGRATOR_PC (stages 0-2) or SPH2D_INTEGRATOR_RK (0-4)");
  if (kind != ST_SAVE_BEFORE && !c->evaluated) return fail(SPH2D_ERR_STATE, "no force evaluation yet");
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_stage(c->stream, c->P, c->B, c->cap_local, kind);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_download_rates(sph2d_ctx* c, int32_t n, double* drhodt, double* acc, int32_t ld) {
  int rc = need_state(c);
  if (rc) return rc;
  if (n != c->n_fluid) return fail(SPH2D_ERR_ARGUMENT, "n must equal the number of fluid particles");
  if (ld < n) return fail(SPH2D_ERR_ARGUMENT, "ld < n");
  if (!c->evaluated) return fail(SPH2D_ERR_STATE, "no force evaluation yet");
  int cnt[C_COUNT];
  rc = read_counters(c, cnt);
  if (rc) return rc;
  const int nl = cnt[C_NLOCAL];
  std::vector<int> name(nl);
  std::vector<double> d(nl);
  std::vector<double2> a(nl);
  HIP_OR_FAIL(hipMemcpyAsync(name.data(), c->B.name, nl * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(d.data(), c->B.drhodt, nl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(a.data(), c->B.acc, nl * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  for (int s = 0; s < nl; ++s) {
    const int g = name[s];
    if (g < 0 || g >= n) continue;                   // ghosts, wavemaker and wall particles
    if (drhodt) drhodt[g] = d[s];
    if (acc) { acc[g] = a[s].x; acc[(size_t)ld + g] = a[s].y; }
  }
  return SPH2D_OK;
}

// What both forms of the reverse evaluation check; then the work space.
static int vjp_prepare(sph2d_ctx* c, int32_t n, int32_t ld, const void* lam_drhodt, const void* lam_acc) {
  int rc = need_state(c);
  if (rc) return rc;
  if (n != c->n_fluid) return fail(SPH2D_ERR_ARGUMENT, "n must equal the number of fluid particles");
  if (ld < n) return fail(SPH2D_ERR_ARGUMENT, "ld < n");
  if (!lam_drhodt || !lam_acc) return fail(SPH2D_ERR_ARGUMENT, "lam_drhodt and lam_acc are required");
  if (!c->vjp_ready)
    return fail(SPH2D_ERR_STATE, "the reverse evaluation needs sph2d_single_step_compute as the context's last "
                                 "state-changing call");
  if (c->P.mixed) return fail(SPH2D_ERR_UNSUPPORTED, "the reverse evaluation is built for the FP64 form");
  if (c->P.renorm) return fail(SPH2D_ERR_UNSUPPORTED, "the reverse evaluation is built for the identity renormalisation");
  if (!c->P.sps) return fail(SPH2D_ERR_UNSUPPORTED, "the reverse evaluation is built for the _SPS solvers");
  if (c->P.solver != SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && c->P.solver != SPH2D_SOLVER_LINEAR_RIEMANN)
    return fail(SPH2D_ERR_UNSUPPORTED, "the reverse evaluation is built for the two linear Riemann solvers");
  if (c->P.walls || c->P.damping || c->n_wm)
    return fail(SPH2D_ERR_UNSUPPORTED, "the reverse evaluation is built for the periodic box");
  if (!(c->prm.box_x > 2 * c->P.three_h) || !(c->prm.box_y > 2 * c->P.three_h))
    return fail(SPH2D_ERR_UNSUPPORTED, "the box must be wider than two kernel supports");
  const size_t nf = (size_t)c->cap_fluid, nl = (size_t)c->cap_local, ng = (size_t)c->cap_global;
  if (!c->adj_buf) {                                // kept until the context is destroyed
    rc = c->alloc(&c->adj_buf, 5 * nf + 20 * nl + 5 * ng);
    if (rc) return rc;
  }
  // a clamped ghost emission leaves slots unwritten that the fold would read
  int cnt[C_COUNT];
  return read_counters(c, cnt);
}

int sph2d_single_step_compute_vjp_device(sph2d_ctx* c, int32_t n, const double* lam_drhodt, const double* lam_acc,
                                         const double* lam_avg, int32_t ld, double* grad_pos, double* grad_vel,
                                         double* grad_rho) {
  int rc = vjp_prepare(c, n, ld, lam_drhodt, lam_acc);
  if (rc) return rc;
  const size_t nf = (size_t)c->cap_fluid, nl = (size_t)c->cap_local;
  double *d_lam = c->adj_buf, *w = d_lam + 5 * nf, *gs = w + 10 * nl, *ga = gs + 5 * nl, *G = ga + 5 * nl;
  launch_vjp_pack(c->stream, n, ld, lam_drhodt, lam_acc, lam_avg, d_lam);
  launch_eval_vjp(c->stream, c->P, c->B, c->cap_local, c->cap_fluid, c->cap_ghost, c->cap_global, d_lam, w, gs, ga,
                  G);
  launch_vjp_unpack(c->stream, n, ld, G, grad_pos, grad_vel, grad_rho);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_single_step_compute_vjp(sph2d_ctx* c, int32_t n, const double* lam_drhodt, const double* lam_acc,
                                  const double* lam_avg, int32_t ld, double* grad_pos, double* grad_vel,
                                  double* grad_rho) {
  int rc = vjp_prepare(c, n, ld, lam_drhodt, lam_acc);
  if (rc) return rc;
  std::vector<double> lam(5 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    lam[5 * (size_t)i + 0] = lam_drhodt[i];
    lam[5 * (size_t)i + 1] = lam_acc[i];
    lam[5 * (size_t)i + 2] = lam_acc[ld + i];
    lam[5 * (size_t)i + 3] = lam_avg ? lam_avg[i] : 0.0;
    lam[5 * (size_t)i + 4] = lam_avg ? lam_avg[ld + i] : 0.0;
  }
  const size_t nf = (size_t)c->cap_fluid, nl = (size_t)c->cap_local;
  double *d_lam = c->adj_buf, *w = d_lam + 5 * nf, *gs = w + 10 * nl, *ga = gs + 5 * nl, *G = ga + 5 * nl;
  std::vector<double> out(5 * (size_t)n);
  HIP_OR_FAIL(hipMemcpyAsync(d_lam, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice, c->stream));
  launch_eval_vjp(c->stream, c->P, c->B, c->cap_local, c->cap_fluid, c->cap_ghost, c->cap_global, d_lam, w, gs, ga,
                  G);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(out.data(), G, sizeof(double) * out.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    const double* g = &out[5 * (size_t)i];
    if (grad_pos) { grad_pos[i] = g[0]; grad_pos[ld + i] = g[1]; }
    if (grad_vel) { grad_vel[i] = g[2]; grad_vel[ld + i] = g[3]; }
    if (grad_rho) grad_rho[i] = g[4];
  }
  return SPH2D_OK;
}

int sph2d_reinitialize_particle_density(sph2d_ctx* c, int32_t it) {
  (void)it;
  int rc = need_state(c);
  if (rc) return rc;
  if (!c->evaluated) return fail(SPH2D_ERR_STATE, "no force evaluation in this step yet");
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_fill_i32(c->stream, c->B.counters + C_MLSFB, 1, 0);
  launch_mls(c->stream, c->P, c->B, c->cap_local);
  launch_shift_apply_mode(c->stream, c->P, c->B, c->cap_local, 2);   // commit rho
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_shift_particle_position(sph2d_ctx* c, int32_t it) {
  (void)it;
  int rc = need_state(c);
  if (rc) return rc;
  if (!c->evaluated) return fail(SPH2D_ERR_STATE, "no force evaluation in this step yet");
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_shift_apply_mode(c->stream, c->P, c->B, c->cap_local, 0);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_refresh_domain(sph2d_ctx* c, int32_t it) {
  (void)it;
  int rc = need_state(c);
  if (rc) return rc;
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_refresh(c->stream, c->P, c->B, c->cap_local);
  if (c->P.walls) launch_wall_refresh(c->stream, c->P, c->B, c->cap_local);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

// MS SPH_Time_Iteration.f90:58-540 with Iteration_Method = 2: Real_time += dt, the body
// motion, ill fluid, the local order, the predictor-corrector (an evaluation first on
// step 1), the density MLS every mls_every steps, shifting and the refresh.
static int enqueue_ms_step(sph2d_ctx* c, int it) {
  c->vjp_ready = false;
  c->eval_part = 0;
  hipStream_t s = c->stream;
  c->ms_time = c->ms_time + c->prm.dt;
  int rc = sph2d_body_motion(c, c->ms_time);
  if (rc) return rc;
  launch_ms_ill(s, c->P, c->B, c->cap_fluid);
  rc = devide_block(c, it, 0);
  if (rc) return rc;
  if (it == 1) do_eval_ms(c, it);
  launch_ms_stage(s, c->P, c->B, c->cap_local, MS_PREDICT);
  do_eval_ms(c, it);
  launch_ms_stage(s, c->P, c->B, c->cap_local, MS_CORRECT);
  const bool mls = c->prm.mls_on && (it % c->prm.mls_every) == 0;
  if (mls) {
    launch_fill_i32(s, c->B.counters + C_MLSFB, 1, 0);
    launch_ms_mls(s, c->P, c->B, c->cap_local);
  }
  const int mode = mls ? (c->prm.shift_on ? 1 : 2) : 0;
  launch_shift_refresh(s, c->P, c->B, c->cap_local, mode, c->prm.shift_on);
  launch_wall_refresh(s, c->P, c->B, c->cap_local);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

// time_iteration.f90 PC (:509-657) / RK-variant (:779-1049), then :1064-1078.
static int enqueue_step(sph2d_ctx* c, int it) {
  if (c->prm.ms) return enqueue_ms_step(c, it);
  c->vjp_ready = false;
  c->eval_part = 0;
  int rc = sph2d_distribute_period_boundary(c);
  if (rc) return rc;
  const bool pc = c->prm.integrator == SPH2D_INTEGRATOR_PC;
  const bool mls = c->prm.mls_on && (it % c->prm.mls_every) == 0;
  // After step 1 the predictor needs no evaluation first: it runs inside the local copy
  // (which keeps before_* only for an MLS step, whose corrector is k_stage's). In the
  // mixed form the FP64 local v, p, c, rho, volume are read only by MLS, the vorticity
  // of a sampling step and XSPH; other steps skip writing them.
  const bool sampling = it % c->prm.sample_every == 0;
  const bool lean = c->P.mixed && !mls && !sampling && c->prm.epsilon_xsph == 0.0 && !c->prm.walls;
  // before_* kept for k_stage's corrector: MLS steps, and every step with a damping zone
  rc = devide_block(c, it, pc && it > 1 ? ((mls || c->relax_on) ? 2 : (lean ? 3 : 1)) : 0);
  if (rc) return rc;
  hipStream_t s = c->stream;
  if (pc) {
    if (it == 1) {
      do_eval(c, false, it);
      launch_stage(s, c->P, c->B, c->cap_local, ST_PC_PREDICT);
    }
    do_eval(c, false, it);
    if (!mls && !c->relax_on) {
      // corrector, Shift_particle_position and Refresh_domain in one pass
      launch_correct_refresh(s, c->P, c->B, c->cap_local, c->prm.shift_on);
      if (c->P.walls) launch_wall_refresh(s, c->P, c->B, c->cap_local);
      HIP_OR_FAIL(hipGetLastError());
      return SPH2D_OK;
    }
    launch_stage(s, c->P, c->B, c->cap_local, ST_PC_CORRECT);
    if (c->relax_on) launch_relax_blend(s, c->P, c->B, c->relax, c->cap_local);   // time_iteration.f90:660-667
  } else {
    if (it == 1) do_eval(c, false, it);
    launch_stage(s, c->P, c->B, c->cap_local, ST_RK_HALF0);
    do_eval(c, false, it);
    launch_stage(s, c->P, c->B, c->cap_local, ST_RK_HALF1);
    do_eval(c, false, it);
    launch_stage(s, c->P, c->B, c->cap_local, ST_RK_FULL2);
    do_eval(c, false, it);
    launch_stage(s, c->P, c->B, c->cap_local, ST_RK_FINAL);
    if (c->relax_on) launch_relax_blend(s, c->P, c->B, c->relax, c->cap_local);   // :1052-1059
  }
  if (mls) {
    launch_fill_i32(s, c->B.counters + C_MLSFB, 1, 0);
    launch_mls(s, c->P, c->B, c->cap_local);
  }
  // Shift_particle_position (+ the MLS density) and Refresh_domain in one pass
  const int mode = mls ? (c->prm.shift_on ? 1 : 2) : 0;
  launch_shift_refresh(s, c->P, c->B, c->cap_local, mode, c->prm.shift_on);
  if (c->P.walls) launch_wall_refresh(s, c->P, c->B, c->cap_local);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_time_step(sph2d_ctx* c, int32_t it) {
  int rc = need_state(c);
  if (rc) return rc;
  if (it < 1) return fail(SPH2D_ERR_ARGUMENT, "i_time_step starts at 1");
  return enqueue_step(c, it);
}

int sph2d_run(sph2d_ctx* c, int32_t first, int32_t nsteps) {
  int rc = need_state(c);
  if (rc) return rc;
  if (first < 1 || nsteps < 0) return fail(SPH2D_ERR_ARGUMENT, "bad step range");
  for (int k = 0; k < nsteps; ++k) {
    rc = enqueue_step(c, first + k);
    if (rc) return rc;
  }
  return SPH2D_OK;
}

int sph2d_sample(sph2d_ctx* c, double* out14) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!out14) return fail(SPH2D_ERR_ARGUMENT, "out is NULL");
  launch_sample(c->stream, c->P, c->B, c->cap_fluid, c->d_out);
  HIP_OR_FAIL(hipGetLastError());
  double v[16];
  HIP_OR_FAIL(hipMemcpyAsync(v, c->d_out, sizeof(double) * 14, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  int cnt[C_COUNT];
  rc = read_counters(c, cnt);
  if (rc) return rc;
  const double N = (double)c->n_fluid;
  out14[0] = v[0] / c->ke0;
  out14[1] = v[1] / c->prm.U;
  for (int a = 0; a < 4; ++a) {
    out14[2 + 3 * a] = v[2 + 3 * a];
    out14[3 + 3 * a] = v[3 + 3 * a] / N;
    out14[4 + 3 * a] = std::sqrt(v[4 + 3 * a] / N);
  }
  return SPH2D_OK;
}

// ---- output -----------------------------------------------------------------------
int sph2d_get_scales(sph2d_ctx* c, sph2d_scales* out) {
  if (!c || !out) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  out->domain_size_y = c->prm.box_y;
  out->velocity_0 = c->prm.U;
  out->pressure_0 = c->prm.rho0 * c->prm.U * c->prm.U;
  out->time_0 = c->prm.box_x / c->prm.U;
  return SPH2D_OK;
}

int sph2d_download_output_fields(sph2d_ctx* c, int32_t n, double* vort_z, double* ana, int32_t ld, int32_t* step) {
  int rc = need_state(c);
  if (rc) return rc;
  if (n != c->n_fluid) return fail(SPH2D_ERR_ARGUMENT, "n must equal the number of fluid particles");
  if (ld < n) return fail(SPH2D_ERR_ARGUMENT, "ld < n");
  if (c->vort_step == 0 || c->vort_step != c->ana_step)
    return fail(SPH2D_ERR_STATE, "no sampling step (i_time_step a multiple of sample_every) has run yet");
  if (vort_z) HIP_OR_FAIL(hipMemcpyAsync(vort_z, c->B.g_vort, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> a;
  if (ana) {
    a.resize((size_t)4 * n);
    HIP_OR_FAIL(hipMemcpyAsync(a.data(), c->B.ana, a.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  if (ana)
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < n; ++i) ana[(size_t)k * ld + i] = a[(size_t)4 * i + k];
  if (step) *step = c->vort_step;
  return SPH2D_OK;
}

// Surface particles of the last evaluation in gid order (Calculation_For_PostProceeding.f90:96-102
// builds surface_particle_name this way) with their positions and normals.
static int surface_particles(sph2d_ctx* c, std::vector<double>& pos, std::vector<double>& nrm,
                             std::vector<int>* gids = nullptr) {
  int cnt[C_COUNT];
  int rc = read_counters(c, cnt);
  if (rc) return rc;
  const int nl = cnt[C_NLOCAL];
  std::vector<int> surf(nl), name(nl);
  std::vector<double2> lp(nl), ln(nl);
  HIP_OR_FAIL(hipMemcpyAsync(surf.data(), c->B.surf, nl * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(name.data(), c->B.name, nl * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(ln.data(), c->B.nrm, nl * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  std::vector<std::pair<int, int>> sel;              // (gid, local slot)
  for (int l = 0; l < nl; ++l)
    if (surf[l] == 1 && name[l] < c->n_fluid) sel.emplace_back(name[l], l);
  std::sort(sel.begin(), sel.end());
  const int m = (int)sel.size();
  std::vector<double2> gp(c->n_fluid);
  HIP_OR_FAIL(hipMemcpyAsync(gp.data(), c->B.g_pos, c->n_fluid * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  pos.assign((size_t)2 * m, 0.0);
  nrm.assign((size_t)2 * m, 0.0);
  if (gids) {
    gids->resize(m);
    for (int k = 0; k < m; ++k) (*gids)[k] = sel[k].first;
  }
  for (int k = 0; k < m; ++k) {
    pos[k] = gp[sel[k].first].x; pos[m + k] = gp[sel[k].first].y;
    nrm[k] = ln[sel[k].second].x; nrm[m + k] = ln[sel[k].second].y;
  }
  return SPH2D_OK;
}

int sph2d_output_for_tecplot(sph2d_ctx* c, int32_t it, const char* video_path, const char* surface_path,
                             const char* binary_path) {
  int rc = need_state(c);
  if (rc) return rc;
  if (it != c->vort_step || it != c->ana_step)
    return fail(SPH2D_ERR_STATE, "i_time_step must be the last step run and a multiple of sample_every");
  const int n = c->n_fluid;
  sph2d_scales sc;
  sph2d_get_scales(c, &sc);
  std::vector<double> pos((size_t)2 * n), vel((size_t)2 * n), press(n);
  rc = sph2d_download_state(c, n, pos.data(), vel.data(), nullptr, press.data(), nullptr, nullptr, nullptr, n);
  if (rc) return rc;
  const double t = it * c->prm.dt;                   // (i_time_step - standing_time_step) * dt
  if (video_path) {
    std::vector<double> vort(n), ana((size_t)4 * n);
    rc = sph2d_download_output_fields(c, n, vort.data(), ana.data(), n, nullptr);
    if (rc) return rc;
    rc = sph2d_write_tecplot_zone(video_path, t, n, pos.data(), vel.data(), press.data(), vort.data(), ana.data(), n,
                                  &sc);
    if (rc) return rc;
  }
  if (surface_path) {
    std::vector<double> sp, sn;
    rc = surface_particles(c, sp, sn);
    if (rc) return rc;
    const int m = (int)(sp.size() / 2);
    rc = sph2d_write_surface_zone(surface_path, t, m, sp.data(), sn.data(), m, &sc);
    if (rc) return rc;
  }
  if (binary_path) {
    rc = sph2d_write_tecplot_binary(binary_path, t, n, pos.data(), vel.data(), press.data(), n, &sc);
    if (rc) return rc;
  }
  return SPH2D_OK;
}

int sph2d_grid_size(sph2d_ctx* c, int32_t* element_number_x, int32_t* element_number_y, double* element_dx) {
  if (!c || !element_number_x || !element_number_y || !element_dx) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  // information_module.f90:179-183 with Factor_Grid_Point = 1 and the box as boundary_size
  const int nex = (int)(c->prm.box_x / (1.0 * c->prm.dx));
  const int ney = (int)(c->prm.box_y / (1.0 * c->prm.dx));
  if (nex < 1 || ney < 1) return fail(SPH2D_ERR_ARGUMENT, "box smaller than one particle spacing");
  *element_number_x = nex;
  *element_number_y = ney;
  *element_dx = c->prm.box_x / nex;
  return SPH2D_OK;
}

int sph2d_calculation_for_postproceeding(sph2d_ctx* c, int32_t it, int32_t n_nodes, double* vel, double* press,
                                         double* vort_z, double* fai, double* ana, int32_t ld) {
  int rc = need_state(c);
  if (rc) return rc;
  int32_t nex, ney;
  double edx;
  rc = sph2d_grid_size(c, &nex, &ney, &edx);
  if (rc) return rc;
  const int nnx = nex + 1, nny = ney + 1, nn = nnx * nny;
  if (n_nodes != nn) return fail(SPH2D_ERR_ARGUMENT, "n_nodes must be (element_number_x+1)*(element_number_y+1) = " +
                                                        std::to_string(nn));
  if (ld < nn) return fail(SPH2D_ERR_ARGUMENT, "ld < n_nodes");
  if (it != c->vort_step || it != c->ana_step)
    return fail(SPH2D_ERR_STATE, "i_time_step must be the last step run and a multiple of sample_every");
  std::vector<double> sp, sn;
  std::vector<int> sg;                               // surface_particle_name (gid order)
  rc = surface_particles(c, sp, sn, &sg);
  if (rc) return rc;
  const int ns = (int)sg.size();
  if ((size_t)ns > c->surf_cap) {
    rc = c->alloc(&c->surf_gid, ns);
    if (!rc) rc = c->alloc(&c->surf_nrm, ns);
    if (rc) return rc;
    c->surf_cap = ns;
  }
  if (ns > 0) {
    std::vector<double2> nr(ns);
    for (int q = 0; q < ns; ++q) nr[q] = make_double2(sn[q], sn[ns + q]);
    HIP_OR_FAIL(hipMemcpyAsync(c->surf_gid, sg.data(), ns * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(hipMemcpyAsync(c->surf_nrm, nr.data(), ns * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  }
  if ((size_t)9 * nn > c->grid_cap) {
    rc = c->alloc(&c->grid_out, (size_t)9 * nn);
    if (rc) return rc;
    c->grid_cap = (size_t)9 * nn;
  }
  GridArgs G;
  G.nnx = nnx; G.nny = nny; G.n_nodes = nn; G.n_surf = ns;
  G.edx = edx;
  G.t_nondim = it * c->prm.dt / (c->prm.box_x / c->prm.U);
  G.Re = c->prm.U * c->prm.box_x / c->prm.nu;
  G.surf_gid = c->surf_gid; G.surf_nrm = c->surf_nrm; G.out = c->grid_out;
  launch_grid_mls(c->stream, c->P, c->B, G);
  HIP_OR_FAIL(hipGetLastError());
  std::vector<double> o((size_t)9 * nn);
  HIP_OR_FAIL(hipMemcpyAsync(o.data(), c->grid_out, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  const size_t N = nn, L = ld;
  for (size_t k = 0; k < N; ++k) {
    if (vel) { vel[k] = o[k]; vel[L + k] = o[N + k]; }
    if (press) press[k] = o[2 * N + k];
    if (vort_z) vort_z[k] = o[3 * N + k];
    if (fai) fai[k] = o[4 * N + k];
    if (ana)
      for (int q = 0; q < 4; ++q) ana[q * L + k] = o[(5 + q) * N + k];
  }
  return SPH2D_OK;
}

int sph2d_output_for_grid_postproceding(sph2d_ctx* c, int32_t it, const char* grid_path, const char* binary_path) {
  int32_t nex, ney;
  double edx;
  int rc = sph2d_grid_size(c, &nex, &ney, &edx);
  if (rc) return rc;
  const int nn = (nex + 1) * (ney + 1);
  std::vector<double> vel((size_t)2 * nn), press(nn), vort(nn), fai(nn), ana((size_t)4 * nn);
  rc = sph2d_calculation_for_postproceeding(c, it, nn, vel.data(), press.data(), vort.data(), fai.data(), ana.data(),
                                            nn);
  if (rc) return rc;
  sph2d_scales sc;
  sph2d_get_scales(c, &sc);
  const double t = it * c->prm.dt;
  if (grid_path) {
    rc = sph2d_write_grid_zone(grid_path, t, nex, ney, edx, vel.data(), press.data(), vort.data(), fai.data(),
                               ana.data(), nn, &sc);
    if (rc) return rc;
  }
  if (binary_path) {
    rc = sph2d_write_grid_binary(binary_path, t, nex, ney, edx, vel.data(), press.data(), fai.data(), nn, &sc);
    if (rc) return rc;
  }
  return SPH2D_OK;
}

namespace {
int probe_space(sph2d_ctx* c, size_t n) {
  if (n > c->probe_cap) {
    const int rc = c->alloc(&c->probe_buf, n);
    if (rc) return rc;
    c->probe_cap = n;
  }
  return SPH2D_OK;
}
}  // namespace

int sph2d_sampling(sph2d_ctx* c, int32_t it, int32_t n, const double* pos, int32_t ld, double gravity,
                   double* average_press, double* mls_press, double* vel) {
  int rc = need_state(c);
  if (rc) return rc;
  if (n < 0 || (n > 0 && !pos)) return fail(SPH2D_ERR_ARGUMENT, "n < 0 or pos is NULL");
  if (ld < n) return fail(SPH2D_ERR_ARGUMENT, "ld < n");
  if (it != c->ana_step || it != c->vort_step)
    return fail(SPH2D_ERR_STATE, "i_time_step must be the last step run and a multiple of sample_every");
  if (n == 0) return SPH2D_OK;
  const double lo_x = c->P.chain_ox, lo_y = c->P.chain_oy;
  const double hi_x = lo_x + c->P.chain_nx * c->P.chain_d, hi_y = lo_y + c->P.chain_ny * c->P.chain_d;
  std::vector<double> h((size_t)2 * n);
  for (int k = 0; k < n; ++k) {
    const double x = pos[k], y = pos[(size_t)ld + k];
    if (!(x >= lo_x && x < hi_x && y >= lo_y && y < hi_y))
      return fail(SPH2D_ERR_ARGUMENT, "sampling point " + std::to_string(k + 1) + " lies outside the background grid");
    h[2 * k] = x;
    h[2 * k + 1] = y;
  }
  rc = probe_space(c, (size_t)6 * n);
  if (rc) return rc;
  HIP_OR_FAIL(hipMemcpyAsync(c->probe_buf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  ProbeArgs Q;
  Q.n = n;
  Q.gravity = gravity;
  Q.pos = reinterpret_cast<const double2*>(c->probe_buf);
  Q.out = c->probe_buf + (size_t)2 * n;
  launch_probe_mls(c->stream, c->P, c->B, Q);
  HIP_OR_FAIL(hipGetLastError());
  std::vector<double> o((size_t)4 * n);
  HIP_OR_FAIL(hipMemcpyAsync(o.data(), Q.out, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  for (int k = 0; k < n; ++k) {
    if (average_press) average_press[k] = o[k];
    if (mls_press) mls_press[k] = o[(size_t)n + k];
    if (vel) { vel[k] = o[(size_t)2 * n + k]; vel[(size_t)ld + k] = o[(size_t)3 * n + k]; }
  }
  return SPH2D_OK;
}

int sph2d_wave_probes(sph2d_ctx* c, int32_t n_probes, const double* probe_x, double* wave_front, double* elevation) {
  int rc = need_state(c);
  if (rc) return rc;
  if (n_probes < 0 || (n_probes > 0 && (!probe_x || !elevation)))
    return fail(SPH2D_ERR_ARGUMENT, "n_probes < 0 or a NULL array");
  rc = probe_space(c, (size_t)2 * n_probes + 1);
  if (rc) return rc;
  // probe x, then the 1 + n_probes maxima as bit patterns (zero = +0.0)
  std::vector<double> h((size_t)2 * n_probes + 1, 0.0);
  for (int j = 0; j < n_probes; ++j) h[j] = probe_x[j];
  HIP_OR_FAIL(hipMemcpyAsync(c->probe_buf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  auto* bits = reinterpret_cast<unsigned long long*>(c->probe_buf + n_probes);
  launch_wave_probes(c->stream, c->P, c->B, c->cap_fluid, c->probe_buf, n_probes, bits);
  HIP_OR_FAIL(hipGetLastError());
  std::vector<double> o((size_t)n_probes + 1);
  HIP_OR_FAIL(hipMemcpyAsync(o.data(), bits, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  if (wave_front) *wave_front = o[0];
  for (int j = 0; j < n_probes; ++j) elevation[j] = o[1 + j];
  return SPH2D_OK;
}

int sph2d_output_sampling_lines(sph2d_ctx* c, int32_t it, const char* dir, int32_t n_lines, const double* first,
                                const double* last, int32_t n_points) {
  if (!c || !dir || n_lines < 0 || (n_lines > 0 && (!first || !last)) || n_points < 2)
    return fail(SPH2D_ERR_ARGUMENT, "bad argument (n_points must be at least 2)");
  sph2d_scales sc;
  int rc = sph2d_get_scales(c, &sc);
  if (rc) return rc;
  const double tn = it * c->prm.dt / (c->prm.box_x / c->prm.U);   // Nondimensionalized_real_time
  const double Re = c->prm.U * c->prm.box_x / c->prm.nu;
  const size_t n = n_points;
  std::vector<double> pos(2 * n), avg(n), mls(n), vel(2 * n), ana_v(2 * n), ana_p(n);
  for (int k = 0; k < n_lines; ++k) {
    rc = sph2d_sampling_line_points(n_points, first + 2 * k, last + 2 * k, pos.data(), n_points);
    if (rc) return rc;
    rc = sph2d_sampling(c, it, n_points, pos.data(), n_points, 0.0, avg.data(), mls.data(), vel.data());
    if (rc) return rc;
    // Sampling_subroutine.f90:481-485
    const double pi = kPiModule, U = c->prm.U;
    for (size_t i = 0; i < n; ++i) {
      const double x = pos[i], y = pos[n + i];
      ana_v[i] = U * std::sin(2 * pi * x) * std::cos(2 * pi * y) * std::exp(-8 * (pi * pi) / Re * tn);
      ana_v[n + i] = -U * std::cos(2 * pi * x) * std::sin(2 * pi * y) * std::exp(-8 * (pi * pi) / Re * tn);
      ana_p[i] = 0.25 * c->prm.rho0 * (U * U) * (std::cos(4 * pi * x) + std::cos(4 * pi * y)) *
                 std::exp(-16 * (pi * pi) / Re * tn);
    }
    rc = sph2d_write_sampling_line_results(dir, k + 1, tn, n_points, pos.data(), ana_v.data(), ana_p.data(),
                                           vel.data(), mls.data(), avg.data(), n_points, &sc);
    if (rc) return rc;
  }
  return SPH2D_OK;
}

// ---- wave-theory wavemaker and its damping (relaxation) zone ---------------------------
int sph2d_relaxation_init(sph2d_ctx* c, double zone_length) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!c->prm.walls) return fail(SPH2D_ERR_UNSUPPORTED, "the wavemaker zone needs a walled tank");
  if (!std::isfinite(zone_length)) return fail(SPH2D_ERR_ARGUMENT, "zone_length must be finite");
  if (!c->relax_ready) {
    const size_t G = (size_t)c->cap_global;
    if ((rc = c->alloc(&c->relax.in_zone, G))) return rc;
    if ((rc = c->alloc(&c->relax.coef, G))) return rc;
    if ((rc = c->alloc(&c->relax.an_pos, G))) return rc;
    if ((rc = c->alloc(&c->relax.an_vel, G))) return rc;
    if ((rc = c->alloc(&c->relax.init_pos, G))) return rc;
    if ((rc = c->alloc(&c->relax.an_p, G))) return rc;
    if ((rc = c->alloc(&c->relax.an_rho, G))) return rc;
    c->relax_ready = true;
  }
  c->vjp_ready = false;
  launch_relax_init(c->stream, c->P, c->B, c->relax, c->n_fluid + c->n_wm, zone_length);
  HIP_OR_FAIL(hipGetLastError());
  c->relax_on = zone_length > 0.0;
  return SPH2D_OK;
}

int sph2d_wave_theory_update(sph2d_ctx* c, const sph2d_wavemaker* w, int32_t it) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!w) return fail(SPH2D_ERR_ARGUMENT, "w is NULL");
  if (w->wave_type != SPH2D_WAVE_STOKES2 && w->wave_type != SPH2D_WAVE_AIRY)
    return fail(SPH2D_ERR_UNSUPPORTED, "the wave-theory wavemaker is built for Stokes_2nd and Airy waves");
  if (w->soft_steps < 1) return fail(SPH2D_ERR_ARGUMENT, "soft_steps must be >= 1");
  if (!c->relax_ready) return fail(SPH2D_ERR_STATE, "sph2d_relaxation_init first (it keeps the initial positions)");
  WaveArgs W;
  W.wave_type = w->wave_type;
  const double pi = kPiModule;
  // Wavemaker_Subroutine.f90:280-287: (i_time_step - standing_time_step) / wave_maker_soft_time
  // divides integers
  const int rel = it - w->standing_steps;
  W.soft = it <= w->soft_steps + w->standing_steps ? 0.5 * (std::sin((-0.5 + rel / w->soft_steps) * pi) + 1) : 1.0;
  W.t = w->dt * rel;
  W.g = w->g; W.d = w->water_depth; W.H = w->wave_height; W.k = w->wave_number; W.om = w->wave_frequency;
  W.phase = w->phase;
  W.T = 2 * pi / W.om; W.L = 2 * pi / W.k; W.A = W.H / 2.0;
  W.skd = std::sinh(W.k * W.d); W.ckd = std::cosh(W.k * W.d);
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_wave_theory(c->stream, c->P, c->B, c->relax, W, c->cap_fluid, c->n_wm, c->relax_on ? 1 : 0);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_wave_maker_damping_zone(sph2d_ctx* c, int32_t it) {
  (void)it;
  int rc = need_state(c);
  if (rc) return rc;
  if (!c->relax_on) return fail(SPH2D_ERR_STATE, "no wavemaker damping zone (sph2d_relaxation_init with L > 0)");
  c->vjp_ready = false;
  launch_relax_blend(c->stream, c->P, c->B, c->relax, c->cap_local);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

// ---- piston wavemaker -------------------------------------------------------------
int sph2d_wavemaker_update(sph2d_ctx* c, double position, double velocity) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!c->prm.wavemaker_layers) return fail(SPH2D_ERR_STATE, "the context has no wavemaker particles");
  if (!std::isfinite(position) || !std::isfinite(velocity))
    return fail(SPH2D_ERR_ARGUMENT, "paddle position and velocity must be finite");
  rc = probe_space(c, 1);
  if (rc) return rc;
  c->vjp_ready = false;
  c->eval_part = 0;
  auto* bits = reinterpret_cast<unsigned long long*>(c->probe_buf);
  launch_fill_i32(c->stream, reinterpret_cast<int*>(bits), 2, 0);
  launch_wavemaker_place(c->stream, c->P, c->B, c->cap_fluid, c->n_wm, position, velocity, bits);
  launch_wavemaker_pull_back(c->stream, c->P, c->B, c->cap_fluid, position, velocity);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_wavemaker_pull_back(sph2d_ctx* c, double position, double velocity) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!std::isfinite(position) || !std::isfinite(velocity))
    return fail(SPH2D_ERR_ARGUMENT, "paddle position and velocity must be finite");
  c->vjp_ready = false;
  c->eval_part = 0;
  launch_wavemaker_pull_back(c->stream, c->P, c->B, c->cap_fluid, position, velocity);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

int sph2d_output_for_paraview(sph2d_ctx* c, int32_t it, const char* path) {
  (void)it;                                          // the caller names the file by the step's time
  int rc = need_state(c);
  if (rc) return rc;
  const int n = c->n_fluid;
  sph2d_scales sc;
  sph2d_get_scales(c, &sc);
  std::vector<double> pos((size_t)2 * n), vel((size_t)2 * n), press(n);
  rc = sph2d_download_state(c, n, pos.data(), vel.data(), nullptr, press.data(), nullptr, nullptr, nullptr, n);
  if (rc) return rc;
  return sph2d_write_paraview(path, n, pos.data(), vel.data(), press.data(), n, &sc);
}

// ---- inspection -------------------------------------------------------------------
int sph2d_set_debug(sph2d_ctx* c, int32_t on) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  if (on && !c->B.dbg) {
    int rc = c->alloc(&c->B.dbg, (size_t)10 * c->cap_local);
    if (rc) return rc;
  }
  c->debug = on != 0;
  return SPH2D_OK;
}

int sph2d_download_local(sph2d_ctx* c, const char* name, void* out, int32_t n) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!name || !out) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  int cnt[C_COUNT];
  rc = read_counters(c, cnt);
  if (rc) return rc;
  const int nl = cnt[C_NLOCAL];
  if (n != nl) return fail(SPH2D_ERR_ARGUMENT, "n must equal the local particle count " + std::to_string(nl));
  Buffers& B = c->B;
  const std::string s(name);
  auto d2 = [&](double2* src, int comp) -> int {
    std::vector<double2> t(nl);
    HIP_OR_FAIL(hipMemcpyAsync(t.data(), src, nl * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    double* o = static_cast<double*>(out);
    for (int i = 0; i < nl; ++i) o[i] = comp ? t[i].y : t[i].x;
    return SPH2D_OK;
  };
  auto d1 = [&](const double* src, int stride, int comp) -> int {
    std::vector<double> t((size_t)nl * stride);
    HIP_OR_FAIL(hipMemcpyAsync(t.data(), src, (size_t)nl * stride * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    double* o = static_cast<double*>(out);
    for (int i = 0; i < nl; ++i) o[i] = t[(size_t)i * stride + comp];
    return SPH2D_OK;
  };
  auto i1 = [&](const int* src) -> int {
    HIP_OR_FAIL(hipMemcpyAsync(out, src, nl * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    return SPH2D_OK;
  };
  if (s == "x") return d2(B.l_pos, 0);
  if (s == "y") return d2(B.l_pos, 1);
  if (s == "u") return d2(B.l_vel, 0);
  if (s == "v") return d2(B.l_vel, 1);
  if (s == "rho") return d2(B.l_rv, 0);
  if (s == "volume") return d2(B.l_rv, 1);
  if (s == "p") return d2(B.l_pc, 0);
  if (s == "c") return d2(B.l_pc, 1);
  if (s == "mass") return d1(B.l_mass, 1, 0);
  if (s == "fra_v") return d1(B.fra_v, 1, 0);
  if (s == "sum_dwdx_x") return d2(B.sdw, 0);
  if (s == "sum_dwdx_y") return d2(B.sdw, 1);
  if (s == "normal_x") return d2(B.nrm, 0);
  if (s == "normal_y") return d2(B.nrm, 1);
  // the FP32-mixed pass A keeps tau and grad(rho) in its records only
  auto f4 = [&](const float4* src, int comp) -> int {
    std::vector<float4> t(nl);
    HIP_OR_FAIL(hipMemcpyAsync(t.data(), src, nl * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    double* o = static_cast<double*>(out);
    for (int i = 0; i < nl; ++i) o[i] = comp == 0 ? t[i].x : (comp == 1 ? t[i].y : t[i].z);
    return SPH2D_OK;
  };
  const bool mx = c->P.mixed != 0;
  if (s == "tau_xx") return mx ? f4(B.m_r1, 0) : d1(B.tau, 3, 0);
  if (s == "tau_xy") return mx ? f4(B.m_r1, 1) : d1(B.tau, 3, 1);
  if (s == "tau_yy") return mx ? f4(B.m_r1, 2) : d1(B.tau, 3, 2);
  if (s == "grad_rho_x") return mx ? f4(B.m_r2, 0) : d2(B.grad, 0);
  if (s == "grad_rho_y") return mx ? f4(B.m_r2, 1) : d2(B.grad, 1);
  static const char* mg_names[6] = {"grad_p_x", "grad_p_y", "grad_u_x", "grad_u_y", "grad_v_x", "grad_v_y"};
  for (int k = 0; k < 6; ++k)
    if (s == mg_names[k]) {
      if (!B.mg) return fail(SPH2D_ERR_STATE, "pressure and velocity gradients exist for the MUSCL solver only");
      std::vector<double2> t((size_t)nl * 3);
      HIP_OR_FAIL(hipMemcpyAsync(t.data(), B.mg, (size_t)nl * 3 * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
      HIP_OR_FAIL(hipStreamSynchronize(c->stream));
      double* o = static_cast<double*>(out);
      for (int i = 0; i < nl; ++i) o[i] = (k & 1) ? t[(size_t)i * 3 + k / 2].y : t[(size_t)i * 3 + k / 2].x;
      return SPH2D_OK;
    }
  if (s == "drhodt") return d1(B.drhodt, 1, 0);
  if (s == "ax") return d2(B.acc, 0);
  if (s == "ay") return d2(B.acc, 1);
  if (s == "shift_x") return d2(B.shift, 0);
  if (s == "shift_y") return d2(B.shift, 1);
  if (s == "avg_x") return d2(B.avg, 0);
  if (s == "avg_y") return d2(B.avg, 1);
  if (s == "type") return i1(B.l_type);
  if (s == "name") return i1(B.name);
  if (s == "surface") return i1(B.surf);
  if (s == "near_surface") return i1(B.near_);
  if (s == "neighbour_count") {                 // the last pass A's recorded neighbours
    std::vector<int> t((size_t)nl);
    HIP_OR_FAIL(hipMemcpyAsync(t.data(), B.nb_count, nl * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
    int* o = static_cast<int*>(out);
    for (int i = 0; i < nl; ++i) o[i] = t[i] & (kNbOverflow - 1);
    return SPH2D_OK;
  }
  static const char* dbg_names[10] = {"internal_x", "internal_y", "visc_x", "visc_y", "art_internal_x",
                                      "art_internal_y", "sps_x", "sps_y", "lap_x", "lap_y"};
  for (int k = 0; k < 10; ++k)
    if (s == dbg_names[k]) {
      if (!B.dbg || !c->debug || c->P.mixed)
        return fail(SPH2D_ERR_STATE, "acceleration components need the FP64 form and sph2d_set_debug(ctx,1) before the evaluation");
      return d1(B.dbg, 10, k);
    }
  return fail(SPH2D_ERR_ARGUMENT, "unknown array name '" + s + "'");
}

int sph2d_download_pairs(sph2d_ctx* c, int64_t capacity, int32_t* lo, int32_t* hi, int64_t* n_pairs) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!n_pairs) return fail(SPH2D_ERR_ARGUMENT, "n_pairs is NULL");
  launch_count_pairs(c->stream, c->P, c->B, c->cap_local);
  int cnt[C_COUNT];
  rc = read_counters(c, cnt);
  if (rc) return rc;
  const int64_t np = (int64_t)(unsigned)cnt[C_NPAIR_LO];
  *n_pairs = np;
  if (!lo || !hi || capacity <= 0) return SPH2D_OK;
  if (capacity < np) return fail(SPH2D_ERR_CAPACITY, "pair capacity too small");
  Buffers& B = c->B;
  int *dlo = nullptr, *dhi = nullptr;
  HIP_OR_FAIL(hipMalloc(&dlo, std::max<int64_t>(np, 1) * sizeof(int)));
  HIP_OR_FAIL(hipMalloc(&dhi, std::max<int64_t>(np, 1) * sizeof(int)));
  B.pair_lo = dlo; B.pair_hi = dhi;
  launch_pairs(c->stream, c->P, B, c->cap_local, np, nullptr);
  std::vector<int> a(np), b(np);
  HIP_OR_FAIL(hipMemcpyAsync(a.data(), dlo, np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipMemcpyAsync(b.data(), dhi, np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  (void)hipFree(dlo); (void)hipFree(dhi);
  B.pair_lo = nullptr; B.pair_hi = nullptr;
  std::vector<int64_t> key(np);
  for (int64_t k = 0; k < np; ++k) key[k] = ((int64_t)a[k] << 32) | (uint32_t)b[k];
  std::sort(key.begin(), key.end());
  for (int64_t k = 0; k < np; ++k) { lo[k] = (int32_t)(key[k] >> 32); hi[k] = (int32_t)(key[k] & 0xffffffff); }
  return SPH2D_OK;
}

int sph2d_get_diagnostics(sph2d_ctx* c, sph2d_diag* d) {
  if (!c || !d) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  if (c->have_state) launch_max_cell(c->stream, c->B, c->n_cells);
  int cnt[C_COUNT];
  int rc = read_counters(c, cnt);
  if (rc) return rc;
  std::memset(d, 0, sizeof(*d));
  d->n_fluid = cnt[C_NFLUID];
  d->n_ghost = cnt[C_NGHOST];
  d->n_local = cnt[C_NLOCAL];
  d->n_ill = cnt[C_NILL];
  d->n_surface = cnt[C_NSURF];
  d->n_singular = cnt[C_NSING];
  d->n_mls_fallback = cnt[C_MLSFB];
  d->max_cell_count = cnt[C_MAXCELL];
  d->rebin_needed = cnt[C_REBIN];
  d->error_flags = cnt[C_ERR];
  d->n_pairs = -1;
  return SPH2D_OK;
}

int sph2d_time_pass(sph2d_ctx* c, int32_t pass, int32_t reps, double* avg_ms) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!avg_ms || reps <= 0 || pass < 0 || pass > 2) return fail(SPH2D_ERR_ARGUMENT, "bad argument");
  if (!c->evaluated) return fail(SPH2D_ERR_STATE, "no force evaluation yet");
  hipEvent_t e0, e1;
  HIP_OR_FAIL(hipEventCreate(&e0));
  HIP_OR_FAIL(hipEventCreate(&e1));
  c->P.write_debug = 0;
  HIP_OR_FAIL(hipEventRecord(e0, c->stream));
  for (int k = 0; k < reps; ++k) launch_eval_pass(c->stream, c->P, c->B, c->cap_local, pass);
  HIP_OR_FAIL(hipEventRecord(e1, c->stream));
  HIP_OR_FAIL(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_OR_FAIL(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / reps;
  return SPH2D_OK;
}

int sph2d_set_pass_timing(sph2d_ctx* c, int32_t on) {
  if (!c) return fail(SPH2D_ERR_ARGUMENT, "ctx is NULL");
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  c->pass_timing = on != 0;
  c->events_used = 0;
  return SPH2D_OK;
}

int sph2d_get_pass_timing(sph2d_ctx* c, double* out3, int32_t* n_evaluations) {
  if (!c || !out3 || !n_evaluations) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  const int n = (int)(c->events_used / 4);
  double fa = 0.0, fc = 0.0, fe = 0.0;
  for (int k = 0; k < n; ++k) {
    float a = 0.f, b = 0.f, e = 0.f;
    HIP_OR_FAIL(hipEventElapsedTime(&a, c->events[4 * k], c->events[4 * k + 1]));
    HIP_OR_FAIL(hipEventElapsedTime(&b, c->events[4 * k + 2], c->events[4 * k + 3]));
    HIP_OR_FAIL(hipEventElapsedTime(&e, c->events[4 * k], c->events[4 * k + 3]));
    fa += a;
    fc += b;
    fe += e;
  }
  *n_evaluations = n;
  out3[0] = n ? fc / n : 0.0;
  out3[1] = n ? fa / n : 0.0;
  out3[2] = n ? fe / n : 0.0;
  c->events_used = 0;
  return SPH2D_OK;
}

int sph2d_get_stream(sph2d_ctx* c, void** stream) {
  if (!c || !stream) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  *stream = (void*)c->stream;
  return SPH2D_OK;
}

}  // extern "C"

// Wavemaker_Subroutine.f90:66-138 (PistonType) and wavemaker_plate_motion.f90 (Goring's
// solitary-wave paddle). Host arithmetic in double; the reference's REAL(4) literals
// (EZ = 0.002, CERR = 0.0001) keep their single-precision values.
int sph2d_wavemaker_motion(const sph2d_wavemaker* w, int32_t it, double* position, double* velocity,
                           double* acceleration) {
  if (!w || !position || !velocity || !acceleration) return fail(SPH2D_ERR_ARGUMENT, "NULL argument");
  if (w->wave_type < SPH2D_WAVE_SOLITARY || w->wave_type > SPH2D_WAVE_AIRY)
    return fail(SPH2D_ERR_ARGUMENT, "wave_type must be one of SPH2D_WAVE_*");
  if (!(w->dt > 0) || !(w->water_depth > 0) || w->soft_steps <= 0)
    return fail(SPH2D_ERR_ARGUMENT, "dt, water_depth and soft_steps must be positive");
  const double pi = kPiModule;
  // :69-73. (i - standing) / soft is an integer division, so the coefficient is
  // 0.5 (sin(-pi/2) + 1) = 0 before the soft time ends and 1 from then on.
  double soft = 1.0;
  if (it <= w->soft_steps + w->standing_steps) {
    const int q = (it - w->standing_steps) / w->soft_steps;
    soft = 0.5 * (std::sin((double)(-0.5f + (float)q) * pi) + 1);
  }
  double x = 0.0, u = 0.0, a = 0.0;
  if (it > w->standing_steps) {
    const double t = w->dt * (it - w->standing_steps);
    const double d = w->water_depth, H = w->wave_height;
    if (w->wave_type == SPH2D_WAVE_SOLITARY) {
      // wavemaker_plate_motion.f90:22-58, called with WMX = WMV = 0 (:75-77)
      const double ez = (double)0.002f, g = w->g;
      const double cl = std::log((1. / std::pow(ez, 0.5)) * (1 + std::sqrt(1. - ez)));
      const double ck = std::sqrt(3. * H / (4. * d));
      const double c = std::sqrt(g * (d + H));
      const double clamda = cl / ck * d;
      const double cerr = (double)0.0001f;
      double xp = 0.0 + 0.0 * w->dt, xt, err;
      int iter = 0;
      do {
        xt = ck / d * (c * t - xp - clamda);
        const double xpi = H / ck * (std::tanh(xt) + std::tanh(ck * clamda / d));
        err = std::fabs(xpi - xp);
        xp = xpi;
        if (++iter > 100000) return fail(SPH2D_ERR_NONFINITE, "solitary paddle iteration did not converge");
      } while (err > cerr);
      const double ch = std::cosh(xt), sh = std::sinh(xt);
      const double up = H * c / (d * (std::pow(ch, 2.) + H / d));
      double vp = 2 * H * c * ch * sh * ck * c;
      const double den = d * (std::pow(ch, 2.) + H / d);
      vp = vp / (den * den);
      x = xp; u = up; a = vp;
    } else {
      // :93-135 (Stokes_2nd and Airy, DualSPHysics' long-crested piston laws)
      const double k = w->wave_number, om = w->wave_frequency;
      const double sh = std::sinh(k * d), ch = std::cosh(k * d);
      const double m1 = 2 * (sh * sh) / (sh * ch + k * d);
      const double th = om * t + w->phase;
      const double S = H / m1;
      x = S / 2.0 * std::sin(th);
      u = S / 2.0 * om * std::cos(th);
      a = -S / 2.0 * (om * om) * std::sin(th);
      if (w->wave_type == SPH2D_WAVE_STOKES2) {
        const double s3 = std::sinh(d * k);
        const double b = (H * H / (32 * d)) * (3 * std::cosh(d * k) / (s3 * s3 * s3)) - 2.0 / m1;
        x = x + b * std::sin(2 * th);
        u = u + b * 2 * om * std::cos(2 * th);
        a = a - b * 4 * (om * om) * std::sin(2 * th);
      }
      x = soft * x; u = soft * u; a = soft * a;
    }
  }
  *position = x;
  *velocity = u;
  *acceleration = a;
  return SPH2D_OK;
}

// ---- batched form: independent contexts advanced together ------------------------------
// Steps whose kernel sequence and arguments repeat (after step 1, neither an MLS nor a
// sampling step) are replayed from one hipGraph that holds every instance's step as a
// parallel branch; step 1, MLS and sampling steps are enqueued per context. The
// instances never exchange data, so each one's results are those of running it alone.
struct sph2d_batch {
  std::vector<sph2d_ctx*> ctx;
  hipStream_t stream = nullptr;               // graph launches
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join, sync;          // per context: capture joins, eager/graph ordering
  hipGraph_t graph = nullptr, graph_mls = nullptr;
  hipGraphExec_t exec = nullptr, exec_mls = nullptr;   // exec_mls: MLS steps, fused form only
  size_t nodes = 0;
  int replays = 0;
  bool on_batch = false;                       // the latest work of the batch is on `stream`
  // one launch per kernel over all instances (when their forms allow it)
  bool fused = false;
  int config = 0;                              // the fused launches' configuration (autotuner)
  BatchDev dev;
  std::vector<void*> dev_allocs;
};

namespace {
// 0: enqueued per context; 1: a repeating step; 2: an MLS step (replayed in the fused form)
int step_kind(const sph2d_ctx* c, int it) {
  if (it <= 1 || c->debug || c->pass_timing || it % c->prm.sample_every == 0) return 0;
  return (c->prm.mls_on && it % c->prm.mls_every == 0) ? 2 : 1;
}

// work enqueued on the context streams completes before what follows on the batch stream
int batch_join(sph2d_batch* b) {
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    HIP_OR_FAIL(hipEventRecord(b->sync[i], b->ctx[i]->stream));
    HIP_OR_FAIL(hipStreamWaitEvent(b->stream, b->sync[i], 0));
  }
  return SPH2D_OK;
}

// work on the batch stream completes before what follows on each context stream
int batch_release(sph2d_batch* b) {
  HIP_OR_FAIL(hipEventRecord(b->fork, b->stream));
  for (sph2d_ctx* c : b->ctx) HIP_OR_FAIL(hipStreamWaitEvent(c->stream, b->fork, 0));
  return SPH2D_OK;
}

// The instances can share one launch per kernel when they run the same kernels with
// the same template arguments: predictor-corrector, one form, one solver, one
// shifting switch, no XSPH, no end damping.
bool fusable(const std::vector<sph2d_ctx*>& v) {
  const sph2d_params& z = v[0]->prm;
  for (const sph2d_ctx* c : v) {
    const sph2d_params& a = c->prm;
    if (a.integrator != SPH2D_INTEGRATOR_PC || a.precision != z.precision || a.solver != z.solver ||
        a.shift_on != z.shift_on || a.epsilon_xsph != 0.0 || a.damping != SPH2D_DAMPING_OFF || a.walls ||
        a.renormalization)
      return false;
  }
  return true;
}

// The instance tables of the fused launches, instances in the given order (or by
// local capacity, largest first, when by_size); xcd: XCD-aware block order of the
// neighbour passes; shape: kNbShape entry of the mixed neighbour passes. The device
// tables are allocated on the first call.
int batch_build_dev(sph2d_batch* b, bool by_size = false, int xcd = 1, int shape = 0) {
  const int n = (int)b->ctx.size();
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  if (by_size)
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int z) { return b->ctx[a]->cap_local > b->ctx[z]->cap_local; });
  std::vector<Inst> inst(n);
  for (int k = 0; k < n; ++k) {
    const int i = k;
    const sph2d_ctx* c = b->ctx[order[k]];
    inst[i].P = c->P;
    inst[i].P.write_debug = 0;
    inst[i].B = c->B;
    inst[i].cap_fluid = c->cap_fluid;
    inst[i].cap_ghost = c->cap_ghost;
    inst[i].cap_global = c->cap_global;
    inst[i].cap_local = c->cap_local;
    inst[i].n_cells = c->n_cells;
  }
  std::vector<int> starts((size_t)BC_COUNT * (n + 1));
  for (int cls = 0; cls < BC_COUNT; ++cls) {
    int acc = 0;
    for (int i = 0; i < n; ++i) {
      starts[(size_t)cls * (n + 1) + i] = acc;
      acc += batch_blocks(cls, inst[i], shape);
    }
    starts[(size_t)cls * (n + 1) + n] = acc;
    b->dev.total[cls] = acc;
  }
  std::vector<unsigned char> q((size_t)n * q32_bytes());
  for (int i = 0; i < n; ++i) q32_fill(inst[i].P, q.data() + (size_t)i * q32_bytes());
  if (b->dev_allocs.empty()) {
    void *di = nullptr, *ds = nullptr, *dq = nullptr;
    HIP_OR_FAIL(hipMalloc(&di, sizeof(Inst) * n));
    b->dev_allocs.push_back(di);
    HIP_OR_FAIL(hipMalloc(&ds, sizeof(int) * starts.size()));
    b->dev_allocs.push_back(ds);
    HIP_OR_FAIL(hipMalloc(&dq, q.size()));
    b->dev_allocs.push_back(dq);
  }
  void *di = b->dev_allocs[0], *ds = b->dev_allocs[1], *dq = b->dev_allocs[2];
  HIP_OR_FAIL(hipDeviceSynchronize());             // no launch may still read the old tables
  HIP_OR_FAIL(hipMemcpy(di, inst.data(), sizeof(Inst) * n, hipMemcpyHostToDevice));
  HIP_OR_FAIL(hipMemcpy(ds, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice));
  HIP_OR_FAIL(hipMemcpy(dq, q.data(), q.size(), hipMemcpyHostToDevice));
  b->dev.inst = static_cast<const Inst*>(di);
  b->dev.n = n;
  for (int cls = 0; cls < BC_COUNT; ++cls) b->dev.start[cls] = static_cast<const int*>(ds) + (size_t)cls * (n + 1);
  b->dev.q32 = dq;
  b->dev.xcd = xcd;
  b->dev.shape = shape;
  b->config = (by_size ? 1 : 0) + (xcd ? 0 : 2) + 4 * shape;
  return SPH2D_OK;
}

void batch_drop_graphs(sph2d_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec_mls) (void)hipGraphExecDestroy(b->exec_mls);
  if (b->graph_mls) (void)hipGraphDestroy(b->graph_mls);
  b->exec = b->exec_mls = nullptr;
  b->graph = b->graph_mls = nullptr;
  b->nodes = 0;
}

// configuration index: bit 0 instance order by size, bit 1 linear block order,
// bits 2.. the tile shape of the mixed neighbour passes
int batch_apply_config(sph2d_batch* b, int cfg) {
  batch_drop_graphs(b);
  return batch_build_dev(b, (cfg & 1) != 0, (cfg & 2) ? 0 : 1, cfg >> 2);
}

int batch_capture(sph2d_batch* b, int it, bool mls) {
  HIP_OR_FAIL(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  int rc = SPH2D_OK;
  hipError_t e = hipSuccess;
  if (b->fused) {
    // the step of every instance, one launch per kernel (it > 1, not a sampling step)
    const sph2d_params& z = b->ctx[0]->prm;
    const int mixed = z.precision == SPH2D_FP32_MIXED;
    launch_batch_step(b->stream, b->dev, mixed, mls ? 1 : 0, z.shift_on, z.solver);
    e = hipGetLastError();
    for (sph2d_ctx* c : b->ctx) c->evaluated = true;
  } else {
    e = hipEventRecord(b->fork, b->stream);
  }
  for (size_t i = 0; !b->fused && i < b->ctx.size() && e == hipSuccess && rc == SPH2D_OK; ++i) {
    sph2d_ctx* c = b->ctx[i];
    e = hipStreamWaitEvent(c->stream, b->fork, 0);
    if (e == hipSuccess) rc = enqueue_step(c, it);
    if (e == hipSuccess && rc == SPH2D_OK) e = hipEventRecord(b->join[i], c->stream);
    if (e == hipSuccess && rc == SPH2D_OK) e = hipStreamWaitEvent(b->stream, b->join[i], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (rc != SPH2D_OK || e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc != SPH2D_OK) return rc;
    return fail(SPH2D_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  if (mls) {
    b->graph_mls = g;
    HIP_OR_FAIL(hipGraphInstantiate(&b->exec_mls, b->graph_mls, nullptr, nullptr, 0));
    return SPH2D_OK;
  }
  b->graph = g;
  HIP_OR_FAIL(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  HIP_OR_FAIL(hipGraphGetNodes(b->graph, nullptr, &b->nodes));
  return SPH2D_OK;
}
}  // namespace

int sph2d_batch_create(int32_t n, sph2d_ctx* const* ctxs, sph2d_batch** out) {
  if (!out || n <= 0 || !ctxs) return fail(SPH2D_ERR_ARGUMENT, "n must be positive, ctxs and out non-NULL");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const sph2d_ctx* c = ctxs[i];
    if (!c) return fail(SPH2D_ERR_ARGUMENT, "a context is NULL");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return fail(SPH2D_ERR_ARGUMENT, "a context appears twice");
    const sph2d_params &a = c->prm, &z = ctxs[0]->prm;
    if (a.integrator != z.integrator || a.mls_on != z.mls_on || a.mls_every != z.mls_every ||
        a.sample_every != z.sample_every)
      return fail(SPH2D_ERR_ARGUMENT, "batched contexts must share integrator, mls_on, mls_every and sample_every");
  }
  auto* b = new sph2d_batch;
  b->ctx.assign(ctxs, ctxs + n);
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  b->join.assign(n, nullptr);
  b->sync.assign(n, nullptr);
  for (int i = 0; i < n && e == hipSuccess; ++i) {
    e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->sync[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    sph2d_batch_destroy(b);
    return fail(SPH2D_ERR_HIP, std::string("batch create: ") + hipGetErrorString(e));
  }
  b->fused = fusable(b->ctx);
  if (b->fused) {
    const int rc = batch_build_dev(b);
    if (rc) {
      sph2d_batch_destroy(b);
      return rc;
    }
  }
  *out = b;
  return SPH2D_OK;
}

int sph2d_batch_run(sph2d_batch* b, int32_t first, int32_t nsteps) {
  if (!b) return fail(SPH2D_ERR_ARGUMENT, "batch is NULL");
  if (first < 1 || nsteps < 0) return fail(SPH2D_ERR_ARGUMENT, "bad step range");
  for (sph2d_ctx* c : b->ctx) {
    const int rc = need_state(c);
    if (rc) return rc;
  }
  for (sph2d_ctx* c : b->ctx) { c->vjp_ready = false; c->eval_part = 0; }
  int rc = SPH2D_OK;
  for (int k = 0; k < nsteps && rc == SPH2D_OK; ++k) {
    const int it = first + k;
    int kind = step_kind(b->ctx[0], it);
    for (const sph2d_ctx* c : b->ctx) kind = std::min(kind, step_kind(c, it));
    if (kind == 2 && !b->fused) kind = 0;
    if (kind != 0) {
      hipGraphExec_t& ex = kind == 2 ? b->exec_mls : b->exec;
      if (!ex) rc = batch_capture(b, it, kind == 2);
      if (rc == SPH2D_OK && !b->on_batch) rc = batch_join(b);
      if (rc == SPH2D_OK) {
        const hipError_t e = hipGraphLaunch(ex, b->stream);
        if (e != hipSuccess) rc = fail(SPH2D_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        b->on_batch = true;
        ++b->replays;
      }
    } else {
      if (b->on_batch) {
        rc = batch_release(b);
        b->on_batch = false;
      }
      for (size_t i = 0; i < b->ctx.size() && rc == SPH2D_OK; ++i) rc = enqueue_step(b->ctx[i], it);
    }
  }
  if (b->on_batch) {
    const int r2 = batch_release(b);
    b->on_batch = false;
    if (rc == SPH2D_OK) rc = r2;
  }
  return rc;
}

int sph2d_batch_info(sph2d_batch* b, int32_t* graph_nodes, int32_t* graph_replays) {
  if (!b) return fail(SPH2D_ERR_ARGUMENT, "batch is NULL");
  // (graph_nodes: one per kernel launch of a step in the fused form, one per launch
  // per instance otherwise)
  if (graph_nodes) *graph_nodes = (int32_t)b->nodes;
  if (graph_replays) *graph_replays = b->replays;
  return SPH2D_OK;
}

int sph2d_batch_autotune(sph2d_batch* b, int32_t first, int32_t trial_steps, double* trial_ms, int32_t* chosen) {
  if (!b) return fail(SPH2D_ERR_ARGUMENT, "batch is NULL");
  if (!b->fused) return fail(SPH2D_ERR_UNSUPPORTED, "the contexts of this batch do not share fused launches");
  if (first < 1 || trial_steps < 1) return fail(SPH2D_ERR_ARGUMENT, "bad step range");
  const bool mixed = b->ctx[0]->prm.precision == SPH2D_FP32_MIXED;
  const int n_cfg = mixed ? SPH2D_BATCH_CONFIGS : 4;   // tile shapes apply to the mixed passes
  if (trial_ms)
    for (int c = 0; c < SPH2D_BATCH_CONFIGS; ++c) trial_ms[c] = 0.0;
  double best = 0.0;
  int best_cfg = 0, it = first;
  for (int cfg = 0; cfg < n_cfg; ++cfg) {
    int rc = batch_apply_config(b, cfg);
    if (!rc) rc = sph2d_batch_run(b, it, 1);        // captures the graphs of this configuration
    ++it;
    if (rc) return rc;
    HIP_OR_FAIL(hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    rc = sph2d_batch_run(b, it, trial_steps);
    if (rc) return rc;
    HIP_OR_FAIL(hipDeviceSynchronize());
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                      trial_steps;
    it += trial_steps;
    if (trial_ms) trial_ms[cfg] = ms;
    if (cfg == 0 || ms < best) { best = ms; best_cfg = cfg; }
  }
  const int rc = batch_apply_config(b, best_cfg);
  if (rc) return rc;
  if (chosen) *chosen = best_cfg;
  return SPH2D_OK;
}

int sph2d_batch_configure(sph2d_batch* b, int32_t config) {
  if (!b) return fail(SPH2D_ERR_ARGUMENT, "batch is NULL");
  if (!b->fused) return fail(SPH2D_ERR_UNSUPPORTED, "the contexts of this batch do not share fused launches");
  const bool mixed = b->ctx[0]->prm.precision == SPH2D_FP32_MIXED;
  if (config < 0 || config >= (mixed ? SPH2D_BATCH_CONFIGS : 4))
    return fail(SPH2D_ERR_ARGUMENT, "config out of range (tile shapes apply to the FP32-mixed form only)");
  return batch_apply_config(b, config);
}

int sph2d_batch_config(sph2d_batch* b, int32_t* config) {
  if (!b || !config) return fail(SPH2D_ERR_ARGUMENT, "batch and config must be non-NULL");
  *config = b->config;
  return SPH2D_OK;
}

int sph2d_batch_destroy(sph2d_batch* b) {
  if (!b) return SPH2D_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec_mls) (void)hipGraphExecDestroy(b->exec_mls);
  if (b->graph_mls) (void)hipGraphDestroy(b->graph_mls);
  for (hipEvent_t e : b->join) if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->sync) if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (void* p : b->dev_allocs) (void)hipFree(p);
  delete b;
  return SPH2D_OK;
}

// ---- the differentiable wave tank: forward with checkpoints, reverse sweep ----------
// A snapshot is everything a predictor-corrector step reads from the steps before it:
// the global state of every particle and the previous rates by local slot (read by the
// predictor's local index), with the device counters.
struct AdjTape {
  sph2d_adjoint_cfg cfg{};
  double a[16] = {};
  int first = 0, nsteps = 0, nsamples = 0, nckpt = 0;
  size_t snap_bytes = 0;
  size_t off[10] = {};
  char* ckpt = nullptr;            // nckpt snapshots
  char* seg = nullptr;             // K snapshots of the segment being reversed
  char* fin = nullptr;             // the state after the forward
  size_t ckpt_cap = 0, seg_cap = 0, fin_cap = 0;
  double* probe_dev = nullptr;     // probe positions (2 per probe), samples, their weights
  size_t probe_cap = 0;
  double* ws = nullptr;            // reverse work space
  size_t ws_cap = 0;
  double* pad = nullptr;           // (dL/dx_p, dL/du_p) per step
  size_t pad_cap = 0;
  int* iws = nullptr;              // flags per gid, arg-max, height bits (K sets with images)
  size_t iws_cap = 0;
  // The segment's steps as whole device images (every allocation of the context after
  // the step), so the reverse restores a step instead of computing it again
  char* img = nullptr;
  size_t img_cap = 0;
  bool ready = false;

  void release() {
    for (void* p : {(void*)ckpt, (void*)seg, (void*)fin, (void*)probe_dev, (void*)ws, (void*)pad, (void*)iws,
                    (void*)img})
      if (p) (void)hipFree(p);
    ckpt = seg = fin = nullptr; probe_dev = ws = pad = nullptr; iws = nullptr; img = nullptr;
    ckpt_cap = seg_cap = fin_cap = probe_cap = ws_cap = pad_cap = iws_cap = img_cap = 0;
  }
};

void tape_free(AdjTape* t) {
  if (!t) return;
  t->release();
  delete t;
}

namespace {

template <class T>
int grow(T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return SPH2D_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess)
    return fail(SPH2D_ERR_CAPACITY, std::string("hipMalloc of the adjoint tape: ") + hipGetErrorString(e));
  *p = static_cast<T*>(q);
  *cap = n;
  return SPH2D_OK;
}

void snap_layout(sph2d_ctx* c, AdjTape* t) {
  const size_t G = (size_t)c->cap_global, L = (size_t)c->cap_local;
  const size_t sz[10] = {16 * G, 16 * G, 8 * G, 8 * G, 8 * G, 8 * G, 4 * G, 8 * L, 16 * L, 4 * (size_t)C_COUNT};
  size_t o = 0;
  for (int k = 0; k < 10; ++k) {
    t->off[k] = o;
    o += (sz[k] + 255) / 256 * 256;
  }
  t->snap_bytes = o;
}

int snap_copy(sph2d_ctx* c, AdjTape* t, char* snap, bool save) {
  const size_t G = (size_t)c->cap_global, L = (size_t)c->cap_local;
  void* dev[10] = {c->B.g_pos, c->B.g_vel, c->B.g_rho, c->B.g_p, c->B.g_c, c->B.g_mass, c->B.g_type,
                   c->B.drhodt, c->B.acc, c->B.counters};
  const size_t sz[10] = {16 * G, 16 * G, 8 * G, 8 * G, 8 * G, 8 * G, 4 * G, 8 * L, 16 * L, 4 * (size_t)C_COUNT};
  AdjCopy C;
  C.n = 10;
  for (int k = 0; k < 10; ++k) {
    char* s = snap + t->off[k];
    C.src[k] = save ? dev[k] : (const void*)s;
    C.dst[k] = save ? (void*)s : dev[k];
    C.words[k] = sz[k] / 4;
  }
  launch_adj_copy(c->stream, C);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

// Allocations a step image leaves out: what the predictor-corrector step of the
// walled tank neither reads nor writes (the RK stage slabs, the MLS density, the
// diagnostic pair census, the analytic and vorticity fields and reductions of
// sampling, the post-processing grid).
bool image_skip(const sph2d_ctx* c, const void* q) {
  const Buffers& B = c->B;
  for (const void* x : {(const void*)B.k_drho, (const void*)B.k_vel, (const void*)B.k_acc, (const void*)B.mls_rho,
                        (const void*)B.pair_count, (const void*)B.pair_off, (const void*)B.ana,
                        (const void*)B.g_vort, (const void*)B.red, (const void*)c->grid_out})
    if (q == x) return true;
  return false;
}

// Bytes of a whole device image of the context: its allocations but the skipped ones,
// each from a 256-byte boundary.
size_t image_bytes(const sph2d_ctx* c) {
  size_t b = 0;
  for (size_t k = 0; k < c->allocs.size(); ++k)
    if (!image_skip(c, c->allocs[k])) b += (c->alloc_bytes[k] + 255) & ~(size_t)255;
  return b;
}

// Saves (save) or restores the allocations of the context's image into / from img.
int image_copy(sph2d_ctx* c, char* img, bool save) {
  AdjCopy C;
  C.n = 0;
  size_t off = 0;
  for (size_t k = 0; k < c->allocs.size(); ++k) {
    if (image_skip(c, c->allocs[k])) continue;
    const size_t n = c->alloc_bytes[k];
    char* q = img + off;
    off += (n + 255) & ~(size_t)255;
    if (n % 4) {                                  // not a whole number of words: a plain copy
      HIP_OR_FAIL(hipMemcpyAsync(save ? (void*)q : c->allocs[k], save ? c->allocs[k] : (const void*)q, n,
                                 hipMemcpyDeviceToDevice, c->stream));
      continue;
    }
    C.src[C.n] = save ? c->allocs[k] : (const void*)q;
    C.dst[C.n] = save ? (void*)q : c->allocs[k];
    C.words[C.n] = n / 4;
    if (++C.n == kAdjCopies) {
      launch_adj_copy(c->stream, C);
      C.n = 0;
    }
  }
  if (C.n) launch_adj_copy(c->stream, C);
  HIP_OR_FAIL(hipGetLastError());
  return SPH2D_OK;
}

void paddle(const AdjTape* t, double dt, int it, double& xp, double& up) {
  xp = 0.0; up = 0.0;
  if (it <= t->cfg.standing_steps) return;
  const double tt = dt * (it - t->cfg.standing_steps);
  for (int k = 1; k <= t->cfg.n_harmonics; ++k) {
    const double om = 2.0 * kPiModule * k * t->cfg.f0;
    xp += t->a[k - 1] * std::sin(om * tt);
    up += t->a[k - 1] * om * std::cos(om * tt);
  }
}

int adj_step(sph2d_ctx* c, AdjTape* t, int it) {
  double xp, up;
  paddle(t, c->prm.dt, it, xp, up);
  int rc = sph2d_wavemaker_update(c, xp, up);
  if (rc) return rc;
  return enqueue_step(c, it);
}

}  // namespace

int sph2d_forward(sph2d_ctx* c, const sph2d_adjoint_cfg* cfg, const double* a, int32_t first, int32_t nsteps,
                  double* probe_p) {
  int rc = need_state(c);
  if (rc) return rc;
  if (!cfg || !a) return fail(SPH2D_ERR_ARGUMENT, "cfg and a are required");
  if (cfg->struct_size != (int32_t)sizeof(sph2d_adjoint_cfg))
    return fail(SPH2D_ERR_ARGUMENT, "sph2d_adjoint_cfg.struct_size mismatch");
  if (cfg->n_harmonics < 1 || cfg->n_harmonics > 16) return fail(SPH2D_ERR_ARGUMENT, "n_harmonics must be in [1, 16]");
  if (cfg->n_probes < 1 || cfg->n_probes > 8) return fail(SPH2D_ERR_ARGUMENT, "n_probes must be in [1, 8]");
  if (cfg->sample_every < 1 || cfg->checkpoint_every < 1 || cfg->standing_steps < 1)
    return fail(SPH2D_ERR_ARGUMENT, "sample_every, checkpoint_every and standing_steps must be >= 1");
  if (first < 2 || nsteps < 1) return fail(SPH2D_ERR_ARGUMENT, "the window starts at step 2 or later, nsteps >= 1");
  if (c->P.mixed) return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint runs in the FP64 form");
  if (c->P.renorm) return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint is built for the identity renormalisation");
  if (c->prm.integrator != SPH2D_INTEGRATOR_PC) return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint needs the predictor-corrector");
  if (!c->prm.walls || c->prm.periodic || !c->prm.wavemaker_layers)
    return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint needs a walled tank with wavemaker particles");
  if (c->prm.mls_on) return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint needs the MLS density reinitialisation off");
  if (c->prm.epsilon_xsph != 0.0) return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint needs epsilon_xsph = 0");
  if (!c->P.sps) return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint is built for the _SPS solvers");
  if (c->P.solver != SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && c->P.solver != SPH2D_SOLVER_LINEAR_RIEMANN)
    return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint is built for the two linear Riemann solvers");
  if (c->prm.rate_reuse != SPH2D_RATE_REUSE_REFERENCE)
    return fail(SPH2D_ERR_UNSUPPORTED, "the adjoint follows the reference's rate re-use by local slot");
  const double lo_x = c->P.chain_ox, lo_y = c->P.chain_oy;
  const double hi_x = lo_x + c->P.chain_nx * c->P.chain_d, hi_y = lo_y + c->P.chain_ny * c->P.chain_d;
  for (int k = 0; k < cfg->n_probes; ++k)
    if (!(cfg->probe_x[k] >= lo_x && cfg->probe_x[k] < hi_x && cfg->probe_y[k] >= lo_y && cfg->probe_y[k] < hi_y))
      return fail(SPH2D_ERR_ARGUMENT, "probe " + std::to_string(k + 1) + " lies outside the background grid");
  if (!c->tape) c->tape = new AdjTape();
  AdjTape* t = c->tape;
  t->ready = false;
  t->cfg = *cfg;
  for (int k = 0; k < 16; ++k) t->a[k] = k < cfg->n_harmonics ? a[k] : 0.0;
  t->first = first;
  t->nsteps = nsteps;
  t->nsamples = nsteps / cfg->sample_every;
  const int K = cfg->checkpoint_every;
  t->nckpt = (nsteps + K - 1) / K;
  snap_layout(c, t);
  const size_t np = (size_t)cfg->n_probes;
  if ((rc = grow(&t->ckpt, &t->ckpt_cap, t->snap_bytes * t->nckpt))) return rc;
  if ((rc = grow(&t->fin, &t->fin_cap, t->snap_bytes))) return rc;
  if ((rc = grow(&t->probe_dev, &t->probe_cap, 2 * np + 2 * np * t->nsamples + 8))) return rc;
  std::vector<double> pos(2 * np);
  for (size_t k = 0; k < np; ++k) { pos[2 * k] = cfg->probe_x[k]; pos[2 * k + 1] = cfg->probe_y[k]; }
  HIP_OR_FAIL(hipMemcpyAsync(t->probe_dev, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  double* samples = t->probe_dev + 2 * np;
  ProbeArgs Q;
  Q.n = cfg->n_probes;
  Q.gravity = c->P.g;
  Q.pos = reinterpret_cast<const double2*>(t->probe_dev);
  for (int k = 0; k < nsteps; ++k) {
    const int it = first + k;
    if (k % K == 0 && (rc = snap_copy(c, t, t->ckpt + (size_t)(k / K) * t->snap_bytes, true))) return rc;
    if ((rc = adj_step(c, t, it))) return rc;
    if ((k + 1) % cfg->sample_every == 0) {
      // k_probe_mls writes 4 blocks of n; the MLS pressure block lands on this sample
      rc = probe_space(c, 4 * np);
      if (rc) return rc;
      Q.out = c->probe_buf;
      launch_probe_mls(c->stream, c->P, c->B, Q);
      const int smp = (k + 1) / cfg->sample_every - 1;
      HIP_OR_FAIL(hipMemcpyAsync(samples + (size_t)smp * np, c->probe_buf + np, np * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
    }
  }
  if ((rc = snap_copy(c, t, t->fin, true))) return rc;
  HIP_OR_FAIL(hipGetLastError());
  if (probe_p && t->nsamples > 0)
    HIP_OR_FAIL(hipMemcpyAsync(probe_p, samples, np * t->nsamples * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  int cnt[C_COUNT];
  if ((rc = read_counters(c, cnt))) return rc;
  t->ready = true;
  return SPH2D_OK;
}

int sph2d_backward(sph2d_ctx* c, const double* dl_dprobe, double* dl_da) {
  int rc = need_state(c);
  if (rc) return rc;
  AdjTape* t = c->tape;
  if (!t || !t->ready) return fail(SPH2D_ERR_STATE, "sph2d_backward needs a completed sph2d_forward");
  if (!dl_dprobe || !dl_da) return fail(SPH2D_ERR_ARGUMENT, "dl_dprobe and dl_da are required");
  const int K = t->cfg.checkpoint_every;
  const size_t np = (size_t)t->cfg.n_probes;
  const size_t G = (size_t)c->cap_global, L = (size_t)c->cap_local;
  // Each step of a segment runs once more: its whole device image is kept (K images)
  // for the reverse, with the wavemaker flags of its start state. Where K images do
  // not fit, the segment keeps step-start snapshots and the reverse runs each step
  // again from its snapshot.
  const size_t ib = image_bytes(c);
  // flags per gid, the arg-max, then the height bits from an 8-byte boundary
  const size_t hoff = (G + 2 + 1) & ~(size_t)1, nflag = hoff + 4;
  bool images = t->img_cap >= ib * (size_t)K && t->img;
  if (!images) {
    if (t->img) (void)hipFree(t->img);
    t->img = nullptr;
    t->img_cap = 0;
    void* q = nullptr;
    if (hipMalloc(&q, ib * (size_t)K) == hipSuccess) {
      t->img = static_cast<char*>(q);
      t->img_cap = ib * (size_t)K;
      images = true;
    } else {
      (void)hipGetLastError();                    // the failed allocation leaves no error behind
    }
  }
  if (!images && (rc = grow(&t->seg, &t->seg_cap, t->snap_bytes * (size_t)K))) return rc;
  const size_t nws = 6 * G + 3 * L + 3 * L + 11 * L + 10 * L + 5 * L + 5 * L + 5 * L + L + 3 * G + 3 * 256;
  if ((rc = grow(&t->ws, &t->ws_cap, nws))) return rc;
  if ((rc = grow(&t->pad, &t->pad_cap, 2 * (size_t)t->nsteps))) return rc;
  if ((rc = grow(&t->iws, &t->iws_cap, nflag * (images ? (size_t)K : 1)))) return rc;
  double* lamG = t->ws;
  double* lamR = lamG + 6 * G;
  double* lamRp = lamR + 3 * L;
  double* scr = lamRp + 3 * L;
  double* W = scr + 11 * L;
  double* gs = W + 10 * L;
  double* ga = gs + 5 * L;
  double* Gs = ga + 5 * L;
  double* wsum = Gs + 5 * L;
  double* c3 = wsum + L;
  double* part = c3 + 3 * G;
  int* flags = t->iws;
  int* argmax = t->iws + G;
  auto* hbits = reinterpret_cast<unsigned long long*>(t->iws + hoff);
  double* lam_probe = t->probe_dev + 2 * np + np * t->nsamples;
  if (t->nsamples > 0)
    HIP_OR_FAIL(hipMemcpyAsync(lam_probe, dl_dprobe, np * t->nsamples * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
  launch_adj_zero(c->stream, lamG, 6 * G);
  launch_adj_zero(c->stream, lamR, 3 * L);
  std::vector<double> pos(2 * np);
  HIP_OR_FAIL(hipMemcpyAsync(pos.data(), t->probe_dev, pos.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  for (int sg = t->nckpt - 1; sg >= 0; --sg) {
    const int c0 = sg * K, len = std::min(K, t->nsteps - c0);
    if ((rc = snap_copy(c, t, t->ckpt + (size_t)sg * t->snap_bytes, false))) return rc;
    if (images) {
      // the segment forward again: the wavemaker's branches from each step's start
      // state, the step, its image
      for (int j = 0; j < len; ++j) {
        const int it = t->first + c0 + j;
        int* fl = t->iws + (size_t)j * nflag;
        double xp, up;
        paddle(t, c->prm.dt, it, xp, up);
        auto* hb = reinterpret_cast<unsigned long long*>(fl + hoff);
        launch_adj_wm_init(c->stream, hb, fl + G);
        launch_wavemaker_height(c->stream, c->P, c->B, c->cap_fluid, xp, hb);
        launch_adj_wm_flags(c->stream, c->P, c->B, c->cap_fluid, xp, fl, hb, fl + G);
        if ((rc = adj_step(c, t, it))) return rc;
        if ((rc = image_copy(c, t->img + (size_t)j * ib, true))) return rc;
      }
    } else {
      for (int j = 0; j < len; ++j) {            // the segment forward again, one snapshot per step
        if ((rc = snap_copy(c, t, t->seg + (size_t)j * t->snap_bytes, true))) return rc;
        if (j + 1 < len && (rc = adj_step(c, t, t->first + c0 + j))) return rc;
      }
    }
    for (int j = len - 1; j >= 0; --j) {
      const int k = c0 + j, it = t->first + k;
      if (images) {
        flags = t->iws + (size_t)j * nflag;
        argmax = flags + G;
        hbits = reinterpret_cast<unsigned long long*>(flags + hoff);
        if ((rc = image_copy(c, t->img + (size_t)j * ib, false))) return rc;
      } else {
        if ((rc = snap_copy(c, t, t->seg + (size_t)j * t->snap_bytes, false))) return rc;
        double xp, up;
        paddle(t, c->prm.dt, it, xp, up);
        // the wavemaker's branches from the state before its update, then the step again
        launch_adj_wm_init(c->stream, hbits, argmax);
        launch_wavemaker_height(c->stream, c->P, c->B, c->cap_fluid, xp, hbits);
        launch_adj_wm_flags(c->stream, c->P, c->B, c->cap_fluid, xp, flags, hbits, argmax);
        if ((rc = adj_step(c, t, it))) return rc;
      }
      if ((k + 1) % t->cfg.sample_every == 0) {
        const int smp = (k + 1) / t->cfg.sample_every - 1;
        for (size_t q = 0; q < np; ++q)
          launch_adj_probe(c->stream, c->P, c->B, pos[2 * q], pos[2 * q + 1], c->P.g, lam_probe + smp * np + q, lamG);
      }
      launch_adj_stage1(c->stream, c->P, c->B, c->cap_local, lamG, lamR, W, scr, c->prm.shift_on);
      launch_eval_vjp_slots(c->stream, c->P, c->B, c->cap_local, c->prm.shift_on, W, gs, ga, Gs, wsum);
      launch_adj_stage2(c->stream, c->P, c->B, c->cap_local, Gs, scr, lamR, lamG, lamRp);
      launch_adj_wm_reverse(c->stream, c->P, c->B, c->n_fluid + c->n_wm, c->n_wm, Gs, flags, hbits, lamG, c3, part,
                            argmax, t->pad + 2 * (size_t)k);
      std::swap(lamR, lamRp);
      HIP_OR_FAIL(hipGetLastError());
    }
  }
  std::vector<double> pad(2 * (size_t)t->nsteps);
  HIP_OR_FAIL(hipMemcpyAsync(pad.data(), t->pad, pad.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if ((rc = snap_copy(c, t, t->fin, false))) return rc;
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  for (int q = 0; q < t->cfg.n_harmonics; ++q) {
    const double om = 2.0 * kPiModule * (q + 1) * t->cfg.f0;
    double g = 0.0;
    for (int k = 0; k < t->nsteps; ++k) {
      const int it = t->first + k;
      if (it <= t->cfg.standing_steps) continue;
      const double tt = c->prm.dt * (it - t->cfg.standing_steps);
      g += pad[2 * k] * std::sin(om * tt) + pad[2 * k + 1] * om * std::cos(om * tt);
    }
    dl_da[q] = g;
  }
  c->vjp_ready = false;
  c->eval_part = 0;
  return SPH2D_OK;
}

// ---- device memory through this library's HIP runtime ------------------------------
int sph2d_device_alloc(size_t bytes, void** out) {
  if (!out) return fail(SPH2D_ERR_ARGUMENT, "out is NULL");
  *out = nullptr;
  const hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) return fail(SPH2D_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  return SPH2D_OK;
}

int sph2d_device_free(void* p) {
  const hipError_t e = hipFree(p);
  if (e != hipSuccess) return fail(SPH2D_ERR_HIP, std::string("hipFree: ") + hipGetErrorString(e));
  return SPH2D_OK;
}

int sph2d_device_copy(void* dst, const void* src, size_t bytes, int32_t to_device) {
  if (bytes && (!dst || !src)) return fail(SPH2D_ERR_ARGUMENT, "NULL pointer");
  const hipError_t e = hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(SPH2D_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return SPH2D_OK;
}
