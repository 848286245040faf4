"""Python host mirror of the reference's hot-path interface, over the C-ABI.

The reference (mini-walker/SPH-2D, Taylor_Green_Vortex/) drives its WCSPH step
from Fortran: time_iteration -> Distribute_Period_Boundary_Particle ->
Devide_block_and_send_tasks -> single_step_compute -> stage updates ->
Reinitialize_particle_density -> Shift_particle_position -> Refresh_domain.
`Simulation` exposes the same routines, with the same names and step
numbering, on a context whose arrays live in the GPU's HBM
(libsph2d.so, include/sph2d.h). There is no CPU path: if the HIP library is
missing or no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsph2d.so")

SOLVER_MODIFIED_LINEAR_RIEMANN = 0
SOLVER_LINEAR_RIEMANN = 1
SOLVER_HLLC = 2
SOLVER_MUSCL = 3
SOLVER_MS_MUSCL_HLLC = 4
SOLVER_MS_HLLC = 5
INTEGRATOR_PC = 2
INTEGRATOR_RK = 3
FP64 = 0
FP32_MIXED = 1
DAMPING_OFF = 0
DAMPING_COSINE = 1
DAMPING_OPENFOAM = 2
DAMPING_DUALSPH = 3
DAMPING_LINEAR = 4

ERRORS = {1: "argument", 2: "capacity", 3: "hip", 4: "state", 5: "unsupported", 6: "nonfinite"}
ERR_ARGUMENT, ERR_CAPACITY, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_NONFINITE = 1, 2, 3, 4, 5, 6


class SphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sph2d error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    """sph2d_params (include/sph2d.h); defaults = Taylor_Green_Vortex/information_module.f90."""
    _fields_ = [
        ("struct_size", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("box_x", ctypes.c_double), ("box_y", ctypes.c_double),
        ("rho0", ctypes.c_double), ("c0sq", ctypes.c_double),
        ("nu", ctypes.c_double), ("U", ctypes.c_double),
        ("dx", ctypes.c_double), ("h", ctypes.c_double), ("dt", ctypes.c_double),
        ("integrator", ctypes.c_int32), ("solver", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("rate_reuse", ctypes.c_int32),
        ("shift_on", ctypes.c_int32), ("mls_on", ctypes.c_int32),
        ("mls_every", ctypes.c_int32), ("mls_order", ctypes.c_int32),
        ("sample_every", ctypes.c_int32), ("periodic", ctypes.c_int32),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("fai", ctypes.c_double),
        ("delta", ctypes.c_double), ("delta_alpha", ctypes.c_double),
        ("epsilon_xsph", ctypes.c_double),
        ("shift_beta", ctypes.c_double), ("shift_R", ctypes.c_double),
        ("max_shift", ctypes.c_double), ("mls_alpha", ctypes.c_double),
        ("damping", ctypes.c_int32), ("damping_length", ctypes.c_double),
        ("damping_amplitude", ctypes.c_double), ("water_depth", ctypes.c_double),
        ("walls", ctypes.c_int32), ("gravity", ctypes.c_double),
        ("wavemaker_layers", ctypes.c_int32),
        ("sps_on", ctypes.c_int32), ("renormalization", ctypes.c_int32),
        ("ms", ctypes.c_int32), ("background_p", ctypes.c_double), ("release_time", ctypes.c_double),
        ("body_cx", ctypes.c_double), ("body_cy", ctypes.c_double), ("body_len", ctypes.c_double),
        ("body_speed", ctypes.c_double), ("body_ramp", ctypes.c_double),
    ]


WAVE_SOLITARY = 0
WAVE_STOKES2 = 1
WAVE_AIRY = 2


class Wavemaker(ctypes.Structure):
    """sph2d_wavemaker: the piston wavemaker's inputs (Wavemaker_Subroutine.f90:66-138)."""
    _fields_ = [("wave_type", ctypes.c_int32), ("soft_steps", ctypes.c_int32), ("standing_steps", ctypes.c_int32),
                ("dt", ctypes.c_double), ("g", ctypes.c_double), ("water_depth", ctypes.c_double),
                ("wave_height", ctypes.c_double), ("wave_number", ctypes.c_double),
                ("wave_frequency", ctypes.c_double), ("phase", ctypes.c_double)]


def wavemaker_motion(w: Wavemaker, i_time_step: int):
    """(position, velocity, acceleration) of the paddle at step i_time_step."""
    x, u, a = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check(lib().sph2d_wavemaker_motion(ctypes.byref(w), i_time_step, ctypes.byref(x), ctypes.byref(u),
                                        ctypes.byref(a)))
    return x.value, u.value, a.value


class AdjointCfg(ctypes.Structure):
    """sph2d_adjoint_cfg: the differentiable wave tank's paddle law, probes and checkpoints."""
    _fields_ = [("struct_size", ctypes.c_int32), ("n_harmonics", ctypes.c_int32), ("f0", ctypes.c_double),
                ("standing_steps", ctypes.c_int32), ("sample_every", ctypes.c_int32), ("n_probes", ctypes.c_int32),
                ("probe_x", ctypes.c_double * 8), ("probe_y", ctypes.c_double * 8),
                ("checkpoint_every", ctypes.c_int32)]


def adjoint_cfg(n_harmonics=8, f0=0.5, standing_steps=1, sample_every=10, probes=((2.0, 0.3),),
                checkpoint_every=20) -> AdjointCfg:
    c = AdjointCfg()
    c.struct_size = ctypes.sizeof(AdjointCfg)
    c.n_harmonics, c.f0, c.standing_steps = n_harmonics, f0, standing_steps
    c.sample_every, c.checkpoint_every = sample_every, checkpoint_every
    if not 1 <= len(probes) <= 8:
        raise ValueError("1 to 8 probes")
    c.n_probes = len(probes)
    for k, (x, y) in enumerate(probes):
        c.probe_x[k], c.probe_y[k] = x, y
    return c


class Scales(ctypes.Structure):
    """sph2d_scales: the output routines' normalisation (information_module.f90:56,150-152)."""
    _fields_ = [("domain_size_y", ctypes.c_double), ("velocity_0", ctypes.c_double),
                ("pressure_0", ctypes.c_double), ("time_0", ctypes.c_double)]


class Diag(ctypes.Structure):
    _fields_ = [
        ("n_fluid", ctypes.c_int32), ("n_ghost", ctypes.c_int32), ("n_local", ctypes.c_int32),
        ("n_ill", ctypes.c_int32), ("n_surface", ctypes.c_int32), ("n_singular", ctypes.c_int32),
        ("n_mls_fallback", ctypes.c_int32), ("max_cell_count", ctypes.c_int32),
        ("n_pairs", ctypes.c_int64), ("rebin_needed", ctypes.c_int32), ("error_flags", ctypes.c_int32),
    ]


# every symbol include/sph2d.h declares (checked by tests on CPU)
EXPORTS = (
    "sph2d_last_error", "sph2d_default_params", "sph2d_ctx_create", "sph2d_ctx_destroy",
    "sph2d_init_tgv", "sph2d_upload_state", "sph2d_download_state", "sph2d_compute_kernel",
    "sph2d_compute_kernel_n", "sph2d_status_equation", "sph2d_status_equation_n",
    "sph2d_distribute_period_boundary", "sph2d_devide_block", "sph2d_single_step_compute",
    "sph2d_single_step_compute_vjp", "sph2d_single_step_compute_vjp_device", "sph2d_find_particle_pair",
    "sph2d_download_pair_list",
    "sph2d_define_freesurface", "sph2d_solver", "sph2d_stage_update", "sph2d_download_rates",
    "sph2d_reinitialize_particle_density", "sph2d_shift_particle_position",
    "sph2d_refresh_domain", "sph2d_time_step", "sph2d_run", "sph2d_sample",
    "sph2d_download_local", "sph2d_download_pairs", "sph2d_set_debug",
    "sph2d_get_diagnostics", "sph2d_get_stream", "sph2d_time_pass",
    "sph2d_reorder_particle_index", "sph2d_reorder_particle_index_noalloc", "sph2d_set_pass_timing",
    "sph2d_get_pass_timing",
    "sph2d_get_scales", "sph2d_download_output_fields", "sph2d_output_for_tecplot",
    "sph2d_output_for_paraview", "sph2d_format_fortran", "sph2d_write_final_results",
    "sph2d_write_tecplot_zone", "sph2d_write_surface_zone", "sph2d_write_tecplot_binary",
    "sph2d_write_paraview", "sph2d_write_latest_results_zone", "sph2d_read_latest_results",
    "sph2d_write_sampling", "sph2d_grid_size", "sph2d_calculation_for_postproceeding",
    "sph2d_output_for_grid_postproceding", "sph2d_write_initial_grid", "sph2d_write_grid_zone",
    "sph2d_write_grid_binary", "sph2d_sampling", "sph2d_wave_probes", "sph2d_output_sampling_lines",
    "sph2d_format_list_directed", "sph2d_write_list_directed", "sph2d_write_f20_8", "sph2d_kalman_filter",
    "sph2d_sampling_line_points", "sph2d_write_sampling_line_positions", "sph2d_write_sampling_line_results",
    "sph2d_wavemaker_motion", "sph2d_wavemaker_pull_back", "sph2d_wavemaker_update", "sph2d_batch_create", "sph2d_batch_run",
    "sph2d_batch_info", "sph2d_batch_destroy", "sph2d_batch_autotune", "sph2d_batch_configure", "sph2d_batch_config",
    "sph2d_forward", "sph2d_backward",
    "sph2d_relaxation_init", "sph2d_wave_theory_update", "sph2d_wave_maker_damping_zone",
    "sph2d_device_alloc", "sph2d_device_free", "sph2d_device_copy",
    "sph2d_moving_square_layout", "sph2d_init_moving_square", "sph2d_body_motion", "sph2d_body_force",
)

_lib = None


def lib():
    """Load libsph2d.so (never a fallback: raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp, i32, i64, dp, ip = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, P(ctypes.c_double), P(ctypes.c_int32)
        cs = ctypes.c_char_p
        L.sph2d_last_error.restype = ctypes.c_char_p
        sig = {
            "sph2d_default_params": [P(Params)],
            "sph2d_ctx_create": [P(Params), i32, P(vp)],
            "sph2d_ctx_destroy": [vp],
            "sph2d_init_tgv": [vp],
            "sph2d_upload_state": [vp, i32, dp, dp, dp, dp, dp, dp, ip, i32],
            "sph2d_download_state": [vp, i32, dp, dp, dp, dp, dp, dp, ip, i32],
            "sph2d_compute_kernel": [ip, dp, dp, dp, dp, dp],
            "sph2d_compute_kernel_n": [i32, dp, dp, dp, dp, dp],
            "sph2d_status_equation": [dp, dp, ip, dp, dp],
            "sph2d_status_equation_n": [P(Params), i32, dp, dp, ip, dp, dp],
            "sph2d_distribute_period_boundary": [vp],
            "sph2d_devide_block": [vp, i32],
            "sph2d_single_step_compute": [vp, i32],
            "sph2d_single_step_compute_vjp": [vp, i32, dp, dp, dp, i32, dp, dp, dp],
            "sph2d_single_step_compute_vjp_device": [vp, i32, vp, vp, vp, i32, vp, vp, vp],
            "sph2d_reinitialize_particle_density": [vp, i32],
            "sph2d_shift_particle_position": [vp, i32],
            "sph2d_refresh_domain": [vp, i32],
            "sph2d_time_step": [vp, i32],
            "sph2d_run": [vp, i32, i32],
            "sph2d_sample": [vp, dp],
            "sph2d_download_local": [vp, ctypes.c_char_p, vp, i32],
            "sph2d_download_pairs": [vp, i64, ip, ip, P(i64)],
            "sph2d_set_debug": [vp, i32],
            "sph2d_get_diagnostics": [vp, P(Diag)],
            "sph2d_get_stream": [vp, P(vp)],
            "sph2d_time_pass": [vp, i32, i32, dp],
            "sph2d_set_pass_timing": [vp, i32],
            "sph2d_get_pass_timing": [vp, dp, ip],
            "sph2d_reorder_particle_index": [i32, i32, ctypes.c_double, ctypes.c_double, vp, vp, vp, i32,
                                             P(vp), i32, P(vp), vp, ip, vp],
            "sph2d_get_scales": [vp, P(Scales)],
            "sph2d_download_output_fields": [vp, i32, dp, dp, i32, ip],
            "sph2d_output_for_tecplot": [vp, i32, cs, cs, cs],
            "sph2d_output_for_paraview": [vp, i32, cs],
            "sph2d_format_fortran": [ctypes.c_double, ctypes.c_char, i32, i32, ctypes.c_char_p, i32],
            "sph2d_write_final_results": [cs, i32, dp, dp, dp, dp, i32],
            "sph2d_write_tecplot_zone": [cs, ctypes.c_double, i32, dp, dp, dp, dp, dp, i32, P(Scales)],
            "sph2d_write_surface_zone": [cs, ctypes.c_double, i32, dp, dp, i32, P(Scales)],
            "sph2d_write_tecplot_binary": [cs, ctypes.c_double, i32, dp, dp, dp, i32, P(Scales)],
            "sph2d_write_paraview": [cs, i32, dp, dp, dp, i32, P(Scales)],
            "sph2d_write_latest_results_zone": [cs, i32, i32, dp, dp, dp, dp, dp, dp, dp, ip, ip, i32],
            "sph2d_read_latest_results": [cs, i32, i32, ip, dp, dp, dp, dp, dp, dp, dp, ip, ip, ip, i32],
            "sph2d_write_sampling": [cs, ctypes.c_double, dp],
            "sph2d_grid_size": [vp, ip, ip, dp],
            "sph2d_calculation_for_postproceeding": [vp, i32, i32, dp, dp, dp, dp, dp, i32],
            "sph2d_output_for_grid_postproceding": [vp, i32, cs, cs],
            "sph2d_write_initial_grid": [cs, cs, i32, i32, ctypes.c_double],
            "sph2d_write_grid_zone": [cs, ctypes.c_double, i32, i32, ctypes.c_double, dp, dp, dp, dp, dp, i32,
                                      P(Scales)],
            "sph2d_write_grid_binary": [cs, ctypes.c_double, i32, i32, ctypes.c_double, dp, dp, dp, i32,
                                        P(Scales)],
            "sph2d_sampling": [vp, i32, i32, dp, i32, ctypes.c_double, dp, dp, dp],
            "sph2d_wave_probes": [vp, i32, dp, dp, dp],
            "sph2d_output_sampling_lines": [vp, i32, cs, i32, dp, dp, i32],
            "sph2d_format_list_directed": [ctypes.c_double, ctypes.c_char_p, i32],
            "sph2d_write_list_directed": [cs, i32, dp],
            "sph2d_write_f20_8": [cs, i32, dp],
            "sph2d_kalman_filter": [i32, i32, dp, dp, dp, dp],
            "sph2d_sampling_line_points": [i32, dp, dp, dp, i32],
            "sph2d_write_sampling_line_positions": [cs, i32, i32, dp, i32],
            "sph2d_write_sampling_line_results": [cs, i32, ctypes.c_double, i32, dp, dp, dp, dp, dp, dp, i32,
                                                  P(Scales)],
            "sph2d_wavemaker_motion": [P(Wavemaker), i32, dp, dp, dp],
            "sph2d_wavemaker_pull_back": [vp, ctypes.c_double, ctypes.c_double],
            "sph2d_wavemaker_update": [vp, ctypes.c_double, ctypes.c_double],
            "sph2d_batch_create": [i32, P(vp), P(vp)],
            "sph2d_batch_run": [vp, i32, i32],
            "sph2d_batch_info": [vp, ip, ip],
            "sph2d_batch_destroy": [vp],
            "sph2d_batch_autotune": [vp, i32, i32, dp, ip],
            "sph2d_batch_configure": [vp, i32],
            "sph2d_batch_config": [vp, ip],
            "sph2d_reorder_particle_index_noalloc": [i32, i32, ctypes.c_double, ctypes.c_double, vp, vp, vp, i32,
                                                     P(vp), i32, P(vp), vp, ip, vp],
            "sph2d_find_particle_pair": [vp, i32],
            "sph2d_download_pair_list": [vp, i64, ip, ip, P(i64)],
            "sph2d_define_freesurface": [vp],
            "sph2d_solver": [vp, i32, i32],
            "sph2d_stage_update": [vp, i32, i32],
            "sph2d_download_rates": [vp, i32, dp, dp, i32],
            "sph2d_forward": [vp, P(AdjointCfg), dp, i32, i32, dp],
            "sph2d_backward": [vp, dp, dp],
            "sph2d_relaxation_init": [vp, ctypes.c_double],
            "sph2d_wave_theory_update": [vp, P(Wavemaker), i32],
            "sph2d_wave_maker_damping_zone": [vp, i32],
            "sph2d_device_alloc": [ctypes.c_size_t, P(vp)],
            "sph2d_device_free": [vp],
            "sph2d_device_copy": [vp, vp, ctypes.c_size_t, i32],
            "sph2d_moving_square_layout": [P(Params), i32, dp, dp, ip, ip, ip, ip],
            "sph2d_init_moving_square": [vp],
            "sph2d_body_motion": [vp, ctypes.c_double],
            "sph2d_body_force": [vp, dp, ip],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise SphError(rc, lib().sph2d_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def default_params(**kw) -> Params:
    p = Params()
    _check(lib().sph2d_default_params(ctypes.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def tgv_params(n_side: int, **kw) -> Params:
    """Taylor-Green vortex on an n_side^2 lattice of the unit box; dt keeps the
    reference's c0*dt/h ratio (dt = 1e-4 * 100/n_side) unless given."""
    dx = 1.0 / n_side
    p = default_params(nx=n_side, ny=n_side, dx=dx, h=1.20 * dx, max_shift=0.25 * dx,
                       dt=1e-4 * (100.0 / n_side))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def moving_square_params(n_per_d: int = 100, re: float = 100.0, tank=(10.0, 5.0), layers: int = 3, **kw) -> Params:
    """The moving square (the reference's Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box
    program, ms = 1): a tank of tank[0] x tank[1] body lengths D = 1 with `layers` layers of
    fixed ghosts, the square of side D centred at (2.5, tank[1]/2), dx = D / n_per_d,
    h = 1.15 dx (Information_Module.f90:616), c0 = 10 U (c0sq = 100), rho0 = 1, nu = U D /
    Re, the body's speed ramped linearly to U = 1 over t = 0.5; MUSCL-HLLC + delta-SPH +
    SPS with XSPH epsilon 0.1, shifting (max 0.15 dx), MLS order 3 every 40 steps, the
    predictor-corrector at dt = 0.2 * 0.25 h / (1.1 c0) (the CFL form of
    Initialize_SPH_Calculation.f90), body force every 100 steps."""
    dx = 1.0 / n_per_d
    h = 1.15 * dx
    p = default_params(nx=int(round(tank[0] * n_per_d)), ny=int(round(tank[1] * n_per_d)), box_x=tank[0],
                       box_y=tank[1], dx=dx, h=h, rho0=1.0, c0sq=100.0, nu=1.0 / re, U=1.0,
                       dt=0.2 * 0.25 * h / (1.1 * 10.0), integrator=INTEGRATOR_PC, solver=SOLVER_MS_MUSCL_HLLC,
                       sps_on=1, shift_on=1, max_shift=0.15 * dx, mls_on=1, mls_every=40, mls_order=3,
                       mls_alpha=1e-5, alpha=0.03, beta=2.0, fai=0.1, delta=0.1, delta_alpha=0.02,
                       epsilon_xsph=0.1, sample_every=100, periodic=0, walls=layers, ms=1, background_p=0.0,
                       release_time=10.0, body_cx=2.5, body_cy=tank[1] / 2, body_len=1.0, body_speed=1.0,
                       body_ramp=0.5, water_depth=tank[1])
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def moving_square_layout(params: Params):
    """(x, y, btype, init, n_fluid) of the moving-square tank (sph2d_moving_square_layout)."""
    nf, nt = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().sph2d_moving_square_layout(ctypes.byref(params), 0, None, None, None, None, ctypes.byref(nf),
                                            ctypes.byref(nt)))
    n = nt.value
    x, y = np.zeros(n), np.zeros(n)
    bt, it = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    _check(lib().sph2d_moving_square_layout(ctypes.byref(params), n, _dp(x), _dp(y), _ip(bt), _ip(it),
                                            ctypes.byref(nf), ctypes.byref(nt)))
    return x, y, bt, it, nf.value


def uneven_tgv_instances(count: int = 64, seed: int = 3):
    """Lattice sides and Reynolds numbers of a set of independent periodic Taylor-Green
    instances of uneven size: n_side = round(64 * 8^U) (log-uniform over 64..512) and
    Re = 50 * 20^U' (log-uniform over 50..1000), U, U' uniform from numpy's default
    generator with the given seed (sides drawn first)."""
    rng = np.random.default_rng(seed)
    sides = [int(round(64 * 8 ** u)) for u in rng.uniform(0, 1, count)]
    res = [50 * 20 ** u for u in rng.uniform(0, 1, count)]
    return sides, res


LOCAL_F64 = ("x", "y", "u", "v", "rho", "mass", "p", "c", "volume", "fra_v", "sum_dwdx_x",
             "sum_dwdx_y", "tau_xx", "tau_xy", "tau_yy", "grad_rho_x", "grad_rho_y", "drhodt",
             "ax", "ay", "shift_x", "shift_y", "avg_x", "avg_y", "normal_x", "normal_y", "internal_x", "internal_y",
             "visc_x", "visc_y", "art_internal_x", "art_internal_y", "sps_x", "sps_y", "lap_x",
             "lap_y", "grad_p_x", "grad_p_y", "grad_u_x", "grad_u_y", "grad_v_x", "grad_v_y")
LOCAL_I32 = ("type", "name", "surface", "near_surface", "neighbour_count")


class Simulation:
    """One HBM-resident simulation context (sph2d_ctx)."""

    def __init__(self, params: Params, capacity: Optional[int] = None):
        self.params = params
        cap = capacity if capacity is not None else params.nx * params.ny
        h = ctypes.c_void_p()
        _check(lib().sph2d_ctx_create(ctypes.byref(params), int(cap), ctypes.byref(h)))
        self._c = h
        self.n_fluid = 0

    def close(self):
        if getattr(self, "_c", None):
            lib().sph2d_ctx_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- state ------------------------------------------------------------------
    def init_tgv(self):
        _check(lib().sph2d_init_tgv(self._c))
        self.n_fluid = self.params.nx * self.params.ny

    # --- the moving-square program (ms = 1) --------------------------------------------
    def init_moving_square(self):
        _check(lib().sph2d_init_moving_square(self._c))
        self.n_fluid = moving_square_layout(self.params)[4]

    def body_motion(self, t: float):
        """Body_Motion_Solver at physical time t (time_step calls it itself)."""
        _check(lib().sph2d_body_motion(self._c, float(t)))

    def body_force(self):
        """(pressure x, y, viscous x, y, total x, y) on the water from the last sampling step's
        evaluation, and that step (0: none yet)."""
        out = np.zeros(6)
        st = ctypes.c_int32()
        _check(lib().sph2d_body_force(self._c, _dp(out), ctypes.byref(st)))
        return out, st.value

    def upload_state(self, x, y, u, v, rho, mass, energy=None, type_=None):
        n = len(x)
        pos = np.ascontiguousarray(np.concatenate([x, y]), dtype=np.float64)
        vel = np.ascontiguousarray(np.concatenate([u, v]), dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        e = None if energy is None else np.ascontiguousarray(energy, dtype=np.float64)
        t = None if type_ is None else np.ascontiguousarray(type_, dtype=np.int32)
        _check(lib().sph2d_upload_state(self._c, n, _dp(pos), _dp(vel), _dp(rho), _dp(mass),
                                        _dp(e) if e is not None else None, None,
                                        _ip(t) if t is not None else None, n))
        self.n_fluid = n

    def download_state(self) -> Dict[str, np.ndarray]:
        n = self.n_fluid
        pos, vel = np.zeros(2 * n), np.zeros(2 * n)
        rho, p, c, m = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        t = np.zeros(n, dtype=np.int32)
        _check(lib().sph2d_download_state(self._c, n, _dp(pos), _dp(vel), _dp(rho), _dp(p), _dp(c),
                                          _dp(m), _ip(t), n))
        return {"x": pos[:n], "y": pos[n:], "u": vel[:n], "v": vel[n:], "rho": rho, "p": p,
                "c": c, "mass": m, "type": t}

    # --- the reference's routines -------------------------------------------------
    def distribute_period_boundary(self):
        _check(lib().sph2d_distribute_period_boundary(self._c))

    def devide_block(self, i_time_step: int):
        _check(lib().sph2d_devide_block(self._c, i_time_step))

    def single_step_compute(self, i_time_step: int = 1):
        _check(lib().sph2d_single_step_compute(self._c, i_time_step))

    def single_step_compute_vjp(self, lam_drhodt, lam_ax, lam_ay, lam_avg_x=None, lam_avg_y=None):
        """Reverse of the last single_step_compute: the gradient of
        sum(lam_drhodt*drho/dt + lam_ax*ax + lam_ay*ay [+ lam_avg*XSPH velocity]) with
        respect to every fluid particle's x, y, u, v, rho (gid order), the periodic
        ghosts' shares included. Neighbour set, surface flags and limiter branches
        are held fixed."""
        n = self.n_fluid

        def weight(a, name):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.ndim != 1 or a.shape[0] != n:
                raise ValueError(f"{name} must be 1-D with one entry per fluid particle ({n}), got {a.shape}")
            return a

        if (lam_avg_x is None) != (lam_avg_y is None):
            raise ValueError("lam_avg_x and lam_avg_y go together")
        ld = weight(lam_drhodt, "lam_drhodt")
        la = np.concatenate([weight(lam_ax, "lam_ax"), weight(lam_ay, "lam_ay")])
        lg = None
        if lam_avg_x is not None:
            lg = np.concatenate([weight(lam_avg_x, "lam_avg_x"), weight(lam_avg_y, "lam_avg_y")])
        gp, gv, gr = np.zeros(2 * n), np.zeros(2 * n), np.zeros(n)
        _check(lib().sph2d_single_step_compute_vjp(self._c, n, _dp(ld), _dp(la), _dp(lg) if lg is not None else None,
                                                   n, _dp(gp), _dp(gv), _dp(gr)))
        return {"x": gp[:n], "y": gp[n:], "u": gv[:n], "v": gv[n:], "rho": gr}

    def single_step_compute_vjp_device(self, lam_drhodt: "DeviceArray", lam_acc: "DeviceArray",
                                       grad_pos: "DeviceArray", grad_vel: "DeviceArray", grad_rho: "DeviceArray",
                                       lam_avg: "DeviceArray" = None):
        """sph2d_single_step_compute_vjp_device: the same reverse on device arrays
        (lam_drhodt n, lam_acc (n,2) columns, lam_avg (n,2) or None; gradients of the
        same shapes), enqueued on the context's stream and synchronised here."""
        n = self.n_fluid
        for a, m, name in ((lam_drhodt, 1, "lam_drhodt"), (lam_acc, 2, "lam_acc"), (grad_pos, 2, "grad_pos"),
                           (grad_vel, 2, "grad_vel"), (grad_rho, 1, "grad_rho"), (lam_avg, 2, "lam_avg")):
            if a is not None and a.host.size != m * n:
                raise ValueError(f"{name} must hold {m} x {n} values, got {a.host.size}")
        _check(lib().sph2d_single_step_compute_vjp_device(
            self._c, n, lam_drhodt.ptr, lam_acc.ptr, lam_avg.ptr if lam_avg is not None else None, n,
            grad_pos.ptr, grad_vel.ptr, grad_rho.ptr))
        self.diagnostics()                         # synchronises the context's stream

    # --- the wave-theory wavemaker and its damping zone ---------------------------------
    def relaxation_init(self, zone_length: float):
        _check(lib().sph2d_relaxation_init(self._c, zone_length))

    def wave_theory_update(self, w: "Wavemaker", i_time_step: int):
        _check(lib().sph2d_wave_theory_update(self._c, ctypes.byref(w), i_time_step))

    def wave_maker_damping_zone(self, i_time_step: int):
        _check(lib().sph2d_wave_maker_damping_zone(self._c, i_time_step))

    # --- the differentiable wave tank ------------------------------------------------
    def forward(self, cfg: AdjointCfg, a, first_step: int, n_steps: int):
        """sph2d_forward: steps first_step..first_step+n_steps-1 driven by the stroke
        sum_k a_k sin(2 pi k f0 t); returns the probe pressures, (n_steps //
        sample_every, n_probes)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.ndim != 1 or a.shape[0] != cfg.n_harmonics:
            raise ValueError(f"a must hold n_harmonics = {cfg.n_harmonics} coefficients")
        ns = n_steps // cfg.sample_every
        out = np.zeros((max(ns, 1), cfg.n_probes))
        _check(lib().sph2d_forward(self._c, ctypes.byref(cfg), _dp(a), first_step, n_steps, _dp(out)))
        self._adj_shape = (ns, cfg.n_probes, cfg.n_harmonics)
        return out[:ns]

    def backward(self, dl_dprobe):
        """sph2d_backward: the gradient with respect to the paddle coefficients of
        sum(dl_dprobe * probe pressures) of the last forward."""
        ns, npb, nh = getattr(self, "_adj_shape", (None, None, 16))
        w = np.ascontiguousarray(dl_dprobe, dtype=np.float64)
        if ns is not None and w.shape != (ns, npb):
            raise ValueError(f"dl_dprobe must have shape {(ns, npb)}, got {w.shape}")
        g = np.zeros(nh)
        _check(lib().sph2d_backward(self._c, _dp(w.reshape(-1)) if w.size else _dp(np.zeros(1)), _dp(g)))
        return g

    # --- single_step_compute routine by routine ---------------------------------------
    def find_particle_pair(self, i_time_step: int = 1):
        _check(lib().sph2d_find_particle_pair(self._c, i_time_step))

    def pair_list(self):
        """(pair_i, pair_j): 1-based local indices in the reference's pair order."""
        n = ctypes.c_int64()
        _check(lib().sph2d_download_pair_list(self._c, 0, None, None, ctypes.byref(n)))
        pi = np.zeros(max(n.value, 1), dtype=np.int32)
        pj = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(lib().sph2d_download_pair_list(self._c, n.value, _ip(pi), _ip(pj), ctypes.byref(n)))
        return pi[:n.value], pj[:n.value]

    def define_freesurface(self):
        _check(lib().sph2d_define_freesurface(self._c))

    def solver(self, variant: int, sps_on: int = 1):
        _check(lib().sph2d_solver(self._c, variant, sps_on))

    def stage_update(self, scheme: int, stage: int):
        _check(lib().sph2d_stage_update(self._c, scheme, stage))

    def rates(self) -> Dict[str, np.ndarray]:
        """drhodt and particle_acceleration of the fluid particles in gid order."""
        n = self.n_fluid
        d = np.zeros(n)
        a = np.zeros(2 * n)
        _check(lib().sph2d_download_rates(self._c, n, _dp(d), _dp(a), n))
        return {"drhodt": d, "ax": a[:n], "ay": a[n:]}

    def reinitialize_particle_density(self, i_time_step: int):
        _check(lib().sph2d_reinitialize_particle_density(self._c, i_time_step))

    def shift_particle_position(self, i_time_step: int):
        _check(lib().sph2d_shift_particle_position(self._c, i_time_step))

    def refresh_domain(self, i_time_step: int):
        _check(lib().sph2d_refresh_domain(self._c, i_time_step))

    def time_step(self, i_time_step: int):
        _check(lib().sph2d_time_step(self._c, i_time_step))

    def run(self, first_step: int, nsteps: int):
        _check(lib().sph2d_run(self._c, first_step, nsteps))

    def sample(self) -> np.ndarray:
        out = np.zeros(14)
        _check(lib().sph2d_sample(self._c, _dp(out)))
        return out

    # --- output (the reference's files) --------------------------------------------
    def scales(self) -> Scales:
        sc = Scales()
        _check(lib().sph2d_get_scales(self._c, ctypes.byref(sc)))
        return sc

    def output_fields(self):
        """(vorticity_z, analytic (n,4): u, v, p, |v|, step) of the last sampling step,
        global order (Refresh_domain.f90:320-351, Devide_block_and_send_tasks.f90:88-116)."""
        n = self.n_fluid
        w, a = np.zeros(n), np.zeros(4 * n)
        step = ctypes.c_int32()
        _check(lib().sph2d_download_output_fields(self._c, n, _dp(w), _dp(a), n, ctypes.byref(step)))
        return w, a.reshape(4, n).T.copy(), step.value

    def output_for_tecplot(self, i_time_step: int, video_path=None, surface_path=None, binary_path=None):
        """Output_for_tecplot.f90: ZONEs appended to video_path / surface_path and the
        #!TDV101 file binary_path (None skips a file)."""
        enc = lambda p: None if p is None else os.fspath(p).encode()
        _check(lib().sph2d_output_for_tecplot(self._c, i_time_step, enc(video_path), enc(surface_path),
                                              enc(binary_path)))

    def output_for_paraview(self, i_time_step: int, path):
        _check(lib().sph2d_output_for_paraview(self._c, i_time_step, os.fspath(path).encode()))

    def grid_size(self):
        """(element_number_x, element_number_y, element_dx) of the post-processing grid."""
        nx, ny, dx = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        _check(lib().sph2d_grid_size(self._c, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(dx)))
        return nx.value, ny.value, dx.value

    def calculation_for_postproceeding(self, i_time_step: int) -> Dict[str, np.ndarray]:
        """Calculation_For_PostProceeding.f90 on the grid nodes: u, v, p, vorticity, Fai
        and the analytic u, v, p, |v| (node order of Distribute_Grid_For_Postproceeding)."""
        nex, ney, _ = self.grid_size()
        nn = (nex + 1) * (ney + 1)
        vel, p, w, f, a = np.zeros(2 * nn), np.zeros(nn), np.zeros(nn), np.zeros(nn), np.zeros(4 * nn)
        _check(lib().sph2d_calculation_for_postproceeding(self._c, i_time_step, nn, _dp(vel), _dp(p), _dp(w), _dp(f),
                                                         _dp(a), nn))
        return {"u": vel[:nn], "v": vel[nn:], "p": p, "vorticity": w, "fai": f, "ana_u": a[:nn],
                "ana_v": a[nn:2 * nn], "ana_p": a[2 * nn:3 * nn], "ana_vm": a[3 * nn:]}

    def output_for_grid_postproceding(self, i_time_step: int, grid_path=None, binary_path=None):
        enc = lambda p: None if p is None else os.fspath(p).encode()
        _check(lib().sph2d_output_for_grid_postproceding(self._c, i_time_step, enc(grid_path), enc(binary_path)))

    def sampling(self, i_time_step: int, x, y, gravity: float = 0.0) -> Dict[str, np.ndarray]:
        """Sampling_subroutine.f90:21-177 at the points (x, y): average and MLS pressure, MLS u, v."""
        n = len(x)
        pos = _col2(x, y)
        avg, mp, vel = np.zeros(n), np.zeros(n), np.zeros(2 * n)
        _check(lib().sph2d_sampling(self._c, i_time_step, n, _dp(pos), n, gravity, _dp(avg), _dp(mp), _dp(vel)))
        return {"average_p": avg, "mls_p": mp, "u": vel[:n], "v": vel[n:]}

    def wave_probes(self, probe_x=()):
        """Sampling_subroutine.f90:246-279: (wave front, elevation at each probe x), unscaled."""
        px = _f64(np.asarray(probe_x, dtype=np.float64).reshape(-1))
        el = np.zeros(max(len(px), 1))
        front = ctypes.c_double()
        _check(lib().sph2d_wave_probes(self._c, len(px), _dp(px) if len(px) else None, ctypes.byref(front),
                                       _dp(el)))
        return front.value, el[:len(px)]

    def output_sampling_lines(self, i_time_step: int, dir_path, first, last, n_points: int = 41):
        """Sampling_subroutine.f90:466-512: <k>th_<t>_SamplingLine_Results.dat for each line."""
        f = _f64(np.asarray(first, dtype=np.float64).reshape(-1))
        l = _f64(np.asarray(last, dtype=np.float64).reshape(-1))
        _check(lib().sph2d_output_sampling_lines(self._c, i_time_step, os.fspath(dir_path).encode(), len(f) // 2,
                                                 _dp(f), _dp(l), n_points))

    def wavemaker_update(self, position: float, velocity: float):
        """Wavemaker_Subroutine.f90:140-215 before a step: place the wavemaker particles, pull back."""
        _check(lib().sph2d_wavemaker_update(self._c, position, velocity))

    def wavemaker_pull_back(self, position: float, velocity: float):
        """Wavemaker_Subroutine.f90:196-215: mirror fluid particles behind the paddle."""
        _check(lib().sph2d_wavemaker_pull_back(self._c, position, velocity))

    def output_final_results(self, path):
        """Final_Doamin_Results.dat (Parallel_Taylor_Green_Vortices_2D_Release.f90:185-199)."""
        st = self.download_state()
        write_final_results(path, st["x"], st["y"], st["u"], st["v"], st["rho"], st["p"])

    def output_latest_results(self, i_time_step: int, path):
        """The restart ZONE of Latest_Time_Results.dat (time_iteration.f90:495-498)."""
        st = self.download_state()
        n = self.n_fluid
        write_latest_results_zone(path, i_time_step, st["x"], st["y"], st["u"], st["v"], st["p"], st["rho"],
                                  st["mass"], np.zeros(n), st["c"], st["type"])

    # --- inspection -------------------------------------------------------------------
    def set_debug(self, on: bool = True):
        _check(lib().sph2d_set_debug(self._c, 1 if on else 0))

    def diagnostics(self) -> Diag:
        d = Diag()
        _check(lib().sph2d_get_diagnostics(self._c, ctypes.byref(d)))
        return d

    def n_local(self) -> int:
        return self.diagnostics().n_local

    def local(self, name: str) -> np.ndarray:
        n = self.n_local()
        if name in LOCAL_I32:
            out = np.zeros(n, dtype=np.int32)
        elif name in LOCAL_F64:
            out = np.zeros(n, dtype=np.float64)
        else:
            raise KeyError(name)
        _check(lib().sph2d_download_local(self._c, name.encode(), out.ctypes.data_as(ctypes.c_void_p), n))
        return out

    def pairs(self) -> np.ndarray:
        """Sorted keys lo*2^32+hi of the unordered pair set (0-based gids)."""
        n = ctypes.c_int64()
        _check(lib().sph2d_download_pairs(self._c, 0, None, None, ctypes.byref(n)))
        lo = np.zeros(max(n.value, 1), dtype=np.int32)
        hi = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(lib().sph2d_download_pairs(self._c, n.value, _ip(lo), _ip(hi), ctypes.byref(n)))
        return lo[: n.value].astype(np.int64) * (1 << 32) + hi[: n.value]

    def time_pass(self, pass_id: int, reps: int) -> float:
        """Average ms of one part of the evaluation (0 force pass, 1 pass A, 2 all)."""
        ms = ctypes.c_double()
        _check(lib().sph2d_time_pass(self._c, pass_id, reps, ctypes.byref(ms)))
        return ms.value

    def set_pass_timing(self, on: bool) -> None:
        """Record HIP events around pass A and the force pass of every evaluation."""
        _check(lib().sph2d_set_pass_timing(self._c, 1 if on else 0))

    def pass_timing(self):
        """Average in-step ms since the last call: (force pass, pass A, evaluation), count."""
        out = np.zeros(3)
        n = ctypes.c_int32()
        _check(lib().sph2d_get_pass_timing(self._c, _dp(out), ctypes.byref(n)))
        return float(out[0]), float(out[1]), float(out[2]), n.value

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(lib().sph2d_get_stream(self._c, ctypes.byref(s)))
        return s.value or 0


def compute_kernel(r, dx, h):
    """Array form of compute_kernel.f90 on the GPU: r (n,), dx (n,2), h (n,)."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    n = len(r)
    d = np.ascontiguousarray(np.asarray(dx, dtype=np.float64).T.reshape(-1))  # (n,2) column-major
    h = np.ascontiguousarray(h, dtype=np.float64)
    w = np.zeros(n)
    dw = np.zeros(2 * n)
    _check(lib().sph2d_compute_kernel_n(n, _dp(r), _dp(d), _dp(h), _dp(w), _dp(dw)))
    return w, np.stack([dw[:n], dw[n:]], axis=1)


def status_equation(rho, type_=None, params: Optional[Params] = None):
    """Array form of status_equation.f90 on the GPU: returns (p, c)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = len(rho)
    t = np.full(n, 2, dtype=np.int32) if type_ is None else np.ascontiguousarray(type_, dtype=np.int32)
    p, c = np.zeros(n), np.zeros(n)
    e = np.zeros(n)
    _check(lib().sph2d_status_equation_n(ctypes.byref(params) if params is not None else None, n,
                                         _dp(rho), _dp(e), _ip(t), _dp(p), _dp(c)))
    return p, c


# --- device buffers for the device-array entry points (HIP runtime via ctypes) ----------
_hip = None


def _hiprt():
    """The HIP runtime by name, for device synchronisation in processes that load no
    other HIP runtime (the bench); device memory goes through DeviceArray."""
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip = h
    return _hip


class DeviceArray:
    """A device copy of a 1-D numpy array, allocated and copied through this library's
    HIP runtime (sph2d_device_alloc/_copy/_free: another HIP runtime loaded into the
    process, e.g. a framework's bundled one, is a different instance); .get() copies
    it back."""

    def __init__(self, a: np.ndarray):
        self.host = np.ascontiguousarray(a)
        self.ptr = ctypes.c_void_p()
        L = lib()
        _check(L.sph2d_device_alloc(max(self.host.nbytes, 1), ctypes.byref(self.ptr)))
        _check(L.sph2d_device_copy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1))

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        _check(lib().sph2d_device_copy(out.ctypes.data, self.ptr, out.nbytes, 0))
        return out

    def free(self):
        if self.ptr:
            lib().sph2d_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def reorder_particle_index(n_fluid, box_x, box_y, x, y, type_, f64_cols=(), i32_cols=(), no_alloc=False,
                           with_perm=True):
    """Reorder_Particle_Index (MS/Reorder_Particle_Index.f90:54-287) on the GPU for host
    arrays: uploads, calls sph2d_reorder_particle_index (no_alloc: its variant with
    the library's static scratch, sph2d_reorder_particle_index_noalloc), downloads.
    Returns (x, y, type, f64_cols, i32_cols, perm, counts) with counts = (fluid,
    wavemaker, fixed ghost, inlet, outlet, total kept); arrays hold the kept
    particles only (perm is None when with_perm is False: the call gets NULL)."""
    n = len(x)
    dx = DeviceArray(np.asarray(x, dtype=np.float64))
    dy = DeviceArray(np.asarray(y, dtype=np.float64))
    dt = DeviceArray(np.asarray(type_, dtype=np.int32))
    df = [DeviceArray(np.asarray(c, dtype=np.float64)) for c in f64_cols]
    di = [DeviceArray(np.asarray(c, dtype=np.int32)) for c in i32_cols]
    dp = DeviceArray(np.full(n, -1, dtype=np.int32)) if with_perm else None
    fptrs = (ctypes.c_void_p * max(len(df), 1))(*[d.ptr.value for d in df])
    iptrs = (ctypes.c_void_p * max(len(di), 1))(*[d.ptr.value for d in di])
    counts = np.zeros(6, dtype=np.int32)
    f = lib().sph2d_reorder_particle_index_noalloc if no_alloc else lib().sph2d_reorder_particle_index
    _check(f(n, n_fluid, box_x, box_y, dx.ptr, dy.ptr, dt.ptr, len(df), fptrs, len(di), iptrs,
             dp.ptr if dp is not None else None, _ip(counts), None))
    k = int(counts[5])
    out = (dx.get()[:k], dy.get()[:k], dt.get()[:k], [d.get()[:k] for d in df], [d.get()[:k] for d in di],
           dp.get()[:k] if dp is not None else None, counts)
    for d in [dx, dy, dt] + ([dp] if dp is not None else []) + df + di:
        d.free()
    return out


# --- the reference's file formats, from host arrays (no GPU needed) ---------------------
def _col2(a, b):
    return np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)]))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def format_fortran(value: float, kind: str, w: int, d: int) -> str:
    """One value under the Fortran edit descriptor Fw.d (kind 'F') or Ew.dE3 ('E')."""
    buf = ctypes.create_string_buffer(w + 1)
    _check(lib().sph2d_format_fortran(float(value), kind.encode(), w, d, buf, w + 1))
    return buf.value.decode()


def time_name(t: float, w: int, d: int) -> str:
    """trim(adjustl(char_time)) of write(char_time,'(fW.D)') t: the reference's file and zone names."""
    return format_fortran(t, "F", w, d).strip()


def write_final_results(path, x, y, u, v, rho, p):
    n = len(x)
    _check(lib().sph2d_write_final_results(os.fspath(path).encode(), n, _dp(_col2(x, y)), _dp(_col2(u, v)),
                                           _dp(_f64(rho)), _dp(_f64(p)), n))


def write_tecplot_zone(path, current_time, x, y, u, v, p, vort_z, ana, scales: Scales):
    """ana: (n,4) analytic u, v, p, |v|."""
    n = len(x)
    a = np.ascontiguousarray(np.asarray(ana, dtype=np.float64).T.reshape(-1))
    _check(lib().sph2d_write_tecplot_zone(os.fspath(path).encode(), current_time, n, _dp(_col2(x, y)),
                                          _dp(_col2(u, v)), _dp(_f64(p)), _dp(_f64(vort_z)), _dp(a), n,
                                          ctypes.byref(scales)))


def write_surface_zone(path, current_time, x, y, nx, ny, scales: Scales):
    n = len(x)
    _check(lib().sph2d_write_surface_zone(os.fspath(path).encode(), current_time, n, _dp(_col2(x, y)),
                                          _dp(_col2(nx, ny)), n, ctypes.byref(scales)))


def write_tecplot_binary(path, time, x, y, u, v, p, scales: Scales):
    n = len(x)
    _check(lib().sph2d_write_tecplot_binary(os.fspath(path).encode(), time, n, _dp(_col2(x, y)), _dp(_col2(u, v)),
                                            _dp(_f64(p)), n, ctypes.byref(scales)))


def write_paraview(path, x, y, u, v, p, scales: Scales):
    n = len(x)
    _check(lib().sph2d_write_paraview(os.fspath(path).encode(), n, _dp(_col2(x, y)), _dp(_col2(u, v)),
                                      _dp(_f64(p)), n, ctypes.byref(scales)))


def write_latest_results_zone(path, i_time_step, x, y, u, v, p, rho, mass, energy, c, initial_type,
                              wavemaker_layer=None):
    n = len(x)
    t = np.ascontiguousarray(initial_type, dtype=np.int32)
    wl = None if wavemaker_layer is None else np.ascontiguousarray(wavemaker_layer, dtype=np.int32)
    _check(lib().sph2d_write_latest_results_zone(os.fspath(path).encode(), i_time_step, n, _dp(_col2(x, y)),
                                                 _dp(_col2(u, v)), _dp(_f64(p)), _dp(_f64(rho)), _dp(_f64(mass)),
                                                 _dp(_f64(energy)), _dp(_f64(c)), _ip(t),
                                                 _ip(wl) if wl is not None else None, n))


def read_latest_results(path, zone: int = 0, capacity: int = 1 << 24) -> Dict[str, np.ndarray]:
    """Distribute_Latest_Particle.f90: the restart zone (0 = last), stepping back past NaN zones."""
    cap = int(capacity)
    n = ctypes.c_int32()
    step = ctypes.c_int32()
    pos, vel = np.zeros(2 * cap), np.zeros(2 * cap)
    p, rho, m, e, c = (np.zeros(cap) for _ in range(5))
    t, wl = np.zeros(cap, dtype=np.int32), np.zeros(cap, dtype=np.int32)
    _check(lib().sph2d_read_latest_results(os.fspath(path).encode(), zone, cap, ctypes.byref(n), _dp(pos),
                                           _dp(vel), _dp(p), _dp(rho), _dp(m), _dp(e), _dp(c), _ip(t), _ip(wl),
                                           ctypes.byref(step), cap))
    k = n.value
    return {"x": pos[:k], "y": pos[cap:cap + k], "u": vel[:k], "v": vel[cap:cap + k], "p": p[:k], "rho": rho[:k],
            "mass": m[:k], "energy": e[:k], "c": c[:k], "type": t[:k], "wavemaker_layer": wl[:k], "step": step.value}


def write_sampling(dir_path, nondim_time: float, out14):
    """One line appended to each of Sampling_subroutine.f90's six history files."""
    o = _f64(out14)
    _check(lib().sph2d_write_sampling(os.fspath(dir_path).encode(), nondim_time, _dp(o)))


def write_initial_grid(grid_path, node_path, nex: int, ney: int, element_dx: float):
    """Initial_Post_Proceeding_Grid.dat / _Node.dat (Distribute_Grid_For_Postproceeding.f90:60-95)."""
    enc = lambda p: None if p is None else os.fspath(p).encode()
    _check(lib().sph2d_write_initial_grid(enc(grid_path), enc(node_path), nex, ney, element_dx))


def write_grid_zone(path, current_time, nex, ney, element_dx, u, v, p, vort_z, fai, ana, scales: Scales):
    """ana: (n_nodes,4) analytic u, v, p, |v|."""
    nn = (nex + 1) * (ney + 1)
    a = np.ascontiguousarray(np.asarray(ana, dtype=np.float64).T.reshape(-1))
    _check(lib().sph2d_write_grid_zone(os.fspath(path).encode(), current_time, nex, ney, element_dx,
                                       _dp(_col2(u, v)), _dp(_f64(p)), _dp(_f64(vort_z)), _dp(_f64(fai)), _dp(a), nn,
                                       ctypes.byref(scales)))


def write_grid_binary(path, time, nex, ney, element_dx, u, v, p, fai, scales: Scales):
    nn = (nex + 1) * (ney + 1)
    _check(lib().sph2d_write_grid_binary(os.fspath(path).encode(), time, nex, ney, element_dx, _dp(_col2(u, v)),
                                         _dp(_f64(p)), _dp(_f64(fai)), nn, ctypes.byref(scales)))


def format_list_directed(value: float) -> str:
    """One REAL(8) item of write(unit,*) as ifort writes it (23 characters)."""
    buf = ctypes.create_string_buffer(32)
    _check(lib().sph2d_format_list_directed(float(value), buf, 32))
    return buf.value.decode()


def write_list_directed(path, values):
    v = _f64(np.asarray(values, dtype=np.float64).reshape(-1))
    _check(lib().sph2d_write_list_directed(os.fspath(path).encode(), len(v), _dp(v)))


def write_f20_8(path, values):
    v = _f64(np.asarray(values, dtype=np.float64).reshape(-1))
    _check(lib().sph2d_write_f20_8(os.fspath(path).encode(), len(v), _dp(v)))


def kalman_filter(current_time_number: int, measured, last_x, last_p) -> np.ndarray:
    """Sampling_subroutine.f90:300-322; last_x and last_p (float64 arrays) are updated in place."""
    m = _f64(measured)
    for a in (last_x, last_p):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous and len(a) == len(m)):
            raise ValueError("last_x / last_p must be contiguous float64 arrays of len(measured)")
    out = np.zeros(len(m))
    _check(lib().sph2d_kalman_filter(len(m), current_time_number, _dp(m), _dp(last_x), _dp(last_p), _dp(out)))
    return out


def sampling_line_points(first, last, n_points: int = 41):
    """Initialize_SPH_Calculation.f90:385-398: (x, y) of n_points points from first to last."""
    out = np.zeros(2 * n_points)
    _check(lib().sph2d_sampling_line_points(n_points, _dp(_f64(first)), _dp(_f64(last)), _dp(out), n_points))
    return out[:n_points], out[n_points:]


def write_sampling_line_positions(dir_path, line: int, x, y):
    n = len(x)
    _check(lib().sph2d_write_sampling_line_positions(os.fspath(dir_path).encode(), line, n, _dp(_col2(x, y)), n))


class SamplingProbes:
    """Sampling_subroutine.f90:232-323 for a probe configuration: each call at a
    sampling step appends the wave-front line, one line per wave gauge and, per
    pressure probe, its MLS, average and Kalman-filtered pressure lines, in the
    reference's files under dir_path (Wave_front_results.dat in wave_front_dir).
    The Kalman state is held here between calls."""

    def __init__(self, sim: "Simulation", dir_path, points=(), probe_x=(), gravity: float = 0.0,
                 sampling_time_step: int = 100, wave_front_dir=None):
        self.sim, self.dir = sim, os.fspath(dir_path)
        self.front_dir = os.fspath(wave_front_dir) if wave_front_dir is not None else "."
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 2)
        self.px, self.py = _f64(pts[:, 0]), _f64(pts[:, 1])
        self.probe_x = _f64(np.asarray(probe_x, dtype=np.float64).reshape(-1))
        self.gravity, self.every = gravity, sampling_time_step
        self.last_x, self.last_p = np.zeros(len(self.px)), np.zeros(len(self.px))
        os.makedirs(os.path.join(self.dir, "Press"), exist_ok=True)
        os.makedirs(os.path.join(self.dir, "WaveHeight"), exist_ok=True)

    def __call__(self, i_time_step: int):
        sc = self.sim.scales()
        t = i_time_step * self.sim.params.dt            # current_time (no standing time)
        front, elev = self.sim.wave_probes(self.probe_x)
        dx = self.sim.params.dx
        for j, e in enumerate(elev):
            write_list_directed(os.path.join(self.dir, "WaveHeight", f"wave_gauge_{j + 1}_result.dat"),
                                [t * sc.time_0, (e + dx / 2.0) / sc.domain_size_y])
        write_list_directed(os.path.join(self.front_dir, "Wave_front_results.dat"),
                            [t * sc.time_0, front / sc.domain_size_y])
        if len(self.px) == 0:
            return
        r = self.sim.sampling(i_time_step, self.px, self.py, self.gravity)
        kal = kalman_filter(i_time_step // self.every, r["mls_p"], self.last_x, self.last_p)
        for m in range(len(self.px)):
            base = os.path.join(self.dir, "Press", f"Probe_{m + 1}_")
            write_f20_8(base + "Normalized_Pressure.dat", [t * sc.time_0, r["mls_p"][m] / sc.pressure_0])
            write_f20_8(base + "Average_Pressure.dat", [t * sc.time_0, r["average_p"][m] / sc.pressure_0])
            write_f20_8(base + "Kalman_Pressure.dat", [t * sc.time_0, kal[m] / sc.pressure_0])


class Batch:
    """Independent simulations advanced together (sph2d_batch): replayable steps run
    as one hipGraph launch holding every instance's step."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (ctypes.c_void_p * len(self.sims))(*[s._c for s in self.sims])
        h = ctypes.c_void_p()
        _check(lib().sph2d_batch_create(len(self.sims), arr, ctypes.byref(h)))
        self._b = h

    def run(self, first_step: int, nsteps: int):
        _check(lib().sph2d_batch_run(self._b, first_step, nsteps))

    SHAPES = ((256, 1024, 1088), (128, 640, 704), (64, 384, 416), (512, 1792, 1856))
    CONFIGS = tuple(f"{order}, {blocks}, tile {sh[0]}/{sh[1]}/{sh[2]}" for sh in SHAPES
                    for blocks in ("XCD-aware", "linear") for order in ("given order", "by size"))

    def autotune(self, first_step: int, trial_steps: int = 20):
        """Time the launch configurations (16 for FP32-mixed instances: instance
        order x block order x tile shape of the neighbour passes; 4 for FP64), each on
        trial_steps real steps after one capture step, and keep the fastest; returns
        (ms per step of each, 0 where not tried; chosen)."""
        ms = np.zeros(len(self.CONFIGS))
        ch = ctypes.c_int32()
        _check(lib().sph2d_batch_autotune(self._b, first_step, trial_steps, _dp(ms), ctypes.byref(ch)))
        return ms, ch.value

    def configure(self, config: int):
        _check(lib().sph2d_batch_configure(self._b, int(config)))

    @property
    def config(self) -> int:
        c = ctypes.c_int32()
        _check(lib().sph2d_batch_config(self._b, ctypes.byref(c)))
        return c.value

    def info(self):
        nodes, reps = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().sph2d_batch_info(self._b, ctypes.byref(nodes), ctypes.byref(reps)))
        return nodes.value, reps.value

    def close(self):
        if self._b:
            lib().sph2d_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def paddle_probe_function():
    """torch.autograd.Function over the differentiable wave tank: apply(a, make_sim,
    cfg, first_step, n_steps) takes a context from make_sim() (at the start of the
    window), runs sph2d_forward with the coefficients a and returns the probe
    pressures as a float64 tensor; backward runs sph2d_backward on that context with
    the incoming gradient."""
    import torch

    class PaddleProbes(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a, make_sim, cfg, first_step, n_steps):
            sim = make_sim()
            p = sim.forward(cfg, a.detach().cpu().double().numpy(), first_step, n_steps)
            ctx.sim = sim
            return torch.from_numpy(p.copy()).to(a.device, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            g = ctx.sim.backward(grad.detach().cpu().double().numpy())
            return torch.from_numpy(g).to(grad.device), None, None, None, None

    return PaddleProbes
