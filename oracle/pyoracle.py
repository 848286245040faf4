"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE, never the product).

Only tests/, the smoke check and the CPU-baseline leg of bench.py load this.
The oracle restates Taylor_Green_Vortex/*.f90 of the reference in C
(oracle/sph2d_oracle.c); this module only marshals arrays.
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsph2d_ref.so")

# solver variants (sph2d_oracle.h ORC_SOLVER_*)
ORC_SOLVER_MODIFIED_LINEAR_RIEMANN = 0
ORC_SOLVER_LINEAR_RIEMANN = 1
ORC_SOLVER_HLLC = 2
ORC_SOLVER_MUSCL = 3
ORC_SOLVER_MS_MUSCL_HLLC = 4
ORC_SOLVER_MS_HLLC = 5


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("box_x", ctypes.c_double), ("box_y", ctypes.c_double),
        ("rho0", ctypes.c_double), ("c0sq", ctypes.c_double),
        ("nu", ctypes.c_double), ("U", ctypes.c_double),
        ("h_factor", ctypes.c_double), ("dt", ctypes.c_double),
        ("integrator", ctypes.c_int32), ("shift_on", ctypes.c_int32),
        ("mls_on", ctypes.c_int32), ("mls_every", ctypes.c_int32),
        ("mls_order", ctypes.c_int32), ("solver", ctypes.c_int32),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("fai", ctypes.c_double),
        ("delta", ctypes.c_double), ("delta_alpha", ctypes.c_double),
        ("epsilon_xsph", ctypes.c_double),
        ("shift_beta", ctypes.c_double), ("shift_R", ctypes.c_double),
        ("mls_alpha", ctypes.c_double),
        ("rate_reuse", ctypes.c_int32),
        ("damping", ctypes.c_int32), ("damping_length", ctypes.c_double),
        ("damping_amplitude", ctypes.c_double), ("water_depth", ctypes.c_double),
        ("walls", ctypes.c_int32), ("gravity", ctypes.c_double),
        ("wavemaker_layers", ctypes.c_int32),
        ("sps_on", ctypes.c_int32), ("renorm", ctypes.c_int32),
        ("ms", ctypes.c_int32), ("background_p", ctypes.c_double), ("release_time", ctypes.c_double),
        ("body_cx", ctypes.c_double), ("body_cy", ctypes.c_double), ("body_len", ctypes.c_double),
        ("body_speed", ctypes.c_double), ("body_ramp", ctypes.c_double),
        ("sample_every", ctypes.c_int32), ("max_shift_dx", ctypes.c_double),
    ]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref where the reference is present)."""
    import subprocess
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.orc_default_params.argtypes = [P(OrcParams)]
        L.orc_create.argtypes = [P(OrcParams)]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_init_tgv", "orc_period_boundary", "orc_reinitialize_density",
                     "orc_shift", "orc_shift_sums", "orc_refresh"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("orc_devide_block", "orc_single_step_compute", "orc_time_step"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.orc_sample.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        L.orc_grid_postprocess.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_double, P(ctypes.c_double)]
        L.orc_sampling_points.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_double),
                                          ctypes.c_double, P(ctypes.c_double)]
        L.orc_wave_probes.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_wall_positions.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_wall_positions.restype = ctypes.c_int32
        L.orc_wavemaker_pull_back.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_wavemaker_update.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_relaxation_init.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_wave_theory_update.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32] + \
            [ctypes.c_double] * 7 + [ctypes.c_int32]
        L.orc_relaxation_blend.argtypes = [ctypes.c_void_p]
        L.orc_initial_ke.argtypes = [ctypes.c_void_p]
        L.orc_initial_ke.restype = ctypes.c_double
        for name in ("orc_n_fluid", "orc_n_global", "orc_n_local", "orc_n_singular"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_int32
        L.orc_pair_number.argtypes = [ctypes.c_void_p]
        L.orc_pair_number.restype = ctypes.c_int64
        L.orc_mut_f64.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(ctypes.c_int64)]
        L.orc_mut_f64.restype = P(ctypes.c_double)
        L.orc_mut_i32.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(ctypes.c_int64)]
        L.orc_mut_i32.restype = P(ctypes.c_int32)
        L.orc_compute_kernel.argtypes = [ctypes.c_int32, ctypes.c_double, P(ctypes.c_double),
                                         ctypes.c_double, P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_status_equation.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_brinv.argtypes = [P(ctypes.c_double), ctypes.c_int32]
        L.orc_brinv.restype = ctypes.c_int32
        L.orc_matrix_inversion_less_three.argtypes = [ctypes.c_int32, P(ctypes.c_double),
                                                      P(ctypes.c_double)]
        L.orc_matrix_inversion_less_three.restype = ctypes.c_int32
        L.orc_set_fluid_state.argtypes = [ctypes.c_void_p, ctypes.c_int32] + [P(ctypes.c_double)] * 6 + [P(ctypes.c_int32)]
        L.orc_set_fluid_state.restype = ctypes.c_int32
        D = ctypes.c_double
        L.orc_muscl_limiter.argtypes = [D, D, D, P(D), P(D)]
        VP = ctypes.c_void_p
        L.orc_reorder_particle_index.argtypes = [ctypes.c_int32, ctypes.c_int32, D, D, P(D), P(D),
                                                 P(ctypes.c_int32), ctypes.c_int32, P(VP), ctypes.c_int32,
                                                 P(VP), P(ctypes.c_int32), P(ctypes.c_int32)]
        L.orc_hllc_star.argtypes = [P(D), D, D, D, P(D), D, D, D, P(D), D, P(D), P(D)]
        L.orc_init_moving_square.argtypes = [VP]
        L.orc_ms_time_step.argtypes = [VP, ctypes.c_int32]
        L.orc_ms_body_motion.argtypes = [VP, D]
        L.orc_ms_single_step_compute.argtypes = [VP, ctypes.c_int32, D]
        L.orc_ms_body_force.argtypes = [VP, P(D)]
        _lib = L
    return _lib


def default_params(**kw) -> OrcParams:
    p = OrcParams()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One oracle context: the reference's global + subdomain arrays."""

    def __init__(self, **kw):
        self.params = default_params(**kw)
        self._c = lib().orc_create(ctypes.byref(self.params))

    def close(self):
        if self._c:
            lib().orc_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- steps -----------------------------------------------------------
    def init_tgv(self):
        lib().orc_init_tgv(self._c)

    # --- the moving-square program (ms = 1) ---------------------------------
    def init_moving_square(self):
        lib().orc_init_moving_square(self._c)

    def ms_time_step(self, it: int):
        lib().orc_ms_time_step(self._c, it)

    def ms_body_motion(self, t: float):
        lib().orc_ms_body_motion(self._c, t)

    def ms_single_step_compute(self, it: int, t: float):
        lib().orc_ms_single_step_compute(self._c, it, t)

    def ms_body_force(self) -> np.ndarray:
        """(pressure x, y, viscous x, y, total x, y) of the last evaluation that computed it."""
        out = np.zeros(6)
        lib().orc_ms_body_force(self._c, _dp(out))
        return out

    def set_fluid_state(self, x, y, u, v, rho, mass, type_=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u, v, rho, mass)]
        t = None
        if type_ is not None:
            t = np.ascontiguousarray(type_, dtype=np.int32)
        rc = lib().orc_set_fluid_state(self._c, len(arrs[0]), *[_dp(a) for a in arrs],
                                       t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if t is not None else None)
        if rc != 0:
            raise ValueError("oracle capacity exceeded")

    def period_boundary(self):
        lib().orc_period_boundary(self._c)

    def devide_block(self, it: int):
        lib().orc_devide_block(self._c, it)

    def single_step_compute(self, it: int = 1):
        lib().orc_single_step_compute(self._c, it)

    def time_step(self, it: int):
        lib().orc_time_step(self._c, it)

    def reinitialize_density(self):
        lib().orc_reinitialize_density(self._c)

    def shift(self):
        lib().orc_shift(self._c)

    def shift_sums(self):
        lib().orc_shift_sums(self._c)

    def refresh(self):
        lib().orc_refresh(self._c)

    def grid_postprocess(self, it: int, nex: int, ney: int, edx: float) -> np.ndarray:
        """Calculation_For_PostProceeding on the element nodes: (n_nodes, 9) columns
        u, v, p, vorticity_z, Fai, analytic u, v, p, |v|."""
        nn = (nex + 1) * (ney + 1)
        out = np.zeros(9 * nn)
        lib().orc_grid_postprocess(self._c, it, nex, ney, edx, _dp(out))
        return out.reshape(9, nn).T.copy()

    def sampling_points(self, x, y, gravity: float = 0.0) -> np.ndarray:
        """Sampling (Sampling_subroutine.f90:21-177) at the points: (n, 4) columns
        average p, MLS p, MLS u, MLS v."""
        px = np.ascontiguousarray(x, dtype=np.float64)
        py = np.ascontiguousarray(y, dtype=np.float64)
        n = len(px)
        out = np.zeros(4 * max(n, 1))
        lib().orc_sampling_points(self._c, n, _dp(px), _dp(py), gravity, _dp(out))
        return out[:4 * n].reshape(4, n).T.copy()

    def wave_probes(self, probe_x=()):
        """(wave front, elevations) of Sampling_subroutine.f90:246-279, unscaled."""
        px = np.ascontiguousarray(probe_x, dtype=np.float64).reshape(-1)
        out = np.zeros(1 + len(px))
        lib().orc_wave_probes(self._c, len(px), _dp(px) if len(px) else None, _dp(out))
        return out[0], out[1:]

    def wall_positions(self):
        """(x, y) of the fixed ghost walls (Distribute_Initial_Particle.f90:317-495)."""
        n = lib().orc_wall_positions(self._c, None, None)
        x, y = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        lib().orc_wall_positions(self._c, _dp(x), _dp(y))
        return x[:n], y[:n]

    def wavemaker_pull_back(self, xp: float, up: float):
        lib().orc_wavemaker_pull_back(self._c, xp, up)

    def relaxation_init(self, zone_length: float):
        lib().orc_relaxation_init(self._c, zone_length)

    def wave_theory_update(self, wave_type, soft_steps, standing_steps, dt, g, depth, height, k, omega, phase, it):
        lib().orc_wave_theory_update(self._c, wave_type, soft_steps, standing_steps, dt, g, depth, height, k, omega,
                                     phase, it)

    def relaxation_blend(self):
        lib().orc_relaxation_blend(self._c)

    def wavemaker_update(self, xp: float, up: float):
        lib().orc_wavemaker_update(self._c, xp, up)

    def sample(self) -> np.ndarray:
        out = np.zeros(14)
        lib().orc_sample(self._c, _dp(out))
        return out

    # --- sizes / arrays -----------------------------------------------------
    @property
    def n_fluid(self):
        return lib().orc_n_fluid(self._c)

    @property
    def n_global(self):
        return lib().orc_n_global(self._c)

    @property
    def n_singular(self):
        return lib().orc_n_singular(self._c)

    @property
    def n_local(self):
        return lib().orc_n_local(self._c)

    @property
    def pair_number(self):
        return lib().orc_pair_number(self._c)

    def f64(self, name: str) -> np.ndarray:
        n = ctypes.c_int64()
        ptr = lib().orc_mut_f64(self._c, name.encode(), ctypes.byref(n))
        if n.value < 0:
            raise KeyError(name)
        if n.value == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()

    def i32(self, name: str) -> np.ndarray:
        n = ctypes.c_int64()
        ptr = lib().orc_mut_i32(self._c, name.encode(), ctypes.byref(n))
        if n.value < 0:
            raise KeyError(name)
        if n.value == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()

    def set_f64(self, name: str, values) -> None:
        n = ctypes.c_int64()
        ptr = lib().orc_mut_f64(self._c, name.encode(), ctypes.byref(n))
        arr = np.ctypeslib.as_array(ptr, shape=(n.value,))
        arr[:] = values

    def set_i32(self, name: str, values) -> None:
        n = ctypes.c_int64()
        ptr = lib().orc_mut_i32(self._c, name.encode(), ctypes.byref(n))
        arr = np.ctypeslib.as_array(ptr, shape=(n.value,))
        arr[:] = values

    def pairs_gid(self) -> np.ndarray:
        """Unordered pair set as sorted (min gid, max gid) rows."""
        name = self.i32("name")
        pi, pj = self.i32("pair_i"), self.i32("pair_j")
        a, b = name[pi], name[pj]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        key = lo.astype(np.int64) * (1 << 32) + hi
        key.sort()
        return key


# scalar helpers ---------------------------------------------------------------
def compute_kernel(r: float, dx, h: float):
    d = np.ascontiguousarray(dx, dtype=np.float64)
    w = ctypes.c_double()
    dw = np.zeros(2)
    lib().orc_compute_kernel(2, r, _dp(d), h, ctypes.byref(w), _dp(dw))
    return w.value, dw


def brinv(a: np.ndarray):
    m = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    ok = lib().orc_brinv(m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m.shape[0])
    return ok, m


def matrix_inversion_less_three(a: np.ndarray):
    m = np.asfortranarray(a, dtype=np.float64)
    res = np.zeros_like(m, order="F")
    ok = lib().orc_matrix_inversion_less_three(m.shape[0], m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return ok, res


def muscl_limiter(delphi_i: float, delphi_j: float, d_phi_ij: float):
    li, lj = ctypes.c_double(), ctypes.c_double()
    lib().orc_muscl_limiter(delphi_i, delphi_j, d_phi_ij, ctypes.byref(li), ctypes.byref(lj))
    return li.value, lj.value


def reorder_particle_index(n_fluid, box_x, box_y, x, y, type_, f64_cols=(), i32_cols=()):
    """Reorder_Particle_Index (MS) on copies of host arrays; returns
    (x, y, type, f64_cols, i32_cols, perm, counts) cut to the kept particles."""
    x = np.array(x, dtype=np.float64)
    y = np.array(y, dtype=np.float64)
    t = np.array(type_, dtype=np.int32)
    n = len(x)
    fc = [np.array(c, dtype=np.float64) for c in f64_cols]
    ic = [np.array(c, dtype=np.int32) for c in i32_cols]
    fptr = (ctypes.c_void_p * max(len(fc), 1))(*[c.ctypes.data for c in fc])
    iptr = (ctypes.c_void_p * max(len(ic), 1))(*[c.ctypes.data for c in ic])
    perm = np.full(n, -1, dtype=np.int32)
    counts = np.zeros(6, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    lib().orc_reorder_particle_index(n, n_fluid, box_x, box_y, _dp(x), _dp(y), ip(t), len(fc), fptr, len(ic),
                                     iptr, ip(perm), ip(counts))
    k = int(counts[5])
    return x[:k], y[:k], t[:k], [c[:k] for c in fc], [c[:k] for c in ic], perm[:k], counts


def hllc_star(v_left, p_left, rho_left, c_left, v_right, p_right, rho_right, c_right, e, x_over_t):
    vl = np.ascontiguousarray(v_left, dtype=np.float64)
    vr = np.ascontiguousarray(v_right, dtype=np.float64)
    ev = np.ascontiguousarray(e, dtype=np.float64)
    vs = np.zeros(2)
    ps = ctypes.c_double()
    lib().orc_hllc_star(_dp(vl), p_left, rho_left, c_left, _dp(vr), p_right, rho_right, c_right,
                        _dp(ev), x_over_t, _dp(vs), ctypes.byref(ps))
    return vs, ps.value


# compiled reference pieces (oracle/_ref) --------------------------------------
_ref = None


def ref_available() -> bool:
    """Whether oracle/_ref is built; checks the file only, so test collection never
    maps the compiled reference into a process that does not use it."""
    return os.path.exists(REF_PATH)


def ref_lib():
    """The reference's own compute_kernel/status_equation/function_module, compiled
    from /root/reference by oracle/build_ref.sh; None where it is not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        _ref = ctypes.CDLL(REF_PATH)
    return _ref


def ref_constant(name: str) -> float:
    L = ref_lib()
    return ctypes.c_double.in_dll(L, "_QMinformation_moduleEC" + name.lower()).value


def ref_int_constant(name: str) -> int:
    L = ref_lib()
    return ctypes.c_int32.in_dll(L, "_QMinformation_moduleEC" + name.lower()).value
