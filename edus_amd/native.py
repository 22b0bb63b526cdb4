"""ctypes bindings of the two native libraries.

libedus_host.so (include/edus_host.h): input.txt -> setup tables, time loop, writers.
libedus_hip.so  (include/edus.h):      the GPU engine behind the C-ABI.

The libraries are built in-tree by ``python -m edus_amd.build``; importing this
module raises if they are missing -- there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

STEPPER_RK4, STEPPER_AB, STEPPER_TAYLOR = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED32, PRECISION_BF16X3 = 0, 1, 2, 3
FIELDS_PER_EVAL = 6

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Problem(C.Structure):
    _fields_ = [
        ("nk", C.c_int32), ("nb", C.c_int32), ("nch", C.c_int32), ("ncc", C.c_int32),
        ("n_nbr", C.c_int32), ("stepper", C.c_int32), ("order", C.c_int32), ("precision", C.c_int32),
        ("nbr", _ip), ("nbr_wb", _dp), ("H", _dp), ("D", _dp), ("U", _dp), ("T", _dp),
        ("rho0", _dp), ("gradH", _dp),
        ("wx2", C.c_double), ("dt", C.c_double), ("j_norm", C.c_double), ("tabs_norm", C.c_double),
        ("nktot", C.c_double), ("taylor_tau_inv", _dp), ("coulomb", C.c_void_p), ("nk_dims", C.c_int32 * 3),
        ("reserved1", C.c_int32),
    ]


class SetupInfo(C.Structure):
    _fields_ = [
        ("nk0", C.c_int32), ("nk1", C.c_int32), ("nk2", C.c_int32), ("nk", C.c_int32),
        ("nb", C.c_int32), ("nch", C.c_int32), ("ncc", C.c_int32), ("nbands", C.c_int32 * 3),
        ("n_nbr", C.c_int32), ("n_shells", C.c_int32), ("stepper", C.c_int32), ("order", C.c_int32),
        ("it_resolution", C.c_int32), ("i_current", C.c_int32), ("i_tabs", C.c_int32), ("n_pumps", C.c_int32),
        ("model", C.c_int32), ("n_diss", C.c_int32), ("itfi", C.c_int64),
        ("dt", C.c_double), ("t_fin", C.c_double), ("t_start", C.c_double),
        ("det_a", C.c_double), ("det_b", C.c_double), ("jacobian", C.c_double),
        ("j_norm", C.c_double), ("tabs_norm", C.c_double), ("nktot", C.c_double), ("wx2", C.c_double),
        ("T1", C.c_double), ("T2", C.c_double), ("Tch", C.c_double), ("shell_residual", C.c_double),
        ("ncut", C.c_int32), ("dyn_dt", C.c_int32), ("eps_step", C.c_double),
        ("print_flags", C.c_int32), ("reserved2", C.c_int32), ("print_start", C.c_double), ("print_end", C.c_double),
        ("print_step", C.c_double),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "nbands" else v
        return d


class CoulombGen(C.Structure):
    """edus_coulomb_gen of include/edus.h."""
    _fields_ = [("nb", C.c_int32), ("ncut", C.c_int32), ("n_sum", C.c_int32), ("n_peak", C.c_int32),
                ("rytova_keldysh", C.c_int32), ("n_g", C.c_int32), ("g", C.POINTER(C.c_int32)),
                ("b1", C.c_double * 3), ("b2", C.c_double * 3), ("eps_area", C.c_double), ("qtf", C.c_double),
                ("r0", C.c_double), ("q_peak", C.c_double), ("disp", C.POINTER(C.c_double)), ("verbose", C.c_int32),
                ("reserved", C.c_int32)]


class Pulse(C.Structure):
    _fields_ = [
        ("gaussian", C.c_int32), ("sin2", C.c_int32), ("pw", C.c_int32), ("envelope", C.c_int32),
        ("wl", C.c_double), ("period", C.c_double), ("intensity", C.c_double), ("A0", C.c_double),
        ("E0", C.c_double), ("t0", C.c_double), ("tf", C.c_double), ("w1", C.c_double), ("w2", C.c_double),
        ("ncycle", C.c_double), ("sigma", C.c_double), ("phase", C.c_double),
        ("pol1", C.c_double * 3), ("pol2", C.c_double * 3),
    ]


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m edus_amd.build`")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


hip = _load("libedus_hip.so")
host = _load("libedus_host.so")

# ---- engine
hip.edus_create.argtypes = [C.POINTER(Problem), C.POINTER(C.c_void_p)]
hip.edus_destroy.argtypes = [C.c_void_p]
for _f in (hip.edus_step_rk4, hip.edus_step_ab, hip.edus_step_taylor):
    _f.argtypes = [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_uint8)]
hip.edus_step_dyn.argtypes = [C.c_void_p, C.c_double, _dp, C.c_uint8, _dp, C.c_double, _dp]
hip.edus_observables.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
hip.edus_coulomb_coefficients.argtypes = [C.POINTER(CoulombGen), _dp, _dp, _dp, _dp, _dp, _dp]
hip.edus_eigen_element.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp]
hip.edus_renorm_energies.argtypes = [C.c_void_p, _dp]
hip.edus_rhs.argtypes = [C.c_void_p, _dp, C.c_uint8, _dp]
hip.edus_get_state.argtypes = [C.c_void_p, _dp]
hip.edus_set_state.argtypes = [C.c_void_p, _dp]
hip.edus_last_step_time.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int32)]
hip.edus_launch_schedule.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
hip.edus_autotune.argtypes = [C.c_void_p, _dp, C.c_uint8, C.c_int32, C.POINTER(C.c_int32)]
hip.edus_wannier_hk.argtypes = [_ip, C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, _dp]
hip.edus_wannier_hk_fp32.argtypes = [_ip, C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, _dp]
hip.edus_last_error.restype = C.c_char_p
hip.edus_eigh_batched.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _dp]
hip.edus_batch_create.argtypes = [C.POINTER(C.POINTER(Problem)), C.c_int32, C.POINTER(C.c_void_p)]
hip.edus_batch_destroy.argtypes = [C.c_void_p]
hip.edus_batch_size.argtypes = [C.c_void_p]
hip.edus_batch_step_rk4.argtypes = [C.c_void_p, C.c_int32, C.POINTER(_dp), C.POINTER(C.POINTER(C.c_uint8))]
hip.edus_batch_last_step_time.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int32)]
hip.edus_batch_member.argtypes = [C.c_void_p, C.c_int32]
hip.edus_batch_member.restype = C.c_void_p
hip.edus_stream_probe.argtypes = [C.c_int64, C.c_int32, _dp, _dp, _dp]
hip.edus_set_field_table.argtypes = [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_uint8)]
hip.edus_step_rk4_graph.argtypes = [C.c_void_p, C.c_int32]

# ---- host
host.edus_setup_from_input.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]
host.edus_setup_free.argtypes = [C.c_void_p]
host.edus_setup_get_info.argtypes = [C.c_void_p, C.POINTER(SetupInfo)]
host.edus_setup_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]
host.edus_setup_array.restype = _dp
host.edus_setup_nbr.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
host.edus_setup_nbr.restype = _ip
host.edus_setup_pulse.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Pulse)]
host.edus_setup_model.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, _dp]
host.edus_setup_eval_tb.argtypes = [C.c_void_p, _dp, C.c_int32, _dp, _dp, _dp]
host.edus_setup_problem.argtypes = [C.c_void_p, C.POINTER(Problem)]
host.edus_setup_fields.argtypes = [C.c_void_p, C.c_int64, C.c_int64, _dp, C.POINTER(C.c_uint8)]
host.edus_setup_time.argtypes = [C.c_void_p, C.c_int64]
host.edus_setup_time.restype = C.c_double
host.edus_run.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, C.c_int32]
host.edus_run.restype = C.c_int64
host.edus_write_synthetic_tb.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_uint64]
host.edus_host_last_error.restype = C.c_char_p


class EdusError(RuntimeError):
    pass


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _check_hip(rc, what):
    if rc != 0:
        raise EdusError(f"{what} failed ({rc}): {hip.edus_last_error().decode()}")


class Setup:
    """Setup tables built by the C++ host from a reference input.txt."""

    def __init__(self, input_path, verify_neighbours=False, fp32_contraction=False):
        """fp32_contraction: Wannier H(k), r(k) from the FP32 matrix-core contraction."""
        h = C.c_void_p()
        flags = (1 if verify_neighbours else 0) | (2 if fp32_contraction else 0)
        rc = host.edus_setup_from_input(os.fsencode(input_path), flags, C.byref(h))
        if rc != 0:
            raise EdusError(host.edus_host_last_error().decode())
        self._h = h
        inf = SetupInfo()
        host.edus_setup_get_info(h, C.byref(inf))
        self.info = inf.as_dict()

    def close(self):
        if self._h:
            host.edus_setup_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, name, complex_=False):
        n = C.c_int64()
        p = host.edus_setup_array(self._h, name.encode(), C.byref(n))
        if not p or n.value == 0:
            return np.zeros(0)
        a = np.ctypeslib.as_array(p, shape=(n.value,)).copy()
        return a.view(np.complex128) if complex_ else a

    def nbr(self):
        n = C.c_int64()
        p = host.edus_setup_nbr(self._h, C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy().reshape(self.info["nk"], -1)

    def pulse(self, index):
        p = Pulse()
        if host.edus_setup_pulse(self._h, index, C.byref(p)) != 0:
            raise EdusError("bad pulse index")
        return {n: (list(getattr(p, n)) if n.startswith("pol") else getattr(p, n)) for n, _ in Pulse._fields_}

    def model(self):
        """(name, params dict or None) of the setup's band model."""
        buf = C.create_string_buffer(256)
        par = np.zeros(8)
        rc = host.edus_setup_model(self._h, buf, 256, _dptr(par))
        if rc < 0:
            raise EdusError(host.edus_host_last_error().decode())
        if rc == 1:
            return buf.value.decode(), None
        keys = ["a", "g", "b", "t2", "phi0", "core", "rchv", "rchc"]
        return buf.value.decode(), dict(zip(keys, par.tolist()))

    def eval_tb(self, kcrys):
        kcrys = np.ascontiguousarray(kcrys, dtype=np.float64).reshape(-1, 3)
        n, nb = len(kcrys), self.info["nb"]
        H = np.zeros((n, nb, nb), dtype=np.complex128)
        U = np.zeros_like(H)
        D = np.zeros((n, nb, nb, 3), dtype=np.complex128)
        rc = host.edus_setup_eval_tb(self._h, _dptr(kcrys), n, _dptr(H.view(np.float64)), _dptr(U.view(np.float64)),
                                     _dptr(D.view(np.float64)))
        if rc != 0:
            raise EdusError(host.edus_host_last_error().decode())
        return H, U, D

    def problem(self):
        p = Problem()
        host.edus_setup_problem(self._h, C.byref(p))
        return p

    def fields_per_step(self):
        return 3 if self.info["stepper"] == STEPPER_RK4 else 1

    def fields(self, it0, n):
        nfe = self.fields_per_step()
        f = np.zeros((n, nfe, FIELDS_PER_EVAL))
        g = np.zeros(n, dtype=np.uint8)
        rc = host.edus_setup_fields(self._h, it0, n, _dptr(f), g.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise EdusError(host.edus_host_last_error().decode())
        return f, g

    def time(self, it):
        return host.edus_setup_time(self._h, it)

    def run(self, outdir, max_steps=-1, verbose=False):
        n = host.edus_run(self._h, os.fsencode(outdir), max_steps, 1 if verbose else 0)
        if n < 0:
            raise EdusError(host.edus_host_last_error().decode())
        return n


def eigh_batched(H):
    """Eigenvalues (ascending) and U (rows conj(v_a), largest component of each
    eigenvector real positive) of a stack of Hermitian matrices [n][nb][nb]."""
    H = np.ascontiguousarray(H, dtype=np.complex128)
    n, nb = H.shape[0], H.shape[1]
    w = np.zeros((n, nb))
    U = np.zeros_like(H)
    _check_hip(hip.edus_eigh_batched(n, nb, _dptr(H.view(np.float64)), _dptr(w), _dptr(U.view(np.float64))), "edus_eigh_batched")
    return w, U


class Batch:
    """Independent RK4 problems advanced together (one launch per stage and kernel
    class); members are Engine views for observables and state access."""

    def __init__(self, problems, precision=None):
        if precision is not None:
            for p in problems:
                p.precision = precision
        arr = (C.POINTER(Problem) * len(problems))(*[C.pointer(p) for p in problems])
        h = C.c_void_p()
        _check_hip(hip.edus_batch_create(arr, len(problems), C.byref(h)), "edus_batch_create")
        self._h = h
        self.members = [Engine._member(hip.edus_batch_member(h, i), p) for i, p in enumerate(problems)]
        self.nk = [p.nk for p in problems]

    def step(self, fields, grad_on):
        """fields[i]: [nsteps][3][6], grad_on[i]: [nsteps], as Engine.step for RK4."""
        n = len(self.members)
        fs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        gs = [np.ascontiguousarray(g, dtype=np.uint8) for g in grad_on]
        nsteps = len(gs[0])
        if any(len(g) != nsteps for g in gs):
            raise ValueError("every member steps the same number of steps")
        fp = (_dp * n)(*[_dptr(f) for f in fs])
        gp = (C.POINTER(C.c_uint8) * n)(*[g.ctypes.data_as(C.POINTER(C.c_uint8)) for g in gs])
        _check_hip(hip.edus_batch_step_rk4(self._h, nsteps, fp, gp), "edus_batch_step_rk4")

    def last_step_time(self):
        ms = np.zeros(1)
        n = C.c_int32()
        _check_hip(hip.edus_batch_last_step_time(self._h, _dptr(ms), C.byref(n)), "edus_batch_last_step_time")
        return float(ms[0]), n.value

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            hip.edus_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One GPU context; the arrays passed to create() are copied to HBM."""

    def __init__(self, problem, stepper=None, precision=None):
        self.stepper = problem.stepper if stepper is None else stepper
        self.nk, self.nb = problem.nk, problem.nb
        if precision is not None:
            problem.precision = precision
        h = C.c_void_p()
        _check_hip(hip.edus_create(C.byref(problem), C.byref(h)), "edus_create")
        self._h = h
        self._owned = True

    @classmethod
    def _member(cls, handle, problem):
        """A batch member's context (owned by the batch)."""
        e = cls.__new__(cls)
        e.stepper, e.nk, e.nb = problem.stepper, problem.nk, problem.nb
        e._h = C.c_void_p(handle)
        e._owned = False
        return e

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                hip.edus_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, fields, grad_on):
        fields = np.ascontiguousarray(fields, dtype=np.float64)
        grad_on = np.ascontiguousarray(grad_on, dtype=np.uint8)
        n = len(grad_on)
        fn = {STEPPER_RK4: hip.edus_step_rk4, STEPPER_AB: hip.edus_step_ab, STEPPER_TAYLOR: hip.edus_step_taylor}[self.stepper]
        _check_hip(fn(self._h, n, _dptr(fields), grad_on.ctypes.data_as(C.POINTER(C.c_uint8))), "edus_step")

    def set_field_table(self, fields, grad_on):
        """Uploads the RK4 field schedule of a run (edus_set_field_table)."""
        fields = np.ascontiguousarray(fields, dtype=np.float64)
        grad_on = np.ascontiguousarray(grad_on, dtype=np.uint8)
        _check_hip(hip.edus_set_field_table(self._h, len(grad_on), _dptr(fields), grad_on.ctypes.data_as(C.POINTER(C.c_uint8))),
                   "edus_set_field_table")

    def step_graph(self, nsteps):
        """nsteps RK4 steps from the field table through the captured step graph."""
        _check_hip(hip.edus_step_rk4_graph(self._h, int(nsteps)), "edus_step_rk4_graph")

    def step_dyn(self, dt, fields, grad_on, tau_inv=None, cond=0.0):
        """One adaptive-dt step; returns the step-size measure max_dP."""
        fields = np.ascontiguousarray(fields, dtype=np.float64)
        ti = None if tau_inv is None else np.ascontiguousarray(tau_inv, dtype=np.float64)
        out = np.zeros(1)
        _check_hip(hip.edus_step_dyn(self._h, float(dt), _dptr(fields), 1 if grad_on else 0,
                                     None if ti is None else _dptr(ti), float(cond), _dptr(out)), "edus_step_dyn")
        return float(out[0])

    def eigen_element(self, ic, jc):
        """(conj(U) P U^T)[ic][jc] per k."""
        out = np.zeros(self.nk, dtype=np.complex128)
        _check_hip(hip.edus_eigen_element(self._h, ic, jc, _dptr(out.view(np.float64))), "edus_eigen_element")
        return out

    def renorm_energies(self):
        out = np.zeros((self.nk, self.nb))
        _check_hip(hip.edus_renorm_energies(self._h, _dptr(out)), "edus_renorm_energies")
        return out

    def observables(self):
        pop = np.zeros(self.nb)
        j1, j2, ta, pm = np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(2)
        _check_hip(hip.edus_observables(self._h, _dptr(pop), _dptr(j1), _dptr(j2), _dptr(ta), _dptr(pm)), "edus_observables")
        return {"pop": pop, "j1": j1, "j2": j2, "tabs": ta, "pminmax": pm}

    def rhs(self, field6, grad_on=True):
        field6 = np.ascontiguousarray(field6, dtype=np.float64)
        pv = np.zeros((self.nk, self.nb, self.nb), dtype=np.complex128)
        _check_hip(hip.edus_rhs(self._h, _dptr(field6), 1 if grad_on else 0, _dptr(pv.view(np.float64))), "edus_rhs")
        return pv

    def get_state(self):
        rho = np.zeros((self.nk, self.nb, self.nb), dtype=np.complex128)
        _check_hip(hip.edus_get_state(self._h, _dptr(rho.view(np.float64))), "edus_get_state")
        return rho

    def set_state(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.complex128)
        _check_hip(hip.edus_set_state(self._h, _dptr(rho.view(np.float64))), "edus_set_state")

    def last_step_time(self):
        ms = C.c_double()
        n = C.c_int32()
        hip.edus_last_step_time(self._h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def launch_schedule(self):
        """edus_launch_schedule: {tile, strip_width, stencil_pairs, pair_own_rows, small_threads}."""
        out = (C.c_int32 * 5)()
        _check_hip(hip.edus_launch_schedule(self._h, out), "edus_launch_schedule")
        return dict(zip(("tile", "strip_width", "stencil_pairs", "pair_own_rows", "small_threads"), list(out)))

    def autotune(self, field6, grad_on=True, reps=5):
        """Launch schedule picked by timing (edus_autotune); returns the kept strip
        width (5..32 bands) or workgroup size (1..4 bands)."""
        f = np.ascontiguousarray(field6, dtype=np.float64)
        ch = C.c_int32()
        _check_hip(hip.edus_autotune(self._h, _dptr(f), 1 if grad_on else 0, reps, C.byref(ch)), "edus_autotune")
        return ch.value


def wannier_hk(nk3, R, HR, rR, nw, fp32=False, bf16x3=False):
    """GPU Wannier contraction: returns H [nk][nw][nw], D [nk][nw][nw][3] (complex).
    fp32: the FP32 matrix-core form (edus_wannier_hk_fp32); bf16x3: the BF16x3
    split-product form (edus_wannier_hk_bf16x3)."""
    nk3 = np.ascontiguousarray(nk3, dtype=np.int32)
    R = np.ascontiguousarray(R, dtype=np.int32)
    HR = np.ascontiguousarray(HR, dtype=np.complex128)
    rR = np.ascontiguousarray(rR, dtype=np.complex128)
    nk = int(np.prod(nk3))
    H = np.zeros((nk, nw, nw), dtype=np.complex128)
    D = np.zeros((nk, nw, nw, 3), dtype=np.complex128)
    if bf16x3:
        fn = hip.edus_wannier_hk_bf16x3
        fn.argtypes = [_ip, C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, _dp]
    else:
        fn = hip.edus_wannier_hk_fp32 if fp32 else hip.edus_wannier_hk
    rc = fn(nk3.ctypes.data_as(_ip), len(R), nw, R.ctypes.data_as(_ip), _dptr(HR.view(np.float64)),
            _dptr(rR.view(np.float64)), _dptr(H.view(np.float64)), _dptr(D.view(np.float64)))
    _check_hip(rc, fn.__name__)
    return H, D


def stream_probe(nbytes=1 << 30, reps=10):
    """Copy, triad and read rates (GB/s) of this library's streaming kernels on the
    current device (edus_stream_probe)."""
    cp, tr, rd = C.c_double(), C.c_double(), C.c_double()
    _check_hip(hip.edus_stream_probe(int(nbytes), int(reps), C.byref(cp), C.byref(tr), C.byref(rd)), "edus_stream_probe")
    return {"copy": cp.value, "triad": tr.value, "read": rd.value}


def write_synthetic_tb(path, nw, rmax, a_ang=2.5, sigma_h_ev=1.5, sigma_r_ang=0.3, decay_ang=2.0, seed=1001):
    rc = host.edus_write_synthetic_tb(os.fsencode(path), nw, rmax, a_ang, sigma_h_ev, sigma_r_ang, decay_ang, seed)
    if rc != 0:
        raise EdusError(host.edus_host_last_error().decode())


def coulomb_coefficients(p, verbose=False):
    """Coulomb mean-field coefficients on the GPU (edus_coulomb_coefficients); p: dict
    with nb, ncut, n_sum, n_peak, rytova, g [n_g][2], b1, b2, eps_area, qtf, r0,
    q_peak, disp [nb][3].  Returns A..D [ncut^2][nb^2], screen [nb^2], vh [nb][nb]."""
    g = np.ascontiguousarray(p["g"], dtype=np.int32).reshape(-1, 2)
    d = np.ascontiguousarray(p["disp"], dtype=np.float64).reshape(-1, 3)
    o = CoulombGen()
    o.nb, o.ncut, o.n_sum, o.n_peak, o.rytova_keldysh, o.n_g = p["nb"], p["ncut"], p["n_sum"], p["n_peak"], p["rytova"], len(g)
    o.g = g.ctypes.data_as(C.POINTER(C.c_int32))
    for i in range(3):
        o.b1[i], o.b2[i] = p["b1"][i], p["b2"][i]
    o.eps_area, o.qtf, o.r0, o.q_peak = p["eps_area"], p["qtf"], p["r0"], p["q_peak"]
    o.disp = d.ctypes.data_as(C.POINTER(C.c_double))
    o.verbose = 1 if verbose else 0
    nb2, nmm = p["nb"] ** 2, p["ncut"] ** 2
    out = [np.zeros((nmm, nb2), dtype=np.complex128) for _ in range(4)]
    sc = np.zeros(nb2, dtype=np.complex128)
    vh = np.zeros((p["nb"], p["nb"]), dtype=np.complex128)
    _check_hip(hip.edus_coulomb_coefficients(C.byref(o), *[_dptr(x.view(np.float64)) for x in out], _dptr(sc.view(np.float64)),
                                             _dptr(vh.view(np.float64))), "edus_coulomb_coefficients")
    return {"A": out[0], "B": out[1], "C": out[2], "D": out[3], "screen": sc, "vh": vh}
