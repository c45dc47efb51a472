"""MI355X (gfx950) tree-tensor contraction engine -- Python binding.

Thin ctypes layer over libqutree_amd.so (C ABI in include/qutree_amd.h).
The shared library must have been built (``python -m qutree_amd.build`` or
``__graft_entry__.build()``); there is no CPU fallback: every entry point
raises if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqutree_amd.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64


class QtgError(RuntimeError):
    pass


def lib():
    """Load libqutree_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QtgError(f"{LIB_PATH} is missing: build it with python -m qutree_amd.build")
    L = C.CDLL(LIB_PATH)
    L.qtg_last_error.restype = C.c_char_p
    L.qtg_init.argtypes = [C.c_int]
    L.qtg_device_alloc.argtypes = [C.POINTER(_vp), C.c_size_t]
    L.qtg_device_free.argtypes = [_vp]
    L.qtg_memcpy_h2d.argtypes = [_vp, _vp, C.c_size_t, _vp]
    L.qtg_memcpy_d2h.argtypes = [_vp, _vp, C.c_size_t, _vp]
    L.qtg_memset.argtypes = [_vp, C.c_int, C.c_size_t, _vp]
    L.qtg_synchronize.argtypes = [_vp]
    L.qtg_timer_start.argtypes = [_vp]
    L.qtg_timer_stop.argtypes = [_vp, C.POINTER(C.c_float)]
    for name in ("qtg_matvec_z", "qtg_tmatvec_z", "qtg_ctmatvec_z"):
        getattr(L, name).argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, C.c_int, _vp]
    L.qtg_rhomat_z.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _vp]
    for name in ("qtg_mode_product_z", "qtg_mode_product_c"):
        getattr(L, name).argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, C.c_int, C.c_int, _vp]
    for name in ("qtg_hole_contraction_z", "qtg_hole_contraction_c"):
        getattr(L, name).argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, C.c_int, _vp]
    L.qtg_mode_product_plan_create.argtypes = [C.POINTER(_vp), _vp, _i64, C.c_int]
    L.qtg_hole_contraction_plan_create.argtypes = [C.POINTER(_vp), _vp, _i64, C.c_int]
    L.qtg_plan_execute.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.qtg_plan_destroy.argtypes = [_vp]
    L.qtg_plan_counts.argtypes = [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.qtg_engine_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(_vp)]
    L.qtg_engine_destroy.argtypes = [_vp]
    L.qtg_engine_set_hamiltonian.argtypes = [_vp, _i64, _dp, _lp, _lp, _lp, _lp, _dp]
    L.qtg_engine_psi_size.argtypes = [_vp]
    L.qtg_engine_psi_size.restype = _i64
    L.qtg_engine_num_nodes.argtypes = [_vp]
    L.qtg_engine_num_nodes.restype = _i64
    L.qtg_engine_node_info.argtypes = [_vp, _i64, _lp, _i64]
    L.qtg_engine_fill_random.argtypes = [_vp, C.c_uint64, C.c_int, _dp]
    L.qtg_engine_set_psi.argtypes = [_vp, _dp, _vp]
    L.qtg_engine_set_psi_device.argtypes = [_vp, _vp, _vp]
    L.qtg_engine_build.argtypes = [_vp, _vp]
    L.qtg_engine_derivative.argtypes = [_vp, C.c_double, C.c_double, _vp]
    L.qtg_engine_layer_derivative_all.argtypes = [_vp, C.c_double, C.c_double, _vp]
    L.qtg_engine_get_dpsi.argtypes = [_vp, _dp, C.c_int, _vp]
    L.qtg_engine_energy.argtypes = [_vp, C.POINTER(C.c_double), _vp]
    L.qtg_engine_get_matrix.argtypes = [_vp, C.c_int, _i64, _i64, _dp, _vp]
    L.qtg_engine_counts.argtypes = [_vp, C.POINTER(C.c_double), _lp]
    L.qtg_engine_plan_flops.argtypes = [_vp, C.POINTER(C.c_double)]
    L.qtg_engine_forward.argtypes = [_vp, C.POINTER(C.c_double), _vp]
    L.qtg_engine_backward.argtypes = [_vp, _vp, _vp]
    L.qtg_engine_canonicalize.argtypes = [_vp, _vp, _vp]
    L.qtg_engine_gradient_descent.argtypes = [_vp, _i64, C.c_double, _vp, _vp, _vp]
    L.qtg_rmatvec_d.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, C.c_int, _vp]
    L.qtg_trmatvec_d.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, C.c_int, _vp]
    L.qtg_engine_gradient.argtypes = [_vp, _dp, C.POINTER(C.c_double), _vp]
    L.qtg_ttno_create.argtypes = [_vp, _lp, _dp, C.POINTER(_vp)]
    L.qtg_ttno_destroy.argtypes = [_vp]
    L.qtg_ttno_apply.argtypes = [_vp, _dp, _dp, _dp, _dp, _vp]
    L.qtg_engine_set_option.argtypes = [_vp, C.c_char_p, C.c_double]
    L.qtg_engine_set_timing.argtypes = [_vp, C.c_int]
    L.qtg_engine_synchronize.argtypes = [_vp]
    L.qtg_cmf_create.argtypes = [_vp, C.POINTER(_vp)]
    L.qtg_cmf_destroy.argtypes = [_vp]
    L.qtg_cmf_set_option.argtypes = [_vp, C.c_char_p, C.c_double]
    L.qtg_scf_create.argtypes = [_vp, C.c_int64, C.POINTER(_vp)]
    L.qtg_scf_destroy.argtypes = [_vp]
    L.qtg_scf_sweep_length.argtypes = [_vp]
    L.qtg_scf_sweep_length.restype = C.c_int64
    L.qtg_scf_run.argtypes = [_vp, _dp, C.c_int64, C.c_int64, C.c_double, _dp, C.c_int64]
    L.qtg_scf_set_option.argtypes = [_vp, C.c_char_p, C.c_double]
    L.qtg_scf_info.argtypes = [_vp, _lp]
    L.qtg_apply_fit_create.argtypes = [_vp, C.POINTER(_vp)]
    L.qtg_apply_fit_destroy.argtypes = [_vp]
    L.qtg_apply_fit_run.argtypes = [_vp, _dp, C.c_uint64, C.c_int64, _dp, C.c_int64, C.POINTER(C.c_int64)]
    L.qtg_apply_fit_set_option.argtypes = [_vp, C.c_char_p, C.c_double]
    L.qtg_apply_fit_info.argtypes = [_vp, _lp]
    L.qtg_cmf_integrate.argtypes = [_vp, _dp] + [C.c_double] * 8 + [_i64, _i64, _dp, _dp, _dp, _lp]
    L.qtg_cmf_integrate.restype = _i64
    L.qtg_cmf_eigenstates.argtypes = ([_vp, _dp, C.c_double, C.c_double, _i64, C.c_double, C.c_double, _i64, _dp,
                                       _dp, _dp, _lp, C.POINTER(C.c_double)])
    L.qtg_cmf_eigenstates.restype = _i64
    L.qtg_host_psi_size.argtypes = [C.c_char_p]
    L.qtg_host_psi_size.restype = _i64
    L.qtg_tensor_tree_write.argtypes = [C.c_char_p, _dp, C.c_void_p, _i64]
    L.qtg_tensor_tree_write.restype = _i64
    L.qtg_tensor_tree_read.argtypes = [C.c_char_p, C.c_void_p, _i64, _dp]
    L.qtg_tensor_tree_read.restype = _i64
    L.qtg_matrix_tree_write.argtypes = [C.c_int, _i64, _lp, _dp, C.c_void_p, _i64]
    L.qtg_matrix_tree_write.restype = _i64
    L.qtg_matrix_tree_read.argtypes = [C.c_int, C.c_void_p, _i64, _i64, _lp, _dp, _i64, C.POINTER(_i64)]
    L.qtg_matrix_tree_read.restype = _i64
    L.qtg_engine_read_timing.argtypes = [_vp, _dp, _lp, _dp, _dp]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise QtgError(f"qtg status {rc}: {lib().qtg_last_error().decode()}")


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        check(lib().qtg_init(device))
        _initialised = True


class DeviceBuffer:
    """Device allocation owned by the library's allocator."""

    def __init__(self, nbytes: int):
        init()
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().qtg_device_alloc(C.byref(p), max(self.nbytes, 16)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().qtg_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, None))

    def download(self, dtype, count: Optional[int] = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(n, dtype)
        check(lib().qtg_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, None))
        return out

    def zero(self) -> None:
        check(lib().qtg_memset(self.ptr, 0, self.nbytes, None))
        synchronize()

    def free(self) -> None:
        if self.ptr:
            lib().qtg_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    check(lib().qtg_synchronize(None))


class Timer:
    """HIP events recorded on the library's default stream."""

    def start(self) -> None:
        check(lib().qtg_timer_start(None))

    def stop(self) -> float:
        ms = C.c_float()
        check(lib().qtg_timer_stop(None, C.byref(ms)))
        return float(ms.value)


MP_DESC = np.dtype([("offC", "i8"), ("offM", "i8"), ("offB", "i8"), ("before", "i8"), ("nB", "i8"),
                    ("nC", "i8"), ("after", "i8"), ("op", "i4"), ("add", "i4")])
HC_DESC = np.dtype([("offS", "i8"), ("offA", "i8"), ("offB", "i8"), ("before", "i8"), ("nA", "i8"),
                    ("nB", "i8"), ("after", "i8"), ("accumulate", "i4"), ("pad", "i4")])


class Plan:
    """Grouped node-level contraction plan (descriptors resident on the device)."""

    def __init__(self, kind: str, descs: np.ndarray, complex64: bool = False):
        init()
        descs = np.ascontiguousarray(descs)
        p = _vp()
        if kind == "mode_product":
            assert descs.dtype == MP_DESC
            check(lib().qtg_mode_product_plan_create(C.byref(p), descs.ctypes.data, len(descs), int(complex64)))
        else:
            assert descs.dtype == HC_DESC
            check(lib().qtg_hole_contraction_plan_create(C.byref(p), descs.ctypes.data, len(descs),
                                                          int(complex64)))
        self.ptr = p.value

    def execute(self, out: DeviceBuffer, first: DeviceBuffer, second: DeviceBuffer) -> None:
        check(lib().qtg_plan_execute(self.ptr, out.ptr, first.ptr, second.ptr, None))

    def counts(self):
        f, b = C.c_double(), C.c_double()
        check(lib().qtg_plan_counts(self.ptr, C.byref(f), C.byref(b)))
        return f.value, b.value

    def __del__(self):
        try:
            if self.ptr:
                lib().qtg_plan_destroy(self.ptr)
        except Exception:
            pass


def _ci(z: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(z, dtype=np.complex128).view(np.float64)


class Engine:
    """Device-resident tree + SOP Hamiltonian (qtg_engine_*)."""

    def __init__(self, tree_text: str, H=None, precision: int = 0, compress: bool = True, graph: bool = True,
                 path: int = 1, **options):
        """`options`: further qtg_engine_set_option keys (split_waves,
        split_min_k, rho_eps), applied before the Hamiltonian is planned."""
        init()
        p = _vp()
        check(lib().qtg_engine_create(tree_text.encode(), precision, C.byref(p)))
        self.ptr = p.value
        check(lib().qtg_engine_set_option(self.ptr, b"compress", float(compress)))
        check(lib().qtg_engine_set_option(self.ptr, b"graph", float(graph)))
        check(lib().qtg_engine_set_option(self.ptr, b"path", float(path)))
        for k, v in options.items():
            check(lib().qtg_engine_set_option(self.ptr, k.encode(), float(v)))
        self.precision = precision
        self.psi_size = lib().qtg_engine_psi_size(self.ptr)
        self.nnodes = lib().qtg_engine_num_nodes(self.ptr)
        self.nodes = []
        for a in range(self.nnodes):
            buf = np.zeros(64, np.int64)
            check(lib().qtg_engine_node_info(self.ptr, a, buf, 64))
            order, nch = int(buf[4]), int(buf[5])
            self.nodes.append(dict(parent=int(buf[0]), childIdx=int(buf[1]), bottom=bool(buf[2]),
                                   depth=int(buf[3]), mode=int(buf[6]), off=int(buf[7]),
                                   dims=[int(x) for x in buf[8:8 + order]],
                                   children=[int(x) for x in buf[8 + order:8 + order + nch]]))
        if H is not None:
            self.set_hamiltonian(H)

    def set_hamiltonian(self, H) -> None:
        c, toff, fm, fk, fp, pool = H.arrays()
        check(lib().qtg_engine_set_hamiltonian(self.ptr, H.nterms, c, toff, fm, fk, fp, pool))

    def fill_random(self, seed: int, delta_lowest: bool = True) -> np.ndarray:
        out = np.zeros(2 * self.psi_size)
        check(lib().qtg_engine_fill_random(self.ptr, seed, int(delta_lowest), out))
        return out.view(np.complex128).copy()

    def set_psi(self, psi: np.ndarray) -> None:
        check(lib().qtg_engine_set_psi(self.ptr, _ci(psi), None))

    def build(self) -> None:
        check(lib().qtg_engine_build(self.ptr, None))

    def derivative(self, phase: complex = 1.0, sync: bool = True) -> None:
        check(lib().qtg_engine_derivative(self.ptr, complex(phase).real, complex(phase).imag, None))
        if sync:
            self.synchronize()

    def synchronize(self) -> None:
        check(lib().qtg_engine_synchronize(self.ptr))

    def layer_derivative_all(self, phase: complex = 1.0) -> None:
        check(lib().qtg_engine_layer_derivative_all(self.ptr, complex(phase).real, complex(phase).imag, None))

    def get(self, which: int = 0) -> np.ndarray:
        out = np.zeros(2 * self.psi_size)
        check(lib().qtg_engine_get_dpsi(self.ptr, out, which, None))
        return out.view(np.complex128).copy()

    def energy(self) -> float:
        e = C.c_double()
        check(lib().qtg_engine_energy(self.ptr, C.byref(e), None))
        return e.value

    def matrix(self, which: int, term: int, node: int) -> np.ndarray:
        n = self.nodes[node]["dims"][-1]
        out = np.zeros(2 * n * n)
        check(lib().qtg_engine_get_matrix(self.ptr, which, term, node, out, None))
        return out.view(np.complex128).reshape(n, n, order="F").copy()

    def counts(self):
        f = C.c_double()
        items = np.zeros(3, np.int64)
        check(lib().qtg_engine_counts(self.ptr, C.byref(f), items))
        return f.value, items

    def gradient(self):
        """(dE/dA* at every node, E) for the state set by set_psi (orthonormal
        non-top nodes)."""
        out = np.zeros(2 * self.psi_size)
        e = C.c_double()
        check(lib().qtg_engine_gradient(self.ptr, out, C.byref(e), None))
        return out.view(np.complex128).copy(), e.value

    def forward(self) -> float:
        """The sweep at phase 1 on the current state; returns E (the
        matrices stay on the device for backward())."""
        e = C.c_double()
        check(lib().qtg_engine_forward(self.ptr, C.byref(e), None))
        return e.value

    def backward(self) -> np.ndarray:
        """dE/dA* at every node (after forward(), or with option
        grad_recompute=1 from the current state)."""
        out = np.zeros(2 * self.psi_size)
        check(lib().qtg_engine_backward(self.ptr, out.ctypes.data, None))
        return out.view(np.complex128).copy()

    def canonicalize(self) -> np.ndarray:
        """QR gauge of the current state on the device; returns it."""
        out = np.zeros(2 * self.psi_size)
        check(lib().qtg_engine_canonicalize(self.ptr, out.ctypes.data, None))
        return out.view(np.complex128).copy()

    def gradient_descent(self, psi, steps: int, eta: float):
        """Device gradient descent from psi: (final state, energies before
        every step and after the last)."""
        self.set_psi(psi)
        en = np.zeros(steps + 1)
        out = np.zeros(2 * self.psi_size)
        check(lib().qtg_engine_gradient_descent(self.ptr, steps, eta, en.ctypes.data, out.ctypes.data, None))
        return out.view(np.complex128).copy(), en

    def executed_flops(self) -> float:
        """FLOPs one sweep of this engine's plan executes."""
        f = C.c_double()
        check(lib().qtg_engine_plan_flops(self.ptr, C.byref(f)))
        return f.value

    KERNELS = ("sandwich_kernel", "heev_reginv_kernel", "apply_kernel", "finish_kernel", "to_f32",
               "combine_kernel", "gm_grouped", "gc_grouped", "solve_kernel", "hole_fused_kernel")

    def set_timing(self, on: bool) -> None:
        check(lib().qtg_engine_set_timing(self.ptr, int(on)))

    def read_timing(self):
        ms = np.zeros(len(self.KERNELS))
        n = np.zeros(len(self.KERNELS), np.int64)
        fl = np.zeros(len(self.KERNELS))
        by = np.zeros(len(self.KERNELS))
        check(lib().qtg_engine_read_timing(self.ptr, ms, n, fl, by))
        return {k: dict(ms=float(ms[i]), launches=int(n[i]), flops=float(fl[i]), bytes=float(by[i]))
                for i, k in enumerate(self.KERNELS)}

    def set_option(self, key: str, value: float) -> None:
        check(lib().qtg_engine_set_option(self.ptr, key.encode(), value))

    def node_tensor(self, psi: np.ndarray, a: int) -> np.ndarray:
        n = self.nodes[a]
        size = int(np.prod(n["dims"]))
        return psi[n["off"]:n["off"] + size].reshape(n["dims"], order="F")

    def __del__(self):
        try:
            if self.ptr:
                lib().qtg_engine_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class CMF:
    """Device CMF integrator bound to a complex128 Engine (qtg_cmf_*)."""

    def __init__(self, engine: Engine):
        p = _vp()
        check(lib().qtg_cmf_create(engine.ptr, C.byref(p)))
        self.ptr = p.value
        self.engine = engine

    def set_option(self, key: str, value) -> None:
        """qtg_cmf_set_option (\"graph\": 1 replays ticks from HIP graphs)."""
        check(lib().qtg_cmf_set_option(self.ptr, key.encode(), float(value)))

    def integrate(self, psi, t0, tend, dt, out, acc_cmf=1e-4, acc_bs=1e-5, phase=1.0, max_steps=10 ** 9,
                  max_out=100000):
        p = _ci(psi).copy()
        times = np.zeros(max_out)
        en = np.zeros(max_out)
        ac = np.zeros(2 * max_out)
        stats = np.zeros(5, np.int64)
        n = lib().qtg_cmf_integrate(self.ptr, p, t0, tend, dt, out, acc_cmf, acc_bs, complex(phase).real,
                                    complex(phase).imag, max_steps, max_out, times, en, ac, stats)
        if n < 0:
            check(-n)
        n = min(n, max_out)
        return dict(psi=p.view(np.complex128).copy(), times=times[:n], energies=en[:n],
                    autocorr=ac.view(np.complex128)[:n].copy(), steps=int(stats[0]), rejected=int(stats[1]),
                    builds=int(stats[2]), derivs=int(stats[3]), ticks=int(stats[4]))

    def eigenstates(self, psi, dt, out, iterations, acc_cmf=1e-6, acc_bs=1e-7, max_steps=10 ** 9):
        """Imaginary-time eigenstates (Eigenstates.cpp:60-103) from psi:
        energies (iterations + 1, nstates), propagator energies and converged
        overlaps (iterations, nstates), the final state and the counters."""
        ns = self.engine.nodes[-1]["dims"][-1]
        p = _ci(psi).copy()
        en = np.zeros((iterations + 1) * ns)
        pe = np.zeros(max(iterations, 1) * ns)
        cv = np.zeros(max(iterations, 1) * ns)
        stats = np.zeros(4, np.int64)
        dte = C.c_double()
        n = lib().qtg_cmf_eigenstates(self.ptr, p, dt, out, iterations, acc_cmf, acc_bs, max_steps, en, pe, cv,
                                      stats, C.byref(dte))
        if n < 0:
            check(-n)
        return dict(psi=p.view(np.complex128).copy(), energies=en.reshape(iterations + 1, ns),
                    propagator=pe[:iterations * ns].reshape(iterations, ns),
                    converged=cv[:iterations * ns].reshape(iterations, ns), steps=int(stats[0]),
                    rejected=int(stats[1]), builds=int(stats[2]), derivs=int(stats[3]), dt_end=dte.value)

    def __del__(self):
        try:
            if self.ptr:
                lib().qtg_cmf_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class SCF:
    """Device Krylov-sweep SCF (qtg_scf_*, SCF.cpp:270-348) bound to a
    complex128 Engine with its Hamiltonian set."""

    def __init__(self, engine: Engine, n_krylov: int = 5):
        p = _vp()
        check(lib().qtg_scf_create(engine.ptr, n_krylov, C.byref(p)))
        self.ptr = p.value
        self.engine = engine
        self.sweep_length = int(lib().qtg_scf_sweep_length(self.ptr))

    def set_option(self, key: str, value) -> None:
        check(lib().qtg_scf_set_option(self.ptr, key.encode(), float(value)))

    def info(self):
        v = np.zeros(3, np.int64)
        check(lib().qtg_scf_info(self.ptr, v))
        return dict(sweep_length=int(v[0]), launches=int(v[1]), device_bytes=int(v[2]))

    def run(self, psi, n_iter, n_itp=0, beta=1.0):
        """n_iter iterations (the first n_itp imaginary-time updates with
        step beta): the final state and the lowest Ritz value of every
        microstep, shape (n_iter, sweep length)."""
        p = _ci(psi).copy()
        e = np.zeros(max(n_iter * self.sweep_length, 1))
        check(lib().qtg_scf_run(self.ptr, p, n_iter, n_itp, beta, e, len(e)))
        return dict(psi=p.view(np.complex128).copy(), energies=e[:n_iter * self.sweep_length].reshape(
            n_iter, self.sweep_length))

    def __del__(self):
        try:
            if self.ptr:
                lib().qtg_scf_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class ApplyFit:
    """H|Psi> fitted into a tree of Psi's shape on the device
    (qtg_apply_fit_*, ApplyFirstOrder.cpp:43-147); the operator is the
    engine's Hamiltonian with its coefficients not applied, as in the
    reference."""

    def __init__(self, engine: Engine):
        p = _vp()
        check(lib().qtg_apply_fit_create(engine.ptr, C.byref(p)))
        self.ptr = p.value
        self.engine = engine

    def set_option(self, key: str, value) -> None:
        check(lib().qtg_apply_fit_set_option(self.ptr, key.encode(), float(value)))

    def info(self):
        v = np.zeros(3, np.int64)
        check(lib().qtg_apply_fit_info(self.ptr, v))
        return dict(launches_first=int(v[0]), launches=int(v[1]), device_bytes=int(v[2]))

    def run(self, psi, seed, max_iter=100):
        """The fit of H|psi> and the fidelity after every iteration."""
        p = _ci(psi).copy()
        f = np.zeros(max_iter)
        n = C.c_int64()
        check(lib().qtg_apply_fit_run(self.ptr, p, seed, max_iter, f, max_iter, C.byref(n)))
        return dict(psi=p.view(np.complex128).copy(), fidelities=f[:n.value].copy())

    def __del__(self):
        try:
            if self.ptr:
                lib().qtg_apply_fit_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class TTNO:
    """Tree tensor network operator on an engine's tree (qtg_ttno_*).
    bond: D per node address (1 at the top); B: the operator tensors
    concatenated in node-address order (complex128)."""

    def __init__(self, engine: Engine, bond, B):
        self.engine = engine  # keeps the engine alive
        p = _vp()
        check(lib().qtg_ttno_create(engine.ptr, np.ascontiguousarray(bond, np.int64), _ci(B), C.byref(p)))
        self.ptr = p.value

    def apply(self, bra: np.ndarray, ket: np.ndarray):
        """(TTNO applied at every node to ket's node tensors, <bra|H|ket>)."""
        out = np.zeros(2 * self.engine.psi_size)
        e = np.zeros(2)
        check(lib().qtg_ttno_apply(self.ptr, _ci(bra), _ci(ket), out, e, None))
        return out.view(np.complex128).copy(), complex(e[0], e[1])

    def __del__(self):
        try:
            if self.ptr:
                lib().qtg_ttno_destroy(self.ptr)
        except Exception:
            pass


# ------------------------------------------------------------------ formats
def tensor_tree_bytes(tree_text: str, psi: np.ndarray) -> bytes:
    """TensorTree<cd>::write (TensorTree_Implementation.h:79-92) of psi
    (engine layout) on the tree in tree_text."""
    p = _ci(psi)
    n = lib().qtg_tensor_tree_write(tree_text.encode(), p, None, 0)
    if n < 0:
        check(-n)
    buf = C.create_string_buffer(int(n))
    m = lib().qtg_tensor_tree_write(tree_text.encode(), p, buf, n)
    if m < 0:
        check(-m)
    return buf.raw[:m]


def tensor_tree_from_bytes(tree_text: str, data: bytes, offset: int = 0):
    """TensorTree<cd>::read of the tree starting at data[offset:]; returns
    (psi, bytes consumed)."""
    size = lib().qtg_host_psi_size(tree_text.encode())
    if size < 0:
        check(-size)
    psi = np.zeros(2 * size)
    view = (C.c_char * (len(data) - offset)).from_buffer_copy(data, offset)
    n = lib().qtg_tensor_tree_read(tree_text.encode(), view, len(data) - offset, psi)
    if n < 0:
        check(-n)
    return psi.view(np.complex128).copy(), int(n)


def matrix_tree_bytes(mats, sparse: bool = False) -> bytes:
    """MatrixTree ("MTre") or SparseMatrixTree ("SMTr") write of a list of
    complex matrices, in node order."""
    mats = [np.asarray(m, np.complex128) for m in mats]
    dims = np.array([d for m in mats for d in m.shape], np.int64)
    flat = (np.concatenate([m.ravel(order="F") for m in mats]) if mats else np.zeros(0, np.complex128))
    fv = np.ascontiguousarray(flat).view(np.float64)
    n = lib().qtg_matrix_tree_write(int(sparse), len(mats), dims, fv, None, 0)
    if n < 0:
        check(-n)
    buf = C.create_string_buffer(int(n))
    m = lib().qtg_matrix_tree_write(int(sparse), len(mats), dims, fv, buf, n)
    if m < 0:
        check(-m)
    return buf.raw[:m]


def matrix_tree_from_bytes(data: bytes, sparse: bool = False):
    """Reads one MatrixTree / SparseMatrixTree; returns (matrices, bytes
    consumed)."""
    # a matrix record takes at least 24 bytes, a coefficient 16
    max_count, max_coeffs = len(data) // 24 + 1, len(data) // 16 + 1
    dims = np.zeros(2 * max_count, np.int64)
    coef = np.zeros(2 * max_coeffs)
    cnt = _i64()
    view = (C.c_char * len(data)).from_buffer_copy(data)
    n = lib().qtg_matrix_tree_read(int(sparse), view, len(data), max_count, dims, coef, max_coeffs, C.byref(cnt))
    if n < 0:
        check(-n)
    z = coef.view(np.complex128)
    out, off = [], 0
    for k in range(cnt.value):
        d1, d2 = int(dims[2 * k]), int(dims[2 * k + 1])
        out.append(z[off:off + d1 * d2].reshape(d1, d2, order="F").copy())
        off += d1 * d2
    return out, int(n)
