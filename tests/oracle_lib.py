"""ctypes binding of the FP64 CPU oracle (oracle/qutree_oracle.cpp).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline load it, and only as the checker."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_ROOT, "oracle", "_build", "libqutree_oracle.so")
_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


_VARIANTS = {"strict": os.path.join(_ROOT, "oracle", "_build", "libqutree_oracle_strict.so"),
             "omp": os.path.join(_ROOT, "oracle", "_build", "libqutree_oracle_omp.so")}
_variant_libs = {}
_use = None


class build_variant:
    """Within this context every oracle call (models created inside it
    included) goes to another build of the same source: "strict" (-O2, no
    fast-math: a second summation order of the same algorithm) or "omp"
    (OpenMP over the independent nodes and SOP terms of a sweep: the
    all-core CPU baseline)."""

    def __init__(self, name):
        assert name in _VARIANTS
        self.name = name

    def __enter__(self):
        global _use
        _use = self.name
        return self

    def __exit__(self, *exc):
        global _use
        _use = None
        return False


def strict_build():
    return build_variant("strict")


def lib():
    global _lib
    if _use is not None:
        if _use not in _variant_libs:
            if not os.path.exists(_VARIANTS[_use]):
                subprocess.check_call(["make", "-s", "-C", os.path.join(_ROOT, "oracle")])
            _variant_libs[_use] = _bind(C.CDLL(_VARIANTS[_use]))
        return _variant_libs[_use]
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(_ROOT, "oracle")])
    _lib = _bind(C.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    L.orc_last_error.restype = C.c_char_p
    L.orc_matvec.argtypes = [dp, dp, dp, C.c_long, C.c_long, C.c_long, C.c_int]
    L.orc_tmatvec.argtypes = L.orc_matvec.argtypes
    L.orc_ctmatvec.argtypes = L.orc_matvec.argtypes
    L.orc_rhomat.argtypes = [dp, dp, dp, C.c_long, C.c_long, C.c_long]
    L.orc_mode_product.argtypes = [dp, dp, dp, C.c_long, C.c_long, C.c_long, C.c_long, C.c_int]
    L.orc_hole_contraction.argtypes = [dp, dp, dp, C.c_long, C.c_long, C.c_long, C.c_long, C.c_int]
    L.orc_microbench_batch.argtypes = [C.c_long, lp, lp, lp, lp, dp, dp, dp, dp, dp]
    L.orc_gram_schmidt.argtypes = [dp, C.c_long, C.c_long]
    L.orc_heev.argtypes = [dp, C.c_long, dp, dp]
    L.orc_threads.restype = C.c_int
    L.orc_rmatvec.argtypes = L.orc_matvec.argtypes
    L.orc_trmatvec.argtypes = L.orc_matvec.argtypes
    L.orc_inverse.argtypes = [dp, C.c_long, dp]
    L.orc_qr.argtypes = [dp, C.c_long, C.c_long, dp]
    L.orc_reg_inverse.argtypes = [dp, C.c_long, C.c_double, dp]
    L.orc_project_out.argtypes = [dp, dp, C.c_long, C.c_long, dp]
    L.orc_balanced_tree_text.argtypes = [C.c_long] * 4 + [C.c_double] * 4 + [C.c_char_p, C.c_long]
    L.orc_balanced_tree_text.restype = C.c_long
    L.orc_model_create.argtypes = [C.c_char_p]
    L.orc_model_create.restype = C.c_void_p
    L.orc_model_destroy.argtypes = [C.c_void_p]
    L.orc_model_nnodes.argtypes = [C.c_void_p]
    L.orc_model_nnodes.restype = C.c_long
    L.orc_model_psi_size.argtypes = [C.c_void_p]
    L.orc_model_psi_size.restype = C.c_long
    L.orc_model_node_info.argtypes = [C.c_void_p, C.c_long, lp, C.c_long]
    L.orc_model_node_info.restype = C.c_long
    L.orc_model_leaf.argtypes = [C.c_void_p, C.c_long, dp, dp, dp]
    L.orc_model_set_hamiltonian.argtypes = [C.c_void_p, C.c_long, dp, lp, lp, lp, lp, dp]
    L.orc_fill_random.argtypes = [C.c_void_p, C.c_ulong, C.c_int, dp]
    L.orc_orthogonal.argtypes = [C.c_void_p, dp]
    L.orc_sparse_tree.argtypes = [C.c_void_p, C.c_long, lp]
    L.orc_sparse_tree.restype = C.c_long
    L.orc_sparse_tree_modes.argtypes = [C.c_void_p, lp, C.c_long, C.c_int, lp]
    L.orc_sparse_tree_modes.restype = C.c_long
    L.orc_build.argtypes = [C.c_void_p, dp, C.c_double]
    L.orc_get_matrix.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_long, dp]
    L.orc_derivative.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double, C.c_double]
    L.orc_apply_all.argtypes = [C.c_void_p, dp, dp]
    L.orc_gradient.argtypes = [C.c_void_p, dp, dp, C.c_double, C.POINTER(C.c_double)]
    L.orc_ttno_apply.argtypes = [C.c_void_p, lp, dp, dp, dp, dp, dp]
    L.orc_energy.argtypes = [C.c_void_p, dp]
    L.orc_energy.restype = C.c_double
    L.orc_dot_product_top.argtypes = [C.c_void_p, dp, dp, dp]
    L.orc_cmf.argtypes = ([C.c_void_p, dp] + [C.c_double] * 8 + [C.c_long, C.c_long, dp, dp, dp, lp])
    L.orc_cmf.restype = C.c_long
    L.orc_eigenstates.argtypes = ([C.c_void_p, dp, C.c_double, C.c_double, C.c_long, C.c_double, C.c_double,
                                   C.c_long, dp, dp, dp, lp, C.POINTER(C.c_double)])
    L.orc_eigenstates.restype = C.c_long
    L.orc_top_matrices.argtypes = [C.c_void_p, dp, dp, dp]
    L.orc_tensor_tree_write.argtypes = [C.c_void_p, dp, C.c_void_p, C.c_long]
    L.orc_tensor_tree_write.restype = C.c_long
    L.orc_scf.argtypes = [C.c_void_p, dp, C.c_long, C.c_long, C.c_long, C.c_double, dp, C.c_long]
    L.orc_scf.restype = C.c_long
    L.orc_scf_probe.argtypes = [C.c_void_p, dp, C.c_long, C.c_long, C.c_double, C.c_long, dp, C.c_long]
    L.orc_scf_probe.restype = C.c_long
    L.orc_scf_sweep_order.argtypes = [C.c_void_p, lp, C.c_long]
    L.orc_scf_sweep_order.restype = C.c_long
    L.orc_apply_fit.argtypes = [C.c_void_p, dp, C.c_ulong, C.c_long, dp, C.c_long]
    L.orc_apply_fit.restype = C.c_long
    return L


def cview(z: np.ndarray) -> np.ndarray:
    """complex128 array -> interleaved float64 view (C-contiguous copy if needed)."""
    z = np.ascontiguousarray(z, dtype=np.complex128)
    return z.view(np.float64)


def as_complex(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(x).view(np.complex128)


def matvec(M, B, a, b, c, C=None, add=False):
    """LA_lib.f matvec on Fortran-ordered flat arrays."""
    out = np.zeros(a * b * c, np.complex128) if C is None else np.array(C, np.complex128)
    o = cview(out)
    lib().orc_matvec(o, cview(B.ravel(order="F") if B.ndim > 1 else B),
                     cview(M.ravel(order="F") if M.ndim > 1 else M), a, b, c, int(add))
    return as_complex(o)


def rhomat(bra, ket, a, b, c):
    out = np.zeros(a * a, np.complex128)
    o = cview(out)
    lib().orc_rhomat(cview(bra), cview(ket), o, a, b, c)
    return as_complex(o)


def balanced_tree_text(num_leaves, dim_leaves=2, dim_nodes=2, dim_inc=0, omega=1.0, r0=0.0,
                       wfr0=1.0, wfomega=0.0) -> str:
    """TreeFactory::balancedTree with the C++ defaults of TreeFactory.h:12-14."""
    L = lib()
    n = L.orc_balanced_tree_text(num_leaves, dim_leaves, dim_nodes, dim_inc, omega, r0, wfr0, wfomega,
                                 None, 0)
    buf = C.create_string_buffer(n)
    L.orc_balanced_tree_text(num_leaves, dim_leaves, dim_nodes, dim_inc, omega, r0, wfr0, wfomega, buf, n)
    return buf.value.decode()


class OracleModel:
    """Tree + Hamiltonian + state, FP64 CPU."""

    def __init__(self, tree_text: str, H=None):
        L = lib()
        self._L = L  # the build this model lives in (see strict_build)
        self.h = L.orc_model_create(tree_text.encode())
        if not self.h:
            raise RuntimeError(L.orc_last_error().decode())
        self.nnodes = L.orc_model_nnodes(self.h)
        self.psi_size = L.orc_model_psi_size(self.h)
        self.nodes = [self._info(a) for a in range(self.nnodes)]
        off = 0
        for n in self.nodes:
            n["off"] = off
            off += int(np.prod(n["dims"]))
        if H is not None:
            self.set_hamiltonian(H)

    def _info(self, a):
        buf = np.zeros(64, np.int64)
        k = self._L.orc_model_node_info(self.h, a, buf, 64)
        v = buf[:k]
        order, nch = int(v[4]), int(v[5])
        return dict(address=a, parent=int(v[0]), childIdx=int(v[1]), bottom=bool(v[2]), layer=int(v[3]),
                    mode=int(v[6]), dims=[int(x) for x in v[7:7 + order]],
                    children=[int(x) for x in v[7 + order:7 + order + nch]])

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_L", None) is not None:
            self._L.orc_model_destroy(self.h)
            self.h = None

    def set_hamiltonian(self, H):
        c, toff, fm, fk, fp, pool = H.arrays()
        r = self._L.orc_model_set_hamiltonian(self.h, H.nterms, c, toff, fm, fk, fp, pool)
        if r:
            raise RuntimeError(self._L.orc_last_error().decode())
        self.nterms = H.nterms

    def leaf(self, mode):
        n = next(n for n in self.nodes if n["bottom"] and n["mode"] == mode)
        d = n["dims"][0]
        x = np.zeros(d)
        kin = np.zeros(d * d)
        tr = np.zeros(d * d)
        self._L.orc_model_leaf(self.h, mode, x, kin, tr)
        return x, kin.reshape(d, d, order="F"), tr.reshape(d, d, order="F")

    def fill_random(self, seed, delta_lowest=True):
        psi = np.zeros(2 * self.psi_size)
        self._L.orc_fill_random(self.h, seed, int(delta_lowest), psi)
        return as_complex(psi).copy()

    def orthogonal(self, psi):
        p = cview(psi).copy()
        self._L.orc_orthogonal(self.h, p)
        return as_complex(p).copy()

    def node_tensor(self, psi, a):
        n = self.nodes[a]
        size = int(np.prod(n["dims"]))
        return psi[n["off"]:n["off"] + size].reshape(n["dims"], order="F")

    def sparse_tree(self, l):
        buf = np.zeros(self.nnodes, np.int64)
        k = self._L.orc_sparse_tree(self.h, l, buf)
        return buf[:k].tolist()

    def sparse_tree_modes(self, modes, tail=True):
        buf = np.zeros(self.nnodes, np.int64)
        m = np.asarray(modes, np.int64)
        k = self._L.orc_sparse_tree_modes(self.h, m, len(m), int(tail), buf)
        return buf[:k].tolist()

    def build(self, psi, eps=1e-10):
        if self._L.orc_build(self.h, cview(psi), eps):
            raise RuntimeError(self._L.orc_last_error().decode())

    def matrix(self, which, l, a):
        n = self.nodes[a]["dims"][-1]
        out = np.zeros(2 * n * n)
        r = self._L.orc_get_matrix(self.h, which, l, a, out)
        if r:
            return None
        return as_complex(out).reshape(n, n, order="F").copy()

    def rho(self, a):
        return self.matrix(0, 0, a)

    def rho_inv(self, a):
        return self.matrix(1, 0, a)

    def hmat(self, l, a):
        return self.matrix(2, l, a)

    def hole(self, l, a):
        return self.matrix(3, l, a)

    def derivative(self, psi, phase=1.0, eps=1e-10):
        out = np.zeros(2 * self.psi_size)
        if self._L.orc_derivative(self.h, cview(psi), out, complex(phase).real, complex(phase).imag, eps):
            raise RuntimeError(self._L.orc_last_error().decode())
        return as_complex(out).copy()

    def gradient(self, psi, eps=1e-10):
        """dE/dA* at every node (orthonormal lower nodes) and the energy."""
        out = np.zeros(2 * self.psi_size)
        e = C.c_double()
        if self._L.orc_gradient(self.h, cview(psi), out, eps, C.byref(e)):
            raise RuntimeError(self._L.orc_last_error().decode())
        return as_complex(out).copy(), e.value

    def ttno_apply(self, bond, B, bra, ket):
        """TTNO application at every node to the ket's node tensors (rep with
        bra, ket) and <bra|H|ket>.  bond: D per node address; B: the operator
        tensors concatenated in address order (complex)."""
        out = np.zeros(2 * self.psi_size)
        e = np.zeros(2)
        if self._L.orc_ttno_apply(self.h, np.ascontiguousarray(bond, np.int64), cview(np.ascontiguousarray(B)),
                                cview(bra), cview(ket), out, e):
            raise RuntimeError(self._L.orc_last_error().decode())
        return as_complex(out).copy(), complex(e[0], e[1])

    def apply_all(self, psi):
        out = np.zeros(2 * self.psi_size)
        if self._L.orc_apply_all(self.h, cview(psi), out):
            raise RuntimeError("build first")
        return as_complex(out).copy()

    def energy(self, psi):
        return self._L.orc_energy(self.h, cview(psi))

    def overlap(self, bra, ket):
        out = np.zeros(2)
        self._L.orc_dot_product_top(self.h, cview(bra), cview(ket), out)
        return complex(out[0], out[1])

    def cmf(self, psi, t0, tend, dt, out, acc_cmf=1e-4, acc_bs=1e-5, phase=1.0, max_steps=10 ** 9,
            max_out=100000):
        p = cview(psi).copy()
        times = np.zeros(max_out)
        en = np.zeros(max_out)
        ac = np.zeros(2 * max_out)
        stats = np.zeros(4, np.int64)
        n = self._L.orc_cmf(self.h, p, t0, tend, dt, out, acc_cmf, acc_bs, complex(phase).real,
                          complex(phase).imag, max_steps, max_out, times, en, ac, stats)
        if n < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        n = min(n, max_out)
        return dict(psi=as_complex(p).copy(), times=times[:n], energies=en[:n],
                    autocorr=as_complex(ac)[:n].copy(), steps=int(stats[0]), rejected=int(stats[1]),
                    builds=int(stats[2]), derivs=int(stats[3]))

    def tensor_tree_bytes(self, psi):
        """TensorTree<cd>::write of psi, as the oracle restates it."""
        n = self._L.orc_tensor_tree_write(self.h, cview(psi), None, 0)
        buf = C.create_string_buffer(n)
        self._L.orc_tensor_tree_write(self.h, cview(psi), buf, n)
        return buf.raw[:n]

    def top_matrices(self, psi):
        """<Psi_i|H|Psi_j> and <Psi_i|Psi_j> of the top node's states."""
        ns = self.nodes[-1]["dims"][-1]
        hm = np.zeros(2 * ns * ns)
        sm = np.zeros(2 * ns * ns)
        if self._L.orc_top_matrices(self.h, cview(psi), hm, sm):
            raise RuntimeError(self._L.orc_last_error().decode())
        return (as_complex(hm).reshape(ns, ns, order="F").copy(), as_complex(sm).reshape(ns, ns, order="F").copy())

    def eigenstates(self, psi, dt, out, iterations, acc_cmf=1e-6, acc_bs=1e-7, max_steps=10 ** 9):
        """Imaginary-time eigenstates (Eigenstates.cpp:60-103): energies
        (iterations + 1, nstates), propagator energies and converged overlaps
        (iterations, nstates), the final state and the counters."""
        ns = self.nodes[-1]["dims"][-1]
        p = cview(psi).copy()
        en = np.zeros((iterations + 1) * ns)
        pe = np.zeros(max(iterations, 1) * ns)
        cv = np.zeros(max(iterations, 1) * ns)
        stats = np.zeros(4, np.int64)
        dte = C.c_double()
        n = self._L.orc_eigenstates(self.h, p, dt, out, iterations, acc_cmf, acc_bs, max_steps, en, pe, cv, stats,
                                  C.byref(dte))
        if n < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return dict(psi=as_complex(p).copy(), energies=en.reshape(iterations + 1, ns),
                    propagator=pe[:iterations * ns].reshape(iterations, ns),
                    converged=cv[:iterations * ns].reshape(iterations, ns), steps=int(stats[0]),
                    rejected=int(stats[1]), builds=int(stats[2]), derivs=int(stats[3]), dt_end=dte.value)

    def scf_sweep_order(self):
        buf = np.zeros(4 * self.nnodes, np.int64)
        k = self._L.orc_scf_sweep_order(self.h, buf, len(buf))
        return buf[:k].tolist()

    def scf(self, psi, n_iter, n_krylov, n_itp=0, beta=1.0):
        """Krylov-sweep SCF (SCF.cpp:270-348): the final state and the lowest
        Ritz value of every microstep (n_iter x sweep length)."""
        p = cview(psi).copy()
        steps = len(self.scf_sweep_order())
        e = np.zeros(max(n_iter * steps, 1))
        k = self._L.orc_scf(self.h, p, n_iter, n_krylov, n_itp, beta, e, len(e))
        if k < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return dict(psi=as_complex(p).copy(), energies=e[:k].reshape(n_iter, steps) if n_iter else e[:0])

    def scf_probe(self, psi, n_krylov, step, n_itp=0, beta=1.0):
        """State and dense effective Hamiltonian (N x N) at the node of SCF
        microstep `step`, before its Krylov solve."""
        p = cview(psi).copy()
        N = max(int(np.prod(n["dims"])) for n in self.nodes)
        h = np.zeros(2 * N * N)
        a = self._L.orc_scf_probe(self.h, p, n_krylov, n_itp, beta, step, h, N * N)
        if a < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        Na = int(np.prod(self.nodes[a]["dims"]))
        return a, as_complex(p).copy(), as_complex(h)[:Na * Na].reshape(Na, Na, order="F").copy()

    def apply_fit(self, psi, seed, n_iter=100):
        """H|Psi> fitted into a tree of Psi's shape (ApplyFirstOrder.cpp:43-147,
        the SOP's coefficients not applied): the fit and the fidelity of
        every iteration."""
        p = cview(psi).copy()
        f = np.zeros(max(n_iter, 1))
        k = self._L.orc_apply_fit(self.h, p, seed, n_iter, f, len(f))
        if k < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return dict(psi=as_complex(p).copy(), fidelities=f[:k].copy())

