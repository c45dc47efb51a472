"""GPU parity of the rest of the drop-in boundary against the FP64 oracle:
QuTree's Fortran seam (matvec_ ... rhomat_ with the reference's by-reference
argument order), the real products, the batched FP64 small-matrix ops, and
the engine options' state rules."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(23)
ip = C.POINTER(C.c_int)


def crand(*shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def dev(x):
    return Q.DeviceBuffer.from_array(np.ascontiguousarray(np.asarray(x).ravel(order="F")))


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def seam(name):
    f = getattr(Q.lib(), name)
    f.restype = None
    if name == "rhomat_":
        f.argtypes = [C.c_void_p] * 3 + [ip] * 3
    else:
        f.argtypes = [C.c_void_p] * 3 + [ip] * 4
    return f


SHAPES = [(3, 4, 5), (1, 6, 3), (7, 5, 1), (16, 8, 2), (2, 9, 11)]


@pytest.mark.parametrize("b,a,c", SHAPES)
@pytest.mark.parametrize("name,oname,real", [("matvec_", "orc_matvec", False), ("tmatvec_", "orc_tmatvec", False),
                                              ("ctmatvec_", "orc_ctmatvec", False), ("rmatvec_", "orc_rmatvec", True),
                                              ("trmatvec_", "orc_trmatvec", True)])
@pytest.mark.parametrize("add", [0, 1])
def test_fortran_seam_matches_oracle(b, a, c, name, oname, real, add):
    """The seam takes (C, B, mat, &a, &b, &c, &add) as QuTree's call sites pass
    them (Tensor_Implementation.h:566, Matrix_Implementation.h:478); a, b, c
    all differ, so any permutation of them or of add changes the result."""
    print(f"case seam {name} b={b} a={a} c={c} add={add}", flush=True)
    if real:
        B, M, C0 = rng.normal(size=(b, a, c)), rng.normal(size=(a, a)), rng.normal(size=(b, a, c))
    else:
        B, M, C0 = crand(b, a, c), crand(a, a), crand(b, a, c)
    dB, dM, dC = dev(B), dev(M), dev(C0)
    ia, ib, ic, iadd = C.c_int(a), C.c_int(b), C.c_int(c), C.c_int(add)
    seam(name)(dC.ptr, dB.ptr, dM.ptr, C.byref(ia), C.byref(ib), C.byref(ic), C.byref(iadd))
    assert Q.lib().qtg_seam_status() == 0
    dt = np.float64 if real else np.complex128
    got = dC.download(dt).reshape((b, a, c), order="F")
    ref = C0.ravel(order="F").copy()
    view = ref if real else O.cview(ref)
    getattr(O.lib(), oname)(view, np.ascontiguousarray(B.ravel(order="F")) if real else O.cview(B.ravel(order="F")),
                            np.ascontiguousarray(M.ravel(order="F")) if real else O.cview(M.ravel(order="F")),
                            a, b, c, add)
    assert rel(got, ref.reshape((b, a, c), order="F")) < 1e-12


@pytest.mark.parametrize("b,a,c", SHAPES)
def test_fortran_seam_rhomat(b, a, c):
    print(f"case seam rhomat_ b={b} a={a} c={c}", flush=True)
    A, B = crand(b, a, c), crand(b, a, c)
    dA, dB, dS = dev(A), dev(B), dev(crand(a, a))
    ia, ib, ic = C.c_int(a), C.c_int(b), C.c_int(c)
    seam("rhomat_")(dA.ptr, dB.ptr, dS.ptr, C.byref(ia), C.byref(ib), C.byref(ic))
    got = dS.download(np.complex128).reshape(a, a, order="F")
    ref = O.rhomat(A.ravel(order="F"), B.ravel(order="F"), a, b, c).reshape(a, a, order="F")
    assert rel(got, ref) < 1e-12


def test_seam_reports_a_bad_shape():
    print("case seam bad shape", flush=True)
    d = dev(np.zeros(4, np.complex128))
    ia, ib, ic, iadd = C.c_int(0), C.c_int(1), C.c_int(1), C.c_int(0)
    seam("matvec_")(d.ptr, d.ptr, d.ptr, C.byref(ia), C.byref(ib), C.byref(ic), C.byref(iadd))
    assert Q.lib().qtg_seam_status() == 1


def batch(n, count, herm=True):
    X = crand(count, n, n)
    if herm:
        X = X @ np.conj(np.transpose(X, (0, 2, 1))) / n + 1e-3 * np.eye(n)
    return X


def to_dev(X):
    return Q.DeviceBuffer.from_array(np.ascontiguousarray(np.transpose(X, (0, 2, 1)).reshape(-1)))


def from_dev(buf, count, m, n):
    return np.transpose(buf.download(np.complex128).reshape(count, n, m), (0, 2, 1))


@pytest.mark.parametrize("n", [1, 2, 7, 16, 33, 48, 64])
def test_heev_and_regularised_inverse_batches(n):
    print(f"case heev batch n={n}", flush=True)
    count = 9
    X = batch(n, count)
    dA, dU = to_dev(X), Q.DeviceBuffer(16 * n * n * count)
    dl = Q.DeviceBuffer(8 * n * count)
    L = Q.lib()
    L.qtg_heev_batch_z.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
    Q.check(L.qtg_heev_batch_z(dA.ptr, dU.ptr, dl.ptr, n, count, None))
    U = from_dev(dU, count, n, n)
    lam = dl.download(np.float64).reshape(count, n)
    L.qtg_reg_inverse_batch_z.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_void_p]
    dR = Q.DeviceBuffer(16 * n * n * count)
    eps = 1e-3  # large enough that the regulariser shows
    Q.check(L.qtg_reg_inverse_batch_z(dA.ptr, dR.ptr, n, count, eps, None))
    R = from_dev(dR, count, n, n)
    for i in range(count):
        Uo, lo = np.zeros(2 * n * n), np.zeros(n)
        O.lib().orc_heev(O.cview(X[i].ravel(order="F")), n, Uo, lo)
        Uo = O.as_complex(Uo).reshape(n, n, order="F")
        assert np.max(np.abs(lam[i] - lo)) < 1e-12 * np.max(np.abs(lo))
        assert np.all(np.diff(lam[i]) >= 0)
        # eigenvectors with the same phase convention (first component real >= 0)
        assert np.max(np.abs(U[i][0].imag)) < 1e-13 and np.all(U[i][0].real >= -1e-13)
        gaps = np.diff(lo)
        if n == 1 or np.min(gaps) > 1e-6 * np.max(np.abs(lo)):
            assert np.max(np.abs(U[i] - Uo)) < 1e-9
        assert rel(U[i] @ np.diag(lam[i]) @ U[i].conj().T, X[i]) < 1e-12
        Ro = np.zeros(2 * n * n)
        O.lib().orc_reg_inverse(O.cview(X[i].ravel(order="F")), n, eps, Ro)
        assert rel(R[i], O.as_complex(Ro).reshape(n, n, order="F")) < 1e-10


@pytest.mark.parametrize("n", [1, 5, 16, 48, 64])
def test_inverse_batch(n):
    print(f"case inverse batch n={n}", flush=True)
    count = 7
    X = crand(count, n, n) + 2 * np.sqrt(n) * np.eye(n)
    dA, dR = to_dev(X), Q.DeviceBuffer(16 * n * n * count)
    L = Q.lib()
    L.qtg_inverse_batch_z.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
    Q.check(L.qtg_inverse_batch_z(dA.ptr, dR.ptr, n, count, None))
    R = from_dev(dR, count, n, n)
    for i in range(count):
        Ro = np.zeros(2 * n * n)
        O.lib().orc_inverse(O.cview(X[i].ravel(order="F")), n, Ro)
        assert rel(R[i], O.as_complex(Ro).reshape(n, n, order="F")) < 1e-12


@pytest.mark.parametrize("m,n", [(1, 1), (8, 8), (40, 7), (576, 24), (128, 16), (2304, 48), (300, 64)])
def test_qr_batch_matches_oracle(m, n):
    print(f"case qr batch m={m} n={n}", flush=True)
    count = 5
    X = rng.normal(size=(count, m, n)) + 1j * rng.normal(size=(count, m, n))
    d = to_dev(X)
    L = Q.lib()
    L.qtg_qr_batch_z.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    Q.check(L.qtg_qr_batch_z(d.ptr, d.ptr, m, n, count, None))  # in place
    Qd = from_dev(d, count, m, n)
    for i in range(count):
        Qo = np.zeros(2 * m * n)
        O.lib().orc_qr(O.cview(X[i].ravel(order="F")), m, n, Qo)
        assert np.max(np.abs(Qd[i] - O.as_complex(Qo).reshape(m, n, order="F"))) < 1e-12


def test_plan_options_are_fixed_after_the_hamiltonian():
    print("case options after set_hamiltonian", flush=True)
    text, H = W.cmf8_model()
    eng = Q.Engine(text, H)
    for key, val in (("path", 0.0), ("compress", 0.0), ("split_waves", 64.0), ("split_min_k", 8.0)):
        rc = Q.lib().qtg_engine_set_option(eng.ptr, key.encode(), val)
        assert rc == 5, key  # QTG_ESTATE
    assert Q.lib().qtg_engine_set_option(eng.ptr, b"rho_eps", -1.0) == 3
    # the engine still runs its own plan
    psi = eng.fill_random(3, delta_lowest=False)
    eng.set_psi(psi)
    eng.derivative()
    om = O.OracleModel(text, H)
    assert rel(eng.get(0), om.derivative(psi)) < 1e-11


def test_rho_eps_change_reaches_a_captured_sweep():
    """A Hartree product with a 1e-3 admixture at the top node has density
    eigenvalues ~1e-6 at the top's children, where
    rho^-1 = 1/(l + eps e^{-l/eps}) depends on eps: the derivative moves by
    ~30 % between eps = 1e-6 and 1e-4.  The first sweep is captured into a
    graph at one eps; after the option changes, the next sweep must use the
    new one."""
    print("case rho_eps change with graph", flush=True)
    text, H = W.cmf8_model()
    om = O.OracleModel(text, H)
    psi = om.fill_random(5, delta_lowest=True)
    top = om.nnodes - 1
    n = om.nodes[top]
    sz = int(np.prod(n["dims"]))
    g = np.random.default_rng(1)
    x = psi[n["off"]:n["off"] + sz] + 1e-3 * (g.normal(size=sz) + 1j * g.normal(size=sz))
    psi[n["off"]:n["off"] + sz] = x / np.linalg.norm(x)
    eng = Q.Engine(text, H, rho_eps=1e-6)
    eng.set_psi(psi)
    eng.derivative()
    d1 = eng.get(0)
    assert rel(d1, om.derivative(psi, eps=1e-6)) < 1e-9
    eng.set_option("rho_eps", 1e-4)
    eng.derivative()
    d2 = eng.get(0)
    ref2 = om.derivative(psi, eps=1e-4)
    assert rel(d2, ref2) < 1e-9
    assert rel(d1, ref2) > 1e-2
