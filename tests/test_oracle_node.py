"""Pins the oracle's node kernels (LA_lib.f matvec/ctmatvec/tmatvec/rhomat,
matrixTensor1/contraction1, gramSchmidt, small FP64 linear algebra) against
the reference tests' closed forms and an independent numpy brute force."""
import numpy as np
import pytest

import oracle_lib as O

rng = np.random.default_rng(7)


def crand(*shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def as3(x, b, a, c):
    return np.asarray(x).reshape((b, a, c), order="F")


def test_hole_product_closed_form():
    # test_Tensor.cpp:184-194: C(bef, act, aft) = act * i on mode 1 of a
    # (2,2,2) tensor => |S(i,j)| = before*after*i*j.
    C = np.zeros((2, 2, 2), np.complex128)
    for a in range(2):
        C[:, a, :] = a * 1j
    S = O.rhomat(C.ravel(order="F"), C.ravel(order="F"), 2, 2, 2).reshape(2, 2, order="F")
    for i in range(2):
        for j in range(2):
            assert abs(abs(S[i, j]) - 4.0 * i * j) < 1e-12


def test_dot_product_closed_form():
    # test_Tensor.cpp:196-205: C2(bef, act, aft) = aft * i; dotProduct on the last index
    C = np.zeros((2, 2, 2), np.complex128)
    for c in range(2):
        C[:, :, c] = c * 1j
    S = O.rhomat(C.ravel(order="F"), C.ravel(order="F"), 2, 4, 1).reshape(2, 2, order="F")
    for i in range(2):
        for j in range(2):
            assert abs(abs(S[i, j]) - 4.0 * i * j) < 1e-12


SHAPES = [(1, 5, 7), (7, 5, 1), (1, 6, 3), (3, 4, 5), (16, 8, 1), (1, 9, 1), (2, 3, 40), (1, 3, 40)]


@pytest.mark.parametrize("b,a,c", SHAPES)
def test_matvec_bruteforce(b, a, c):
    B = crand(b, a, c)
    M = crand(a, a)
    C0 = crand(b, a, c)
    ref = np.einsum("kl,blc->bkc", M, B)
    out = O.matvec(M, B, a, b, c).reshape((b, a, c), order="F")
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1, np.max(np.abs(ref)))
    out2 = O.matvec(M, B, a, b, c, C=C0.ravel(order="F"), add=True).reshape((b, a, c), order="F")
    assert np.max(np.abs(out2 - (ref + C0))) < 1e-12 * max(1, np.max(np.abs(ref)))


@pytest.mark.parametrize("b,a,c", SHAPES)
def test_tmatvec_ctmatvec_bruteforce(b, a, c):
    L = O.lib()
    B = crand(b, a, c)
    M = crand(a, a)
    for fn, Mx in ((L.orc_tmatvec, M.T), (L.orc_ctmatvec, M.conj().T)):
        out = np.zeros(a * b * c, np.complex128)
        o = O.cview(out)
        fn(o, O.cview(B.ravel(order="F")), O.cview(M.ravel(order="F")), a, b, c, 0)
        ref = np.einsum("kl,blc->bkc", Mx, B)
        got = O.as_complex(o).reshape((b, a, c), order="F")
        assert np.max(np.abs(got - ref)) < 1e-12 * max(1, np.max(np.abs(ref)))


@pytest.mark.parametrize("b,a,c", SHAPES)
def test_rhomat_bruteforce_and_overwrite(b, a, c):
    A = crand(b, a, c)
    B = crand(b, a, c)
    ref = np.einsum("bjc,blc->jl", A.conj(), B)
    S = O.rhomat(A.ravel(order="F"), B.ravel(order="F"), a, b, c).reshape(a, a, order="F")
    assert np.max(np.abs(S - ref)) < 1e-12 * max(1, np.max(np.abs(ref)))
    # Hermiticity of S(A, A)
    SAA = O.rhomat(A.ravel(order="F"), A.ravel(order="F"), a, b, c).reshape(a, a, order="F")
    assert np.max(np.abs(SAA - SAA.conj().T)) < 1e-12 * np.max(np.abs(SAA))


@pytest.mark.parametrize("before,nB,nC,after", [(1, 4, 6, 5), (3, 5, 2, 4), (6, 3, 3, 1), (2, 7, 7, 3)])
def test_rectangular_mode_product_and_contraction(before, nB, nC, after):
    L = O.lib()
    B = crand(before, nB, after)
    M = crand(nC, nB)
    out = np.zeros(before * nC * after, np.complex128)
    o = O.cview(out)
    L.orc_mode_product(o, O.cview(M.ravel(order="F")), O.cview(B.ravel(order="F")), before, nB, nC, after, 0)
    ref = np.einsum("kl,blc->bkc", M, B)
    assert np.max(np.abs(O.as_complex(o).reshape((before, nC, after), order="F") - ref)) < 1e-12 * np.max(np.abs(ref))
    A2 = crand(before, nC, after)
    S = np.zeros(nC * nB, np.complex128)
    s = O.cview(S)
    L.orc_hole_contraction(s, O.cview(A2.ravel(order="F")), O.cview(B.ravel(order="F")), before, nC, nB, after, 0)
    ref = np.einsum("bjc,blc->jl", A2.conj(), B)
    assert np.max(np.abs(O.as_complex(s).reshape((nC, nB), order="F") - ref)) < 1e-12 * np.max(np.abs(ref))


def test_integer_inputs_are_exact():
    # integer-valued operands: every product and partial sum is an exact integer in FP64
    b, a, c = 5, 6, 7
    B = rng.integers(-9, 10, size=(b, a, c)) + 1j * rng.integers(-9, 10, size=(b, a, c))
    M = rng.integers(-9, 10, size=(a, a)) + 1j * rng.integers(-9, 10, size=(a, a))
    out = O.matvec(M.astype(np.complex128), B.astype(np.complex128), a, b, c).reshape((b, a, c), order="F")
    ref = np.einsum("kl,blc->bkc", M, B)
    assert np.array_equal(out, ref)
    S = O.rhomat(B.astype(np.complex128).ravel(order="F"), B.astype(np.complex128).ravel(order="F"), a, b, c)
    assert np.array_equal(S.reshape(a, a, order="F"), np.einsum("bjc,blc->jl", B.conj(), B))


def test_identity_and_permutation_matrices_are_exact():
    b, a, c = 4, 5, 3
    B = crand(b, a, c)
    I = np.eye(a, dtype=np.complex128)
    assert np.array_equal(O.matvec(I, B, a, b, c).reshape((b, a, c), order="F"), B)
    perm = rng.permutation(a)
    P = I[perm]
    assert np.array_equal(O.matvec(P, B, a, b, c).reshape((b, a, c), order="F"), B[:, perm, :])


def test_gram_schmidt_matches_qr_span():
    A = crand(9, 5)
    Q = A.ravel(order="F").copy()
    q = O.cview(Q)
    O.lib().orc_gram_schmidt(q, 9, 5)
    Q = O.as_complex(q).reshape(9, 5, order="F")
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(5))) < 1e-13
    # same flag as column-by-column QR: R = Q^H A upper triangular with positive real diagonal
    R = Q.conj().T @ A
    assert np.max(np.abs(np.tril(R, -1))) < 1e-12
    assert np.all(np.real(np.diag(R)) > 0) and np.max(np.abs(np.imag(np.diag(R)))) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 5, 16, 24])
def test_heev_against_numpy(n):
    X = crand(n, n)
    H = X + X.conj().T
    U = np.zeros(2 * n * n)
    ev = np.zeros(n)
    O.lib().orc_heev(O.cview(H.ravel(order="F")), n, U, ev)
    U = O.as_complex(U).reshape(n, n, order="F")
    assert np.max(np.abs(ev - np.linalg.eigvalsh(H))) < 1e-12 * max(1, np.max(np.abs(ev)))
    assert np.max(np.abs(U @ np.diag(ev) @ U.conj().T - H)) < 1e-12 * np.max(np.abs(H))
    assert np.all(np.real(U[0]) >= 0)


def test_heev_real_symmetric_phase_convention():
    # rDiag convention (Matrix_Implementation.h:271-278): first row non-negative
    n = 8
    X = rng.normal(size=(n, n))
    H = (X + X.T).astype(np.complex128)
    U = np.zeros(2 * n * n)
    ev = np.zeros(n)
    O.lib().orc_heev(O.cview(H.ravel(order="F")), n, U, ev)
    U = O.as_complex(U).reshape(n, n, order="F")
    w, V = np.linalg.eigh(H.real)
    V = V * np.sign(V[0])
    assert np.max(np.abs(U - V)) < 1e-11


def test_regularized_inverse():
    # test_MatrixTree.cpp:105-128: H^-1 H = I for H = X X^H; plus the explicit
    # regulariser on a singular matrix.
    n = 6
    X = crand(n, n)
    H = X @ X.conj().T
    R = np.zeros(2 * n * n)
    O.lib().orc_reg_inverse(O.cview(H.ravel(order="F")), n, 1e-10, R)
    R = O.as_complex(R).reshape(n, n, order="F")
    assert np.max(np.abs(R @ H - np.eye(n))) < 1e-8
    rho = np.diag([1.0, 0.0, 0.0]).astype(np.complex128)
    R = np.zeros(18)
    O.lib().orc_reg_inverse(O.cview(rho.ravel(order="F")), 3, 1e-10, R)
    R = O.as_complex(R).reshape(3, 3, order="F")
    # 1/(0 + 1e-10 * e^0) = 1e10 for the empty directions, 1/(1 + 1e-10 e^{-1e10}) = 1
    assert abs(R[0, 0] - 1.0) < 1e-15 and abs(R[1, 1] - 1e10) < 1e-3 and abs(R[2, 2] - 1e10) < 1e-3


def test_project_out():
    before, n = 12, 4
    A = crand(before, n)
    B = crand(before, n)
    out = np.zeros(2 * before * n)
    O.lib().orc_project_out(O.cview(A.ravel(order="F")), O.cview(B.ravel(order="F")), before, n, out)
    out = O.as_complex(out).reshape(before, n, order="F")
    # (1 - P_B) A annihilates the span of B and is idempotent
    assert np.max(np.abs(B.conj().T @ out)) < 1e-12 * np.max(np.abs(A))
    ref = A - B @ np.linalg.solve(B.conj().T @ B, B.conj().T @ A)
    assert np.max(np.abs(out - ref)) < 1e-12 * np.max(np.abs(A))


@pytest.mark.parametrize("b,a,c", SHAPES)
def test_real_matvec_bruteforce(b, a, c):
    """rmatvec / trmatvec (LA_lib.f:314-379, 402-467) against numpy einsum on
    real data, overwrite and accumulate."""
    L = O.lib()
    B = rng.normal(size=(b, a, c))
    M = rng.normal(size=(a, a))
    C0 = rng.normal(size=(b, a, c))
    for fn, ref in ((L.orc_rmatvec, np.einsum("kl,blc->bkc", M, B)), (L.orc_trmatvec, np.einsum("lk,blc->bkc", M, B))):
        for add in (0, 1):
            out = C0.ravel(order="F").copy()
            fn(out, np.ascontiguousarray(B.ravel(order="F")), np.ascontiguousarray(M.ravel(order="F")), a, b, c, add)
            want = ref + (C0 if add else 0)
            assert np.max(np.abs(out.reshape((b, a, c), order="F") - want)) < 1e-12 * max(1, np.max(np.abs(want)))


def oracle_qr(A):
    m, n = A.shape
    Q = np.zeros(2 * m * n)
    O.lib().orc_qr(O.cview(np.ascontiguousarray(A.ravel(order="F"))), m, n, Q)
    return O.as_complex(Q).reshape(m, n, order="F")


@pytest.mark.parametrize("m,n", [(1, 1), (5, 1), (6, 6), (40, 7), (576, 24), (9, 8)])
def test_householder_qr_matches_lapack(m, n):
    """The oracle's thin Q equals LAPACK's (numpy.linalg.qr: zgeqrf + zungqr),
    whose reflector convention Eigen's HouseholderQR shares, column for
    column; QᴴQ = I and Qᴴ A is upper triangular with the diagonal sign
    -sign(Re a) the convention fixes."""
    A = crand(m, n)
    Q = oracle_qr(A)
    Qn, Rn = np.linalg.qr(A)
    assert np.max(np.abs(Q - Qn)) < 1e-12
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(n))) < 1e-12
    R = Q.conj().T @ A
    assert np.max(np.abs(np.tril(R, -1))) < 1e-11 * np.max(np.abs(A))
    # the first reflector's beta: -sign(Re a00) |A[:, 0]|
    assert abs(R[0, 0] - (-np.sign(A[0, 0].real) * np.linalg.norm(A[:, 0]))) < 1e-11 * np.linalg.norm(A[:, 0])


def test_householder_qr_special_cases():
    # a column already zero below a real diagonal gets the identity reflector
    # (tau = 0): Q keeps it and R keeps the sign of the diagonal
    A = np.diag([2.0, -3.0, 0.5]).astype(np.complex128)
    Q = oracle_qr(A)
    assert np.max(np.abs(Q - np.eye(3))) < 1e-15
    # a complex diagonal is not skipped (Eigen and LAPACK make R's diagonal real)
    A = np.diag([1j, 2.0]).astype(np.complex128)
    Q = oracle_qr(A)
    R = Q.conj().T @ A
    assert abs(R[0, 0].imag) < 1e-15 and R[0, 0].real < 0


@pytest.mark.parametrize("n", [1, 3, 16, 48])
def test_inverse_matches_numpy(n):
    A = crand(n, n) + n * np.eye(n)
    R = np.zeros(2 * n * n)
    O.lib().orc_inverse(O.cview(np.ascontiguousarray(A.ravel(order="F"))), n, R)
    X = O.as_complex(R).reshape(n, n, order="F")
    assert np.max(np.abs(X - np.linalg.inv(A))) < 1e-12 * np.max(np.abs(np.linalg.inv(A)))
