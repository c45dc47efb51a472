"""GPU parity of the node-level C ABI (matvec/tmatvec/ctmatvec/rhomat
replacements, rectangular mode products and hole contractions, grouped plans)
against the FP64 oracle.  complex128: 1e-12 relative; complex64: 2e-6."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(11)


def crand(*shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def dev(z, c64=False):
    z = np.ascontiguousarray(z.ravel(order="F"), np.complex128)
    return Q.DeviceBuffer.from_array(z.astype(np.complex64) if c64 else z)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


SHAPES = [(1, 5, 7), (7, 5, 1), (1, 6, 3), (3, 4, 5), (16, 8, 1), (1, 9, 1), (2, 3, 40), (1, 3, 40),
          (64, 64, 1), (1, 64, 64), (70, 33, 3), (300, 12, 5)]


@pytest.mark.parametrize("b,a,c", SHAPES)
@pytest.mark.parametrize("fn,op", [("qtg_matvec_z", "N"), ("qtg_tmatvec_z", "T"), ("qtg_ctmatvec_z", "H")])
@pytest.mark.parametrize("add", [0, 1])
def test_matvec_family(b, a, c, fn, op, add):
    print(f"case {fn} b={b} a={a} c={c} add={add}", flush=True)
    B = crand(b, a, c)
    M = crand(a, a)
    C0 = crand(b, a, c)
    dB, dM, dC = dev(B), dev(M), dev(C0)
    Q.check(getattr(Q.lib(), fn)(dC.ptr, dB.ptr, dM.ptr, a, b, c, add, None))
    got = dC.download(np.complex128).reshape((b, a, c), order="F")
    flat = C0.ravel(order="F").copy()
    oracle = {"N": O.lib().orc_matvec, "T": O.lib().orc_tmatvec, "H": O.lib().orc_ctmatvec}[op]
    f = O.cview(flat)
    oracle(f, O.cview(B.ravel(order="F")), O.cview(M.ravel(order="F")), a, b, c, add)
    ref = O.as_complex(f).reshape((b, a, c), order="F")
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("b,a,c", SHAPES)
def test_rhomat_overwrites(b, a, c):
    print(f"case rhomat b={b} a={a} c={c}", flush=True)
    A, B = crand(b, a, c), crand(b, a, c)
    dA, dB = dev(A), dev(B)
    dS = dev(crand(a, a))  # garbage that must be overwritten
    Q.check(Q.lib().qtg_rhomat_z(dA.ptr, dB.ptr, dS.ptr, a, b, c, None))
    got = dS.download(np.complex128).reshape(a, a, order="F")
    ref = O.rhomat(A.ravel(order="F"), B.ravel(order="F"), a, b, c).reshape(a, a, order="F")
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("before,nB,nC,after", [(1, 4, 6, 5), (3, 5, 2, 4), (6, 3, 3, 1), (2, 7, 7, 3),
                                                (65, 17, 9, 2)])
@pytest.mark.parametrize("c64", [False, True])
def test_rectangular_products(before, nB, nC, after, c64):
    print(f"case rect {before},{nB},{nC},{after} c64={c64}", flush=True)
    L = Q.lib()
    B = crand(before, nB, after)
    M = crand(nC, nB)
    C0 = crand(before, nC, after)
    dB, dM, dC = dev(B, c64), dev(M, c64), dev(C0, c64)
    fn = L.qtg_mode_product_c if c64 else L.qtg_mode_product_z
    Q.check(fn(dC.ptr, dM.ptr, dB.ptr, before, nB, nC, after, 0, 1, None))
    dt = np.complex64 if c64 else np.complex128
    got = dC.download(dt).astype(np.complex128).reshape((before, nC, after), order="F")
    ref = np.einsum("kl,blc->bkc", M, B) + C0
    assert rel(got, ref) < (2e-6 if c64 else 1e-12)
    # hole contraction with accumulate
    A2 = crand(before, nC, after)
    S0 = crand(nC, nB)
    dA2, dS = dev(A2, c64), dev(S0, c64)
    fn = L.qtg_hole_contraction_c if c64 else L.qtg_hole_contraction_z
    Q.check(fn(dS.ptr, dA2.ptr, dB.ptr, before, nC, nB, after, 1, None))
    got = dS.download(dt).astype(np.complex128).reshape((nC, nB), order="F")
    Sv = np.zeros(nC * nB, np.complex128)
    s = O.cview(Sv)
    O.lib().orc_hole_contraction(s, O.cview(A2.ravel(order="F")), O.cview(B.ravel(order="F")), before, nC, nB,
                                 after, 0)
    ref = O.as_complex(s).reshape((nC, nB), order="F") + S0
    assert rel(got, ref) < (2e-6 if c64 else 1e-12)


def test_errors_are_status_codes():
    L = Q.lib()
    d = Q.DeviceBuffer(1024)
    assert L.qtg_mode_product_z(d.ptr, d.ptr, d.ptr, 1, 0, 1, 1, 0, 0, None) == 1  # QTG_ESHAPE
    assert L.qtg_mode_product_z(d.ptr, d.ptr, d.ptr, 1, 2, 2, 1, 5, 0, None) == 1
    assert L.qtg_hole_contraction_z(None, d.ptr, d.ptr, 1, 2, 2, 1, 0, None) == 3  # QTG_EARG


def _microbench_problem(count, seed):
    shapes, modes = W.microbench_shapes(count, seed)
    sizes = np.prod(shapes, axis=1)
    offT = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    nk = shapes[np.arange(count), modes]
    offM = np.concatenate([[0], np.cumsum(nk * nk)[:-1]]).astype(np.int64)
    tot, totM = int(sizes.sum()), int((nk * nk).sum())
    g = np.random.default_rng(seed + 1)
    A = g.normal(size=tot) + 1j * g.normal(size=tot)
    B = g.normal(size=tot) + 1j * g.normal(size=tot)
    M = g.normal(size=totM) + 1j * g.normal(size=totM)
    mp = np.zeros(count, Q.MP_DESC)
    hc = np.zeros(count, Q.HC_DESC)
    for t in range(count):
        d = shapes[t]
        k = modes[t]
        before = int(np.prod(d[:k]))
        after = int(np.prod(d[k + 1:]))
        mp[t] = (offT[t], offM[t], offT[t], before, d[k], d[k], after, 0, 0)
        hc[t] = (offM[t], offT[t], offT[t], before, d[k], d[k], after, 0, 0)
    return shapes, modes, offT, offM, A, B, M, mp, hc


@pytest.mark.parametrize("c64", [False, True])
def test_grouped_microbench_subset(c64):
    count = 96
    print(f"case grouped microbench count={count} c64={c64}", flush=True)
    shapes, modes, offT, offM, A, B, M, mp, hc = _microbench_problem(count, 5)
    Cref = np.zeros_like(A)
    Sref = np.zeros_like(M)
    cr, sr = O.cview(Cref), O.cview(Sref)
    O.lib().orc_microbench_batch(count, np.ascontiguousarray(shapes.ravel()), modes, offT, offM, O.cview(A),
                                 O.cview(B), O.cview(M), cr, sr)
    Cref, Sref = O.as_complex(cr), O.as_complex(sr)
    dt = np.complex64 if c64 else np.complex128
    dA, dB, dM = (Q.DeviceBuffer.from_array(x.astype(dt)) for x in (A, B, M))
    dC = Q.DeviceBuffer(A.size * np.dtype(dt).itemsize)
    dS = Q.DeviceBuffer(M.size * np.dtype(dt).itemsize)
    pm = Q.Plan("mode_product", mp, c64)
    ph = Q.Plan("hole", hc, c64)
    pm.execute(dC, dM, dA)
    ph.execute(dS, dA, dB)
    Q.synchronize()
    got_C = dC.download(dt).astype(np.complex128)
    got_S = dS.download(dt).astype(np.complex128)
    tol = 3e-6 if c64 else 1e-12
    for t in range(count):
        n = int(np.prod(shapes[t]))
        sl = slice(offT[t], offT[t] + n)
        assert rel(got_C[sl], Cref[sl]) < tol, t
        nk = shapes[t][modes[t]]
        sm = slice(offM[t], offM[t] + nk * nk)
        assert rel(got_S[sm], Sref[sm]) < tol, t
