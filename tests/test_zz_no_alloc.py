"""The no-allocation grouped hole contraction
(qtg_hole_contraction_grouped_noalloc_z) against the plan it mirrors
(qtg_hole_contraction_plan_create + qtg_plan_execute) and the oracle: two
calls in a row, the second with fewer and smaller descriptors, each equal to
the plan's result bit for bit and within 1e-12 of the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q

pytestmark = pytest.mark.gpu


def problem(seed, count, legs):
    g = np.random.default_rng(seed)
    descs = np.zeros(count, Q.HC_DESC)
    offA, offS = 0, 0
    shapes = []
    for t in range(count):
        d = g.choice(legs, size=3)
        k = int(g.integers(0, 3))
        before, after = int(np.prod(d[:k])), int(np.prod(d[k + 1:]))
        descs[t] = (offS, offA, offA, before, d[k], d[k], after, 0, 0)
        shapes.append((before, int(d[k]), after, offA, offS))
        offA += int(np.prod(d))
        offS += int(d[k]) ** 2
    A = g.normal(size=offA) + 1j * g.normal(size=offA)
    B = g.normal(size=offA) + 1j * g.normal(size=offA)
    return descs, A, B, offS, shapes


def noalloc(descs, dS, dA, dB):
    L = Q.lib()
    L.qtg_hole_contraction_grouped_noalloc_z.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                                         C.c_void_p]
    Q.check(L.qtg_hole_contraction_grouped_noalloc_z(descs.ctypes.data, len(descs), dS.ptr, dA.ptr, dB.ptr, None))


def test_noalloc_matches_plan_on_two_calls():
    results = []
    for seed, count, legs in ((1, 300, [8, 16, 24, 32, 48, 64]), (2, 40, [4, 8, 12])):
        print(f"case noalloc hole contraction seed={seed} count={count}", flush=True)
        descs, A, B, nS, shapes = problem(seed, count, legs)
        dA, dB = Q.DeviceBuffer.from_array(A), Q.DeviceBuffer.from_array(B)
        dS1 = Q.DeviceBuffer.from_array(np.full(nS, 7.0 + 7.0j))  # overwritten (accumulate = 0)
        dS2 = Q.DeviceBuffer.from_array(np.full(nS, 7.0 + 7.0j))
        noalloc(descs, dS1, dA, dB)
        got = dS1.download(np.complex128)
        Q.Plan("hole", descs).execute(dS2, dA, dB)
        Q.synchronize()
        ref_plan = dS2.download(np.complex128)
        assert np.array_equal(got, ref_plan)
        for before, n, after, oa, os_ in shapes:
            N = before * n * after
            ref = O.rhomat(A[oa:oa + N], B[oa:oa + N], n, before, after)
            assert np.max(np.abs(got[os_:os_ + n * n] - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
        results.append(got)
    assert len(results[1]) < len(results[0])


def test_noalloc_rejects_sizes_beyond_its_maximum():
    print("case noalloc limits", flush=True)
    descs = np.zeros(4097, Q.HC_DESC)
    for t in range(len(descs)):
        descs[t] = (0, 0, 0, 1, 2, 2, 1, 0, 0)
    d = Q.DeviceBuffer(64)
    L = Q.lib()
    L.qtg_hole_contraction_grouped_noalloc_z.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                                         C.c_void_p]
    assert L.qtg_hole_contraction_grouped_noalloc_z(descs.ctypes.data, len(descs), d.ptr, d.ptr, d.ptr, None) == 1
    one = np.zeros(1, Q.HC_DESC)
    one[0] = (0, 0, 0, 1, 8200, 8200, 1, 0, 0)  # 513^2 > 2^18 tiles
    assert L.qtg_hole_contraction_grouped_noalloc_z(one.ctypes.data, 1, d.ptr, d.ptr, d.ptr, None) == 1
