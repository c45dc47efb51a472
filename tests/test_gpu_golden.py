"""The HIP path against the committed golden fixtures: the node-level C ABI
on the numpy-einsum node products and hole contractions, and the engine's
H|Psi> sweep on the 8-mode tree (dense brute force)."""
import numpy as np
import pytest

import golden_data as G
import qutree_amd as Q

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def dev(z):
    return Q.DeviceBuffer.from_array(np.ascontiguousarray(z.ravel(order="F"), np.complex128))


@pytest.mark.parametrize("i", range(6))
def test_node_abi_matches_golden(i):
    g = G.node_ops()
    b, a, c = (int(x) for x in g["shapes"][i])
    print(f"case golden node ops shape {(b, a, c)}", flush=True)
    B, M, C0, A = g[f"{i}_B"], g[f"{i}_M"], g[f"{i}_C0"], g[f"{i}_A"]
    dB, dM = dev(B), dev(M)
    for op, fn in (("N", "qtg_matvec_z"), ("T", "qtg_tmatvec_z"), ("H", "qtg_ctmatvec_z")):
        for add, key in ((0, "set"), (1, "add")):
            dC = dev(C0)
            Q.check(getattr(Q.lib(), fn)(dC.ptr, dB.ptr, dM.ptr, a, b, c, add, None))
            got = dC.download(np.complex128).reshape((b, a, c), order="F")
            assert rel(got, g[f"{i}_C_{op}_{key}"]) < 1e-12, (op, add)
    dA, dS = dev(A), dev(np.ones((a, a), np.complex128))
    Q.check(Q.lib().qtg_rhomat_z(dA.ptr, dB.ptr, dS.ptr, a, b, c, None))
    assert rel(dS.download(np.complex128).reshape(a, a, order="F"), g[f"{i}_S"]) < 1e-12


@pytest.mark.parametrize("path,compress", [(1, True), (1, False), (0, True)])
def test_engine_sweep8_matches_golden(path, compress):
    text, H, d = G.sweep8()
    print(f"case golden sweep8 path={path} compress={compress}", flush=True)
    eng = Q.Engine(text, H, compress=compress, path=path)
    eng.set_psi(d["psi"])
    eng.derivative()
    e = eng.energy()
    assert abs(e - float(d["energy"])) < 1e-12 * abs(float(d["energy"]))
    assert rel(eng.get(1), d["apply"]) < 1e-11
    assert rel(eng.get(0), d["derivative"]) < 1e-10
