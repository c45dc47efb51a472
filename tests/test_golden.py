"""The oracle against the committed golden fixtures: node products and hole
contractions (numpy einsum of the LA_lib.f definitions) and one H|Psi> sweep
on an 8-mode tree (dense brute force over all 4^8 amplitudes)."""
import numpy as np
import pytest

import golden_data as G
import oracle_lib as O


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


@pytest.mark.parametrize("i", range(6))
def test_oracle_node_ops_match_golden(i):
    g = G.node_ops()
    b, a, c = (int(x) for x in g["shapes"][i])
    B, M, C0, A = g[f"{i}_B"], g[f"{i}_M"], g[f"{i}_C0"], g[f"{i}_A"]
    fns = {"N": O.lib().orc_matvec, "T": O.lib().orc_tmatvec, "H": O.lib().orc_ctmatvec}
    for op, fn in fns.items():
        for add, key in ((0, "set"), (1, "add")):
            flat = C0.ravel(order="F").copy()
            f = O.cview(flat)
            fn(f, O.cview(B.ravel(order="F")), O.cview(M.ravel(order="F")), a, b, c, add)
            got = O.as_complex(f).reshape((b, a, c), order="F")
            assert rel(got, g[f"{i}_C_{op}_{key}"]) < 1e-13, (op, add)
    S = O.rhomat(A.ravel(order="F"), B.ravel(order="F"), a, b, c).reshape(a, a, order="F")
    assert rel(S, g[f"{i}_S"]) < 1e-13


def test_oracle_sweep8_matches_golden():
    text, H, d = G.sweep8()
    m = O.OracleModel(text, H)
    psi = d["psi"]
    m.build(psi)
    assert abs(m.energy(psi) - float(d["energy"])) < 1e-12 * abs(float(d["energy"]))
    assert rel(m.apply_all(psi), d["apply"]) < 1e-11
    assert rel(m.derivative(psi), d["derivative"]) < 1e-10
