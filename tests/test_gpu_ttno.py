"""GPU TTNO application (qtg_ttno_*) against the oracle's restatement of
TTOrepresentation / TTOcontraction / apply(): random operator tensors with
bra != ket, and the exact TTNO of an SOP, on small trees, a tree with
24 x 24 x 24 nodes and the 64-mode tree."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W
from test_oracle_ttno import leaf_matrix_fn
from ttno_util import random_ttno, sop_ttno

pytestmark = pytest.mark.gpu


def _models():
    return {
        "f4": (W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 11, c_sigma=0.1)),
        "cmf8": W.cmf8_model(),
        "f8big": (W.balanced_tree_text(8, 8, [8, 24, 24]), W.coupled_quartic(8, 5, c_sigma=0.05)),
        "sweep64": W.sweep64_model(),
    }


@pytest.mark.parametrize("name", ["f4", "cmf8", "f8big", "sweep64"])
def test_random_ttno_matches_oracle(name):
    text, H = _models()[name]
    print(f"case ttno random {name}", flush=True)
    om = O.OracleModel(text, H)
    rng = np.random.default_rng(17)
    bond = [1 if n["parent"] < 0 else int(rng.integers(1, 4)) for n in om.nodes]
    B = random_ttno(om.nodes, bond, 23)
    bra = om.fill_random(3, delta_lowest=False)
    ket = om.fill_random(4, delta_lowest=False)
    ref, e_ref = om.ttno_apply(bond, B, bra, ket)
    eng = Q.Engine(text, H)
    got, e = Q.TTNO(eng, bond, B).apply(bra, ket)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  max rel err {err:.3e}  <bra|H|ket> {e:.6e}", flush=True)
    assert err < 1e-11
    assert abs(e - e_ref) < 1e-11 * max(1.0, abs(e_ref))


@pytest.mark.parametrize("name", ["f4", "cmf8"])
def test_sop_ttno_matches_oracle(name):
    text, H = _models()[name]
    print(f"case ttno sop {name}", flush=True)
    om = O.OracleModel(text, H)
    bond, B = sop_ttno(om.nodes, H.coeff, leaf_matrix_fn(om, H))
    psi = om.fill_random(5, delta_lowest=False)
    ref, e_ref = om.ttno_apply(bond, B, psi, psi)
    eng = Q.Engine(text, H)
    got, e = Q.TTNO(eng, bond, B).apply(psi, psi)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
    assert abs(e - e_ref) < 1e-11 * abs(e_ref)
