"""CPU checks of the mctdh job runner (qutree_amd.mctdh): the balanced tree
generator against the oracle's restatement of TreeFactory::balancedTree,
CoupledHO's terms, the shared random generator of the wavefunction jobs
(seed, create twice), and save / read of TensorTree files."""
import os

import numpy as np
import pytest

import oracle_lib as O
from qutree_amd import mctdh as M
from qutree_amd import workloads as W


def parsed(text):
    """(node / leaf lines as integers, PhysPar lines as floats): nodes are
    "n -k", leaves "dim type mode", PhysPar lines have four numbers."""
    lines = [ln.split() for ln in text.strip().splitlines() if ln.strip()]
    tree = [tuple(int(x) for x in ln) for ln in lines if len(ln) in (2, 3)]
    phys = [tuple(float(x) for x in ln) for ln in lines if len(ln) == 4]
    return tree, phys


@pytest.mark.parametrize("args", [(6, 3, 2, 1), (4, 2, 4, 0), (8, 8, 4, 0), (5, 4, 3, 2), (128, 10, 4, 4),
                                  (3, 5, 7, 0), (1, 3, 2, 0), (16, 6, 3, 1)])
def test_balanced_tree_matches_oracle(args):
    n, d, dn, di = args
    mine = M.balanced_tree(n, d, dn, di, 0, 1.0, 0.0, 1.0, 0.0)
    ref = O.balanced_tree_text(n, d, dn, di, 1.0, 0.0, 1.0, 0.0)
    assert parsed(mine) == parsed(ref)


def test_balanced_tree_yaml_width_key():
    # the reference reads the Gaussian width from "wfomoega" (mctdh.cpp:87)
    t1 = M.tree_from_job(dict(type="balanced", number_leaves=4, dimension_leaves=3, wfomega=0.3))
    t2 = M.tree_from_job(dict(type="balanced", number_leaves=4, dimension_leaves=3, wfomoega=0.3))
    assert parsed(t1)[1][0][3] == 1.0 and parsed(t2)[1][0][3] == 0.3


def test_coupled_ho_terms():
    H = M.coupled_ho(5)
    assert H.nterms == 15
    lam, om = 2000.0 / M.CM, 4000.0 / M.CM
    assert np.allclose([c.real for c in H.coeff[:2]], [1.0, 0.5 * om * om])
    pairs = [tuple(H.fmode[H.term_off[l]:H.term_off[l + 1]]) for l in range(10, 15)]
    assert pairs == [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]
    assert np.allclose([c.real for c in H.coeff[10:]], lam * lam)


def test_seeded_creates_continue_one_generator(tmp_path):
    text = M.balanced_tree(8, 6, 3, 1, 0, 1.0, 0.0, 1.0, 1.0)
    cfg = {"run": [{"job": "tree", "type": "compact", "tree": text}, {"job": "hamiltonian", "name": "coupled_ho"},
                   {"job": "seed", "seed": 11},
                   {"job": "wavefunction", "name": "A", "type": "create", "Hartree": False},
                   {"job": "wavefunction", "name": "B", "type": "create", "Hartree": False}]}
    st = M.run(cfg)
    om = O.OracleModel(text)
    # the first state is TensorTree(mt19937(11), tree) ...
    ref = om.fill_random(11, delta_lowest=False)
    assert np.max(np.abs(st.wavefunctions["A"] - ref)) < 1e-13
    # ... and the generator moved on by two 32-bit outputs per uniform draw
    # (generate_canonical<double, 53> on a 32-bit engine) of every upper node
    draws = sum(int(np.prod(n["dims"])) for n in om.nodes if not n["bottom"])
    assert st.drawn == 2 * 2 * draws
    assert np.max(np.abs(st.wavefunctions["B"] - ref)) > 1e-3
    # a new seed job restarts the stream
    st2 = M.run({"run": cfg["run"][:4]})
    assert np.array_equal(st2.wavefunctions["A"], st.wavefunctions["A"])


def test_save_and_read_last_tree(tmp_path):
    text = M.balanced_tree(4, 5, 3, 0, 0, 1.0, 0.0, 1.0, 1.0)
    f = str(tmp_path / "psi.dat")
    base = [{"job": "tree", "type": "compact", "tree": text}, {"job": "hamiltonian", "name": "coupled_ho"},
            {"job": "seed", "seed": 3}]
    st = M.run({"run": base + [{"job": "wavefunction", "name": "Psi", "type": "create"},
                               {"job": "wavefunction", "name": "Psi", "type": "save", "filename": f}]})
    st2 = M.run({"run": base + [{"job": "wavefunction", "name": "Psi", "type": "read", "filename": f}]})
    assert np.array_equal(st.wavefunctions["Psi"], st2.wavefunctions["Psi"])
    # a file with several appended trees: read takes the last one
    from qutree_amd import tensor_tree_bytes
    other = st.wavefunctions["Psi"] * 2.0
    with open(f, "ab") as fh:
        fh.write(tensor_tree_bytes(text, other))
    st3 = M.run({"run": base + [{"job": "wavefunction", "name": "Psi", "type": "read", "filename": f}]})
    assert np.array_equal(st3.wavefunctions["Psi"], other)


def test_jobs_off_the_path_are_refused():
    with pytest.raises(ValueError):
        M.run({"run": [{"job": "potential", "name": "ch3"}]})
    with pytest.raises(ValueError):
        M.run({"run": [{"job": "overlaps"}]})
