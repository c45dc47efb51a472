"""GPU parity of the tree engine (HamiltonianRepresentation::build, Apply,
LayerDerivative, Derivative, energy) against the FP64 oracle on the same
seeded trees, Hamiltonians and states.

complex128: 1e-10 relative per node (random full-rank states keep rho^-1
well conditioned); complex64 bulk: energy within 1e-6 relative of FP64."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu


def _models():
    return {
        "f4": (W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 11, c_sigma=0.1)),
        "f8quartic": (W.balanced_tree_text(8, 2, [2, 3, 3]), W.coupled_quartic(8, 12, c_sigma=0.1)),
        "cmf8": W.cmf8_model(),
        "f16": (W.balanced_tree_text(16, 16, [8, 10, 11, 13]), W.coupled_ho(16, 16, c_sigma=0.01,
                                                                             one_mode_matrix=True)),
        "sweep64": W.sweep64_model(),
        # upper nodes of 24 x 24 x 24 (221 KB in complex128): beyond the
        # LDS-resident kernels, grouped path only
        "f8big": (W.balanced_tree_text(8, 8, [8, 24, 24]), W.coupled_quartic(8, 5, c_sigma=0.05)),
    }


def node_rel(eng, a_vec, b_vec):
    """Worst per-node relative difference.  A node whose SPFs span its whole
    configuration space has (1-P) = 0, so its derivative is rounding noise on
    both sides; the scale floor (1e-6 of the largest entry) covers those."""
    floor = 1e-6 * np.max(np.abs(b_vec))
    worst = 0.0
    for a in range(eng.nnodes):
        x = eng.node_tensor(a_vec, a)
        y = eng.node_tensor(b_vec, a)
        worst = max(worst, np.max(np.abs(x - y)) / max(floor, np.max(np.abs(y))))
    return worst


def derivative_err(eng, om, got, ref, app_ref):
    """Per-node derivative difference in units of the projector input:
    |d_gpu - d_ref| / (max|Apply_v| * ||rho_v^-1||_2).  (1-P) subtracts the
    component of Apply inside span(A_v), so both implementations lose the
    same leading digits; this scale is what the rounding of either side is
    proportional to."""
    worst = 0.0
    for a in range(eng.nnodes):
        d = np.max(np.abs(eng.node_tensor(got, a) - om.node_tensor(ref, a)))
        scale = np.max(np.abs(om.node_tensor(app_ref, a)))
        if eng.nodes[a]["parent"] >= 0:
            scale *= np.linalg.norm(om.rho_inv(a), 2)
        worst = max(worst, d / max(scale, 1e-300))
    return worst


@pytest.mark.parametrize("path", [1, 0])
@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("name", ["f4", "f8quartic", "cmf8", "f16", "sweep64", "f8big"])
def test_derivative_matches_oracle(name, compress, path):
    if name == "f8big" and path == 0:
        pytest.skip("node tensors beyond the LDS-resident kernels")
    text, H = _models()[name]
    print(f"case engine derivative {name} compress={compress} path={path}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(3, delta_lowest=False)
    ref = om.derivative(psi)
    e_ref = om.energy(psi)
    app_ref = om.apply_all(psi)
    eng = Q.Engine(text, H, compress=compress, path=path)
    eng.set_psi(psi)
    eng.derivative()
    got = eng.get(0)
    assert node_rel(eng, eng.get(1), app_ref) < 1e-11
    err = derivative_err(eng, om, got, ref, app_ref)
    print(f"  derivative err {err:.3e}", flush=True)
    assert err < 1e-13
    assert abs(eng.energy() - e_ref) < 1e-12 * abs(e_ref)


@pytest.mark.parametrize("path", [1, 0])
@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("name", ["f4", "cmf8", "f16", "f8big"])
def test_representation_matrices_match_oracle(name, compress, path):
    if name == "f8big" and path == 0:
        pytest.skip("node tensors beyond the LDS-resident kernels")
    text, H = _models()[name]
    print(f"case engine matrices {name} compress={compress} path={path}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(4, delta_lowest=False)
    om.build(psi)
    eng = Q.Engine(text, H, compress=compress, path=path)
    eng.set_psi(psi)
    eng.build()
    Q.synchronize()
    rng = np.random.default_rng(0)
    for a in range(eng.nnodes):
        if eng.nodes[a]["parent"] < 0:
            continue
        r = om.rho(a)
        assert np.max(np.abs(eng.matrix(0, 0, a) - r)) < 1e-12 * max(1, np.max(np.abs(r)))
        ri = om.rho_inv(a)
        assert np.max(np.abs(eng.matrix(1, 0, a) - ri)) < 1e-9 * np.max(np.abs(ri))
    for l in rng.choice(H.nterms, size=min(H.nterms, 12), replace=False):
        for a in om.sparse_tree(int(l)):
            if eng.nodes[a]["parent"] < 0:
                continue
            h = om.hmat(int(l), a)
            g = om.hole(int(l), a)
            assert np.max(np.abs(eng.matrix(2, int(l), a) - h)) < 1e-11 * max(1, np.max(np.abs(h)))
            assert np.max(np.abs(eng.matrix(3, int(l), a) - g)) < 1e-11 * max(1, np.max(np.abs(g)))


def test_hartree_start_energy_and_top_derivative():
    # Hartree product start: rho = diag(1,0,..) so rho^-1 carries 1e10 in the
    # empty directions; the gauge-invariant energy and the top-node derivative
    # stay comparable.
    text, H = W.cmf8_model()
    print("case engine hartree cmf8", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(2024, delta_lowest=True)
    ref = om.derivative(psi)
    eng = Q.Engine(text, H)
    eng.set_psi(psi)
    eng.derivative()
    top = eng.nnodes - 1
    got = eng.get(0)
    x, y = eng.node_tensor(got, top), om.node_tensor(ref, top)
    assert np.max(np.abs(x - y)) < 1e-12 * np.max(np.abs(y))
    assert abs(eng.energy() - om.energy(psi)) < 1e-12 * abs(om.energy(psi))


@pytest.mark.parametrize("name", ["cmf8", "sweep64"])
def test_complex64_bulk_error_against_fp64(name):
    text, H = _models()[name]
    print(f"case engine c64 {name}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(5, delta_lowest=False)
    ref = om.derivative(psi)
    e64 = om.energy(psi)
    eng = Q.Engine(text, H, precision=1)
    eng.set_psi(psi)
    eng.derivative()
    e32 = eng.energy()
    app_ref = om.apply_all(psi)
    err = derivative_err(eng, om, eng.get(0), ref, app_ref)
    print(f"  energy rel err {abs(e32 - e64) / abs(e64):.3e}  derivative err {err:.3e}", flush=True)
    assert abs(e32 - e64) < 1e-6 * abs(e64)
    assert err < 1e-5


@pytest.mark.parametrize("name", ["cmf8", "sweep64"])
def test_split_precision_energy_against_fp64(name):
    """complex64 bulk with the large-coefficient terms (and the densities) in
    FP64: the energy error drops from ~1e-8 (all-FP32 bulk) to ~1e-10 on
    random states (1e-12 on the 64-mode tree), checked against 1e-9, ten
    times below the 1e-8 relative-energy target of the reduced forms."""
    text, H = _models()[name]
    print(f"case engine split precision {name}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(5, delta_lowest=False)
    ref = om.derivative(psi)
    e64 = om.energy(psi)
    eng = Q.Engine(text, H, precision=2)
    eng.set_psi(psi)
    eng.derivative()
    e = eng.energy()
    app_ref = om.apply_all(psi)
    err = derivative_err(eng, om, eng.get(0), ref, app_ref)
    print(f"  energy rel err {abs(e - e64) / abs(e64):.3e}  derivative err {err:.3e}", flush=True)
    assert abs(e - e64) < 1e-9 * abs(e64)
    assert err < 1e-5


def test_repeated_derivative_is_deterministic():
    text, H = W.sweep64_model()
    print("case engine determinism sweep64", flush=True)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(9, delta_lowest=False)
    eng.set_psi(psi)
    eng.derivative()
    a = eng.get(0)
    eng.derivative()
    b = eng.get(0)
    assert np.array_equal(a, b)


def test_graph_replays_match_eager():
    """The captured sweep is replayed several times with new states and eager
    work in between; every replay must equal the eager result bit for bit."""
    text, H = W.sweep64_model()
    print("case engine graph replays sweep64", flush=True)
    g = Q.Engine(text, H, graph=True)
    e = Q.Engine(text, H, graph=False)
    for k, (seed, phase) in enumerate([(1, 1.0), (2, 1.0), (3, 1.0), (3, -1j), (4, -1j)]):
        psi = e.fill_random(seed, delta_lowest=False)
        g.set_psi(psi)
        e.set_psi(psi)
        g.derivative(phase)
        e.derivative(phase)
        a, b = g.get(0), e.get(0)
        assert np.array_equal(a, b), k
        assert g.energy() == e.energy()
        # eager work on the graph engine between replays
        g.build()
        g.synchronize()


@pytest.mark.parametrize("graph", [False, True])
def test_inline_eigensolver_equals_side_stream(graph):
    """heev_side 0 runs the density eigensolver in line on the engine stream
    instead of on the side stream: the same kernels on the same data, so the
    derivative and the energy agree bit for bit."""
    text, H = W.sweep64_model()
    print(f"case engine heev_side 0 vs 1 graph={graph}", flush=True)
    a = Q.Engine(text, H, graph=graph, heev_side=0)
    b = Q.Engine(text, H, graph=graph, heev_side=1)
    for seed in (5, 6):
        psi = a.fill_random(seed, delta_lowest=False)
        a.set_psi(psi)
        b.set_psi(psi)
        a.derivative()
        b.derivative()
        assert np.array_equal(a.get(0), b.get(0)), seed
        assert a.energy() == b.energy()


@pytest.mark.parametrize("amp", [0.3, 0.003])
@pytest.mark.parametrize("name", ["cmf8", "f16"])
def test_derivative_non_orthonormal_state(name, amp):
    """Lower node tensors A (I + amp R) with R random.  amp = 0.3: S = A^H A
    is far from the identity and the projector solve eliminates (the
    Newton-Schulz steps need n max|I - S| < 1/2); amp = 0.003: several
    Newton-Schulz steps.  Same bounds as the orthonormal cases for Apply and
    the derivative; the energy (the top node's expectation value) assumes
    orthonormal lower nodes and is not compared."""
    text, H = _models()[name]
    print(f"case engine derivative non-orthonormal {name} amp={amp}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(5, delta_lowest=False)
    rng = np.random.default_rng(11)
    for nd in om.nodes:
        if nd["parent"] < 0:
            continue
        dims = nd["dims"]
        size, k = int(np.prod(dims)), int(dims[-1])
        X = psi[nd["off"]:nd["off"] + size].reshape(-1, k, order="F")
        R = rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k))
        psi[nd["off"]:nd["off"] + size] = (X @ (np.eye(k) + amp * R)).ravel(order="F")
    ref = om.derivative(psi)
    app_ref = om.apply_all(psi)
    eng = Q.Engine(text, H)
    eng.set_psi(psi)
    eng.derivative()
    got = eng.get(0)
    assert node_rel(eng, eng.get(1), app_ref) < 1e-11
    err = derivative_err(eng, om, got, ref, app_ref)
    print(f"  derivative err {err:.3e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("name", ["sweep64", "cmf8"])
def test_fused_hole_layers_match_two_launch_layers(name):
    """fuse_holes 1 (default: each hole layer one launch, ket tensors built
    in LDS beside their hole contractions, bras of all layers first) and 0
    (a mode-product and a hole-contraction launch per layer) compute the same
    products with different K chunking: derivative within 1e-12 of its scale,
    energy within 1e-13 relative."""
    text, H = W.sweep64_model() if name == "sweep64" else W.cmf8_model()
    print(f"case engine fuse_holes 1 vs 0 {name}", flush=True)
    a = Q.Engine(text, H, fuse_holes=1)
    b = Q.Engine(text, H, fuse_holes=0)
    for seed in (3, 4):
        psi = a.fill_random(seed, delta_lowest=False)
        a.set_psi(psi)
        b.set_psi(psi)
        a.derivative()
        b.derivative()
        da, db = a.get(0), b.get(0)
        scale = np.max(np.abs(db))
        err = np.max(np.abs(da - db)) / scale
        print(f"  seed {seed} max |dD| / max |D| {err:.2e}  energy {a.energy():.15f} {b.energy():.15f}", flush=True)
        assert err < 1e-12
        assert abs(a.energy() - b.energy()) <= 1e-13 * abs(b.energy())
