"""QuTree's binary tree files through the C ABI (host only): TensorTree
("TTre"; TensorTree_Implementation.h:79-132, Tensor_Implementation.h:204-258,
TensorShape.cpp:68-111) and MatrixTree / SparseMatrixTree ("MTre", "SMTr";
MatrixTree.cpp:52-90, SparseMatrixTree.cpp:27-70, Matrix_Implementation.h:
365-416).  The expected bytes are assembled here field by field with
``struct`` from the reference's write functions; no file written by the
reference exists to compare with."""
import struct

import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W


def expected_ttre(om, psi):
    out = [b"TTre", struct.pack("<i", om.nnodes)]
    for n in om.nodes:
        dims = n["dims"]
        size = int(np.prod(dims))
        out += [b"TENS", b"TDIM", struct.pack("<i", len(dims)), struct.pack(f"<{len(dims)}i", *dims),
                struct.pack("<i", 16), np.ascontiguousarray(psi[n["off"]:n["off"] + size]).tobytes()]
    return b"".join(out)


@pytest.mark.parametrize("model", ["cmf8", "eig8x3"])
def test_tensor_tree_bytes_and_round_trip(model):
    text = W.cmf8_model()[0] if model == "cmf8" else W.eig8_model(3)[0]
    om = O.OracleModel(text)
    psi = om.fill_random(5, delta_lowest=False)
    data = Q.tensor_tree_bytes(text, psi)
    assert data == expected_ttre(om, psi)
    back, used = Q.tensor_tree_from_bytes(text, data)
    assert used == len(data)
    assert np.array_equal(back, psi)


def test_tensor_tree_appended_file():
    """TensorTree::write(file, append) leaves trees back to back."""
    text = W.cmf8_model()[0]
    om = O.OracleModel(text)
    a, b = om.fill_random(1, True), om.fill_random(2, False)
    data = Q.tensor_tree_bytes(text, a) + Q.tensor_tree_bytes(text, b)
    x, n1 = Q.tensor_tree_from_bytes(text, data)
    y, n2 = Q.tensor_tree_from_bytes(text, data, n1)
    assert n1 + n2 == len(data)
    assert np.array_equal(x, a) and np.array_equal(y, b)


def test_tensor_tree_errors():
    text = W.cmf8_model()[0]
    om = O.OracleModel(text)
    psi = om.fill_random(1, True)
    data = Q.tensor_tree_bytes(text, psi)
    with pytest.raises(Q.QtgError, match="TTre"):
        Q.tensor_tree_from_bytes(text, b"MTre" + data[4:])
    with pytest.raises(Q.QtgError, match="truncated"):
        Q.tensor_tree_from_bytes(text, data[:-1])
    other = W.balanced_tree_text(8, 8, [5, 4, 4])  # bottom nodes with 5 SPFs
    with pytest.raises(Q.QtgError, match="dimension 1"):
        Q.tensor_tree_from_bytes(other, data)
    fewer = W.balanced_tree_text(4, 8, [4, 4])
    with pytest.raises(Q.QtgError, match="tensors for a tree of"):
        Q.tensor_tree_from_bytes(fewer, data)
    with pytest.raises(Q.QtgError, match="16 bytes|complex<double>"):
        # a complex<float> file (coefficient size 8) is refused
        first = expected_ttre(om, psi)
        hdr = 4 + 4 + 4 + 4 + 4 + 4 * len(om.nodes[0]["dims"])
        bad = first[:hdr] + struct.pack("<i", 8) + first[hdr + 4:]
        Q.tensor_tree_from_bytes(text, bad)


@pytest.mark.parametrize("sparse", [False, True])
def test_matrix_tree_bytes_and_round_trip(sparse):
    text, H = W.cmf8_model()
    om = O.OracleModel(text, H)
    psi = om.fill_random(3, False)
    om.build(psi)
    mats = [om.rho(a) if om.nodes[a]["parent"] >= 0 else np.eye(1, dtype=np.complex128) for a in range(om.nnodes)]
    mats.append(np.arange(6, dtype=np.complex128).reshape(2, 3) * (1 + 2j))  # rectangular
    data = Q.matrix_tree_bytes(mats, sparse=sparse)
    exp = [b"SMTr" if sparse else b"MTre", struct.pack("<i", len(mats))]
    for m in mats:
        exp += [b"MATR", struct.pack("<QQ", *m.shape), struct.pack("<i", 16),
                np.asfortranarray(m).ravel(order="F").astype(np.complex128).tobytes()]
    assert data == b"".join(exp)
    back, used = Q.matrix_tree_from_bytes(data, sparse=sparse)
    assert used == len(data) and len(back) == len(mats)
    for x, y in zip(back, mats):
        assert x.shape == y.shape and np.array_equal(x, y)
    with pytest.raises(Q.QtgError, match="SMTr|MTre"):
        Q.matrix_tree_from_bytes(data, sparse=not sparse)
    with pytest.raises(Q.QtgError, match="truncated"):
        Q.matrix_tree_from_bytes(data[:-8], sparse=sparse)


def test_matrix_tree_read_limits():
    """The reader stores no more matrices or coefficients than it was given
    room for."""
    import ctypes as C

    mats = [np.eye(3, dtype=np.complex128), np.ones((2, 2), np.complex128)]
    data = Q.matrix_tree_bytes(mats)
    view = (C.c_char * len(data)).from_buffer_copy(data)
    dims = np.zeros(4, np.int64)
    coef = np.zeros(2 * 13)
    cnt = C.c_int64()
    assert Q.lib().qtg_matrix_tree_read(0, view, len(data), 1, dims, coef, 13, C.byref(cnt)) == -3
    assert cnt.value == 2
    assert Q.lib().qtg_matrix_tree_read(0, view, len(data), 2, dims, coef, 12, C.byref(cnt)) == -3
    assert Q.lib().qtg_matrix_tree_read(0, view, len(data), 2, dims, coef, 13, C.byref(cnt)) == len(data)
    assert list(dims) == [3, 3, 2, 2]


def test_tensor_tree_bytes_match_oracle_writer():
    """The oracle's own restatement of TensorTree::write produces the same
    bytes, and the C ABI reads them back."""
    text, _, _ = W.eig8_model(3)
    om = O.OracleModel(text)
    psi = om.fill_random(9, delta_lowest=False)
    data = om.tensor_tree_bytes(psi)
    assert data == Q.tensor_tree_bytes(text, psi)
    back, used = Q.tensor_tree_from_bytes(text, data)
    assert used == len(data) and np.array_equal(back, psi)
