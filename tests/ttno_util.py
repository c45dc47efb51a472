"""Builders of tree tensor network operators (TTNO) for the tests, in the
layout of TensorTreeOperator (bottom (dim^2, D) with leaf matrices as columns,
upper (D_child0, ..., D)), concatenated in node-address order."""
import numpy as np


def shapes(nodes, bond):
    out = []
    for n in nodes:
        if n["bottom"]:
            d = n["dims"][0]
            out.append((d * d, bond[n["address"]]))
        else:
            out.append(tuple(bond[c] for c in n["children"]) + (bond[n["address"]],))
    return out


def random_ttno(nodes, bond, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for s in shapes(nodes, bond):
        size = int(np.prod(s))
        parts.append(rng.uniform(-1, 1, size) + 1j * rng.uniform(-1, 1, size))
    return np.concatenate(parts)


def sop_ttno(nodes, coeffs, leaf_mats):
    """Exact TTNO of sum_l c_l prod_m M_{l,m}: bond index l carries term l
    on every edge (diagonal upper tensors), the top collects c_l.
    leaf_mats(l, mode) -> d x d matrix (identity where term l does not act)."""
    L = len(coeffs)
    bond = [1 if n["parent"] < 0 else L for n in nodes]
    parts = []
    for n in nodes:
        if n["bottom"]:
            d = n["dims"][0]
            B = np.zeros((d * d, L), np.complex128)
            for l in range(L):
                B[:, l] = leaf_mats(l, n["mode"]).reshape(-1, order="F")
        else:
            C = len(n["children"])
            B = np.zeros((L,) * C + (bond[n["address"]],), np.complex128)
            for l in range(L):
                idx = (l,) * C + ((0,) if n["parent"] < 0 else (l,))
                B[idx] = coeffs[l] if n["parent"] < 0 else 1.0
        parts.append(B.reshape(-1, order="F"))
    return bond, np.concatenate(parts)


def dense_operator(nodes, bond, B):
    """Dense matrix of a TTNO over the modes in leaf order (mode 0 slowest,
    the C-order of a dense psi[x0, x1, ...]); tiny trees only."""
    sh = shapes(nodes, bond)
    offs = np.concatenate([[0], np.cumsum([int(np.prod(s)) for s in sh])])
    ops = {}
    for n in nodes:  # children first
        a = n["address"]
        T = B[offs[a]:offs[a + 1]].reshape(sh[a], order="F")
        if n["bottom"]:
            d = n["dims"][0]
            ops[a] = [T[:, l].reshape((d, d), order="F") for l in range(bond[a])]
        else:
            ch = n["children"]
            res = []
            for l in range(bond[a]):
                acc = None
                for idx in np.ndindex(*[bond[c] for c in ch]):
                    coef = T[idx + (l,)]
                    K = ops[ch[0]][idx[0]]
                    for j in range(1, len(ch)):
                        K = np.kron(K, ops[ch[j]][idx[j]])
                    acc = coef * K if acc is None else acc + coef * K
                res.append(acc)
            ops[a] = res
    return ops[len(nodes) - 1][0]
