"""Gauge helpers for tree states on the host (numpy): orthonormalise a node by
QR and push R into the parent, which leaves the represented state unchanged.
The engine's energy, derivative and gradient assume orthonormal non-top
nodes; these restore that gauge after an update such as a gradient step.
`nodes` is Engine.nodes (dims, off, parent, childIdx per node address)."""
import numpy as np


def node_view(psi, node):
    dims = node["dims"]
    N = int(np.prod(dims))
    return psi[node["off"]:node["off"] + N].reshape(dims, order="F")


def push_up(psi, nodes, a):
    """QR of node a (rows: all indices but the last), R into the parent on
    a's leg.  Returns the parent address (or -1 at the top)."""
    n = nodes[a]
    p = n["parent"]
    if p < 0:
        return -1
    dims = n["dims"]
    N, k = int(np.prod(dims)), dims[-1]
    A = psi[n["off"]:n["off"] + N].reshape((N // k, k), order="F")
    Q, R = np.linalg.qr(A)
    psi[n["off"]:n["off"] + N] = Q.reshape(-1, order="F")
    P = node_view(psi, nodes[p]).copy()
    leg = n["childIdx"]
    P = np.moveaxis(np.tensordot(R, P, axes=([1], [leg])), 0, leg)
    pn = nodes[p]
    psi[pn["off"]:pn["off"] + int(np.prod(pn["dims"]))] = P.reshape(-1, order="F")
    return p


def canonicalize_path(psi, nodes, a):
    """Orthonormalise node a and its ancestors below the top (in place)."""
    while a >= 0 and nodes[a]["parent"] >= 0:
        a = push_up(psi, nodes, a)
    return psi


def canonicalize_all(psi, nodes):
    """Orthonormalise every non-top node, children before parents (in place)."""
    for a in range(len(nodes)):
        if nodes[a]["parent"] >= 0:
            push_up(psi, nodes, a)
    return psi


def gradient_descent(engine, psi, steps, eta):
    """Plain gradient descent on the energy functional: A <- A - eta dE/dA*
    at every node, then QR re-orthonormalisation.  Returns the final state
    and the energy before every step and after the last one."""
    energies = []
    for _ in range(steps):
        engine.set_psi(psi)
        g, E = engine.gradient()
        energies.append(E)
        psi = canonicalize_all(psi - eta * g, engine.nodes)
    engine.set_psi(psi)
    engine.derivative()
    energies.append(engine.energy())
    return psi, energies
