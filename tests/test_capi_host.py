"""CPU checks of the product library: every symbol of include/qutree_amd.h is
exported, and the host-side logic (tree parsing, random states, work-item
planning) agrees with the oracle and with an independent count."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "qutree_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s+(qtg_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) > 30
    L = Q.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _host():
    L = Q.lib()
    L.qtg_host_psi_size.argtypes = [C.c_char_p]
    L.qtg_host_psi_size.restype = C.c_int64
    L.qtg_host_fill_random.argtypes = [C.c_char_p, C.c_uint64, C.c_int, np.ctypeslib.ndpointer(np.float64)]
    L.qtg_host_plan_counts.argtypes = [C.c_char_p, C.c_int64] + [np.ctypeslib.ndpointer(np.float64)] + \
        [np.ctypeslib.ndpointer(np.int64)] * 4 + [np.ctypeslib.ndpointer(np.float64), C.c_int,
                                                  C.POINTER(C.c_double), np.ctypeslib.ndpointer(np.int64)]
    return L


@pytest.mark.parametrize("model", ["cmf8", "sweep64", "ref8"])
def test_random_state_matches_oracle(model):
    if model == "cmf8":
        text, H = W.cmf8_model()
    elif model == "sweep64":
        text, H = W.sweep64_model()
    else:
        text = O.balanced_tree_text(8, 2, 2)
    L = _host()
    n = L.qtg_host_psi_size(text.encode())
    m = O.OracleModel(text)
    assert n == m.psi_size
    for delta in (True, False):
        out = np.zeros(2 * n)
        assert L.qtg_host_fill_random(text.encode(), 77, int(delta), out) == 0
        ref = m.fill_random(77, delta_lowest=delta)
        # same mt19937 stream; Gram-Schmidt rounding differs between the two builds
        assert np.max(np.abs(out.view(np.complex128) - ref)) < 1e-11


def test_bad_tree_text_is_an_error():
    L = _host()
    assert L.qtg_host_psi_size(b"4 -2\n 3 0 0\n") == -1
    assert L.qtg_host_psi_size(b"garbage") == -1


@pytest.mark.parametrize("model", ["cmf8", "sweep64", "small"])
def test_plan_counts_against_independent_count(model):
    if model == "cmf8":
        text, H = W.cmf8_model()
    elif model == "sweep64":
        text, H = W.sweep64_model()
    else:
        text, H = W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 1)
    m = O.OracleModel(text, H)
    L = _host()
    c, toff, fm, fk, fp, pool = H.arrays()
    flops = C.c_double()
    items = np.zeros(3, np.int64)
    assert L.qtg_host_plan_counts(text.encode(), H.nterms, c, toff, fm, fk, fp, pool, 0, C.byref(flops),
                                  items) == 0
    # independent: one represent + one hole item per (term, active non-top node),
    # one density item per non-top node, one apply term per (term, active node)
    nodes = m.nodes
    pairs = 0
    apply_terms = 0
    fl = 0.0
    for l in range(H.nterms):
        act = set(m.sparse_tree(l))
        for a in act:
            n = nodes[a]
            N = int(np.prod(n["dims"]))
            nv = n["dims"][-1]
            apply_terms += 1
            if n["parent"] < 0:
                fl += sum(8.0 * N * nodes[c]["dims"][-1] for c in n["children"] if c in act)
                continue
            pairs += 1
            # represent
            if n["bottom"]:
                fl += 8.0 * N * n["dims"][0]
            else:
                fl += sum(8.0 * N * nodes[c]["dims"][-1] for c in n["children"] if c in act)
            fl += 8.0 * N * nv
            # hole
            p = nodes[n["parent"]]
            Np = int(np.prod(p["dims"]))
            fl += sum(8.0 * Np * nodes[s]["dims"][-1] for s in p["children"] if s != a and s in act)
            if p["parent"] >= 0:
                fl += 8.0 * Np * p["dims"][-1]
            fl += 8.0 * Np * nv
            # apply
            if n["bottom"]:
                fl += 8.0 * N * n["dims"][0]
            else:
                fl += sum(8.0 * N * nodes[c]["dims"][-1] for c in n["children"] if c in act)
            fl += 8.0 * N * nv
    nontop = [n for n in nodes if n["parent"] >= 0]
    for n in nontop:
        p = nodes[n["parent"]]
        Np = int(np.prod(p["dims"]))
        fl += 8.0 * Np * n["dims"][-1] + (8.0 * Np * p["dims"][-1] if p["parent"] >= 0 else 0.0)
        fl += 4 * 8.0 * int(np.prod(n["dims"])) * n["dims"][-1]
    assert items[0] == 2 * pairs + len(nontop)
    assert items[2] == apply_terms
    assert abs(flops.value - fl) < 1e-9 * fl
    if model == "sweep64":
        assert pairs == 22752  # 64 + 2016 terms over the 64-mode tree


def _subtree_modes(nodes):
    sub = {}
    for n in nodes:  # children precede parents
        if n["bottom"]:
            sub[n["address"]] = {n["mode"]}
        else:
            sub[n["address"]] = set().union(*(sub[c] for c in n["children"]))
    return sub


@pytest.mark.parametrize("model", ["cmf8", "sweep64", "quartic16"])
def test_compressed_plan_counts_distinct_restrictions(model):
    """With compression one represent item exists per distinct restriction of
    a term to a node's subtree, and one hole item per distinct restriction to
    its complement (plus the density matrix, the empty restriction)."""
    if model == "cmf8":
        text, H = W.cmf8_model()
    elif model == "sweep64":
        text, H = W.sweep64_model()
    else:
        text, H = W.balanced_tree_text(16, 8, [4, 6, 6, 8]), W.coupled_quartic(16, 3)
    m = O.OracleModel(text, H)
    sub = _subtree_modes(m.nodes)
    # term factors as (mode, kind, pool values) -> hashable per mode, merged per mode in order
    terms = []
    for l in range(H.nterms):
        per = {}
        for f in range(H.term_off[l], H.term_off[l + 1]):
            vals = tuple(H.pool[2 * H.fpool[f]: 2 * H.fpool[f] + 4]) if H.fkind[f] == W.FACTOR_POLY else ()
            per.setdefault(H.fmode[f], []).append((H.fkind[f], vals))
        terms.append({k: tuple(v) for k, v in per.items()})
    nrep = nhole = 0
    for n in m.nodes:
        if n["parent"] < 0:
            continue
        a = n["address"]
        ins, outs = set(), {()}
        for l, t in enumerate(terms):
            if not (set(t) & sub[a]):
                continue
            ins.add(tuple(sorted((k, v) for k, v in t.items() if k in sub[a])))
            outs.add(tuple(sorted((k, v) for k, v in t.items() if k not in sub[a])))
        nrep += len(ins)
        nhole += len(outs)
    L = _host()
    c, toff, fm, fk, fp, pool = H.arrays()
    flops = C.c_double()
    items = np.zeros(3, np.int64)
    assert L.qtg_host_plan_counts(text.encode(), H.nterms, c, toff, fm, fk, fp, pool, 1, C.byref(flops),
                                  items) == 0
    assert items[0] == nrep + nhole
    if model == "sweep64":
        assert nrep + nhole < 45630 // 4


def seam_functions():
    text = open(os.path.join(ROOT, "include", "qutree_seam.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:void|int)\s+([a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_library_exports_the_fortran_seam():
    names = seam_functions()
    assert set(names) == {"matvec_", "tmatvec_", "ctmatvec_", "rmatvec_", "trmatvec_", "rhomat_",
                          "qtg_seam_status"}
    L = Q.lib()
    assert all(hasattr(L, n) for n in names)


# The reference's own declarations of the Fortran kernels, as QuTree's call
# sites see them (include/Core/Tensor_Implementation.h:450-459,
# include/Core/Matrix_Implementation.h:454-461).  A translation unit that
# declares them and then includes the seam header compiles only if every
# prototype agrees, and links only if the library defines each one.
REFERENCE_SEAM = """
extern "C" {
void matvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
void ctmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
void rmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
void rhomat_(double *Bra, double *Ket, double *M, int *a, int *b, int *c);
void tmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
void trmatvec_(double *C, double *B, double *mat, int *a, int *b, int *c, int *add);
}
#include "qutree_seam.h"
// call sites shaped like Tensor_Implementation.h:566 and Matrix_Implementation.h:478
void call_sites(double* C, double* B, double* A, int a, int b, int c, int add) {
  matvec_(C, B, A, &a, &b, &c, &add);
  ctmatvec_(C, B, A, &a, &b, &c, &add);
  rmatvec_(C, B, A, &a, &b, &c, &add);
  tmatvec_(C, A, B, &a, &b, &c, &add);
  trmatvec_(C, A, B, &a, &b, &c, &add);
  rhomat_(B, A, C, &a, &b, &c);
}
int main() { return qtg_seam_status(); }
"""


def test_reference_call_sites_compile_and_link_against_the_seam(tmp_path):
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "seam_user.cpp"
    src.write_text(REFERENCE_SEAM)
    exe = tmp_path / "seam_user"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           Q.LIB_PATH, "-Wl,--unresolved-symbols=ignore-in-shared-libs"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # a prototype that disagrees with the seam is a compile error
    bad = tmp_path / "seam_bad.cpp"
    bad.write_text(REFERENCE_SEAM.replace("void matvec_(double *C, double *B, double *mat, int *a, int *b, int *c, "
                                          "int *add);", "void matvec_(double *C, double *B, double *mat, int *add, "
                                          "int *a, int *b, long *c);", 1))
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(bad)],
                       capture_output=True, text=True)
    assert r.returncode != 0


def test_shape_guards_reject_32bit_overflow_before_any_launch():
    """Hole contractions index K = before*after and a*before*n in 32 bits;
    mode products index matrices in 32 bits: shapes beyond that are QTG_ESHAPE
    (checked on the host, before any device call)."""
    L = Q.lib()
    fake = 4096  # never dereferenced: the shape check comes first
    big = 1 << 16
    assert L.qtg_hole_contraction_z(fake, fake, fake, big, 2, 2, big // 2, 0, None) == 1   # before*after = 2^31
    assert L.qtg_hole_contraction_z(fake, fake, fake, 1 << 20, 1 << 11, 2, 1, 0, None) == 1  # before*nA = 2^31
    assert L.qtg_hole_contraction_z(fake, fake, fake, 1 << 20, 2, 1 << 11, 1, 0, None) == 1  # before*nB = 2^31
    assert L.qtg_rhomat_z(fake, fake, fake, 2, big, big // 2, None) == 1
    assert L.qtg_mode_product_z(fake, fake, fake, 1, 1 << 16, 1 << 15, 1, 0, 0, None) == 1  # nB*nC = 2^31
    descs = np.zeros(1, Q.HC_DESC)
    descs[0] = (0, 0, 0, big, 2, 2, big // 2, 0, 0)
    p = C.c_void_p()
    assert L.qtg_hole_contraction_plan_create(C.byref(p), descs.ctypes.data, 1, 0) == 1
    # just inside the limit is accepted by the check (no launch: null pointer)
    assert L.qtg_hole_contraction_z(None, fake, fake, big, 2, 2, big // 2 - 1, 0, None) == 3


def test_small_matrix_shape_checks():
    L = Q.lib()
    fake = 4096
    for fn in ("qtg_heev_batch_z", "qtg_inverse_batch_z"):
        getattr(L, fn).argtypes = None
    assert L.qtg_heev_batch_z(C.c_void_p(fake), C.c_void_p(fake), None, C.c_int64(65), C.c_int64(1), None) == 1
    assert L.qtg_inverse_batch_z(C.c_void_p(fake), C.c_void_p(fake), C.c_int64(0), C.c_int64(1), None) == 1
    L.qtg_qr_batch_z.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    assert L.qtg_qr_batch_z(fake, fake, 3, 4, 1, None) == 1  # m < n
    L.qtg_reg_inverse_batch_z.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_void_p]
    assert L.qtg_reg_inverse_batch_z(fake, fake, 4, 1, 0.0, None) == 3  # eps must be positive


@pytest.mark.parametrize("model", ["small", "cmf8", "grad32", "sweep64", "ho128"])
def test_scf_plan_stays_inside_its_buffers(model):
    """The SCF sweep is planned on the host (no device) and every launch of
    one iteration is checked: reads and writes inside their buffers, outputs
    of one launch disjoint, no launch writing what it reads."""
    from qutree_amd import mctdh as M
    if model == "small":
        text, H, k = W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 11, c_sigma=0.1), 4
    elif model == "cmf8":
        (text, H), k = W.cmf8_model()[:2], 10
    elif model == "grad32":
        (text, H), k = W.grad32_model()[:2], 5
    elif model == "sweep64":
        (text, H), k = W.sweep64_model()[:2], 5
    else:
        text, H, k = M.balanced_tree(128, 10, 4, 4, 0, 0.01822, 0.0, 0.0, 0.01822), M.coupled_ho(128), 10
    L = _host()
    L.qtg_host_scf_plan.argtypes = [C.c_char_p, C.c_int64, np.ctypeslib.ndpointer(np.float64)] + \
        [np.ctypeslib.ndpointer(np.int64)] * 4 + [np.ctypeslib.ndpointer(np.float64), C.c_int64,
                                                  np.ctypeslib.ndpointer(np.int64)]
    c, toff, fm, fk, fp, pool = H.arrays()
    info = np.zeros(6, np.int64)
    rc = L.qtg_host_scf_plan(text.encode(), H.nterms, c, toff, fm, fk, fp, pool, k, info)
    assert rc == 0, L.qtg_last_error().decode()
    nnodes = O.OracleModel(text).nnodes
    assert info[0] == 2 * nnodes - 1
    assert info[1] > info[0] * k  # at least one product and one Lanczos step per Krylov vector



@pytest.mark.parametrize("model", ["contractive4", "quartic8", "cmf8", "sweep64", "single"])
def test_apply_fit_plan_stays_inside_its_buffers(model):
    """The fitted SOP application is planned on the host and every launch of
    the first h-matrix pass and of one iteration is checked like the SCF
    plan."""
    import test_oracle_apply_fit as TA
    if model == "contractive4":
        text, H = W.balanced_tree_text(4, 3, [2, 3]), TA.contraction_sop(4, 3, 4, 1)
    elif model == "quartic8":
        text, H = W.balanced_tree_text(8, 2, [2, 3, 3]), TA.contraction_sop(8, 2, 4, 2)
    elif model == "cmf8":
        text, H = W.cmf8_model()[:2]
    elif model == "sweep64":
        text, H = W.sweep64_model()[:2]
    else:
        text = W.balanced_tree_text(4, 3, [2, 3])
        H = W.Hamiltonian()
        H.add_term(1.0, [(1, W.FACTOR_X), (2, W.FACTOR_KIN)])
    L = _host()
    L.qtg_host_apply_fit_plan.argtypes = [C.c_char_p, C.c_int64, np.ctypeslib.ndpointer(np.float64)] + \
        [np.ctypeslib.ndpointer(np.int64)] * 4 + [np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.int64)]
    c, toff, fm, fk, fp, pool = H.arrays()
    info = np.zeros(4, np.int64)
    rc = L.qtg_host_apply_fit_plan(text.encode(), H.nterms, c, toff, fm, fk, fp, pool, info)
    assert rc == 0, L.qtg_last_error().decode()
    assert info[0] >= 1
