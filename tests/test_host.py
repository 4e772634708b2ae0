"""Host-side logic (CPU only): tree, lists, flattened layout, precomputation,
the precompute cache format, the Python API surface and the C ABI exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import exafmm
import oracle as O
from exafmm import abi
from fmm_cases import CACHE, check_matrix, kernel_np, make_problem, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", params=[("laplace", "cube"), ("laplace", "plummer"), ("modified_helmholtz", "sphere"),
                                        ("helmholtz", "clusters")])
def problem(request):
    kernel, dist = request.param
    return make_problem(kernel, 6000, 4, ncrit=24, dist=dist, seed=2)


def test_tree_every_body_once(problem):
    # reference tests/tree.cpp:35-43
    f = problem["flat"]
    n = len(problem["x"])
    assert sorted(f["src_ibody"].tolist()) == list(range(n))
    assert sorted(f["trg_ibody"].tolist()) == list(range(n))
    # sorted coordinates are the input coordinates
    assert np.array_equal(f["src_coord"].reshape(-1, 3), problem["x"][f["src_ibody"]])
    assert np.array_equal(problem["qs"], problem["q"][f["src_ibody"]])


def test_tree_leaf_capacity_and_balance(problem):
    f = problem["flat"]
    ncrit = problem["fmm"].ncrit
    leaf = f["child0"] < 0
    assert np.all(f["src_count"][leaf] <= ncrit) and np.all(f["trg_count"][leaf] <= ncrit)
    assert np.all(f["src_count"][~leaf] == 0)
    # 2:1 balance: touching leaves differ by at most one level
    cen, lev, r0, x0 = f["center"].reshape(-1, 3), f["level"], f["r0"], np.array(f["x0"])
    L = np.nonzero(leaf)[0]
    half = r0 / 2.0 ** lev[L]
    lo, hi = cen[L] - half[:, None], cen[L] + half[:, None]
    tol = 1e-9 * r0
    for i in range(len(L)):
        touch = np.all((lo <= hi[i] + tol) & (hi >= lo[i] - tol), axis=1)
        assert np.max(np.abs(lev[L][touch] - lev[L[i]])) <= 1
    # node levels: 8 children per non-leaf, level l+1, inside the parent box
    c0 = f["child0"]
    for nd in np.nonzero(~leaf)[0]:
        ch = np.arange(c0[nd], c0[nd] + 8)
        assert np.all(lev[ch] == lev[nd] + 1)
        off = cen[ch] - cen[nd]
        oct_bits = (off > 0).astype(int) @ np.array([1, 2, 4])
        assert np.array_equal(oct_bits, np.arange(8))
    assert np.array_equal(np.diff(f["level_offset"]) > 0, np.ones(f["depth"] + 1, bool))
    assert np.all(np.diff(lev) >= 0)


@pytest.fixture(scope="module", params=[("laplace", "cube", 3), ("laplace", "clusters", 4),
                                        ("helmholtz", "sphere", 3)])
def uniform_problem(request):
    kernel, dist, depth = request.param
    return make_problem(kernel, 5000, 4, dist=dist, seed=4, depth=depth)


def test_non_adaptive_tree_structure(uniform_problem):
    """Non-adaptive tree (reference build_non_adaptive_tree.h:95-194): every
    body once, all bodies in leaves at `depth`, an empty node is a leaf at any
    level and never split, a node with bodies above `depth` is split."""
    f = uniform_problem["flat"]
    depth = uniform_problem["fmm"].depth
    n = len(uniform_problem["x"])
    assert f["depth"] == depth
    assert sorted(f["src_ibody"].tolist()) == list(range(n))
    assert sorted(f["trg_ibody"].tolist()) == list(range(n))
    leaf = f["child0"] < 0
    lev = f["level"]
    has = (f["src_count"] > 0) | (f["trg_count"] > 0)
    assert np.all(lev[leaf & has] == depth)
    c0 = f["child0"]
    sub = f["src_count"].astype(np.int64) + f["trg_count"]
    for nd in sorted(np.nonzero(~leaf)[0], key=lambda i: -lev[i]):
        sub[nd] = sub[c0[nd]:c0[nd] + 8].sum()
    assert np.all(sub[~leaf] > 0)              # split nodes hold bodies
    assert np.all(lev[~leaf] < depth)
    assert np.all(sub[leaf & (lev < depth)] == 0)  # leaves above depth are ghosts
    assert np.array_equal(f["src_coord"].reshape(-1, 3), uniform_problem["x"][f["src_ibody"]])


def test_non_adaptive_lists_complete(uniform_problem):
    test_lists_complete_and_disjoint(uniform_problem)


def test_lists_complete_and_disjoint(problem):
    """Counting FMM (reference tests/list.cpp:44 with test.h DummyFmm): every
    target leaf receives every source exactly once through P2P, M2P, P2L (at
    itself or an ancestor) and the non-adjacent child pairs of M2L."""
    f = problem["flat"]
    nn = f["nnodes"]
    c0, lev = f["child0"], f["level"]
    cen = f["center"].reshape(-1, 3)
    parent = np.full(nn, -1)
    for nd in np.nonzero(c0 >= 0)[0]:
        parent[c0[nd]:c0[nd] + 8] = nd
    sub = f["src_count"].astype(np.int64).copy()
    for nd in sorted(np.nonzero(c0 >= 0)[0], key=lambda i: -lev[i]):
        sub[nd] = sub[c0[nd]:c0[nd] + 8].sum()
    nsrc = len(f["src_ibody"])

    def csr(name, i):
        return f[name + "_source"][f[name + "_offset"][i]:f[name + "_offset"][i + 1]]

    side = 2 * f["r0"] / 2.0 ** lev
    far = np.zeros(nn, np.int64)  # sources reaching node a through P2L(a) and M2L at its parent
    for a in range(nn):
        s = int(f["src_count"][csr("p2l", a)].sum())
        P = parent[a]
        if P >= 0:
            for r in range(26):
                S = f["m2l_source"][P * 26 + r]
                if S < 0:
                    continue
                for b in range(8):
                    sc = c0[S] + b
                    rel = np.round((cen[sc] - cen[a]) / side[a]).astype(int)
                    if np.max(np.abs(rel)) > 1:
                        s += sub[sc]
        far[a] = s
    for t in np.nonzero((c0 < 0) & (f["trg_count"] > 0))[0]:
        tot = int(f["src_count"][csr("p2p", t)].sum()) + int(sub[csr("m2p", t)].sum())
        a = t
        while a >= 0:
            tot += far[a]
            a = parent[a]
        assert tot == nsrc, (t, tot, nsrc)


def test_m2l_operators_match_oracle():
    # product precompute (C++ own DFT) vs the oracle's independent build
    for kernel, kid, wk in (("laplace", O.LAPLACE, 0.0), ("modified_helmholtz", O.YUKAWA, 5.0),
                            ("helmholtz", O.HELMHOLTZ, 5.0 + 0j)):
        pr = make_problem(kernel, 3000, 4, ncrit=32, dist="cube", seed=4)
        ops, f = pr["ops"], pr["flat"]
        lap = kid == O.LAPLACE
        slots = 1 if lap else f["depth"]
        ref = np.concatenate([O.m2l_operator(kid, 4, f["r0"], 0 if lap else l + 1, lap, wk) for l in range(slots)])
        assert rel_err(ops["m2l"], ref) < 1e-12, kernel


@pytest.mark.parametrize("kernel,kid,wk", [("laplace", 0, 0.0), ("modified_helmholtz", 2, 5.0),
                                           ("helmholtz", 1, 5.0 + 0j)])
def test_pseudo_inverse_and_translations(kernel, kid, wk):
    p = 6
    pr = make_problem(kernel, 3000, p, ncrit=32, dist="cube", seed=6)
    ops, f = pr["ops"], pr["flat"]
    ns = 6 * (p - 1) ** 2 + 2
    r0 = f["r0"]
    rng = np.random.default_rng(0)
    for lvl in range(ops["nlevels"]):
        A = check_matrix(kid, wk, p, r0, lvl, True)          # check <- equivalent (upward)
        sl = slice(lvl * ns * ns, (lvl + 1) * ns * ns)
        pinv = ops["uc2e_v"][sl].reshape(ns, ns) @ ops["uc2e_u"][sl].reshape(ns, ns)
        # Moore-Penrose: A pinv A = A
        assert rel_err(A @ pinv @ A, A) < 1e-8
        # downward pinv is the pinv of the transposed (downward) matrix
        Ad = check_matrix(kid, wk, p, r0, lvl, False)
        assert rel_err(Ad, A.T) < 1e-14
        pinvd = ops["dc2e_v"][sl].reshape(ns, ns) @ ops["dc2e_u"][sl].reshape(ns, ns)
        assert rel_err(Ad @ pinvd @ Ad, Ad) < 1e-8
        if lvl + 1 >= ops["nlevels"] and ops["nlevels"] > 1:
            continue
        # M2M: the parent's equivalent density reproduces the child's field on
        # the parent check surface
        c = np.zeros(3)
        pchk = O.surface(p, r0, lvl, c, 2.95)
        pequ = O.surface(p, r0, lvl, c, 1.05)
        for o in range(8):
            cc = np.array([(o >> d) & 1 for d in range(3)]) * 2 - 1.0
            cequ = O.surface(p, r0, lvl + 1, cc * r0 / 2 ** (lvl + 1), 1.05)
            x = rng.random(ns) - 0.5
            M = ops["m2m"][(lvl * 8 + o) * ns * ns:(lvl * 8 + o + 1) * ns * ns].reshape(ns, ns)
            direct_field = kernel_np(kid, wk, pchk[:, None] - cequ[None]) @ x
            via_m2m = kernel_np(kid, wk, pchk[:, None] - pequ[None]) @ (M @ x)
            assert rel_err(via_m2m, direct_field) < 1e-9
            # L2L = K(child check <- parent equiv) DC2E: a parent check potential
            # generated by far sources maps to the child check potential
            L = ops["l2l"][(lvl * 8 + o) * ns * ns:(lvl * 8 + o + 1) * ns * ns].reshape(ns, ns)
            far = rng.random((20, 3)) * 0.5 * r0 / 2 ** lvl + 6.0 * r0 / 2 ** lvl
            qf = rng.random(20)
            pdck = O.surface(p, r0, lvl, c, 1.05)
            cdck = O.surface(p, r0, lvl + 1, cc * r0 / 2 ** (lvl + 1), 1.05)
            chk_p = kernel_np(kid, wk, pdck[:, None] - far[None]) @ qf
            chk_c = kernel_np(kid, wk, cdck[:, None] - far[None]) @ qf
            assert rel_err(L @ chk_p, chk_c) < 5e-5  # truncation error of a p = 6 expansion (a wrong octant or transpose gives O(1))


def test_precompute_cache_reference_layout(tmp_path):
    # fmm_scale_invariant.h:86-148: r0 | UC2E_U UC2E_V DC2E_U DC2E_V | M2M[8] | L2L[8] | M2L[26]
    L = exafmm.laplace
    x = np.random.default_rng(0).random((2000, 3))
    q = np.ones(2000)
    fn = str(tmp_path / "lap.dat")
    f1 = L.LaplaceFmm(4, 32, fn)
    t1 = L.setup(L.init_sources(x, q), L.init_targets(x), f1)
    ns, nf = f1.nsurf, f1.nfreq
    assert os.path.getsize(fn) == (2 * 8 + 4) * ns * ns * 8 + 26 * nf * 128 * 8 + 8
    raw = np.fromfile(fn, dtype=np.float64)
    assert raw[0] == f1.r0
    ops = f1.operators()
    assert np.array_equal(raw[1:1 + ns * ns], ops["uc2e_u"])
    assert np.array_equal(raw[1 + 4 * ns * ns:1 + 5 * ns * ns], ops["m2m"][:ns * ns])
    assert np.array_equal(raw[1 + 12 * ns * ns:1 + 13 * ns * ns], ops["l2l"][:ns * ns])
    assert np.array_equal(raw[1 + 20 * ns * ns:], ops["m2l"])
    # reload gives identical operators; a truncated file is recomputed
    f2 = L.LaplaceFmm(4, 32, fn)
    L.setup(L.init_sources(x, q), L.init_targets(x), f2)
    assert all(np.array_equal(ops[k], f2.operators()[k]) for k in ("uc2e_v", "m2l", "l2l"))
    with open(fn, "r+b") as fh:
        fh.truncate(1000)
    f3 = L.LaplaceFmm(4, 32, fn)
    L.setup(L.init_sources(x, q), L.init_targets(x), f3)
    assert rel_err(f3.operators()["m2l"], ops["m2l"]) < 1e-15
    assert os.path.getsize(fn) == (2 * 8 + 4) * ns * ns * 8 + 26 * nf * 128 * 8 + 8


def test_precompute_cache_rejects_other_wave_number(tmp_path):
    # The reference layout does not record the wave number (fmm.h:103-138): a
    # file written for k = 5 has the size a k = 2 run expects.  The loaded
    # pseudo-inverses are checked against this kernel's check matrix, so the
    # k = 2 run recomputes instead of evaluating with k = 5 operators.
    Y = exafmm.modified_helmholtz
    x = np.random.default_rng(1).random((1500, 3))
    q = np.ones(1500)
    fn = str(tmp_path / "yuk.dat")
    f5 = Y.ModifiedHelmholtzFmm(4, 64, 5.0, fn)
    Y.setup(Y.init_sources(x, q), Y.init_targets(x), f5)
    size5 = os.path.getsize(fn)
    f2 = Y.ModifiedHelmholtzFmm(4, 64, 2.0, fn)
    Y.setup(Y.init_sources(x, q), Y.init_targets(x), f2)
    assert os.path.getsize(fn) == size5
    fresh = Y.ModifiedHelmholtzFmm(4, 64, 2.0, str(tmp_path / "fresh.dat"))
    Y.setup(Y.init_sources(x, q), Y.init_targets(x), fresh)
    for k in ("uc2e_v", "m2m", "m2l"):
        assert rel_err(f2.operators()[k], fresh.operators()[k]) < 1e-12, k
        assert rel_err(f2.operators()[k], f5.operators()[k]) > 1e-3, k
    # the file now holds k = 2 operators: a k = 2 run loads them unchanged
    f2b = Y.ModifiedHelmholtzFmm(4, 64, 2.0, fn)
    Y.setup(Y.init_sources(x, q), Y.init_targets(x), f2b)
    assert np.array_equal(f2b.operators()["m2l"], f2.operators()["m2l"])


def test_python_api_errors():
    L = exafmm.laplace
    with pytest.raises(RuntimeError):
        L.init_sources(np.zeros((5, 2)), np.zeros(5))
    with pytest.raises(RuntimeError):
        L.init_sources(np.zeros((5, 3)), np.zeros(4))
    with pytest.raises(RuntimeError):
        L.init_targets(np.zeros(5))
    # modified Helmholtz reuses the Laplace body/tree types (python/exafmm.cpp:390-537)
    assert exafmm.modified_helmholtz.Body is L.Body
    import importlib
    assert importlib.import_module("exafmm.laplace") is L  # submodule import as with the reference module


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "exafmm_t_amd.h")).read()
    names = sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(xfmm_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 25
    lib = ctypes.CDLL(abi.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    libo = ctypes.CDLL(O.LIB_PATH)
    for n in ("orc_direct", "orc_evaluate", "orc_m2l_operator", "orc_dft3_r2c", "orc_dft3_c2r", "orc_surface"):
        assert hasattr(libo, n)


def test_c_abi_argument_errors_without_device():
    # errors come back as status codes with a message, never as a crash
    lib = abi.load()
    h = ctypes.c_void_p()
    assert lib.xfmm_create(7, 4, 0.0, 0.0, 0, 0, ctypes.byref(h)) == 1
    assert b"kernel" in lib.xfmm_last_error(None)
    assert lib.xfmm_create(0, 1, 0.0, 0.0, 0, 0, ctypes.byref(h)) == 1
    assert lib.xfmm_create(0, 4, 0.0, 0.0, 5, 0, ctypes.byref(h)) == 1
    assert lib.xfmm_upward_pass(None) == 1
    if abi.device_count() == 0:
        rc = lib.xfmm_create(0, 4, 0.0, 0.0, 0, 0, ctypes.byref(h))
        assert rc != 0 and lib.xfmm_last_error(None)
