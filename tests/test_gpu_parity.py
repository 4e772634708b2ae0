"""GPU parity: the gfx950 engine (through the C ABI) against the oracle.

Bar (FP64 path): relative L2 difference <= 1e-10 against the oracle on the same
tree, operators and charges, after every operator of the evaluation phase and
for the full upward + downward pass.  Mixed path and FLOAT mode (pseudo-inverses
truncated at 4e-8 max(S)): potential and gradient within 1e-5 of the oracle
evaluating the same operators in FP64, and as close to the FP64 direct sum as
the FP64 FMM (or within 1e-5 of it).
"""
import numpy as np
import pytest

import oracle as O
from exafmm import abi
from fmm_cases import equiv_rel_diff, make_problem, rel_err, truncated_ops

pytestmark = pytest.mark.gpu

KERNELS = [("laplace", 0, 0.0), ("modified_helmholtz", 2, 5.0), ("helmholtz", 1, 5.0 + 2.0j)]
STAGE_ORDER = ["P2M", "M2M", "P2L", "M2P", "P2P", "M2L", "L2L", "L2P"]
FP64_TOL = 1e-10
MIXED_TOL = 1e-5  # potential and gradient, mixed and FLOAT paths


def close(a, b, tol):
    nb = np.linalg.norm(b)
    if nb == 0:
        return np.linalg.norm(a) == 0
    return np.linalg.norm(a - b) <= tol * nb


@pytest.mark.parametrize("name,kid,wk", KERNELS)
@pytest.mark.parametrize("gradient", [0, 1])
def test_direct_fp64(require_gpu, name, kid, wk, gradient):
    # reference tests/p2p_*.cpp: vectorised P2P vs the scalar loop, 1e-12 in FP64
    rng = np.random.default_rng(7 + kid)
    src, trg = rng.random((1500, 3)), rng.random((700, 3))
    trg[:50] = src[:50]  # coincident pairs must contribute nothing
    q = rng.random(1500) - 0.5
    if kid == 1:
        q = q + 1j * (rng.random(1500) - 0.5)
    g = abi.direct(kid, src, q, trg, wk, abi.FP64, bool(gradient))
    r = O.direct(kid, src, q, trg, wk, bool(gradient))
    assert rel_err(g, r) < 1e-12


@pytest.mark.parametrize("name,kid,wk", KERNELS)
def test_direct_fp32(require_gpu, name, kid, wk):
    # reference tests/p2p_laplace.cpp:82-84: 1e-6 in single precision
    rng = np.random.default_rng(11 + kid)
    src, trg = rng.random((1200, 3)), rng.random((600, 3))
    q = rng.random(1200) - 0.5
    if kid == 1:
        q = q + 1j * (rng.random(1200) - 0.5)
    g = abi.direct(kid, src, q, trg, wk, abi.MIXED, True)
    r = O.direct(kid, src, q, trg, wk, True)
    assert rel_err(g[:, 0], r[:, 0]) < 1e-6
    assert rel_err(g[:, 1:], r[:, 1:]) < 1e-6


CASES = [
    ("laplace", 4, "cube", 6000, 64),
    ("laplace", 6, "plummer", 12000, 32),
    ("laplace", 8, "clusters", 8000, 64),
    ("modified_helmholtz", 5, "cube", 5000, 64),
    ("modified_helmholtz", 6, "sphere", 8000, 32),
    ("helmholtz", 5, "cube", 5000, 64),
    ("helmholtz", 5, "clusters", 6000, 32),
]


def _engine(pr, precision, ops=None):
    eng = abi.Engine(pr["kid"], pr["p"], pr["wavek"], precision)
    eng.set_operators(pr["ops"] if ops is None else ops)
    eng.set_tree(pr["flat"])
    eng.set_charges(pr["qs"])
    return eng


@pytest.mark.parametrize("kernel,p,dist,n,ncrit", CASES)
def test_operators_fp64(require_gpu, kernel, p, dist, n, ncrit):
    pr = make_problem(kernel, n, p, ncrit=ncrit, dist=dist, seed=3)
    print("case", kernel, p, dist, n, ncrit, "nodes", pr["flat"]["nnodes"], flush=True)
    check_stages_fp64(pr)


def check_stages_fp64(pr):
    """FP64 engine vs oracle after every operator of the reference order, then
    two full passes (each pass zeroes what it produces): all within 1e-10."""
    p = pr["p"]
    eng = _engine(pr, abi.FP64)
    orc = O.Oracle(pr["kid"], p, pr["flat"], pr["ops"], pr["wavek"])
    orc.set_charges(pr["qs"])
    f = pr["flat"]
    lev = f["level"]
    leaf = f["child0"] < 0
    conv = np.nonzero(leaf & ((f["src_count"] > 0) | (f["trg_count"] > 0)))[0]  # leaves L2P converts
    other = np.setdiff1d(np.arange(f["nnodes"]), conv)
    allrows = np.arange(f["nnodes"])
    eng.clear()
    orc.clear()
    for st in STAGE_ORDER:
        eng.op(st)
        orc.run(O.STAGES[st])
        up, dn = eng.equiv()
        v = eng.values()
        # equivalent densities come out of the ill-conditioned pseudo-inverse:
        # compare the check potentials they generate (raw difference printed)
        e_up = equiv_rel_diff(pr["kid"], pr["wavek"], p, f["r0"], lev, up, orc.up, allrows, True)
        print(st, "up raw %.2e check-metric %.2e" % (rel_err(up, orc.up) if np.any(orc.up) else 0, e_up))
        assert e_up <= FP64_TOL, (st, "up", e_up)
        if st == "L2P":
            e_dn = equiv_rel_diff(pr["kid"], pr["wavek"], p, f["r0"], lev, dn, orc.dn, conv, False)
            assert e_dn <= FP64_TOL, (st, "dn equiv", e_dn)
            assert close(dn[other], orc.dn[other], FP64_TOL), (st, "dn", rel_err(dn[other], orc.dn[other]))
        else:
            assert close(dn, orc.dn, FP64_TOL), (st, "dn", rel_err(dn, orc.dn))
        assert close(v, orc.val, FP64_TOL), (st, "values", rel_err(v, orc.val))
    # full passes (each pass zeroes what it produces): twice, identical results
    ref = orc.evaluate()
    for _ in range(2):
        eng.upward_pass()
        eng.downward_pass()
        v = eng.values()
        assert rel_err(v[:, 0], ref[:, 0]) < FP64_TOL
        assert rel_err(v[:, 1:], ref[:, 1:]) < FP64_TOL
    eng.close()
    return ref


@pytest.mark.parametrize("kernel,p,dist,n,ncrit", [CASES[0], CASES[2], CASES[3], CASES[5]])
def test_mixed_precision(require_gpu, kernel, p, dist, n, ncrit):
    pr = make_problem(kernel, n, p, ncrit=ncrit, dist=dist, seed=5, precision=abi.MIXED)
    eng = _engine(pr, abi.MIXED)
    eng.upward_pass()
    eng.downward_pass()
    v = eng.values()
    orc = O.Oracle(pr["kid"], p, pr["flat"], pr["ops"], pr["wavek"])
    orc.set_charges(pr["qs"])
    ref = orc.evaluate()
    tc = pr["flat"]["trg_coord"].reshape(-1, 3)
    idx = np.arange(0, tc.shape[0], 7)
    d = O.direct(pr["kid"], pr["flat"]["src_coord"].reshape(-1, 3), pr["qs"], tc[idx], pr["wavek"])
    for c in (slice(0, 1), slice(1, 4)):
        e64 = rel_err(ref[idx, c], d[:, c])
        emx = rel_err(v[idx, c], d[:, c])
        print(kernel, "fp64 err", e64, "mixed err", emx, "vs oracle", rel_err(v[:, c], ref[:, c]))
        assert rel_err(v[:, c], ref[:, c]) < MIXED_TOL
        assert emx < max(2 * e64, MIXED_TOL)


@pytest.mark.parametrize("kernel,p,dist,depth", [("laplace", 6, "clusters", 4), ("modified_helmholtz", 5, "sphere", 4),
                                                ("helmholtz", 5, "cube", 3)])
def test_non_adaptive_tree(require_gpu, kernel, p, dist, depth):
    """Non-adaptive tree (the reference's NON_ADAPTIVE build: uniform to
    `depth`, ghost empty nodes, target-only child lists): FP64 full passes
    equal the oracle to 1e-10, the mixed path stays within 1e-5 of them."""
    pr = make_problem(kernel, 12000, p, dist=dist, seed=8, depth=depth)
    print("non-adaptive case", kernel, p, dist, depth, "nodes", pr["flat"]["nnodes"], flush=True)
    orc = O.Oracle(pr["kid"], p, pr["flat"], pr["ops"], pr["wavek"])
    orc.set_charges(pr["qs"])
    ref = orc.evaluate()
    ops_t = truncated_ops(pr)
    orc_t = O.Oracle(pr["kid"], p, pr["flat"], ops_t, pr["wavek"])
    orc_t.set_charges(pr["qs"])
    ref_t = orc_t.evaluate()
    for prec, tol, ops, r in ((abi.FP64, FP64_TOL, None, ref), (abi.MIXED, MIXED_TOL, ops_t, ref_t)):
        eng = _engine(pr, prec, ops)
        eng.upward_pass()
        eng.downward_pass()
        v = eng.values()
        print("precision", prec, "phi %.2e grad %.2e" % (rel_err(v[:, 0], r[:, 0]), rel_err(v[:, 1:], r[:, 1:])))
        assert rel_err(v[:, 0], r[:, 0]) < tol
        assert rel_err(v[:, 1:], r[:, 1:]) < tol


@pytest.mark.parametrize("kernel,p,dist,n", [("laplace", 10, "clusters", 30000), ("laplace", 8, "plummer", 20000),
                                            ("modified_helmholtz", 6, "sphere", 20000), ("helmholtz", 5, "cube", 10000)])
def test_float_mode(require_gpu, kernel, p, dist, n):
    """FLOAT mode (the reference's -DFLOAT build: pseudo-inverses truncated at
    4e-8 max(S), every evaluation stage including the translation GEMMs in
    FP32): against the oracle evaluating the same operators in FP64 it carries
    only FP32 rounding, and against the direct sum it stays near the FP64
    path's accuracy."""
    pr = make_problem(kernel, n, p, dist=dist, seed=12, precision=abi.FLOAT)
    print("float case", kernel, p, dist, n, flush=True)
    orc = O.Oracle(pr["kid"], p, pr["flat"], pr["ops"], pr["wavek"])
    orc.set_charges(pr["qs"])
    ref = orc.evaluate()
    eng = _engine(pr, abi.FLOAT)
    eng.upward_pass()
    eng.downward_pass()
    v = eng.values()
    e0, e1 = rel_err(v[:, 0], ref[:, 0]), rel_err(v[:, 1:], ref[:, 1:])
    tc = pr["flat"]["trg_coord"].reshape(-1, 3)
    idx = np.arange(0, tc.shape[0], 13)
    d = O.direct(pr["kid"], pr["flat"]["src_coord"].reshape(-1, 3), pr["qs"], tc[idx], pr["wavek"])
    d0, d64 = rel_err(v[idx, 0], d[:, 0]), rel_err(ref[idx, 0], d[:, 0])
    print("vs oracle phi %.2e grad %.2e; vs direct %.2e (oracle %.2e)" % (e0, e1, d0, d64))
    # same bar as the mixed path (test_mixed_precision)
    assert e0 < MIXED_TOL and e1 < MIXED_TOL
    assert d0 < max(3 * d64, MIXED_TOL)


@pytest.mark.parametrize("kernel,p", [("laplace", 6), ("modified_helmholtz", 5), ("helmholtz", 5)])
def test_pseudo_inverse_rank(require_gpu, kernel, p):
    """P2M and L2P apply each pseudo-inverse over its first r singular
    directions only, r = 1 + the last column of C2E_V (per operator level)
    that is not exactly zero.  Operators whose C2E_V columns past odd,
    level-dependent ranks are zeroed: the FP64 path equals the oracle on the
    same operators to 1e-10 and the FLOAT path stays within the FP32 bar."""
    pr = make_problem(kernel, 5000, p, dist="cube", seed=21)
    ns = 6 * (p - 1) ** 2 + 2
    ops = dict(pr["ops"])
    for key, r0 in (("uc2e_v", 37), ("dc2e_v", 53)):
        m = np.array(ops[key]).reshape(ops["nlevels"], ns, ns)
        for L in range(m.shape[0]):
            m[L, :, r0 + 6 * L:] = 0
        ops[key] = m.reshape(-1)
    pr = dict(pr, ops=ops)
    orc = O.Oracle(pr["kid"], p, pr["flat"], ops, pr["wavek"])
    orc.set_charges(pr["qs"])
    ref = orc.evaluate()
    for prec, tol in ((abi.FP64, FP64_TOL), (abi.FLOAT, MIXED_TOL)):
        eng = _engine(pr, prec)
        eng.upward_pass()
        eng.downward_pass()
        v = eng.values()
        eng.close()
        e0, e1 = rel_err(v[:, 0], ref[:, 0]), rel_err(v[:, 1:], ref[:, 1:])
        print(kernel, "precision", prec, "phi %.2e grad %.2e" % (e0, e1), flush=True)
        assert e0 < tol and e1 < tol


@pytest.mark.parametrize("kernel,p,dist", [("laplace", 10, "clusters"), ("helmholtz", 5, "cube")])
def test_mixed_hadamard_fp16_vs_fp32(require_gpu, monkeypatch, kernel, p, dist):
    """The default mixed Hadamard (split FP16 on the f16 matrix cores) and the
    plain FP32 one (selected with XFMM_MIXED_STAGES) both stay within the
    mixed tolerance of the FP64 evaluation on the same (truncated) operators,
    and close to each other."""
    pr = make_problem(kernel, 20000, p, ncrit=64, dist=dist, seed=11, precision=abi.MIXED)
    print("hadamard variants", kernel, p, dist, flush=True)
    out = {}
    for name, stages in (("fp64", None), ("had16", "p2p,m2l,had16,fft"), ("had32", "p2p,m2l,fft")):
        if stages is None:
            monkeypatch.delenv("XFMM_MIXED_STAGES", raising=False)
        else:
            monkeypatch.setenv("XFMM_MIXED_STAGES", stages)
        eng = _engine(pr, abi.FP64 if stages is None else abi.MIXED)
        eng.upward_pass()
        eng.downward_pass()
        out[name] = eng.values()
    monkeypatch.delenv("XFMM_MIXED_STAGES", raising=False)
    for name in ("had16", "had32"):
        e0 = rel_err(out[name][:, 0], out["fp64"][:, 0])
        e1 = rel_err(out[name][:, 1:], out["fp64"][:, 1:])
        print(name, "phi %.2e grad %.2e" % (e0, e1))
        assert e0 < MIXED_TOL and e1 < MIXED_TOL, (name, e0, e1)
    assert rel_err(out["had16"][:, 0], out["had32"][:, 0]) < 1e-5


@pytest.mark.parametrize("kernel,p", [("laplace", 8), ("helmholtz", 5)])
def test_hadamard_tuning_same_values(require_gpu, kernel, p):
    """The launch-shape parameters change no value: the split-FP16 Hadamard with 1, 2 or 4 target blocks
    per workgroup, 1 or 2 target tiles per wave and 4 or 32 frequencies per workgroup gives bit-identical
    results (each target's sum is computed by the same lanes in the same order), and the autotuner picks a
    valid setting."""
    pr = make_problem(kernel, 20000, p, ncrit=64, dist="cube", seed=12, precision=abi.MIXED)
    eng = _engine(pr, abi.MIXED)
    out = []
    for groups, kch, tiles in ((1, 16, 1), (2, 16, 1), (4, 16, 1), (2, 4, 1), (2, 32, 1), (1, 16, 2), (2, 16, 2),
                               (4, 8, 2)):
        eng.set_tuning(abi.TUNE_HAD_GROUPS, groups)
        eng.set_tuning(abi.TUNE_HAD_KCH, kch)
        eng.set_tuning(abi.TUNE_HAD_TILES, tiles)
        eng.upward_pass()
        eng.downward_pass()
        out.append(eng.values())
    for v in out[1:]:
        assert np.array_equal(v, out[0])
    tuning, ms = eng.autotune(2)
    print("autotune", tuning, ms, flush=True)
    assert tuning["had_groups"] in (1, 2) and tuning["had_kch"] in (4, 8, 16, 32, 64)
    assert tuning["had_tiles"] in (1, 2)
    eng.set_tuning(abi.TUNE_P2P_SPLIT, 32)  # the near field's split changes its summation order
    eng.upward_pass()
    eng.downward_pass()
    assert np.array_equal(eng.values(), out[0])
    with pytest.raises(abi.XfmmError):
        eng.set_tuning(abi.TUNE_HAD_GROUPS, 3)
    with pytest.raises(abi.XfmmError):
        eng.set_tuning(abi.TUNE_HAD_TILES, 4)
    eng.close()


def test_python_api_evaluate_and_verify(require_gpu):
    import exafmm
    L = exafmm.laplace
    rng = np.random.default_rng(2)
    n = 20000
    x = rng.random((n, 3))
    q = rng.random(n)
    import os
    from fmm_cases import CACHE
    fmm = L.LaplaceFmm(p=8, ncrit=64, filename=os.path.join(CACHE, "api_laplace_p8.dat"))
    tree = L.setup(L.init_sources(x, q), L.init_targets(x), fmm)
    out = L.evaluate(tree, fmm)
    assert out.shape == (n, 4)
    ep, eg = fmm.verify(tree.leafs, True)
    assert ep < 1e-6 and eg < 1e-5, (ep, eg)
    ref = O.direct(0, x, q, x[::97])
    assert rel_err(out[::97, 0], ref[:, 0]) < 1e-6
    # BEM-style loop: update charges, clear, evaluate again (python/exafmm.cpp:275-323)
    q2 = rng.random(n) - 0.5
    L.update_charges(tree, q2)
    L.clear_values(tree)
    out2 = L.evaluate(tree, fmm)
    ref2 = O.direct(0, x, q2, x[::97])
    assert rel_err(out2[::97, 0], ref2[:, 0]) < 1e-6


def test_linearity_and_repeatability(require_gpu):
    pr = make_problem("laplace", 30000, 6, ncrit=64, dist="cube", seed=9)
    eng = _engine(pr, abi.FP64)
    rng = np.random.default_rng(4)
    q1, q2 = rng.random(pr["qs"].size) - 0.5, rng.random(pr["qs"].size) - 0.5
    res = []
    for q in (q1, q2, q1 + 2.0 * q2):
        eng.set_charges(q)
        eng.upward_pass()
        eng.downward_pass()
        res.append(eng.values())
    assert rel_err(res[2], res[0] + 2.0 * res[1]) < 1e-12
    eng.set_charges(q1)
    eng.upward_pass()
    eng.downward_pass()
    assert np.array_equal(eng.values(), res[0])
    # a second downward pass on the same upward state recomputes from zero
    eng.downward_pass()
    assert np.array_equal(eng.values(), res[0])
    # ... also after the caller overwrote the local expansions
    eng.set_equiv(None, np.ones_like(eng.equiv()[1]))
    eng.downward_pass()
    assert np.array_equal(eng.values(), res[0])


@pytest.mark.parametrize("kernel,precision", [("laplace", abi.MIXED), ("modified_helmholtz", abi.FP64),
                                              ("helmholtz", abi.MIXED), ("laplace", abi.FLOAT)])
def test_graph_replay_matches_eager(require_gpu, kernel, precision):
    """A captured evaluation replays to the eager values, replay after replay, with eager evaluations and new
    charges in between (the graph holds kernels only: its clears are fill kernels).  The last charge set is
    of one sign: the mixed path then runs M2M / L2L and the transforms in FP64, and a replay after the
    change of class captures the graph again by itself."""
    pr = make_problem(kernel, 6000, 6, ncrit=32, dist="plummer", seed=21, precision=precision)
    print("graph case", kernel, precision, flush=True)
    eng = _engine(pr, precision)
    rng = np.random.default_rng(8)
    qs = [pr["qs"]] + [pr["qs"] * (rng.random(pr["qs"].size) + 0.5) for _ in range(3)]
    qs.append((np.abs(pr["qs"]) + 0.1).astype(pr["qs"].dtype))
    eager = []
    for q in qs:
        eng.set_charges(q)
        eng.upward_pass()
        eng.downward_pass()
        eager.append(eng.values())
    eng.graph_capture()
    for rep in range(2):
        for i, q in enumerate(qs):
            eng.set_charges(q)
            eng.graph_launch()
            eng.synchronize()
            assert np.array_equal(eng.values(), eager[i]), (rep, i)
            # eager work between replays
            eng.set_charges(qs[(i + 1) % len(qs)])
            eng.upward_pass()
            eng.downward_pass()
            assert np.array_equal(eng.values(), eager[(i + 1) % len(qs)]), (rep, i)
    # a new tree drops the graph
    eng.set_tree(pr["flat"])
    with pytest.raises(abi.XfmmError):
        eng.graph_launch()


def test_wave_number_change_between_evaluations(require_gpu, tmp_path):
    """Python exposes wavek read-write (reference python/exafmm.cpp:383-386):
    after a change and a new setup, every operator, including the near field
    and the surface evaluations, uses the new wave number, also when the
    precompute file of the old one is still there."""
    Y = __import__("exafmm").modified_helmholtz
    rng = np.random.default_rng(31)
    n = 12000
    x = rng.random((n, 3))
    q = rng.random(n) - 0.5
    fmm = Y.ModifiedHelmholtzFmm(6, 64, 5.0, str(tmp_path / "yuk_p6.dat"))
    for k in (5.0, 2.0, 5.0):
        fmm.wavek = k
        tree = Y.setup(Y.init_sources(x, q), Y.init_targets(x), fmm)
        out = Y.evaluate(tree, fmm)
        ref = O.direct(2, x, q, x[::41], k)
        e0, e1 = rel_err(out[::41, 0], ref[:, 0]), rel_err(out[::41, 1:], ref[:, 1:])
        print("wavek", k, "phi %.2e grad %.2e" % (e0, e1), flush=True)
        assert e0 < 1e-5 and e1 < 1e-4, (k, e0, e1)
