"""Batched Laplace problems as one forest (exafmm.batch).

CPU: the merged forest keeps every problem's tree and lists (node renumbering
is a bijection per problem; children stay a contiguous block of the next
level; the lists map through it), the coordinates land in the shared root
box, and the oracle run on the merged forest equals the oracle run on each
problem's own tree (scaled by s and s^2) on small problems.
GPU: Plummer problems of uneven size (the batched workload's distribution,
scale radius 1, charges 1/N) evaluated as one forest, each problem against
the oracle on its own tree and operators (FP64 1e-10), against the FP64
direct sum, through the captured graph (two replays) and after the
autotuner, and the mixed path against the oracle on the truncated operators.
"""
import numpy as np
import pytest

import exafmm
import oracle as O
from exafmm import abi
from exafmm.batch import LaplaceBatch, merge_forest, R_BOX, X_BOX
import os

from fmm_cases import CACHE, points, rel_err

L = exafmm.laplace


def plummer_problems(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        x = points(n, "plummer", rng)
        out.append((x, np.full(n, 1.0 / n), x))
    return out


def own_tree(x, q, p, ncrit, precision=abi.FP64):
    fmm = L.LaplaceFmm(p, ncrit, os.path.join(CACHE, "batch_own_p%d_%d.dat" % (p, precision)))
    fmm.precision = precision
    tree = L.setup(L.init_sources(x, q), L.init_targets(x), fmm)
    return fmm, tree, fmm.flat_tree(tree), fmm.operators(), fmm.sorted_charges(tree)


def test_merge_forest_keeps_every_tree():
    probs = plummer_problems([1500, 4000, 800], seed=3)
    flats = []
    for x, q, _ in probs:
        fmm = L.LaplaceFmm(4, 32)
        tree = L.setup(L.init_sources(x, q), L.init_targets(x), fmm, precompute=False)
        flats.append(fmm.flat_tree(tree))
    F, scale, soff, toff = merge_forest(flats)
    nn = F["nnodes"]
    assert nn == sum(int(f["nnodes"]) for f in flats)
    lo = F["level_offset"]
    assert lo[0] == 0 and lo[1] == len(flats) and lo[-1] == nn and np.all(np.diff(lo) >= 0)
    lev = np.repeat(np.arange(len(lo) - 1), np.diff(lo))
    c0 = F["child0"]
    for n in np.nonzero(c0 >= 0)[0]:
        assert lev[c0[n]] == lev[n] + 1 and c0[n] + 8 <= lo[lev[n] + 2]
    # every node of every problem appears once, with its lists mapped
    seen = np.zeros(nn, dtype=int)
    for i, f in enumerate(flats):
        lvl = np.asarray(f["level"])
        flo = np.asarray(f["level_offset"])
        perm = lo[lvl] + (np.arange(f["nnodes"]) - flo[lvl])
        perm += np.array([sum(int(np.diff(g["level_offset"])[l]) if l <= g["depth"] else 0 for g in flats[:i])
                          for l in lvl])
        seen[perm] += 1
        for name in ("p2p", "m2p", "p2l"):
            off, lst = np.asarray(f[name + "_offset"]), np.asarray(f[name + "_source"])
            for n in range(f["nnodes"]):
                got = F[name + "_source"][F[name + "_offset"][perm[n]]:F[name + "_offset"][perm[n] + 1]]
                assert np.array_equal(got, perm[lst[off[n]:off[n + 1]]])
        ml = np.asarray(f["m2l_source"]).reshape(-1, 26)
        mg = F["m2l_source"].reshape(-1, 26)[perm]
        assert np.array_equal(mg, np.where(ml >= 0, perm[np.maximum(ml, 0)], -1))
        assert np.array_equal(F["src_count"][perm], f["src_count"])
        assert np.all(F["src_begin"][perm][np.asarray(f["src_count"]) > 0]
                      == np.asarray(f["src_begin"])[np.asarray(f["src_count"]) > 0] + soff[i])
        assert scale[i] == pytest.approx(R_BOX / f["r0"])
    assert np.all(seen == 1)
    xs = F["src_coord"].reshape(-1, 3)
    assert np.all(np.abs(xs - np.asarray(X_BOX)) <= R_BOX)


def test_oracle_on_forest_equals_each_tree():
    p, ncrit = 4, 32
    probs = plummer_problems([1200, 2500], seed=8)
    flats, qs, refs = [], [], []
    for x, q, _ in probs:
        fmm, tree, flat, ops, qsi = own_tree(x, q, p, ncrit)
        orc = O.Oracle(O.LAPLACE, p, flat, ops)
        orc.set_charges(qsi)
        refs.append(orc.evaluate())
        flats.append(flat)
        qs.append(qsi)
    F, scale, soff, toff = merge_forest(flats)
    opf = L.LaplaceFmm(p, ncrit, os.path.join(CACHE, "batch_forest_p%d.dat" % p))
    opf.set_bounds(list(X_BOX), R_BOX)
    opf.precompute()
    orc = O.Oracle(O.LAPLACE, p, F, opf.operators())
    orc.set_charges(np.concatenate(qs))
    v = orc.evaluate()
    for i, ref in enumerate(refs):
        vi = v[toff[i]:toff[i + 1]].copy()
        vi[:, 0] *= scale[i]
        vi[:, 1:] *= scale[i] ** 2
        assert rel_err(vi, ref) < 1e-11, (i, rel_err(vi, ref))


@pytest.mark.gpu
def test_batch_plummer_vs_oracle(require_gpu):
    p, ncrit = 8, 64
    sizes = [12000, 3000, 25000, 6000]
    probs = plummer_problems(sizes, seed=21)
    print("batch of %d Plummer problems, N = %s, p=%d" % (len(sizes), sizes, p), flush=True)
    b = LaplaceBatch(probs, p=p, ncrit=ncrit, precision=abi.FP64, autotune=True, graph=True)
    print("forest: %d nodes, depth %d, tuning %s %s" % (b.forest.flat["nnodes"], b.depth, b.tuning, b.tune_ms), flush=True)
    vals = [b.evaluate(), b.evaluate()]  # two graph replays
    for i, (x, q, _) in enumerate(probs):
        fmm, tree, flat, ops, qsi = own_tree(x, q, p, ncrit)
        orc = O.Oracle(O.LAPLACE, p, flat, ops)
        orc.set_charges(qsi)
        ref_sorted = orc.evaluate()
        ref = np.empty_like(ref_sorted)
        ref[np.asarray(flat["trg_ibody"])] = ref_sorted
        for r, v in enumerate(vals):
            e = rel_err(v[i], ref)
            print("problem %d (N=%d) replay %d: vs oracle %.2e" % (i, sizes[i], r, e), flush=True)
            assert e < 1e-10, (i, r, e)
        idx = np.random.default_rng(i).choice(sizes[i], 300, replace=False)
        d = O.direct(0, x, q, x[idx], 0.0)
        ed = (rel_err(vals[0][i][idx, 0], d[:, 0]), rel_err(vals[0][i][idx, 1:], d[:, 1:]))
        print("  vs direct phi %.2e grad %.2e" % ed, flush=True)
        assert ed[0] < 1e-6 and ed[1] < 1e-5
    # new charges through the same graph
    q2 = [np.random.default_rng(50 + i).uniform(-1, 1, n) for i, n in enumerate(sizes)]
    b.update_charges(q2)
    v2 = b.evaluate()
    for i, (x, _, _) in enumerate(probs):
        idx = np.random.default_rng(70 + i).choice(sizes[i], 200, replace=False)
        d = O.direct(0, x, q2[i], x[idx], 0.0)
        e = rel_err(v2[i][idx], d)
        print("problem %d new charges vs direct %.2e" % (i, e), flush=True)
        assert e < 1e-5
    b.close()


@pytest.mark.gpu
def test_batch_mixed_vs_oracle(require_gpu):
    p, ncrit = 8, 64
    sizes = [9000, 20000, 4000]
    probs = plummer_problems(sizes, seed=22)
    b = LaplaceBatch(probs, p=p, ncrit=ncrit, precision=abi.MIXED, autotune=False, graph=True)
    vals = b.evaluate()
    b.close()
    for i, (x, q, _) in enumerate(probs):
        fmm, tree, flat, ops, qsi = own_tree(x, q, p, ncrit, precision=abi.MIXED)
        orc = O.Oracle(O.LAPLACE, p, flat, ops)
        orc.set_charges(qsi)
        ref_sorted = orc.evaluate()
        ref = np.empty_like(ref_sorted)
        ref[np.asarray(flat["trg_ibody"])] = ref_sorted
        e = (rel_err(vals[i][:, 0], ref[:, 0]), rel_err(vals[i][:, 1:], ref[:, 1:]))
        print("mixed problem %d (N=%d) vs oracle (truncated operators) phi %.2e grad %.2e" % (i, sizes[i], *e),
              flush=True)
        # charges 1/N (one sign): M2M / L2L and the transforms run in FP64 on the mixed path
        assert e[0] < 1e-5 and e[1] < 1e-5
