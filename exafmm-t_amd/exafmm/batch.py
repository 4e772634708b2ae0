"""Many independent Laplace FMM problems evaluated as one forest on one engine.

No counterpart in the reference: exafmm-t evaluates one tree per FMM object
and recomputes (or reloads) its operators whenever the root box changes
(fmm_scale_invariant.h:113-148, the r0 check of the .dat cache).  The Laplace
operators are scale invariant, so here every problem is mapped into one shared
root box (centre X_BOX, half side R_BOX): problem i, whose own root box has
centre x0_i and half side r0_i, is moved by x' = (x - x0_i) s_i + X_BOX with
s_i = R_BOX / r0_i.  Its octree, 2:1 balance and interaction lists are built
by the same host code as a single problem (`laplace.setup`), in its own
coordinates, so the lists are exactly those of the single problem.  The
flattened trees are then merged level by level into one forest (level l of the
forest = the level-l nodes of problem 0, then problem 1, ...; the roots of all
problems form level 0), which the engine evaluates with ONE operator set and
one launch per operator for all problems, optionally replayed from one
captured HIP graph.  Potentials come back multiplied by s_i and gradients by
s_i^2 on the device (phi(x) = s phi'(x'), grad phi(x) = s^2 grad' phi'(x')).

    batch = LaplaceBatch([(src0, q0, trg0), (src1, q1, trg1), ...], p=8, ncrit=64)
    values = batch.evaluate()          # list of (ntrg_i, 4) arrays, original target order
    batch.update_charges([q0, q1, ...])

The engine's autotuner (xfmm_autotune) picks the P2P tail split and the
Hadamard frequency chunk for the merged forest before the graph is captured.
"""
import numpy as np

import exafmm
from . import abi

R_BOX = 1.0
X_BOX = (0.0, 0.0, 0.0)


def _merge_csr(parts, nnew):
    """CSR lists of several trees in the merged node numbering; parts = [(perm, off, lst)]."""
    lens = np.zeros(nnew, dtype=np.int64)
    for perm, off, _ in parts:
        lens[perm] = np.diff(off)
    off_new = np.zeros(nnew + 1, dtype=np.int64)
    np.cumsum(lens, out=off_new[1:])
    out = np.empty(int(off_new[-1]), dtype=np.int32)
    for perm, off, lst in parts:
        n = len(off) - 1
        cnt = np.diff(off)
        if off[-1] == 0:
            continue
        owner = np.repeat(np.arange(n), cnt)
        pos = off_new[perm[owner]] + (np.arange(int(off[-1])) - off[owner])
        out[pos] = perm[lst]
    return off_new.astype(np.int32), out


def merge_forest(flats, x0=X_BOX, r0=R_BOX):
    """Merge flattened trees (dicts of Fmm.flat_tree) into one forest in the shared root box (x0, r0).

    Returns (forest dict for abi.Engine.set_tree, per-problem scale s_i, source offsets, target offsets)."""
    P = len(flats)
    if P == 0:
        raise ValueError("merge_forest: no problems")
    D = max(int(f["depth"]) for f in flats)
    cnt = np.zeros((P, D + 1), dtype=np.int64)
    for i, f in enumerate(flats):
        lo = np.asarray(f["level_offset"], dtype=np.int64)
        for lv in range(int(f["depth"]) + 1):
            cnt[i, lv] = lo[lv + 1] - lo[lv]
    lvl_off = np.zeros(D + 2, dtype=np.int64)
    np.cumsum(cnt.sum(axis=0), out=lvl_off[1:])
    before = np.cumsum(cnt, axis=0) - cnt
    nn = int(lvl_off[-1])
    X = np.asarray(x0, dtype=np.float64)
    center = np.zeros((nn, 3))
    child0 = np.full(nn, -1, dtype=np.int32)
    sb = np.zeros(nn, dtype=np.int32)
    sc = np.zeros(nn, dtype=np.int32)
    tb = np.zeros(nn, dtype=np.int32)
    tc = np.zeros(nn, dtype=np.int32)
    m2l = np.full((nn, 26), -1, dtype=np.int32)
    src, trg, scale, soff, toff = [], [], [], [0], [0]
    p2p, m2p, p2l = [], [], []
    for i, f in enumerate(flats):
        lo = np.asarray(f["level_offset"], dtype=np.int64)
        lev = np.asarray(f["level"], dtype=np.int64)
        n = int(f["nnodes"])
        perm = lvl_off[lev] + before[i, lev] + (np.arange(n) - lo[lev])
        s = r0 / float(f["r0"])
        xi = np.asarray(f["x0"], dtype=np.float64)
        center[perm] = (np.asarray(f["center"]).reshape(-1, 3) - xi) * s + X
        c0 = np.asarray(f["child0"])
        child0[perm] = np.where(c0 >= 0, perm[np.maximum(c0, 0)], -1)
        sb[perm] = np.asarray(f["src_begin"]) + soff[-1]
        sc[perm] = f["src_count"]
        tb[perm] = np.asarray(f["trg_begin"]) + toff[-1]
        tc[perm] = f["trg_count"]
        ml = np.asarray(f["m2l_source"]).reshape(n, 26)
        m2l[perm] = np.where(ml >= 0, perm[np.maximum(ml, 0)], -1)
        for name, acc in (("p2p", p2p), ("m2p", m2p), ("p2l", p2l)):
            acc.append((perm, np.asarray(f[name + "_offset"], dtype=np.int64), np.asarray(f[name + "_source"])))
        src.append((np.asarray(f["src_coord"]).reshape(-1, 3) - xi) * s + X)
        trg.append((np.asarray(f["trg_coord"]).reshape(-1, 3) - xi) * s + X)
        scale.append(s)
        soff.append(soff[-1] + src[-1].shape[0])
        toff.append(toff[-1] + trg[-1].shape[0])
    forest = dict(nnodes=nn, depth=D, r0=float(r0), x0=[float(v) for v in X],
                  level_offset=lvl_off.astype(np.int32), center=center.reshape(-1), child0=child0,
                  src_begin=sb, src_count=sc, trg_begin=tb, trg_count=tc,
                  src_coord=np.concatenate(src).reshape(-1), trg_coord=np.concatenate(trg).reshape(-1),
                  m2l_source=m2l.reshape(-1))
    for name, parts in (("p2p", p2p), ("m2p", m2p), ("p2l", p2l)):
        forest[name + "_offset"], forest[name + "_source"] = _merge_csr(parts, nn)
    return forest, np.asarray(scale), np.asarray(soff), np.asarray(toff)


class LaplaceForest:
    """Host side of a batch: each problem's tree, 2:1 balance and lists (built in its own coordinates by
    laplace.setup, without operators), merged into one forest in the shared root box.  Independent of the
    precision mode, so one forest can feed several engines."""

    def __init__(self, problems, p=8, ncrit=64):
        L = exafmm.laplace
        self.p, self.ncrit = p, ncrit
        flats, self.src_ibody, self.trg_ibody, self.ntrg, self.nsrc = [], [], [], [], []
        qs = []
        for xs, q, xt in problems:
            xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
            xt = np.ascontiguousarray(xt, dtype=np.float64).reshape(-1, 3)
            q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
            if q.shape[0] != xs.shape[0]:
                raise ValueError("LaplaceForest: %d charges for %d sources" % (q.shape[0], xs.shape[0]))
            fmm = L.LaplaceFmm(p, ncrit)
            tree = L.setup(L.init_sources(xs, q), L.init_targets(xt), fmm, precompute=False)
            f = fmm.flat_tree(tree)
            if np.any(np.asarray(f["src_ibody"]) < 0) or np.any(np.asarray(f["trg_ibody"]) < 0):
                raise ValueError("LaplaceForest: tree with ghost bodies (adaptive trees only)")
            flats.append(f)
            self.src_ibody.append(np.asarray(f["src_ibody"]))
            self.trg_ibody.append(np.asarray(f["trg_ibody"]))
            self.nsrc.append(xs.shape[0])
            self.ntrg.append(xt.shape[0])
            qs.append(q[self.src_ibody[-1]])
            del tree, fmm
        self.flat, self.scale, self.soff, self.toff = merge_forest(flats)
        self.depth = int(self.flat["depth"])
        self.sorted_charges = np.concatenate(qs)

    @property
    def nproblems(self):
        return len(self.ntrg)

    def value_scale(self):
        """(ntrg_total, 2) factors (s_i, s_i^2) of the merged sorted targets."""
        vs = np.empty((int(self.toff[-1]), 2))
        for i, s in enumerate(self.scale):
            vs[self.toff[i]:self.toff[i + 1], 0] = s
            vs[self.toff[i]:self.toff[i + 1], 1] = s * s
        return vs

    def sort_charges(self, charges):
        """Charges of every problem (original source order) -> merged sorted order."""
        if len(charges) != self.nproblems:
            raise ValueError("%d charge arrays for %d problems" % (len(charges), self.nproblems))
        qs = []
        for i, q in enumerate(charges):
            q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
            if q.shape[0] != self.nsrc[i]:
                raise ValueError("problem %d has %d sources, got %d charges" % (i, self.nsrc[i], q.shape[0]))
            qs.append(q[self.src_ibody[i]])
        return np.concatenate(qs)

    def split_values(self, v):
        """Merged sorted (ntrg_total, 4) values -> list of (ntrg_i, 4) in each problem's target order."""
        out = []
        for i in range(self.nproblems):
            o = np.empty((self.ntrg[i], 4))
            o[self.trg_ibody[i]] = v[self.toff[i]:self.toff[i + 1]]
            out.append(o)
        return out


class LaplaceBatch:
    """Independent Laplace problems [(sources (n,3), charges (n,), targets (m,3)), ...] of uneven size (or a
    LaplaceForest of them), evaluated together: one operator set, one engine, one launch per operator, one
    captured graph."""

    def __init__(self, problems, p=8, ncrit=64, precision=abi.FP64, filename="", device=0, autotune=True,
                 graph=True):
        L = exafmm.laplace
        self.forest = problems if isinstance(problems, LaplaceForest) else LaplaceForest(problems, p, ncrit)
        self.p, self.ncrit, self.precision = self.forest.p, self.forest.ncrit, precision
        opf = L.LaplaceFmm(self.p, self.ncrit, filename)
        opf.precision = precision
        opf.set_bounds(list(X_BOX), R_BOX)
        opf.precompute()
        self.engine = abi.Engine(abi.LAPLACE, self.p, 0.0, precision, device)
        self.engine.set_operators(opf.operators())
        self.engine.set_tree(self.forest.flat)
        self.engine.set_charges(self.forest.sorted_charges)
        self.engine.set_value_scale(self.forest.value_scale())
        self.tuning = self.tune_ms = None
        if autotune:
            self.tuning, self.tune_ms = self.engine.autotune()
        self.graph = bool(graph)
        if self.graph:
            self.engine.graph_capture()

    @property
    def nproblems(self):
        return self.forest.nproblems

    @property
    def depth(self):
        return self.forest.depth

    def update_charges(self, charges):
        """New charges for every problem (original source order); the forest and graph are kept."""
        self.engine.set_charges(self.forest.sort_charges(charges))

    def run(self):
        """One evaluation of every problem (asynchronous on the engine's stream)."""
        if self.graph:
            self.engine.graph_launch()
        else:
            self.engine.upward_pass()
            self.engine.downward_pass()

    def sorted_values(self):
        """(ntrg_total, 4) values of the merged sorted targets (problem by problem)."""
        return self.engine.values()

    def values(self):
        """List of (ntrg_i, 4) arrays in each problem's original target order."""
        return self.forest.split_values(self.sorted_values())

    def evaluate(self):
        self.run()
        self.engine.synchronize()
        return self.values()

    def close(self):
        self.engine.close()
