"""Seeded problem generators shared by the tests (host side only)."""
import os
import tempfile

import numpy as np

import exafmm

# precompute cache files (reference layout) go to a private temp directory,
# keyed by kernel, p and wave number (the file format itself does not record
# the wave number)
CACHE = tempfile.mkdtemp(prefix="exafmm_ops_")

KERNELS = {
    "laplace": (exafmm.laplace, "LaplaceFmm", 0, None),
    "helmholtz": (exafmm.helmholtz, "HelmholtzFmm", 1, 5.0 + 0.0j),
    "modified_helmholtz": (exafmm.modified_helmholtz, "ModifiedHelmholtzFmm", 2, 5.0),
}


def points(n, dist, rng):
    if dist == "cube":
        return rng.random((n, 3))
    if dist == "sphere":
        x = rng.normal(size=(n, 3))
        return x / np.linalg.norm(x, axis=1)[:, None]
    if dist == "plummer":
        # Plummer sphere (scale radius 1), radius truncated at 10
        m = rng.random(n) * 0.999
        r = 1.0 / np.sqrt(m ** (-2.0 / 3.0) - 1.0)
        r = np.minimum(r, 10.0)
        v = rng.normal(size=(n, 3))
        return v / np.linalg.norm(v, axis=1)[:, None] * r[:, None]
    if dist == "pairs":
        # uniform, a tenth of the points with a partner at distance 10^U(-7, -2)
        m = n // 10
        x = rng.random((n - m, 3))
        d = rng.normal(size=(m, 3))
        d *= (10.0 ** rng.uniform(-7.0, -2.0, m) / np.linalg.norm(d, axis=1))[:, None]
        return np.concatenate([x, np.clip(x[:m] + d, 0.0, 1.0)])
    if dist == "clusters":
        c = rng.random((4, 3))
        k = rng.integers(0, 4, n)
        return c[k] + 0.01 * rng.normal(size=(n, 3)) + (rng.random((n, 1)) < 0.2) * rng.random((n, 3))
    raise ValueError(dist)


def charges(kind, n, rng, cplx):
    """Charge distributions of the workloads: "centered" U[-0.5,0.5), "u01" U[0,1), "u11" U[-1,1), "normal"
    N(0,1); complex kernels get an independent imaginary part of the same distribution."""
    def draw():
        if kind == "centered":
            return rng.random(n) - 0.5
        if kind == "u01":
            return rng.random(n)
        if kind == "u11":
            return rng.uniform(-1.0, 1.0, n)
        if kind == "normal":
            return rng.normal(size=n)
        raise ValueError(kind)
    q = draw()
    return q + 1j * draw() if cplx else q


def make_problem(kernel, n, p, ncrit=64, dist="cube", seed=0, wavek=None, precision=0, depth=None,
                 charge="centered"):
    """depth=None: adaptive tree (setup); depth=d: the non-adaptive tree of that
    depth (setup_non_adaptive, the reference's NON_ADAPTIVE build)."""
    mod, cls, kid, wk_default = KERNELS[kernel]
    wk = wk_default if wavek is None else wavek
    rng = np.random.default_rng(seed)
    x = points(n, dist, rng)
    q = charges(charge, n, rng, kid == 1)
    src = mod.init_sources(x, q)
    trg = mod.init_targets(x)
    Cls = getattr(mod, cls)
    fname = os.path.join(CACHE, "%s_p%d_k%s_%d.dat" % (kernel, p, str(wk).replace(" ", ""), seed))
    fmm = Cls(p, ncrit, filename=fname) if wk is None else Cls(p, ncrit, wk, filename=fname)
    fmm.precision = precision
    tree = mod.setup(src, trg, fmm) if depth is None else mod.setup_non_adaptive(src, trg, fmm, depth)
    return dict(mod=mod, fmm=fmm, tree=tree, kid=kid, wavek=0.0 if wk is None else wk, x=x, q=q, p=p,
                flat=fmm.flat_tree(tree), ops=fmm.operators(), qs=fmm.sorted_charges(tree))


def truncated_ops(pr):
    """Operators of the mixed and FLOAT modes for problem `pr`: the same precompute with the pseudo-inverses
    truncated at 4e-8 max(S) (the operators of pr itself are kept)."""
    fmm = pr["fmm"]
    save = fmm.precision
    fmm.precision = 1
    fmm.precompute()
    ops = fmm.operators()
    fmm.precision = save
    fmm.precompute()
    return ops


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def kernel_np(kid, wavek, d):
    """Unit-charge potential 1/(4 pi r) (0), e^{ikr}/(4 pi r) (1), e^{-kr}/(4 pi r) (2)."""
    r = np.sqrt(np.sum(d * d, axis=-1))
    with np.errstate(divide="ignore", invalid="ignore"):
        if kid == 0:
            v = 1.0 / r
        elif kid == 2:
            v = np.exp(-np.real(wavek) * r) / r
        else:
            v = np.exp(1j * complex(wavek) * r) / r
    v = np.where(r > 0, v, 0.0)
    return v / (4 * np.pi)


def check_matrix(kid, wavek, p, r0, level, up):
    """Kernel matrix from an equivalent surface to its check surface at `level`:
    upward (equivalent 1.05 -> check 2.95) or downward (2.95 -> 1.05)."""
    import oracle as O
    c = np.zeros(3)
    a_chk, a_eq = (2.95, 1.05) if up else (1.05, 2.95)
    chk = O.surface(p, r0, level, c, a_chk)
    eq = O.surface(p, r0, level, c, a_eq)
    return kernel_np(kid, wavek, chk[:, None, :] - eq[None, :, :])


def equiv_rel_diff(kid, wavek, p, r0, levels, a, b, rows, up):
    """Relative difference of equivalent densities a, b (rows `rows`) measured by
    the check potentials they generate: |K (a-b)| / |K b|."""
    num = den = 0.0
    for lv in np.unique(levels[rows]):
        r = rows[levels[rows] == lv]
        K = check_matrix(kid, wavek, p, r0, int(lv), up)
        num += np.linalg.norm((a[r] - b[r]) @ K.T) ** 2
        den += np.linalg.norm(b[r] @ K.T) ** 2
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))
