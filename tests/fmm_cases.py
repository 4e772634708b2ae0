"""Seeded problem generators shared by the tests (host side only)."""
import numpy as np

import exafmm

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
    if dist == "clusters":
        c = rng.random((4, 3))
        k = rng.integers(0, 4, n)
        return c[k] + 0.01 * rng.normal(size=(n, 3)) + (rng.random((n, 1)) < 0.2) * rng.random((n, 3))
    raise ValueError(dist)


def make_problem(kernel, n, p, ncrit=64, dist="cube", seed=0, wavek=None, precision=0):
    mod, cls, kid, wk_default = KERNELS[kernel]
    wk = wk_default if wavek is None else wavek
    rng = np.random.default_rng(seed)
    x = points(n, dist, rng)
    q = rng.random(n) - 0.5
    if kid == 1:
        q = q + 1j * (rng.random(n) - 0.5)
    src = mod.init_sources(x, q)
    trg = mod.init_targets(x)
    Cls = getattr(mod, cls)
    fmm = Cls(p, ncrit, filename="") if wk is None else Cls(p, ncrit, wk, filename="")
    fmm.precision = precision
    tree = mod.setup(src, trg, fmm)
    return dict(mod=mod, fmm=fmm, tree=tree, kid=kid, wavek=0.0 if wk is None else wk, x=x, q=q, p=p,
                flat=fmm.flat_tree(tree), ops=fmm.operators(), qs=fmm.sorted_charges(tree))


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))
